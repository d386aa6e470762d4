// input_info.cpp -- muphfasa_inputinfo <inputs> [key=value ...]
//
// Reads an inputs file with the solver's own parser and unit converter and
// prints what a run would start from: lattice and storage, cells and device
// memory, physical scales, lattice units (dx, dt, u_lat, nu_lat, tau, Ma),
// boundary faces in lattice units, the step count the run loop will take,
// output counts, and advisories.  Out-of-range tau / u_lat are reported with
// the u_lat (or tau) that would bring them into the ideal range at this
// resolution, or the refinement needed when none does.  Nothing is allocated
// on a GPU and nothing is run.
//
// Plays the role of the reference's util/input_info/input_info.cpp for this
// engine: same inputs keys and unit derivation; dimensionality follows
// lbm.lattice_type instead of a compile-time AMREX_SPACEDIM.  Exit status 1
// when the solver itself would refuse the inputs, 0 otherwise.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <exception>
#include <fstream>
#include <string>
#include <vector>

#include "InputParser.hpp"
#include "ParmParse.hpp"
#include "PlotIO.hpp"
#include "Units.hpp"

namespace {

using namespace muphfasa;
namespace lim = muphfasa::unit_limits;

std::string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char* f, ...)
{
    char buf[1024];
    va_list a;
    va_start(a, f);
    std::vsnprintf(buf, sizeof(buf), f, a);
    va_end(a);
    return buf;
}

void row(const char* k, const std::string& v) { std::printf("  %-18s %s\n", k, v.c_str()); }

std::string vec(const double* v, int D, const char* spec = "%.6g")
{
    std::string s = "(";
    for (int i = 0; i < D; ++i) s += (i ? ", " : "") + fmt(spec, v[i]);
    return s + ")";
}

std::string bytes(double b)
{
    const char* u[] = {"B", "KiB", "MiB", "GiB", "TiB"};
    int i = 0;
    while (b >= 1024.0 && i < 4) { b /= 1024.0; ++i; }
    return fmt("%.3g %s", b, u[i]);
}

struct Notes {
    std::vector<std::string> warn, info;
    bool fatal = false;
};

int lattice_q(const std::string& name)
{
    if (name == "D2Q9") return 9;
    if (name == "D3Q15") return 15;
    if (name == "D3Q19") return 19;
    if (name == "D3Q27") return 27;
    return 0;
}

int storage_bytes(const std::string& s)
{
    if (s == "f64") return 8;
    if (s == "f32") return 4;
    if (s == "f16dev" || s == "bf16dev") return 2;
    return 0;
}

// tau - 1/2 = (N_ref / (Re cs2)) * u_lat: the pair (u_lat, tau) moves along
// this line when only the time step changes.
void tau_advice(const UnitConverter& u, double cs2, double bc_ratio, const std::string& pinned,
                Notes& notes)
{
    const double tau = u.relaxationTime(), ul = u.latticeVelocity(), Re = u.reynoldsNumber();
    const double Nref = u.toLatticeLength(u.refLength());
    const double s = Nref / (Re * cs2);
    const double u_cap = lim::u_lat_ideal_hi / (bc_ratio > 1.0 ? bc_ratio : 1.0);
    const double u_lo = (lim::tau_ideal_lo - 0.5) / s;
    const double u_hi = std::fmin(u_cap, (lim::tau_ideal_hi - 0.5) / s);
    const bool tau_out = tau < lim::tau_ideal_lo || tau > lim::tau_ideal_hi;
    const bool u_out = !(ul > 0.0) || ul * (bc_ratio > 1.0 ? bc_ratio : 1.0) > lim::u_lat_ideal_hi;
    if (tau_out)
        notes.warn.push_back(fmt("tau = %.4g is outside the ideal range %.2g < tau < %.2g%s", tau,
                                 lim::tau_ideal_lo, lim::tau_ideal_hi,
                                 (tau > lim::tau_hard_lo && tau < lim::tau_hard_hi)
                                     ? ""
                                     : "; the solver refuses it (hard limit 0.5 < tau < 2)"));
    if (u_out)
        notes.warn.push_back(fmt("fastest lattice speed %.4g (Ma %.3g) exceeds the low-Mach guideline %.2g%s",
                                 ul * (bc_ratio > 1.0 ? bc_ratio : 1.0), u.machNumber(),
                                 lim::u_lat_ideal_hi,
                                 ul < lim::u_lat_hard_hi ? "" : "; the solver refuses u_lat >= 0.3"));
    if (!tau_out && !u_out) return;
    if (u_lo <= u_hi) {
        const double pref = std::fmin(lim::u_lat_default, u_cap);
        const double u_rec = (pref >= u_lo && pref <= u_hi) ? pref : 0.5 * (u_lo + u_hi);
        const double t_rec = 0.5 + s * u_rec;
        notes.info.push_back(fmt("at this resolution (L_ref = %.4g cells, Re = %.4g) the ideal range is "
                                 "u_lat in [%.4g, %.4g], i.e. tau in [%.4g, %.4g]",
                                 Nref, Re, u_lo, u_hi, 0.5 + s * u_lo, 0.5 + s * u_hi));
        notes.info.push_back(pinned == "lbm.tau" ? fmt("suggest lbm.tau = %.4g (u_lat %.4g)", t_rec, u_rec)
                                                 : fmt("suggest lbm.u_lat = %.4g (tau %.4g)", u_rec, t_rec));
        if (ul > 0.0)
            notes.info.push_back(fmt("that changes dt by x%.3g: x%.3g steps for the same physical time",
                                     u_rec / ul, ul / u_rec));
    } else {
        // no time step satisfies both limits: refine until tau_ideal_lo is
        // reachable at u_cap (s grows linearly with the cell count)
        const double f = ((lim::tau_ideal_lo - 0.5) / u_cap) / s;
        notes.info.push_back(fmt("no (u_lat, tau) pair meets both limits at this resolution; refine every "
                                 "direction by at least x%.3g (then u_lat = %.3g gives tau = %.3g)",
                                 f, u_cap, lim::tau_ideal_lo));
    }
}

int report(const std::string& file, const ParmParse& pp)
{
    const InputParser in(pp);
    Notes notes;
    const std::string lat = in.latticeType();
    const int Q = lattice_q(lat);
    const int D = in.dim();
    const double cs2 = 1.0 / 3.0;
    if (!Q) { notes.warn.push_back("lbm.lattice_type '" + lat + "' is not D2Q9, D3Q15, D3Q19 or D3Q27"); notes.fatal = true; }
    if (in.collisionModel() != "BGK") {
        notes.warn.push_back("lbm.collision_model '" + in.collisionModel() + "': only BGK exists");
        notes.fatal = true;
    }
    const std::string st = in.storage();
    if (!storage_bytes(st)) { notes.warn.push_back("lbm.storage '" + st + "' is not f64, f32, f16dev or bf16dev"); notes.fatal = true; }

    const std::vector<int> N = in.ncells();
    const std::vector<double> lo = in.geometryProbLo(), hi = in.geometryProbHi();
    const auto tauIn = in.relaxationTime();
    const auto uIn = in.latticeVelocity();
    const auto ReIn = in.reynoldsNumber();
    const auto nuIn = in.physViscosity();
    const UnitConverter u(in.refLength(), in.refVelocity(), in.refDensity(), N, lo, hi, cs2, ReIn,
                          nuIn, uIn, tauIn, /*inspect_only=*/true);
    const std::string pinned = tauIn ? "lbm.tau" : (uIn ? "lbm.u_lat" : "default");
    const bool hard_bad = !(u.relaxationTime() > lim::tau_hard_lo && u.relaxationTime() < lim::tau_hard_hi) ||
                          !(u.latticeVelocity() < lim::u_lat_hard_hi);
    if (hard_bad) notes.fatal = true;

    double cells = 1;
    for (int i = 0; i < D; ++i) cells *= N[i];
    const double mem = cells * (2.0 * Q * storage_bytes(st) + 8.0 * (1 + D));

    std::printf("muphfasa input summary\n");
    row("input file", file);
    std::printf("\nSimulation\n");
    row("lattice", fmt("%s  (D = %d, Q = %d, cs^2 = 1/3)", lat.c_str(), D, Q));
    row("collision", in.collisionModel());
    row("storage", st + (st == "f64" ? "" : "  (shifted deviations f - w, FP32 collision)"));
    std::string dom;
    for (int i = 0; i < D; ++i) dom += fmt("%s[%.6g, %.6g]", i ? " x " : "", lo[i], hi[i]);
    row("domain [m]", dom);
    std::string nc;
    for (int i = 0; i < D; ++i) nc += fmt("%s%d", i ? " x " : "", N[i]);
    row("cells", fmt("%s = %.0f", nc.c_str(), cells));
    row("device memory", fmt("~%s  (2 x %d populations + rho, u; no ghost layer)", bytes(mem).c_str(), Q));
    const std::string restart = in.restartFile();
    if (restart.empty()) {
        row("initial state", in.initName());
    } else {
        // the checks restartFromCheckpoint makes before it reads the populations
        BoxGeom g;
        g.dim = D;
        for (int i = 0; i < D && i < (int)N.size(); ++i) g.n[i] = N[i];
        CheckpointHeader h;
        std::string err = readCheckpointHeader(restart, h, D);
        std::vector<std::string> problems;
        if (err.empty() && Q) problems = checkpointMismatches(restart, h, g, Q);
        if (!err.empty()) problems.insert(problems.begin(), err);
        if (problems.empty()) {
            // the populations as the restart reads them: every FAB of VisMF
            // "f" in the checkpoint's own box layout and ghost width
            const std::string verr = readVisMF(restart + "/f", g, Q, nullptr);
            if (!verr.empty()) problems.push_back(verr);
        }
        if (problems.empty()) {
            row("initial state", fmt("restart from %s (written at step %d, t_lat = %.17g); the run resumes at "
                                     "step %d", restart.c_str(), h.step, h.time, h.step + 1));
        } else {
            row("initial state", "restart from " + restart + ": unusable");
            for (const auto& m : problems) notes.warn.push_back("amr.restart: " + m);
            notes.fatal = true;
        }
    }

    std::printf("\nPhysical scales\n");
    row("L_ref", fmt("%.6g m", u.refLength()));
    row("U_ref", fmt("%.6g m/s", u.refVelocity()));
    row("rho_ref", fmt("%.6g kg/m^3", u.refDensity()));
    row("Re", fmt("%.6g   [%s]", u.reynoldsNumber(), ReIn ? "units.Re" : "from units.nu_phys"));
    row("nu_phys", fmt("%.6g m^2/s   [%s]", u.physViscosity(), nuIn ? "units.nu_phys" : "from units.Re"));

    std::printf("\nLattice units\n");
    row("dt pinned by", pinned == "default" ? fmt("default u_lat = %.3g", lim::u_lat_default) : pinned);
    row("dx", fmt("%.6g m", u.dx()));
    row("dt", fmt("%.6g s", u.dt()));
    row("L_ref", fmt("%.6g cells", u.toLatticeLength(u.refLength())));
    row("u_lat", fmt("%.6g   (Ma = %.4g)", u.latticeVelocity(), u.machNumber()));
    row("nu_lat", fmt("%.6g", u.latticeViscosity()));
    row("tau", fmt("%.6g   (omega = %.6g)", u.relaxationTime(), 1.0 / u.relaxationTime()));
    const auto acc = in.forcingAccel();
    bool forced = false;
    double accl[3];
    for (int k = 0; k < 3; ++k) {
        accl[k] = u.toLatticeAcceleration(acc[k]);
        forced = forced || (k < D && acc[k] != 0.0);
    }
    row("forcing", forced ? fmt("%s m/s^2 -> lattice %s", vec(acc.data(), D).c_str(), vec(accl, D, "%.5g").c_str())
                          : std::string("none"));

    std::printf("\nBoundary faces\n");
    const auto names = in.bcNames();
    const auto Uw = in.Uwall(), Ui = in.Uinflow();
    const auto rf = in.faceDensity();
    const char* face[6] = {"x-lo", "x-hi", "y-lo", "y-hi", "z-lo", "z-hi"};
    double bc_speed = 0.0;
    bool outflow = false;
    for (int f = 0; f < 2 * D; ++f) {
        const std::string& t = names[f];
        std::string detail;
        if (t == "MovingWall" || t == "Inflow") {
            const auto& v = t == "Inflow" ? Ui[f] : Uw[f];
            double vl[3], sp = 0.0;
            for (int k = 0; k < 3; ++k) { vl[k] = u.toLatticeVelocity(v[k]); sp += v[k] * v[k]; }
            bc_speed = std::fmax(bc_speed, std::sqrt(sp));
            detail = fmt("U = %s m/s -> lattice %s", vec(v.data(), D).c_str(), vec(vl, D, "%.5g").c_str());
        } else if (t == "PressureOutlet") {
            detail = fmt("rho = %.6g -> lattice %.6g", rf[f], u.toLatticeDensity(rf[f]));
        }
        outflow = outflow || t == "Outflow";
        std::printf("  %-5s %-15s %s\n", face[f], t.c_str(), detail.c_str());
        if (f % 2 == 1 && (names[f - 1] == "Periodic") != (t == "Periodic"))
            notes.warn.push_back(fmt("%s: one side Periodic, the other '%s'; periodicity is taken from the lo face",
                                     face[f], (t == "Periodic" ? names[f - 1] : t).c_str()));
    }
    if (forced && outflow)
        notes.warn.push_back("body force with a zero-gradient Outflow face: nothing anchors the density; "
                             "a PressureOutlet face does");

    std::printf("\nTime stepping\n");
    const long long max_step = in.maxStep();
    const double stop = in.stopTime();
    // the run loop advances for step = 0..max_step unless t_lat passes stop_time
    long long n_adv = max_step + 1;
    std::string limit = "sim.max_step";
    if (stop > 0.0 && u.dt() > 0.0) {
        const long long by_time = (long long)std::floor(u.toLatticeTime(stop)) + 1;
        if (by_time < n_adv) { n_adv = by_time; limit = "sim.stop_time"; }
    }
    row("steps", fmt("%lld  (limit: %s)", n_adv, limit.c_str()));
    row("physical time", fmt("%.6g s", n_adv * u.dt()));
    const int pi = in.plotInterval(), ci = in.checkpointInterval();
    row("plotfiles", pi > 0 ? fmt("%lld  (%s, every %d steps)", n_adv / pi, in.plotfilePrefix().c_str(), pi)
                            : std::string("none"));
    row("checkpoints", ci > 0 ? fmt("%lld  (%s, every %d steps)", n_adv / ci, in.checkpointPrefix().c_str(), ci)
                              : std::string("none"));
    if (Q) {
        const double b_lu = 2.0 * Q * storage_bytes(st);
        row("work", fmt("%.4g lattice updates, %s of population traffic", cells * n_adv,
                        bytes(cells * n_adv * b_lu).c_str()));
    }

    tau_advice(u, cs2, bc_speed / u.refVelocity(), pinned, notes);
    if (!notes.warn.empty() || !notes.info.empty()) {
        std::printf("\nAdvisories\n");
        for (const auto& w : notes.warn) std::printf("  [!] %s\n", w.c_str());
        for (const auto& i : notes.info) std::printf("      %s\n", i.c_str());
    }
    std::printf("\n%s\n", notes.fatal ? "The solver would refuse these inputs." : "The solver accepts these inputs.");
    return notes.fatal ? 1 : 0;
}

}  // namespace

int main(int argc, char* argv[])
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: muphfasa_inputinfo <inputs> [key=value ...]\n");
        return 2;
    }
    try {
        muphfasa::ParmParse pp;
        pp.addFile(argv[1]);
        for (int i = 2; i < argc; ++i) pp.addOverride(argv[i]);
        return report(argv[1], pp);
    } catch (const std::exception& e) {
        std::printf("\nERROR: cannot summarise %s:\n  - %s\n", argv[1], e.what());
        return 1;
    }
}
