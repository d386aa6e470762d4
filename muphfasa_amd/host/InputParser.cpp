// InputParser.cpp -- see InputParser.hpp.  Defaults and checks follow
// src/IO/InputParser.cpp:23-323 of the reference.
#include "InputParser.hpp"

#include <stdexcept>

namespace muphfasa {

static void require(bool ok, const std::string& msg)
{
    if (!ok) throw std::runtime_error(msg);
}

int InputParser::dim() const { return latticeType() == "D2Q9" ? 2 : 3; }

std::vector<int> InputParser::ncells() const
{
    std::vector<int> n;
    m_pp.getarr("grid.ncells", n);
    require((int)n.size() >= dim(), "InputParser: grid.ncells must list one count per dimension");
    n.resize(dim());
    for (int v : n) require(v > 0, "InputParser: grid.ncells must be positive");
    return n;
}

int InputParser::maxGridSize() const
{
    int v = 64;
    m_pp.query("grid.max_grid_size", v);
    return v;
}

int InputParser::maxStep() const
{
    int v = 0;
    m_pp.get("sim.max_step", v);
    return v;
}

double InputParser::stopTime() const
{
    double v = -1.0;
    m_pp.query("sim.stop_time", v);
    return v;
}

std::string InputParser::latticeType() const
{
    std::string v = "D3Q19";
    m_pp.query("lbm.lattice_type", v);
    return v;
}

std::string InputParser::collisionModel() const
{
    std::string v = "BGK";
    m_pp.query("lbm.collision_model", v);
    return v;
}

std::optional<double> InputParser::relaxationTime() const
{
    double v;
    if (m_pp.query("lbm.tau", v)) return v;
    return std::nullopt;
}

std::optional<double> InputParser::latticeVelocity() const
{
    double v;
    if (m_pp.query("lbm.u_lat", v)) return v;
    return std::nullopt;
}

std::optional<double> InputParser::reynoldsNumber() const
{
    double v;
    if (m_pp.query("units.Re", v)) return v;
    return std::nullopt;
}

std::optional<double> InputParser::physViscosity() const
{
    double v;
    if (m_pp.query("units.nu_phys", v)) return v;
    return std::nullopt;
}

double InputParser::refLength() const
{
    double v = 1.0;
    m_pp.query("units.ref_length", v);
    return v;
}

double InputParser::refVelocity() const
{
    double v = 1.0;
    m_pp.query("units.ref_velocity", v);
    return v;
}

double InputParser::refDensity() const
{
    double v = 1.0;
    m_pp.query("units.ref_density", v);
    return v;
}

std::string InputParser::geometryType() const
{
    std::string v = "box";
    m_pp.query("geometry.type", v);
    return v;
}

std::vector<double> InputParser::geometryProbLo() const
{
    std::vector<double> v;
    m_pp.getarr("geometry.prob_lo", v);
    require((int)v.size() >= dim(), "InputParser: geometry.prob_lo needs one value per dimension");
    v.resize(dim());
    return v;
}

std::vector<double> InputParser::geometryProbHi() const
{
    std::vector<double> v;
    m_pp.getarr("geometry.prob_hi", v);
    require((int)v.size() >= dim(), "InputParser: geometry.prob_hi needs one value per dimension");
    v.resize(dim());
    return v;
}

bool InputParser::useEB() const
{
    int v = 0;
    m_pp.query("geometry.use_eb", v);
    return v != 0;
}

std::array<std::string, 6> InputParser::bcNames() const
{
    // src/IO/InputParser.cpp:225-249
    std::vector<std::string> lo, hi;
    m_pp.getarr("bc.lo_type", lo);
    m_pp.getarr("bc.hi_type", hi);
    const int D = dim();
    require((int)lo.size() == D && (int)hi.size() == D,
            "InputParser::bcNames: bc.lo_type and bc.hi_type must each list exactly one name per "
            "direction");
    std::array<std::string, 6> names{};
    for (int d = 0; d < 3; ++d) {
        names[2 * d] = d < D ? lo[d] : "Periodic";
        names[2 * d + 1] = d < D ? hi[d] : "Periodic";
    }
    return names;
}

std::array<std::array<double, 3>, 6> InputParser::perFaceVelocity(const char* const keys[6],
                                                                  bool normal_zero,
                                                                  bool tangent_zero) const
{
    // readPerFaceVelocity, src/IO/InputParser.cpp:23-65
    std::array<std::array<double, 3>, 6> U{};
    const int D = dim();
    for (int f = 0; f < 2 * D; ++f) {
        std::vector<double> u;
        if (!m_pp.queryarr(keys[f], u) || u.empty()) continue;
        require((int)u.size() == D, std::string("InputParser: ") + keys[f] +
                                        " must list exactly one component per dimension");
        const int normal = f / 2;
        for (int d = 0; d < D; ++d) {
            if (normal_zero && d == normal)
                require(u[d] == 0., std::string("InputParser: ") + keys[f] +
                                        " has a nonzero component along the wall normal; only "
                                        "tangential wall velocities are supported for Moving walls.");
            else if (tangent_zero && d != normal)
                require(u[d] == 0., std::string("InputParser: ") + keys[f] +
                                        " has a nonzero component tangent to the wall; only "
                                        "velocities normal to the face are supported for Inflows.");
            U[f][d] = u[d];
        }
    }
    return U;
}

std::array<std::array<double, 3>, 6> InputParser::Uwall() const
{
    static const char* const keys[6] = {"bc.lox_Uwall", "bc.hix_Uwall", "bc.loy_Uwall",
                                        "bc.hiy_Uwall", "bc.loz_Uwall", "bc.hiz_Uwall"};
    return perFaceVelocity(keys, true, false);
}

std::array<std::array<double, 3>, 6> InputParser::Uinflow() const
{
    static const char* const keys[6] = {"bc.lox_Uin", "bc.hix_Uin", "bc.loy_Uin",
                                        "bc.hiy_Uin", "bc.loz_Uin", "bc.hiz_Uin"};
    return perFaceVelocity(keys, false, true);
}

std::array<double, 6> InputParser::faceDensity() const
{
    static const char* const keys[6] = {"bc.lox_rho", "bc.hix_rho", "bc.loy_rho",
                                        "bc.hiy_rho", "bc.loz_rho", "bc.hiz_rho"};
    std::array<double, 6> rho{};
    for (int i = 0; i < 6; ++i) {
        rho[i] = refDensity();
        m_pp.query(keys[i], rho[i]);
        require(rho[i] > 0., "InputParser: bc.<face>_rho must be positive");
    }
    return rho;
}

std::array<double, 3> InputParser::forcingAccel() const
{
    std::array<double, 3> a{0., 0., 0.};
    std::vector<double> v;
    if (!m_pp.queryarr("forcing.accel", v) || v.empty()) return a;
    require((int)v.size() == dim(),
            "InputParser: forcing.accel must list exactly one component per dimension");
    for (int d = 0; d < dim(); ++d) a[d] = v[d];
    return a;
}

int InputParser::plotInterval() const
{
    int v = 100;
    m_pp.query("amr.plot_int", v);
    return v;
}

int InputParser::checkpointInterval() const
{
    int v = 1000;
    m_pp.query("amr.chk_int", v);
    return v;
}

std::string InputParser::plotfilePrefix() const
{
    std::string v = "plt";
    m_pp.query("amr.plot_file", v);
    return v;
}

std::string InputParser::checkpointPrefix() const
{
    std::string v = "chk";
    m_pp.query("amr.check_file", v);
    return v;
}

std::string InputParser::restartFile() const
{
    std::string v;
    m_pp.query("amr.restart", v);
    return v;
}

std::string InputParser::storage() const
{
    std::string v = "f64";
    m_pp.query("lbm.storage", v);
    return v;
}

std::string InputParser::initName() const
{
    std::string v = "rest";
    m_pp.query("lbm.init", v);
    return v;
}

double InputParser::initAmplitude() const
{
    double v = 1e-3;
    m_pp.query("lbm.init_amp", v);
    return v;
}

long long InputParser::seed() const
{
    int v = 2024;
    m_pp.query("lbm.seed", v);
    return v;
}

}  // namespace muphfasa
