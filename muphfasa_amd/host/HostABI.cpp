// HostABI.cpp -- C ABI of the host (include/mfa_host.h) and the run loop of
// the solver executable (src/main.cpp:14-38 of the reference).
#include <algorithm>
#include <cstdio>
#include <memory>
#include <string>

#include "../../include/mfa_host.h"
#include "Simulation.hpp"

using namespace muphfasa;

struct mfa_sim {
    ParmParse pp;
    std::unique_ptr<InputParser> params;
    std::unique_ptr<ISimulation> sim;
};

static thread_local std::string g_host_err;

template <class Fn>
static int guarded(Fn&& fn)
{
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        g_host_err = e.what();
        return -1;
    }
}

extern "C" {

const char* mfa_sim_error(void) { return g_host_err.c_str(); }

mfa_sim* mfa_sim_create(const char* inputs_text, int noverrides, const char* const* overrides)
{
    auto s = std::make_unique<mfa_sim>();
    const int rc = guarded([&] {
        if (inputs_text) s->pp.addText(inputs_text);
        for (int i = 0; i < noverrides; ++i) s->pp.addOverride(overrides[i]);
        s->params = std::make_unique<InputParser>(s->pp);
        s->sim = makeSimulation(*s->params);
    });
    return rc == 0 ? s.release() : nullptr;
}

void mfa_sim_destroy(mfa_sim* s) { delete s; }

int mfa_sim_initialize(mfa_sim* s) { return guarded([&] { s->sim->initialize(); }); }

int mfa_sim_advance(mfa_sim* s, int nsteps)
{
    return guarded([&] { s->sim->advanceSteps(nsteps); });
}

int mfa_sim_update_macroscopic(mfa_sim* s, int step)
{
    return guarded([&] { s->sim->updateMacroscopic(step); });
}

int mfa_sim_write_output(mfa_sim* s, int step) { return guarded([&] { s->sim->writeOutput(step); }); }

int mfa_sim_stage_times(mfa_sim* s, double ms[5], long long calls[5])
{
    return guarded([&] {
        if (!ms || !calls) throw std::runtime_error("mfa_sim_stage_times: null argument");
        s->sim->stageTimes(ms, calls);
    });
}

int mfa_sim_health(mfa_sim* s, double out[2])
{
    return guarded([&] {
        if (!out) throw std::runtime_error("mfa_sim_health: null output");
        const std::array<double, 2> h = s->sim->health();
        out[0] = h[0];
        out[1] = h[1];
    });
}

int mfa_sim_sync(mfa_sim* s)
{
    return guarded([&] { hip_check(hipStreamSynchronize(s->sim->stream()), "sync"); });
}

int mfa_sim_info(mfa_sim* s, int* lattice, int n[3], double* omega, double force[3], double* time,
                 int* start_step)
{
    return guarded([&] {
        if (lattice) *lattice = s->sim->lattice();
        if (n)
            for (int d = 0; d < 3; ++d) n[d] = s->sim->geom().n[d];
        if (omega) *omega = s->sim->omega();
        if (force)
            for (int d = 0; d < 3; ++d) force[d] = s->sim->forceLattice()[d];
        if (time) *time = s->sim->time();
        if (start_step) *start_step = s->sim->startStep();
    });
}

int mfa_sim_copy_f(mfa_sim* s, double* f)
{
    return guarded([&] { s->sim->f().copyValidToHost(f, s->sim->stream()); });
}

int mfa_sim_copy_macro(mfa_sim* s, double* rho, double* vel)
{
    return guarded([&] {
        if (rho) s->sim->rho().copyValidToHost(rho, s->sim->stream());
        if (vel) s->sim->vel().copyValidToHost(vel, s->sim->stream());
    });
}

int mfa_inputs_units(const char* inputs_text, int noverrides, const char* const* overrides,
                     double out[12])
{
    return guarded([&] {
        ParmParse pp;
        if (inputs_text) pp.addText(inputs_text);
        for (int i = 0; i < noverrides; ++i) pp.addOverride(overrides[i]);
        InputParser p(pp);
        UnitConverter u(p.refLength(), p.refVelocity(), p.refDensity(), p.ncells(), p.geometryProbLo(),
                        p.geometryProbHi(), 1.0 / 3.0, p.reynoldsNumber(), p.physViscosity(),
                        p.latticeVelocity(), p.relaxationTime());
        const auto a = p.forcingAccel();
        // validate the remaining keys the simulation would read
        (void)p.bcNames();
        (void)p.Uwall();
        (void)p.Uinflow();
        (void)p.faceDensity();
        (void)p.maxStep();
        const double v[12] = {u.dx(), u.dt(), u.latticeVelocity(), u.latticeViscosity(),
                              u.relaxationTime(), 1.0 / u.relaxationTime(), u.reynoldsNumber(),
                              u.physViscosity(), u.machNumber(), u.toLatticeAcceleration(a[0]),
                              u.toLatticeAcceleration(a[1]), u.toLatticeAcceleration(a[2])};
        for (int i = 0; i < 12; ++i) out[i] = v[i];
    });
}

int mfa_checkpoint_read(const char* dir, int dim, const int n[3], int ncomp, double* f, int* step,
                        double* time)
{
    return guarded([&] {
        if (!dir || !n || !f || !step || !time) throw std::runtime_error("mfa_checkpoint_read: null argument");
        if (dim != 2 && dim != 3) throw std::runtime_error("mfa_checkpoint_read: dim must be 2 or 3");
        BoxGeom g;
        g.dim = dim;
        for (int d = 0; d < 3; ++d) g.n[d] = d < dim ? n[d] : 1;
        std::vector<double> v;
        readCheckpoint(dir, v, g, ncomp, *step, *time);
        std::copy(v.begin(), v.end(), f);
    });
}

int mfa_run_main(int argc, const char* const* argv)
{
    return guarded([&] {
        if (argc < 2) throw std::runtime_error("usage: muphfasa <inputs> [key=value ...]");
        ParmParse pp;
        pp.addFile(argv[1]);
        for (int i = 2; i < argc; ++i) pp.addOverride(argv[i]);
        InputParser params(pp);
        auto sim = makeSimulation(params);
        std::fputs(sim->units().summary().c_str(), stdout);
        sim->initialize();
        const int max_step = params.maxStep();
        const double stop_time_lat = sim->units().toLatticeTime(params.stopTime());
        std::printf("%g\n", stop_time_lat);
        // max_step + 1 iterations of advance(); writeOutput(step), as in the
        // reference's loop.  Runs of steps that write nothing and stay inside
        // the stop time go to advanceSteps (graph replay) in one call.
        int step = sim->startStep();
        while (step <= max_step) {
            if (stop_time_lat > 0 && stop_time_lat < sim->time()) break;
            int k = 1;
            while (step + k <= max_step && !sim->outputDue(step + k - 1) &&
                   !(stop_time_lat > 0 && stop_time_lat < sim->time() + k))
                ++k;
            sim->advanceSteps(k);
            // fail fast on a diverged state, before anything is written
            const std::array<double, 2> h = sim->health();
            if (h[0] > 0 || h[1] > 0) {
                char msg[200];
                std::snprintf(msg, sizeof msg,
                              "step %d: the state has diverged: %.0f cells with non-finite "
                              "populations, %.0f cells with non-positive density",
                              step + k - 1, h[0], h[1]);
                throw std::runtime_error(msg);
            }
            sim->writeOutput(step + k - 1);
            step += k;
        }
        hip_check(hipStreamSynchronize(sim->stream()), "sync");
        sim->reportStageTiming();
    });
}

}  // extern "C"
