// Simulation.cpp -- SimulationImpl: owns the device fields and drives the
// gfx950 kernels through the C ABI.  Orchestration follows
// src/Simulation/SimulationSetup.cpp:32-299 of the reference.
#include "Simulation.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>

namespace muphfasa {

const char* const kStageNames[ISimulation::kStages] = {"advance", "updateMacroscopic", "writePlotfile",
                                                      "writeCheckpoint", "initialize"};

BoxGeom Geometry::box() const
{
    BoxGeom b;
    b.dim = dim;
    for (int d = 0; d < 3; ++d) {
        b.n[d] = n[d];
        b.prob_lo[d] = prob_lo[d];
        b.prob_hi[d] = prob_hi[d];
    }
    return b;
}

static int lattice_id(const std::string& name, int& Q)
{
    if (name == "D2Q9") { Q = 9; return MFA_D2Q9; }
    if (name == "D3Q15") { Q = 15; return MFA_D3Q15; }
    if (name == "D3Q19") { Q = 19; return MFA_D3Q19; }
    if (name == "D3Q27") { Q = 27; return MFA_D3Q27; }
    throw std::runtime_error("Unknown or unsupported lbm.lattice_type '" + name +
                             "' (expected D2Q9 | D3Q15 | D3Q19 | D3Q27).");
}

static int bc_id(const std::string& name)
{
    // parselbmBCtype, src/LBM/BoundaryConditions.cpp:8-19
    if (name == "MovingWall") return MFA_BC_MOVING_WALL;
    if (name == "NoSlip") return MFA_BC_NO_SLIP;
    if (name == "Inflow") return MFA_BC_INFLOW;
    if (name == "Outflow") return MFA_BC_OUTFLOW;
    if (name == "PressureOutlet") return MFA_BC_PRESSURE_OUTLET;
    if (name == "Periodic") return MFA_BC_PERIODIC;
    throw std::runtime_error("parselbmBCtype: unrecognized BC name \"" + name +
                             "\" (expected MovingWall, NoSlip, Inflow, Outflow, PressureOutlet or "
                             "Periodic)");
}

void initData(const InputParser& params, const UnitConverter& units, const Geometry& g,
              std::vector<double>& rho, std::vector<double>& vel)
{
    const long long nc = (long long)g.n[0] * g.n[1] * g.n[2];
    rho.assign(nc, 1.0);
    vel.assign((size_t)g.dim * nc, 0.0);
    const std::string name = params.initName();
    if (name == "rest") return;  // src/Simulation/InitData.cpp:9-16
    if (name == "noise") {
        // i.i.d. uniform components in [-amp, amp] (lattice units); component
        // outermost, then z, y, x.
        std::mt19937_64 rng((unsigned long long)params.seed());
        const double amp = params.initAmplitude();
        std::uniform_real_distribution<double> uni(-amp, amp);
        for (auto& v : vel) v = uni(rng);
        return;
    }
    if (name == "tgv3d") {
        if (g.dim != 3) throw std::runtime_error("lbm.init = tgv3d needs a 3-D lattice");
        // Taylor-Green vortex: u = u0 sin x cos y cos z, v = -u0 cos x sin y cos z, w = 0,
        // p = rho0 u0^2/16 (cos 2x + cos 2y)(cos 2z + 2), rho = 1 + p/cs2 (lattice units).
        const double u0 = params.initAmplitude();
        const double two_pi = 2.0 * M_PI;
        for (int k = 0; k < g.n[2]; ++k)
            for (int j = 0; j < g.n[1]; ++j)
                for (int i = 0; i < g.n[0]; ++i) {
                    const double x = two_pi * (i + 0.5) / g.n[0];
                    const double y = two_pi * (j + 0.5) / g.n[1];
                    const double z = two_pi * (k + 0.5) / g.n[2];
                    const long long c = i + (long long)g.n[0] * (j + (long long)g.n[1] * k);
                    vel[c] = u0 * std::sin(x) * std::cos(y) * std::cos(z);
                    vel[nc + c] = -u0 * std::cos(x) * std::sin(y) * std::cos(z);
                    vel[2 * nc + c] = 0.0;
                    const double p = u0 * u0 / 16.0 * (std::cos(2 * x) + std::cos(2 * y)) *
                                     (std::cos(2 * z) + 2.0);
                    rho[c] = 1.0 + 3.0 * p;
                }
        return;
    }
    if (name == "tgv2d") {
        // tests/regression/tgv2d/InitData.cpp:26-52 with TgvAnalytic.H:34-50 at t = 0: one
        // vortex pair per side, k = 2 pi / n_x, amplitude units.ref_velocity in lattice units,
        // cell centres at i + 0.5,
        //   u = -u0 cos(kx) sin(ky),  v = u0 sin(kx) cos(ky),
        //   rho = 1 - u0^2 / (4 cs2) (cos 2kx + cos 2ky).
        if (g.dim != 2) throw std::runtime_error("lbm.init = tgv2d needs a 2-D lattice");
        const double u0 = units.toLatticeVelocity(params.refVelocity());
        const double k = 2.0 * M_PI / g.n[0];
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                const double x = i + 0.5, y = j + 0.5;
                const long long c = i + (long long)g.n[0] * j;
                vel[c] = -u0 * std::cos(k * x) * std::sin(k * y);
                vel[nc + c] = u0 * std::sin(k * x) * std::cos(k * y);
                rho[c] = 1.0 - u0 * u0 / (4.0 / 3.0) * (std::cos(2 * k * x) + std::cos(2 * k * y));
            }
        return;
    }
    throw std::runtime_error("unknown lbm.init '" + name +
                             "' (expected rest | noise | tgv3d | tgv2d)");
}

class SimulationImpl final : public ISimulation {
public:
    explicit SimulationImpl(const InputParser& params);
    ~SimulationImpl() override
    {
        for (auto& g : m_graph)
            if (g) (void)hipGraphExecDestroy(g);
        if (m_stream) (void)hipStreamSynchronize(m_stream);
        for (auto& p : m_pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : m_event_pool) (void)hipEventDestroy(e);
        if (m_scratch) (void)hipFree(m_scratch);
        if (m_stream) (void)hipStreamDestroy(m_stream);
    }

    void initialize() override;
    void advance() override;
    void advanceSteps(int n) override;
    bool outputDue(int step) const noexcept override
    {
        return (m_params.plotInterval() > 0 && step % m_params.plotInterval() == 0) ||
               (m_params.checkpointInterval() > 0 && step % m_params.checkpointInterval() == 0);
    }
    void updateMacroscopic(int step) override;
    void writeOutput(int step) override;

    const DeviceField& f() const noexcept override { return *m_fin; }
    const DeviceField& rho() const noexcept override { return m_rho; }
    const DeviceField& vel() const noexcept override { return m_vel; }
    double time() const noexcept override { return m_time; }
    int startStep() const noexcept override { return m_start_step; }
    double omega() const noexcept override { return m_omega; }
    const UnitConverter& units() const noexcept override { return *m_units; }
    const Geometry& geom() const noexcept override { return m_geom; }
    int lattice() const noexcept override { return m_lattice; }
    const mfa_geom_bc& boundary() const noexcept override { return m_bc; }
    const std::array<double, 3>& forceLattice() const noexcept override { return m_force; }
    bool hasForcing() const noexcept override { return m_has_forcing; }
    hipStream_t stream() const noexcept override { return m_stream; }
    double totalMass() override;
    std::array<double, 2> health() override;
    void stageTimes(double ms[kStages], long long calls[kStages]) override;
    void reportStageTiming() override;

private:
    enum Stage { kAdvance = 0, kMacro = 1, kPlot = 2, kCheckpoint = 3, kInit = 4 };
    // device stages: a pair of events recorded on the stream around the
    // stage's launches, summed once the second has completed
    struct Pending {
        int stage;
        hipEvent_t a, b;
    };
    hipEvent_t takeEvent();
    Pending deviceStageBegin(int stage);
    void deviceStageEnd(Pending p, long long calls);
    void harvest(bool wait);
    void stepOnce();

    void setInitialConditions();
    void restartFromCheckpoint(const std::string& dir);
    void buildBCManager();

    const InputParser& m_params;
    std::unique_ptr<UnitConverter> m_units;
    double m_omega = 0;
    std::array<double, 3> m_force{0., 0., 0.};
    bool m_has_forcing = false;
    Geometry m_geom;
    int m_lattice = MFA_D3Q19, m_Q = 19;
    mfa_geom_bc m_bc{};
    DeviceField m_fa, m_fb, m_rho, m_vel;
    DeviceField* m_fin = &m_fa;
    DeviceField* m_fout = &m_fb;
    double* m_scratch = nullptr;
    double m_time = 0.0;
    int m_start_step = 0;
    int m_macro_update_step = -1;
    hipStream_t m_stream = nullptr;
    // graphs of kGraphSteps steps starting from m_fa (0) and from m_fb (1)
    static constexpr int kGraphSteps = 64;
    hipGraphExec_t m_graph[2] = {nullptr, nullptr};
    // steps in pairs through mfa_step2 until it refuses the box
    bool m_pairs = true;
    double m_stage_ms[kStages] = {0, 0, 0, 0, 0};
    long long m_stage_calls[kStages] = {0, 0, 0, 0, 0};
    std::vector<Pending> m_pending;
    std::vector<hipEvent_t> m_event_pool;
};

hipEvent_t SimulationImpl::takeEvent()
{
    if (m_event_pool.empty()) {
        hipEvent_t e = nullptr;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }
    hipEvent_t e = m_event_pool.back();
    m_event_pool.pop_back();
    return e;
}

SimulationImpl::Pending SimulationImpl::deviceStageBegin(int stage)
{
    Pending p{stage, takeEvent(), takeEvent()};
    hip_check(hipEventRecord(p.a, m_stream), "hipEventRecord");
    return p;
}

void SimulationImpl::deviceStageEnd(Pending p, long long calls)
{
    hip_check(hipEventRecord(p.b, m_stream), "hipEventRecord");
    m_stage_calls[p.stage] += calls;
    m_pending.push_back(p);
    // keep the list short on long runs of single steps: fold in what has
    // completed, and wait for the oldest pairs past a bound
    if (m_pending.size() > 256) harvest(m_pending.size() > 1024);
}

void SimulationImpl::harvest(bool wait)
{
    size_t keep = 0;
    for (size_t i = 0; i < m_pending.size(); ++i) {
        Pending& p = m_pending[i];
        hipError_t q = wait ? hipEventSynchronize(p.b) : hipEventQuery(p.b);
        if (q == hipErrorNotReady) {
            m_pending[keep++] = p;
            continue;
        }
        hip_check(q, "hipEventQuery");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
        m_stage_ms[p.stage] += ms;
        m_event_pool.push_back(p.a);
        m_event_pool.push_back(p.b);
    }
    m_pending.resize(keep);
}

void SimulationImpl::stageTimes(double ms[kStages], long long calls[kStages])
{
    harvest(true);
    for (int i = 0; i < kStages; ++i) {
        ms[i] = m_stage_ms[i];
        calls[i] = m_stage_calls[i];
    }
}

void SimulationImpl::reportStageTiming()
{
    double ms[kStages];
    long long calls[kStages];
    stageTimes(ms, calls);
    std::printf("Stage timing (device time for advance and updateMacroscopic, host time otherwise)\n");
    std::printf("  %-20s %14s %12s %14s\n", "stage", "total ms", "calls", "ms per call");
    for (int i = 0; i < kStages; ++i)
        std::printf("  %-20s %14.3f %12lld %14.6f\n", kStageNames[i], ms[i], calls[i],
                    calls[i] ? ms[i] / (double)calls[i] : 0.0);
}

SimulationImpl::SimulationImpl(const InputParser& params) : m_params(params)
{
    m_lattice = lattice_id(params.latticeType(), m_Q);
    const int D = params.dim();
    m_units = std::make_unique<UnitConverter>(
        params.refLength(), params.refVelocity(), params.refDensity(), params.ncells(),
        params.geometryProbLo(), params.geometryProbHi(), 1.0 / 3.0, params.reynoldsNumber(),
        params.physViscosity(), params.latticeVelocity(), params.relaxationTime());
    m_omega = 1.0 / m_units->relaxationTime();
    const auto accel = params.forcingAccel();
    for (int d = 0; d < D; ++d) {
        m_force[d] = m_units->toLatticeAcceleration(accel[d]);
        if (accel[d] != 0.0) m_has_forcing = true;
    }
    if (params.useEB())
        throw std::runtime_error("inputs: geometry.use_eb=1 but this engine has no embedded boundaries");
    if (params.storage() != "f64")
        throw std::runtime_error("lbm.storage '" + params.storage() +
                                 "' is run by the batched/precision engine, not SimulationImpl");

    m_geom.dim = D;
    const auto n = params.ncells();
    const auto lo = params.geometryProbLo();
    const auto hi = params.geometryProbHi();
    const auto names = params.bcNames();
    for (int d = 0; d < 3; ++d) {
        m_geom.n[d] = d < D ? n[d] : 1;
        m_geom.prob_lo[d] = d < D ? lo[d] : 0.0;
        m_geom.prob_hi[d] = d < D ? hi[d] : 1.0;
        // periodicity comes from bc.lo_type only (src/Geometry/GeometryHandler.cpp:40-47)
        m_geom.periodic[d] = d < D ? (names[2 * d] == "Periodic") : 1;
    }
    // populations in the row-interleaved order (the step streams it faster);
    // moments stay in the MultiFab's planes, which plotfiles write as is
    m_fa.define(m_geom.n, m_Q, 0, D, MFA_F64, Layout::Rows);
    m_fb.define(m_geom.n, m_Q, 0, D, MFA_F64, Layout::Rows);
    m_rho.define(m_geom.n, 1, 0, D);
    m_vel.define(m_geom.n, D, 0, D);
    hip_check(hipStreamCreate(&m_stream), "hipStreamCreate");
    hip_check(hipMalloc((void**)&m_scratch, 8 * sizeof(double)), "hipMalloc(scratch)");
    buildBCManager();
}

void SimulationImpl::buildBCManager()
{
    // src/Simulation/SimulationSetup.cpp:236-299
    const auto names = m_params.bcNames();
    const double rhoWall = m_units->toLatticeDensity(m_params.refDensity());
    auto Uwall = m_params.Uwall();
    auto Uin = m_params.Uinflow();
    const auto faceRho = m_params.faceDensity();
    const int D = m_geom.dim;
    int ft[6];
    double fr[6], fu[18];
    for (int f = 0; f < 6; ++f) {
        const int dir = f / 2;
        ft[f] = MFA_BC_PERIODIC;
        fr[f] = rhoWall;
        fu[3 * f] = fu[3 * f + 1] = fu[3 * f + 2] = 0.0;
        if (dir >= D) continue;
        const std::string& name = names[f];
        if (name == "Periodic") {
            if (!m_geom.periodic[dir])
                throw std::runtime_error("buildBCManager: bc input requests \"Periodic\" on a face whose "
                                         "Geometry is not periodic in that direction");
            continue;
        }
        ft[f] = bc_id(name);
        if (name == "MovingWall")
            for (int d = 0; d < D; ++d) fu[3 * f + d] = m_units->toLatticeVelocity(Uwall[f][d]);
        else if (name == "Inflow")
            for (int d = 0; d < D; ++d) fu[3 * f + d] = m_units->toLatticeVelocity(Uin[f][d]);
        else if (name == "PressureOutlet")
            fr[f] = m_units->toLatticeDensity(faceRho[f]);
    }
    mfa_check(mfa_resolve_bc(m_lattice, ft, fr, fu, &m_bc), "buildBCManager");
    for (int d = 0; d < 3; ++d) m_bc.periodic[d] = m_geom.periodic[d];
}

void SimulationImpl::initialize()
{
    const auto t0 = std::chrono::steady_clock::now();
    const std::string restart = m_params.restartFile();
    if (restart.empty()) setInitialConditions();
    else restartFromCheckpoint(restart);
    hip_check(hipStreamSynchronize(m_stream), "sync");
    m_stage_ms[kInit] +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ++m_stage_calls[kInit];
}

void SimulationImpl::setInitialConditions()
{
    // src/Simulation/SimulationSetup.cpp:200-219
    std::vector<double> rho, vel;
    initData(m_params, *m_units, m_geom, rho, vel);
    m_rho.copyValidFromHost(rho.data(), m_stream);
    m_vel.copyValidFromHost(vel.data(), m_stream);
    mfa_check(mfa_init_equilibrium(m_rho.view(), m_vel.view(), m_fin->view(), m_lattice,
                                   m_has_forcing ? m_force.data() : nullptr, m_stream),
              "setInitialConditions");
    hip_check(hipStreamSynchronize(m_stream), "sync");
}

void SimulationImpl::restartFromCheckpoint(const std::string& dir)
{
    // src/Simulation/SimulationSetup.cpp:106-117
    std::vector<double> f;
    int step = 0;
    double t = 0;
    readCheckpoint(dir, f, m_geom.box(), m_Q, step, t);
    m_fin->copyValidFromHost(f.data(), m_stream);
    m_time = t;
    m_start_step = step + 1;
}

void SimulationImpl::stepOnce()
{
    // src/Simulation/SimulationSetup.cpp:119-139: fused step (pull, collide,
    // periodic wrap and wall ghost values inline), swap, time += 1.
    mfa_check(mfa_step(m_fin->view(), m_fout->view(), &m_bc, m_omega,
                       m_has_forcing ? m_force.data() : nullptr, m_stream),
              "advance");
    std::swap(m_fin, m_fout);
    m_time += 1.0;
}

void SimulationImpl::advance()
{
    const Pending p = deviceStageBegin(kAdvance);
    stepOnce();
    deviceStageEnd(p, 1);
}

void SimulationImpl::advanceSteps(int n)
{
    if (n <= 0) return;
    const Pending p = deviceStageBegin(kAdvance);
    const int total = n;
    // Whole graphs first: kGraphSteps (even) steps leave the state in the
    // buffer they started from, so m_fin / m_fout keep their roles.
    while (n >= kGraphSteps) {
        const int which = (m_fin == &m_fa) ? 0 : 1;
        if (!m_graph[which]) {
            hipGraph_t graph = nullptr;
            hip_check(hipStreamBeginCapture(m_stream, hipStreamCaptureModeThreadLocal), "capture");
            DeviceField* a = m_fin;
            DeviceField* b = m_fout;
            const double* F = m_has_forcing ? m_force.data() : nullptr;
            // the steps go in pairs (mfa_step2, bit-identical to two mfa_step
            // calls; FP64 boxes of every lattice, periodic or walled), or one
            // at a time where mfa_step2 refuses the box (a z-plane beyond
            // 4 GiB); it refuses before it launches anything
            int rc = MFA_OK;
            for (int s = 0; s < kGraphSteps && rc == MFA_OK;) {
                if (m_pairs) {
                    rc = mfa_step2(a->view(), b->view(), &m_bc, m_omega, F, m_stream);
                    if (rc == MFA_ERR_UNSUPPORTED && s == 0) {
                        m_pairs = false;
                        rc = MFA_OK;
                        continue;
                    }
                    s += 2;
                } else {
                    rc = mfa_step(a->view(), b->view(), &m_bc, m_omega, F, m_stream);
                    s += 1;
                }
                std::swap(a, b);
            }
            const hipError_t ce = hipStreamEndCapture(m_stream, &graph);
            if (rc != MFA_OK || ce != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                mfa_check(rc, "advanceSteps (capture)");
                hip_check(ce, "hipStreamEndCapture");
            }
            const hipError_t ie = hipGraphInstantiate(&m_graph[which], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            hip_check(ie, "hipGraphInstantiate");
        }
        hip_check(hipGraphLaunch(m_graph[which], m_stream), "hipGraphLaunch");
        m_time += kGraphSteps;
        n -= kGraphSteps;
    }
    while (n-- > 0) stepOnce();
    deviceStageEnd(p, total);
}

void SimulationImpl::updateMacroscopic(int step)
{
    if (step > m_macro_update_step) {
        const Pending p = deviceStageBegin(kMacro);
        mfa_check(mfa_moments(m_fin->view(), m_rho.view(), m_vel.view(), &m_bc,
                              m_has_forcing ? m_force.data() : nullptr, 1, m_stream),
                  "updateMacroscopic");
        deviceStageEnd(p, 1);
        m_macro_update_step = step;
    }
}

double SimulationImpl::totalMass()
{
    mfa_check(mfa_reduce(m_rho.view(), m_lattice, MFA_RED_COMP_SUM, m_scratch, m_stream), "totalMass");
    double v = 0;
    hip_check(hipMemcpyAsync(&v, m_scratch, sizeof(double), hipMemcpyDeviceToHost, m_stream), "copy");
    hip_check(hipStreamSynchronize(m_stream), "sync");
    return v;
}

std::array<double, 2> SimulationImpl::health()
{
    mfa_check(mfa_reduce(m_fin->view(), m_lattice, MFA_RED_HEALTH, m_scratch, m_stream), "health");
    std::array<double, 2> v{0., 0.};
    hip_check(hipMemcpyAsync(v.data(), m_scratch, sizeof(v), hipMemcpyDeviceToHost, m_stream), "copy");
    hip_check(hipStreamSynchronize(m_stream), "sync");
    return v;
}

void SimulationImpl::writeOutput(int step)
{
    // src/Simulation/SimulationSetup.cpp:168-198
    using clock = std::chrono::steady_clock;
    if (m_params.plotInterval() > 0 && step % m_params.plotInterval() == 0) {
        updateMacroscopic(step);
        // the stage starts once the device work queued before it is done
        hip_check(hipStreamSynchronize(m_stream), "sync");
        const auto t0 = clock::now();
        const int D = m_geom.dim;
        const long long nc = m_rho.cells();
        std::vector<double> data((size_t)(1 + D) * nc);
        m_rho.copyValidToHost(data.data(), m_stream);
        m_vel.copyValidToHost(data.data() + nc, m_stream);
        // physical units: rho * rho_ref, u * dx/dt (UnitConverter::toPhysical)
        const double rs = m_units->refDensity(), us = m_units->dx() / m_units->dt();
        for (long long i = 0; i < nc; ++i) data[i] *= rs;
        for (long long i = nc; i < (1 + D) * nc; ++i) data[i] *= us;
        std::vector<std::string> names = {"density", "vx", "vy"};
        if (D == 3) names.push_back("vz");
        const std::string fname = plotfileName(m_params.plotfilePrefix(), step);
        writeSingleLevelPlotfile(fname, names, data.data(), m_geom.box(),
                                 m_units->toPhysicalTime(m_time), step);
        m_stage_ms[kPlot] += std::chrono::duration<double, std::milli>(clock::now() - t0).count();
        ++m_stage_calls[kPlot];
        std::printf("Plotfile: wrote %s  step=%d  time=%.6g\n", fname.c_str(), step,
                    m_units->toPhysicalTime(m_time));
    }
    if (m_params.checkpointInterval() > 0 && step % m_params.checkpointInterval() == 0) {
        hip_check(hipStreamSynchronize(m_stream), "sync");
        const auto t0 = clock::now();
        std::vector<double> f((size_t)m_Q * m_fin->cells());
        m_fin->copyValidToHost(f.data(), m_stream);
        const std::string cname = checkpointName(m_params.checkpointPrefix(), step);
        writeCheckpoint(cname, f.data(), m_Q, m_geom.box(), step, m_time);
        m_stage_ms[kCheckpoint] += std::chrono::duration<double, std::milli>(clock::now() - t0).count();
        ++m_stage_calls[kCheckpoint];
        std::printf("Checkpoint: wrote %s  step=%d  time=%.17g\n", cname.c_str(), step, m_time);
    }
}

std::unique_ptr<ISimulation> makeSimulation(const InputParser& params)
{
    // src/Simulation/SimulationFactory.cpp:18-60
    int Q;
    (void)lattice_id(params.latticeType(), Q);
    const std::string collision = params.collisionModel();
    if (collision == "BGK") return std::make_unique<SimulationImpl>(params);
    if (collision == "TRT" || collision == "MRT")
        throw std::runtime_error("lbm.collision_model '" + collision +
                                 "' is a recognised tag but not yet implemented.");
    throw std::runtime_error("Unknown lbm.collision_model '" + collision + "' (expected BGK | TRT | MRT).");
}

}  // namespace muphfasa
