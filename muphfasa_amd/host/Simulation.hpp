// Simulation.hpp -- the solver's simulation API (the plugin boundary every
// driver and test of the reference uses): ISimulation
// (src/Simulation/ISimulation.H:16-44), SimulationImpl
// (src/Simulation/SimulationSetup.H:36-111) and makeSimulation
// (src/Simulation/SimulationFactory.cpp:36-60).  The lattice is a run-time
// property of the engine (its kernels dispatch on it), so SimulationImpl is
// one class; BGK is the only collision model, as in the reference.
//
// State convention (reference): f() holds post-collision populations; rho()
// and vel() are refreshed only by updateMacroscopic(), from the pulled
// populations with the half-force shift when a body force is set.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "Field.hpp"
#include "InputParser.hpp"
#include "PlotIO.hpp"
#include "Units.hpp"

namespace muphfasa {

struct Geometry {
    int dim = 3;
    int n[3] = {1, 1, 1};
    double prob_lo[3] = {0, 0, 0};
    double prob_hi[3] = {1, 1, 1};
    int periodic[3] = {1, 1, 1};
    BoxGeom box() const;
};

class ISimulation {
public:
    virtual ~ISimulation() = default;
    virtual void initialize() = 0;
    virtual void advance() = 0;
    virtual void updateMacroscopic(int step) = 0;
    virtual void writeOutput(int step) = 0;

    virtual const DeviceField& f() const noexcept = 0;
    virtual const DeviceField& rho() const noexcept = 0;
    virtual const DeviceField& vel() const noexcept = 0;
    virtual double time() const noexcept = 0;
    virtual int startStep() const noexcept = 0;
    virtual double omega() const noexcept = 0;
    virtual const UnitConverter& units() const noexcept = 0;
    virtual const Geometry& geom() const noexcept = 0;

    // engine additions
    // n x advance(), replayed from captured HIP graphs of an even number of
    // steps where n allows (the same kernels and results as n advance()
    // calls, without a host launch per step)
    virtual void advanceSteps(int n) = 0;
    // writeOutput(step) would write a plotfile or a checkpoint
    virtual bool outputDue(int step) const noexcept = 0;
    virtual int lattice() const noexcept = 0;
    virtual const mfa_geom_bc& boundary() const noexcept = 0;
    virtual const std::array<double, 3>& forceLattice() const noexcept = 0;
    virtual bool hasForcing() const noexcept = 0;
    virtual hipStream_t stream() const noexcept = 0;
    // FP64 sum of rho over valid cells after updateMacroscopic (host value)
    virtual double totalMass() = 0;
    // Health of the current populations (host values): {cells with a
    // non-finite f_q, cells whose density sum_q f_q is finite and <= 0}
    virtual std::array<double, 2> health() = 0;
    // Time spent per stage since creation, in kStageNames order: advance
    // (device time of the steps; calls = steps), updateMacroscopic (device
    // time of the moment passes that ran), writePlotfile and writeCheckpoint
    // (host time, device-to-host copies included), initialize (host time).
    // Synchronises the stream.  The reference scopes the same stages with
    // BL_PROFILE and declares reportStageTiming
    // (src/Simulation/SimulationSetup.H:110).
    static constexpr int kStages = 5;
    virtual void stageTimes(double ms[kStages], long long calls[kStages]) = 0;
    // Prints the stage table to stdout.
    virtual void reportStageTiming() = 0;
};

extern const char* const kStageNames[ISimulation::kStages];

std::unique_ptr<ISimulation> makeSimulation(const InputParser& params);

// Initial rho/u by name (lbm.init): the engine's replacement for the
// reference's link-time initData override (src/Simulation/InitData.cpp:9-16).
// rho: n cells, vel: D planes of n cells, x fastest.
void initData(const InputParser& params, const UnitConverter& units, const Geometry& g,
              std::vector<double>& rho, std::vector<double>& vel);

}  // namespace muphfasa
