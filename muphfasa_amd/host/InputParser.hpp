// InputParser.hpp -- typed accessors for the solver's inputs keys, the same
// keys, defaults and validation as the reference's InputParser
// (src/IO/InputParser.H:23-98, src/IO/InputParser.cpp:75-323).  The spatial
// dimension follows lbm.lattice_type (D2Q9 -> 2, D3Qxx -> 3) instead of a
// compile-time AMREX_SPACEDIM.
//
// Keys added by this engine (no reference counterpart):
//   lbm.storage   = f64 | f32 | f16dev | bf16dev   (population storage type)
//   lbm.init      = rest | noise | tgv3d | tgv2d   (named initial condition,
//                   replacing the reference's link-time initData override)
//   lbm.init_amp, lbm.seed                          (parameters of lbm.init)
#pragma once

#include <array>
#include <optional>
#include <string>
#include <vector>

#include "ParmParse.hpp"

namespace muphfasa {

class InputParser {
public:
    explicit InputParser(const ParmParse& pp) : m_pp(pp) {}

    int dim() const;

    // grid
    std::vector<int> ncells() const;
    int maxGridSize() const;
    // time stepping
    int maxStep() const;
    double stopTime() const;
    // LBM
    std::string latticeType() const;
    std::string collisionModel() const;
    std::optional<double> relaxationTime() const;
    std::optional<double> latticeVelocity() const;
    // units
    std::optional<double> reynoldsNumber() const;
    std::optional<double> physViscosity() const;
    double refLength() const;
    double refVelocity() const;
    double refDensity() const;
    // geometry
    std::string geometryType() const;
    std::vector<double> geometryProbLo() const;
    std::vector<double> geometryProbHi() const;
    bool useEB() const;
    // boundaries, faces ordered [xlo,xhi,ylo,yhi,zlo,zhi]
    std::array<std::string, 6> bcNames() const;
    std::array<std::array<double, 3>, 6> Uwall() const;
    std::array<std::array<double, 3>, 6> Uinflow() const;
    std::array<double, 6> faceDensity() const;
    // forcing
    std::array<double, 3> forcingAccel() const;
    // output
    int plotInterval() const;
    int checkpointInterval() const;
    std::string plotfilePrefix() const;
    std::string checkpointPrefix() const;
    std::string restartFile() const;
    // engine additions
    std::string storage() const;
    std::string initName() const;
    double initAmplitude() const;
    long long seed() const;

    const ParmParse& parms() const { return m_pp; }

private:
    std::array<std::array<double, 3>, 6> perFaceVelocity(const char* const keys[6],
                                                         bool normal_zero, bool tangent_zero) const;
    const ParmParse& m_pp;
};

}  // namespace muphfasa
