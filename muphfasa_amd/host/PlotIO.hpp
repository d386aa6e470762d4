// PlotIO.hpp -- on-disk formats either side of the step: the AMReX
// single-level plotfile written by the reference's writeMacroPlotfile
// (src/IO/Plotfile.cpp:16-75: HyperCLaw-V1.1 Header, Level_0/Cell_H,
// Level_0/Cell_D_00000 with native little-endian FP64 FABs) and the
// muphfasa_checkpoint_v1 directory (src/IO/Checkpoint.cpp:26-154: text
// Header + VisMF "f").  Written as one box per level (this engine runs one
// box per GPU); read in any box layout and ghost width.
#pragma once

#include <array>
#include <string>
#include <vector>

namespace muphfasa {

struct BoxGeom {
    int dim = 3;
    int n[3] = {1, 1, 1};
    double prob_lo[3] = {0, 0, 0};
    double prob_hi[3] = {1, 1, 1};
};

// data: ncomp planes of n0*n1*n2 doubles, x fastest.
void writeSingleLevelPlotfile(const std::string& dir, const std::vector<std::string>& names,
                              const double* data, const BoxGeom& g, double time, int step);
std::string plotfileName(const std::string& prefix, int step);

struct CheckpointHeader {
    std::string version;
    int step = 0;
    double time = 0;
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    int ncomp = 0;
};

void writeCheckpoint(const std::string& dir, const double* f, int ncomp, const BoxGeom& g, int step,
                     double time);
// Returns "" on success or the reason the checkpoint cannot be used.
std::string readCheckpointHeader(const std::string& dir, CheckpointHeader& h, int dim);
std::vector<std::string> checkpointMismatches(const std::string& dir, const CheckpointHeader& h,
                                              const BoxGeom& g, int ncomp);
// The valid cells of a VisMF MultiFab (<path>_H and its FAB files) in any box
// layout and ghost width VisMF::Write produces, copied into f as ncomp planes
// of the domain; f == nullptr only checks.  Returns "" or the reason the data
// cannot be used.
std::string readVisMF(const std::string& path, const BoxGeom& g, int ncomp, std::vector<double>* f);
// Reads the populations (valid cells) into f (ncomp planes); throws on error.
void readCheckpoint(const std::string& dir, std::vector<double>& f, const BoxGeom& g, int ncomp,
                    int& step, double& time);
std::string checkpointName(const std::string& prefix, int step);

}  // namespace muphfasa
