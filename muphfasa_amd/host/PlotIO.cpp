// PlotIO.cpp -- see PlotIO.hpp.
#include "PlotIO.hpp"

#include <sys/stat.h>

#include <cerrno>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace muphfasa {

namespace {

void mkdir_p(const std::string& path)
{
    std::string cur;
    std::istringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        if (::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
            throw std::runtime_error("cannot create directory " + cur + ": " + std::strerror(errno));
    }
}

std::string ivect(const int* v, int dim)
{
    std::ostringstream s;
    s << '(';
    for (int d = 0; d < dim; ++d) s << (d ? "," : "") << v[d];
    s << ')';
    return s.str();
}

// AMReX Box text form: ((lo) (hi) (type))
std::string box_text(const int* lo, const int* hi, int dim)
{
    const int zero[3] = {0, 0, 0};
    return "(" + ivect(lo, dim) + " " + ivect(hi, dim) + " " + ivect(zero, dim) + ")";
}

// RealDescriptor of native little-endian IEEE double
const char* kFabDouble = "((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))";

// VisMF single-box MultiFab: <path>_H and <path>_D_00000.
void write_vismf(const std::string& path, const double* data, int ncomp, const BoxGeom& g)
{
    const int lo[3] = {0, 0, 0};
    const int hi[3] = {g.n[0] - 1, g.n[1] - 1, g.n[2] - 1};
    const long long npts = (long long)g.n[0] * g.n[1] * (g.dim == 3 ? g.n[2] : 1);
    const std::string dname = path + "_D_00000";
    {
        std::ofstream d(dname, std::ios::binary);
        if (!d) throw std::runtime_error("cannot write " + dname);
        d << "FAB " << kFabDouble << box_text(lo, hi, g.dim) << ' ' << ncomp << '\n';
        d.write((const char*)data, (std::streamsize)(sizeof(double) * npts * ncomp));
        if (!d) throw std::runtime_error("short write " + dname);
    }
    std::ofstream h(path + "_H");
    if (!h) throw std::runtime_error("cannot write " + path + "_H");
    h << std::setprecision(17);
    h << 1 << '\n' << 0 << '\n' << ncomp << '\n' << 0 << '\n';
    h << "(1 0\n" << box_text(lo, hi, g.dim) << "\n)\n";
    const size_t slash = dname.find_last_of('/');
    h << 1 << '\n' << "FabOnDisk: " << dname.substr(slash + 1) << ' ' << 0 << "\n\n";
    std::vector<double> mn(ncomp, std::numeric_limits<double>::infinity()),
        mx(ncomp, -std::numeric_limits<double>::infinity());
    for (int c = 0; c < ncomp; ++c)
        for (long long i = 0; i < npts; ++i) {
            const double v = data[c * npts + i];
            mn[c] = v < mn[c] ? v : mn[c];
            mx[c] = v > mx[c] ? v : mx[c];
        }
    h << 1 << ',' << ncomp << '\n';
    for (int c = 0; c < ncomp; ++c) h << mn[c] << ',';
    h << "\n\n" << 1 << ',' << ncomp << '\n';
    for (int c = 0; c < ncomp; ++c) h << mx[c] << ',';
    h << "\n\n";
}

}  // namespace

std::string plotfileName(const std::string& prefix, int step)
{
    std::ostringstream ss;
    ss << prefix << std::setw(7) << std::setfill('0') << step;
    return ss.str();
}

std::string checkpointName(const std::string& prefix, int step) { return plotfileName(prefix, step); }

void writeSingleLevelPlotfile(const std::string& dir, const std::vector<std::string>& names,
                              const double* data, const BoxGeom& g, double time, int step)
{
    // WriteSingleLevelPlotfile -> WriteGenericPlotfileHeader layout
    mkdir_p(dir + "/Level_0");
    std::ofstream h(dir + "/Header");
    if (!h) throw std::runtime_error("cannot write " + dir + "/Header");
    h << std::setprecision(17);
    h << "HyperCLaw-V1.1\n" << names.size() << '\n';
    for (auto& nm : names) h << nm << '\n';
    h << g.dim << '\n' << time << '\n' << 0 << '\n';
    for (int d = 0; d < g.dim; ++d) h << g.prob_lo[d] << ' ';
    h << '\n';
    for (int d = 0; d < g.dim; ++d) h << g.prob_hi[d] << ' ';
    h << "\n\n";
    const int lo[3] = {0, 0, 0};
    const int hi[3] = {g.n[0] - 1, g.n[1] - 1, g.n[2] - 1};
    h << box_text(lo, hi, g.dim) << " \n" << step << " \n";
    double dx[3];
    for (int d = 0; d < g.dim; ++d) {
        dx[d] = (g.prob_hi[d] - g.prob_lo[d]) / g.n[d];
        h << dx[d] << ' ';
    }
    h << '\n' << 0 << '\n' << 0 << '\n';
    h << 0 << ' ' << 1 << ' ' << time << '\n' << step << '\n';
    for (int d = 0; d < g.dim; ++d) h << g.prob_lo[d] << ' ' << g.prob_hi[d] << '\n';
    h << "Level_0/Cell\n";
    write_vismf(dir + "/Level_0/Cell", data, (int)names.size(), g);
}

void writeCheckpoint(const std::string& dir, const double* f, int ncomp, const BoxGeom& g, int step,
                     double time)
{
    mkdir_p(dir);
    {
        std::ofstream h(dir + "/Header");
        if (!h) throw std::runtime_error("cannot write " + dir + "/Header");
        const int lo[3] = {0, 0, 0};
        const int hi[3] = {g.n[0] - 1, g.n[1] - 1, g.n[2] - 1};
        const int ng[3] = {0, 0, 0};
        h << "muphfasa_checkpoint_v1\n" << step << '\n'
          << std::setprecision(17) << time << '\n'
          << box_text(lo, hi, g.dim) << '\n'
          << ncomp << '\n'
          << ivect(ng, g.dim) << '\n';
    }
    write_vismf(dir + "/f", f, ncomp, g);
}

std::string readCheckpointHeader(const std::string& dir, CheckpointHeader& h, int dim)
{
    // src/IO/Checkpoint.cpp:61-82
    std::ifstream in(dir + "/Header");
    if (!in) return "no Header file found in " + dir;
    in >> h.version;
    if (h.version != "muphfasa_checkpoint_v1")
        return "unrecognized checkpoint version \"" + h.version + "\" in " + dir +
               " (expected muphfasa_checkpoint_v1)";
    std::string line;
    in >> h.step >> h.time >> std::ws;
    std::getline(in, line);
    if (!in || line.rfind("((", 0) != 0) return "malformed Header in " + dir;
    // ((a,b[,c]) (d,e[,f]) (0,0[,0]))
    std::string t = line;
    for (char& ch : t)
        if (ch == '(' || ch == ')' || ch == ',') ch = ' ';
    std::istringstream bs(t);
    for (int d = 0; d < dim; ++d) bs >> h.lo[d];
    for (int d = 0; d < dim; ++d) bs >> h.hi[d];
    if (!bs) return "malformed Header in " + dir;
    if (!(in >> h.ncomp)) return "malformed Header in " + dir;
    return {};
}

std::vector<std::string> checkpointMismatches(const std::string& dir, const CheckpointHeader& h,
                                              const BoxGeom& g, int ncomp)
{
    std::vector<std::string> problems;
    bool dom_ok = true;
    for (int d = 0; d < g.dim; ++d) dom_ok = dom_ok && h.lo[d] == 0 && h.hi[d] == g.n[d] - 1;
    if (!dom_ok) {
        std::ostringstream m;
        m << "domain mismatch: " << dir << " was written for " << box_text(h.lo, h.hi, g.dim)
          << " but this run has ";
        const int lo[3] = {0, 0, 0};
        const int hi[3] = {g.n[0] - 1, g.n[1] - 1, g.n[2] - 1};
        m << box_text(lo, hi, g.dim) << " (check grid.ncells)";
        problems.push_back(m.str());
    }
    if (h.ncomp != ncomp) {
        std::ostringstream m;
        m << dir << " holds " << h.ncomp << " distribution components but this run needs " << ncomp
          << " (check lbm.lattice_type)";
        problems.push_back(m.str());
    }
    return problems;
}

void readCheckpoint(const std::string& dir, std::vector<double>& f, const BoxGeom& g, int ncomp,
                    int& step, double& time)
{
    CheckpointHeader h;
    const std::string err = readCheckpointHeader(dir, h, g.dim);
    if (!err.empty()) throw std::runtime_error("readCheckpoint: " + err);
    const auto problems = checkpointMismatches(dir, h, g, ncomp);
    if (!problems.empty()) {
        std::string msg = "readCheckpoint: " + problems.front();
        for (size_t i = 1; i < problems.size(); ++i) msg += "; " + problems[i];
        throw std::runtime_error(msg);
    }
    const std::string dname = dir + "/f_D_00000";
    std::ifstream d(dname, std::ios::binary);
    if (!d) throw std::runtime_error("readCheckpoint: data in " + dir + "/f is missing");
    std::string fabhdr;
    std::getline(d, fabhdr);
    if (fabhdr.rfind("FAB ", 0) != 0)
        throw std::runtime_error("readCheckpoint: data in " + dir + "/f is not a FAB file");
    const long long npts = (long long)g.n[0] * g.n[1] * (g.dim == 3 ? g.n[2] : 1);
    f.assign((size_t)npts * ncomp, 0.0);
    d.read((char*)f.data(), (std::streamsize)(sizeof(double) * f.size()));
    if (!d)
        throw std::runtime_error("readCheckpoint: data in " + dir +
                                 "/f does not cover the domain (incomplete checkpoint?)");
    step = h.step;
    time = h.time;
}

}  // namespace muphfasa
