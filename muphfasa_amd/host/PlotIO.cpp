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

namespace {

// Integers of an AMReX Box or IntVect text, "((0,0,0) (7,5,3) (0,0,0))".
std::vector<long long> text_ints(const std::string& text)
{
    std::string t = text;
    for (char& ch : t)
        if (ch == '(' || ch == ')' || ch == ',') ch = ' ';
    std::istringstream s(t);
    std::vector<long long> v;
    long long x;
    while (s >> x) v.push_back(x);
    return v;
}

struct VBox {
    long long lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};

// ((lo) (hi) (type)) in dim dimensions
bool parse_box(const std::string& text, int dim, VBox& b)
{
    const auto v = text_ints(text);
    if ((int)v.size() != 3 * dim) return false;
    for (int d = 0; d < dim; ++d) {
        b.lo[d] = v[d];
        b.hi[d] = v[dim + d];
        if (v[2 * dim + d] != 0) return false;   // cell-centred only
    }
    return true;
}

long long volume(const VBox& b, int dim)
{
    long long v = 1;
    for (int d = 0; d < dim; ++d) v *= b.hi[d] - b.lo[d] + 1;
    return v;
}

}  // namespace

std::string readVisMF(const std::string& path, const BoxGeom& g, int ncomp, std::vector<double>* f)
{
    // VisMF::Write's header (version 1): version, how, ncomp, ghost width
    // (an int, or an IntVect when it differs by direction), the BoxArray of
    // valid boxes, one "FabOnDisk: file offset" per box, then min/max.
    const int dim = g.dim;
    std::ifstream h(path + "_H");
    if (!h) return "no header " + path + "_H";
    int vers = 0, how = 0, nc = 0;
    if (!(h >> vers >> how >> nc)) return "malformed header " + path + "_H";
    if (vers != 1) return "VisMF header version " + std::to_string(vers) + " in " + path + "_H (expected 1)";
    if (nc != ncomp)
        return path + " holds " + std::to_string(nc) + " components but this run needs " + std::to_string(ncomp);
    std::string tok;
    h >> std::ws;
    std::getline(h, tok);
    const auto gv = text_ints(tok);
    long long ng[3] = {0, 0, 0};
    if (gv.size() == 1) {
        for (int d = 0; d < dim; ++d) ng[d] = gv[0];
    } else if ((int)gv.size() == dim) {
        for (int d = 0; d < dim; ++d) ng[d] = gv[d];
    } else {
        return "malformed ghost width in " + path + "_H";
    }
    for (int d = 0; d < dim; ++d)
        if (ng[d] < 0) return "negative ghost width in " + path + "_H";
    // BoxArray: "(N 0" then N boxes then ")"
    std::string open;
    long long nbox = -1, zero = -1;
    if (!(h >> open) || open.rfind("(", 0) != 0) return "malformed box array in " + path + "_H";
    if (open.size() > 1) nbox = std::stoll(open.substr(1));
    else h >> nbox;
    h >> zero;
    if (!h || nbox < 1 || nbox > (1LL << 24)) return "malformed box array in " + path + "_H";
    std::vector<VBox> boxes((size_t)nbox);
    for (auto& b : boxes) {
        std::string line;
        h >> std::ws;
        std::getline(h, line);
        if (!h || !parse_box(line, dim, b)) return "malformed box in " + path + "_H";
    }
    h >> tok;
    if (tok != ")") return "malformed box array in " + path + "_H";
    long long nfod = -1;
    if (!(h >> nfod) || nfod != nbox) return "the number of FABs in " + path + "_H differs from the boxes";
    std::vector<std::pair<std::string, long long>> fods((size_t)nbox);
    for (auto& fo : fods) {
        std::string label;
        if (!(h >> label >> fo.first >> fo.second) || label != "FabOnDisk:" || fo.second < 0)
            return "malformed FabOnDisk entry in " + path + "_H";
    }
    // The boxes must lie in the domain, not overlap, and together hold all
    // of its cells (src/IO/Checkpoint.cpp:131-135 checks the last two).
    long long cells = 0, domain = 1;
    for (int d = 0; d < dim; ++d) domain *= g.n[d];
    for (size_t i = 0; i < boxes.size(); ++i) {
        const VBox& b = boxes[i];
        for (int d = 0; d < dim; ++d)
            if (b.lo[d] < 0 || b.hi[d] >= g.n[d] || b.hi[d] < b.lo[d])
                return "a box of " + path + " lies outside the domain";
        cells += volume(b, dim);
    }
    for (size_t i = 0; i < boxes.size(); ++i)
        for (size_t j = i + 1; j < boxes.size(); ++j) {
            bool meet = true;
            for (int d = 0; d < dim; ++d)
                meet = meet && boxes[i].lo[d] <= boxes[j].hi[d] && boxes[j].lo[d] <= boxes[i].hi[d];
            if (meet) return "two boxes of " + path + " overlap";
        }
    if (cells != domain) return "data in " + path + " does not cover the domain (incomplete checkpoint?)";
    const std::string dir = path.substr(0, path.find_last_of('/') + 1);
    if (f) f->assign((size_t)domain * ncomp, 0.0);
    for (size_t i = 0; i < boxes.size(); ++i) {
        const VBox& b = boxes[i];
        const std::string fname = dir + fods[i].first;
        std::ifstream d(fname, std::ios::binary);
        if (!d) return "data file " + fname + " is missing";
        d.seekg(fods[i].second);
        std::string fh;
        if (!d || !std::getline(d, fh) || fh.rfind("FAB ", 0) != 0)
            return "data in " + path + " is not a FAB file (nothing at offset " +
                   std::to_string(fods[i].second) + " of " + fname + ")";
        // FAB <RealDescriptor><box> <ncomp>: native little-endian IEEE doubles only
        const size_t bs = fh.find(")))");
        if (bs == std::string::npos || fh.compare(4, bs + 3 - 4, kFabDouble) != 0)
            return "FAB in " + fname + " does not hold little-endian IEEE doubles";
        const std::string rest = fh.substr(bs + 3);
        const size_t sp = rest.find_last_of(' ');
        VBox fb;
        if (sp == std::string::npos || !parse_box(rest.substr(0, sp), dim, fb))
            return "malformed FAB header in " + fname;
        long long fnc = -1;
        try {
            fnc = std::stoll(rest.substr(sp + 1));
        } catch (...) {
        }
        if (fnc != ncomp) return "FAB in " + fname + " has another component count";
        for (int dd = 0; dd < dim; ++dd)
            if (fb.lo[dd] != b.lo[dd] - ng[dd] || fb.hi[dd] != b.hi[dd] + ng[dd])
                return "FAB in " + fname + " is not its box grown by the ghost width";
        long long e[3] = {1, 1, 1};
        for (int dd = 0; dd < dim; ++dd) e[dd] = fb.hi[dd] - fb.lo[dd] + 1;
        const long long fpts = e[0] * e[1] * e[2];
        std::vector<double> buf((size_t)(fpts * ncomp));
        d.read((char*)buf.data(), (std::streamsize)(sizeof(double) * buf.size()));
        if (!d)
            return "data in " + path + " does not cover the domain (incomplete checkpoint?): the FAB at offset " +
                   std::to_string(fods[i].second) + " of " + fname + " is short";
        if (!f) continue;
        const long long n0 = g.n[0], n1 = g.n[1];
        const long long bz0 = dim == 3 ? b.lo[2] : 0, bz1 = dim == 3 ? b.hi[2] : 0;
        for (int c = 0; c < ncomp; ++c)
            for (long long z = bz0; z <= bz1; ++z)
                for (long long y = b.lo[1]; y <= b.hi[1]; ++y)
                    for (long long x = b.lo[0]; x <= b.hi[0]; ++x) {
                        const long long lx = x - fb.lo[0], ly = y - fb.lo[1], lz = dim == 3 ? z - fb.lo[2] : 0;
                        (*f)[(size_t)(((c * (dim == 3 ? g.n[2] : 1) + z) * n1 + y) * n0 + x)] =
                            buf[(size_t)(((c * e[2] + lz) * e[1] + ly) * e[0] + lx)];
                    }
    }
    return {};
}

void readCheckpoint(const std::string& dir, std::vector<double>& f, const BoxGeom& g, int ncomp,
                    int& step, double& time)
{
    // src/IO/Checkpoint.cpp:105-141: the Header, then the populations of
    // VisMF "f" in the checkpoint's own box layout, valid cells only
    CheckpointHeader h;
    const std::string err = readCheckpointHeader(dir, h, g.dim);
    if (!err.empty()) throw std::runtime_error("readCheckpoint: " + err);
    const auto problems = checkpointMismatches(dir, h, g, ncomp);
    if (!problems.empty()) {
        std::string msg = "readCheckpoint: " + problems.front();
        for (size_t i = 1; i < problems.size(); ++i) msg += "; " + problems[i];
        throw std::runtime_error(msg);
    }
    const std::string verr = readVisMF(dir + "/f", g, ncomp, &f);
    if (!verr.empty()) throw std::runtime_error("readCheckpoint: " + verr);
    step = h.step;
    time = h.time;
}

}  // namespace muphfasa
