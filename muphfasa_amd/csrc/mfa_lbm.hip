// mfa_lbm.hip -- gfx950 kernels and C ABI of the LBM step engine.
//
// Layout: structure of arrays, x fastest, addressed through strides (see
// include/mfa_lbm.h): the MultiFab's component planes or the row-interleaved
// order in which the Q rows of one (y,z) lie side by side, so that a block's
// 19 (D3Q19) write streams fall in one contiguous region.  A wavefront covers
// 64 consecutive x cells, so every population load and store of a wavefront
// is one contiguous 512-byte (FP64) run; pulls with c_x = +-1 are shifted by
// one element and share cache lines with the neighbouring wavefront.
//
// mfa_step is the hot path: a single pass that reads each population of fin
// once and writes each population of fout once (2*Q*8 bytes per cell in FP64),
// with periodic wrap and wall ghost values computed inline, so no halo pass
// and no extra launch per step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>

#include "kernels_common.hpp"
#include "kernels_step2.hpp"

namespace mfa {

// ---------------------------------------------------------------- errors ---
static thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}


// ----------------------------------------------------------- step kernel --
// One thread per cell column of NC cells along z; block = (bx, by);
// grid = (ceil(nx/bx), ceil(ny/by), ceil(nz/NC)).  All NC*Q loads of a
// thread are issued before the first collision so they are in flight
// together.
template <class L, bool FORCED, int MODE, bool WALLS, int NC, int NT>
__global__ __launch_bounds__(256) void k_step(const double* __restrict__ fin,
                                              double* __restrict__ fout, const Geo g,
                                              const View vi, const View vo, const double omega,
                                              const double F0, const double F1, const double F2)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z0 = blockIdx.z * NC;
    if (x >= g.n[0] || y >= g.n[1]) return;
    double f[NC][L::Q];
#pragma unroll
    for (int k = 0; k < NC; ++k)
        if (NC == 1 || z0 + k < g.n[2]) pull<L, MODE, WALLS, NT>(f[k], fin, x, y, z0 + k, g, vi);
#pragma unroll
    for (int k = 0; k < NC; ++k)
        if (NC == 1 || z0 + k < g.n[2])
            collide_store<L, FORCED>(f[k], fout, x + y * vo.sy + (z0 + k) * vo.sz, vo, omega, F0,
                                     F1, F2);
}

// -------------------------------------------------------- moments kernel --
template <class L, bool FORCED, int MODE>
__global__ __launch_bounds__(256) void k_moments(const double* __restrict__ fin,
                                                 double* __restrict__ rho_out,
                                                 double* __restrict__ vel_out, const Geo g,
                                                 const View vi, const View vr, const View vv,
                                                 const double F0, const double F1,
                                                 const double F2)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= g.n[0] || y >= g.n[1]) return;
    double f[L::Q];
    pull<L, MODE>(f, fin, x, y, z, g, vi);
    const double F[3] = {F0, F1, F2};
    double rho, u0, u1, u2;
    moments<L, FORCED>(f, F, rho, u0, u1, u2);
    rho_out[x + y * vr.sy + z * vr.sz] = rho;
    const long long sv = x + y * vv.sy + z * vv.sz;
    vel_out[sv] = u0;
    vel_out[sv + vv.sq] = u1;
    if constexpr (L::D == 3) vel_out[sv + 2 * vv.sq] = u2;
}

// ----------------------------------------------------------- init kernel --
// setInitialConditions: src/Simulation/SimulationSetup.cpp:200-219,
// src/LBM/Macroscopic.cpp:164-181, src/LBM/Equilibrium.cpp:32-67.
template <class L, bool FORCED>
__global__ __launch_bounds__(256) void k_init(double* __restrict__ rho_io, double* __restrict__ vel_io,
                                              double* __restrict__ fout, const Geo g, const View vr,
                                              const View vv, const View vo, const double F0,
                                              const double F1, const double F2)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= g.n[0] || y >= g.n[1]) return;
    const long long cr = x + y * vr.sy + z * vr.sz;
    const long long cv = x + y * vv.sy + z * vv.sz;
    const double rho = rho_io[cr];
    double u0 = vel_io[cv], u1 = vel_io[cv + vv.sq];
    double u2 = (L::D == 3) ? vel_io[cv + 2 * vv.sq] : 0.0;
    if constexpr (FORCED) {
        u0 -= 0.5 * F0 / rho;
        u1 -= 0.5 * F1 / rho;
        if constexpr (L::D == 3) u2 -= 0.5 * F2 / rho;
    }
    const double usq = usq_of<L>(u0, u1, u2);
    double f[L::Q];
    const long long self = x + y * vo.sy + z * vo.sz;
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = feq<L, q>(rho, u0, u1, u2, usq);
        fout[q * vo.sq + self] = f[q];
    });
    if constexpr (FORCED) {
        const double F[3] = {F0, F1, F2};
        double r2, v0, v1, v2;
        moments<L, true>(f, F, r2, v0, v1, v2);
        rho_io[cr] = r2;
        vel_io[cv] = v0;
        vel_io[cv + vv.sq] = v1;
        if constexpr (L::D == 3) vel_io[cv + 2 * vv.sq] = v2;
    }
}

// ------------------------------------------------------ ghost-fill kernels --
// One thread per cell of the grown box; valid cells return at once.
template <class L>
__global__ __launch_bounds__(256) void k_fill_periodic(double* __restrict__ f, const Geo g,
                                                       const View v)
{
    const int gx = L::D >= 1, gy = 1, gz = (L::D == 3);
    const int x = blockIdx.x * blockDim.x + threadIdx.x - gx;
    const int y = blockIdx.y * blockDim.y + threadIdx.y - gy;
    const int z = (int)blockIdx.z - gz;
    if (x >= g.n[0] + gx || y >= g.n[1] + gy) return;
    const int p[3] = {x, y, z};
    int s[3];
    bool out = false, ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        s[d] = p[d];
        if (p[d] < 0 || p[d] >= g.n[d]) {
            out = true;
            if (!g.per[d]) ok = false;
            s[d] = p[d] < 0 ? p[d] + g.n[d] : p[d] - g.n[d];
        }
    }
    if (!out || !ok) return;
    const long long dst = x + y * v.sy + z * v.sz;
    const long long src = s[0] + s[1] * v.sy + s[2] * v.sz;
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q * v.sq + dst] = f[q * v.sq + src];
    });
}

template <class L>
__global__ __launch_bounds__(256) void k_apply_bc(double* __restrict__ f, const Geo g, const View v)
{
    const int gz = (L::D == 3);
    const int x = blockIdx.x * blockDim.x + threadIdx.x - 1;
    const int y = blockIdx.y * blockDim.y + threadIdx.y - 1;
    const int z = (int)blockIdx.z - gz;
    if (x >= g.n[0] + 1 || y >= g.n[1] + 1) return;
    const int p[3] = {x, y, z};
    int code[3] = {0, 0, 0};
    bool wall = false;
#pragma unroll
    for (int d = 0; d < L::D; ++d) {
        if (p[d] < 0) code[d] = -1;
        else if (p[d] >= g.n[d]) code[d] = 1;
        if (g.per[d]) code[d] = 0;
        if (code[d]) wall = true;
    }
    if (!wall) return;
    const int idx = (code[0] + 1) + 3 * (code[1] + 1) + (L::D == 3 ? 9 * (code[2] + 1) : 0);
    if (g.rtype[idx] == MFA_BC_PERIODIC) return;
    const long long ghost = x + y * v.sy + z * v.sz;
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int sx = x + L::c(q, 0), sy = y + L::c(q, 1), sz = z + L::c(q, 2);
        if (sx < 0 || sx >= g.n[0] || sy < 0 || sy >= g.n[1] || sz < 0 || sz >= g.n[2]) return;
        const long long src = sx + sy * v.sy + sz * v.sz;
        f[q * v.sq + ghost] = wall_value<L, q>(f, src, v, g, idx);
    });
}

// ------------------------------------------------------------- reductions --
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;
constexpr int kRedMaxOut = 27;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Deterministic block sum of NV values per thread into part[blockIdx][k].
template <int NV>
__device__ __forceinline__ void block_store(double (&acc)[NV], double* __restrict__ part)
{
    __shared__ double sh[kRedThreads / 64][NV > 0 ? NV : 1];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double s = wave_sum(acc[k]);
        if (lane == 0) sh[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kRedThreads / 64; ++w) s += sh[w][threadIdx.x];
        part[blockIdx.x * kRedMaxOut + threadIdx.x] = s;
    }
}

template <class L, int WHAT>
__global__ __launch_bounds__(kRedThreads) void k_reduce_f(const double* __restrict__ f, const Geo g,
                                                          const View v, int ncomp,
                                                          double* __restrict__ part)
{
    constexpr int NV = (WHAT == MFA_RED_MASS)       ? 1
                       : (WHAT == MFA_RED_MOMENTUM) ? 3
                       : (WHAT == MFA_RED_HEALTH)   ? 2
                                                    : kRedMaxOut;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    const long long ncell = (long long)g.n[0] * g.n[1] * g.n[2];
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(c % g.n[0]);
        const long long r = c / g.n[0];
        const int y = (int)(r % g.n[1]);
        const int z = (int)(r / g.n[1]);
        const long long self = x + y * v.sy + z * v.sz;
        if constexpr (WHAT == MFA_RED_COMP_SUM) {
            for (int k = 0; k < ncomp && k < kRedMaxOut; ++k) acc[k] += f[k * v.sq + self];
        } else if constexpr (WHAT == MFA_RED_HEALTH) {
            // counts of cells (exact in FP64): a non-finite population, or a
            // finite density that is not positive
            double rho = 0.0;
            bool finite = true;
            for_q<L>([&](auto qc) {
                const double fq = f[decltype(qc)::value * v.sq + self];
                finite = finite && __builtin_isfinite(fq);
                rho += fq;
            });
            acc[0] += finite ? 0.0 : 1.0;
            acc[1] += (finite && rho <= 0.0) ? 1.0 : 0.0;
        } else {
            for_q<L>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const double fq = f[q * v.sq + self];
                if constexpr (WHAT == MFA_RED_MASS) acc[0] += fq;
                else {
                    if constexpr (L::c(q, 0) != 0) acc[0] += L::c(q, 0) * fq;
                    if constexpr (L::c(q, 1) != 0) acc[1] += L::c(q, 1) * fq;
                    if constexpr (L::c(q, 2) != 0) acc[2] += L::c(q, 2) * fq;
                }
            });
        }
    }
    block_store<NV>(acc, part);
}

__global__ __launch_bounds__(kRedThreads) void k_residual(const double* __restrict__ a,
                                                          const double* __restrict__ b, const Geo g,
                                                          const View va, const View vb, int ncomp,
                                                          double* __restrict__ part)
{
    double acc[2] = {0.0, 0.0};
    const long long ncell = (long long)g.n[0] * g.n[1] * g.n[2];
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(c % g.n[0]);
        const long long r = c / g.n[0];
        const int y = (int)(r % g.n[1]);
        const int z = (int)(r / g.n[1]);
        for (int k = 0; k < ncomp; ++k) {
            const double av = a[x + y * va.sy + z * va.sz + k * va.sq];
            const double bv = b[x + y * vb.sy + z * vb.sz + k * vb.sq];
            const double d = av - bv;
            acc[0] += d * d;
            acc[1] += av * av;
        }
    }
    block_store<2>(acc, part);
}

__global__ __launch_bounds__(kRedThreads) void k_reduce_final(const double* __restrict__ part,
                                                              int nblocks, int nv,
                                                              double* __restrict__ out)
{
    // thread t sums blocks t, t+256, ... in order; then a fixed tree.
    for (int k = 0; k < nv; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += part[b * kRedMaxOut + k];
        __shared__ double sh[kRedThreads / 64];
        const double w = wave_sum(s);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = w;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int i = 0; i < kRedThreads / 64; ++i) t += sh[i];
            out[k] = t;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------ two-step kernels --
// 3-D boxes: k_step2p (kernels_step2.hpp).  The 2-D kernel below keeps its
// TX x TY tile grown by one cell in LDS.
template <int TX_, int TY_>
struct Step2Tile {                          // (TX+2) x (TY+2) grown column
    static constexpr int TX = TX_, TY = TY_;
    static constexpr int EX = TX + 2, EY = TY + 2, EC = EX * EY;
    static constexpr int THREADS = (EC + 63) / 64 * 64;
};

// Two steps per pass on a 2-D box, periodic or with walls.  A block owns a
// TX x TY tile of output cells.  Phase A computes the first step on the tile
// grown by one cell (periodic images; grown cells behind a wall are not
// needed) into LDS; phase B pulls the second step from LDS.  At a wall the
// second step's ghost values come from the cell's own intermediate
// populations in LDS, through the same wall_value arithmetic as the
// one-step kernel, so the pair equals two k_step launches bit for bit.
template <class L, bool FORCED, bool WALLS, class T>
__global__ __launch_bounds__(T::THREADS) void k_step2_2d(
    const double* __restrict__ fin, double* __restrict__ fout, const Geo g, const View vi,
    const View vo, const double omega, const double F0, const double F1, const double F2)
{
    __shared__ double lds[L::Q][T::EC];
    const int tid = threadIdx.x;
    const int nx = g.n[0], ny = g.n[1];
    const unsigned nb = gridDim.x * gridDim.y;
    const unsigned b = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned per = nb / kXcds, rem = nb % kXcds, k = b % kXcds;
    const unsigned t = k * per + (k < rem ? k : rem) + b / kXcds;
    const int X0 = (int)(t % gridDim.x) * T::TX, Y0 = (int)(t / gridDim.x) * T::TY;
    if (tid < T::EC) {
        int gx = X0 - 1 + tid % T::EX, gy = Y0 - 1 + tid / T::EX;
        bool in = true;
        if (g.per[0]) gx = (gx % nx + nx) % nx;
        else in = in && gx >= 0 && gx < nx;
        if (g.per[1]) gy = (gy % ny + ny) % ny;
        else in = in && gy >= 0 && gy < ny;
        if (in) {
            double P[L::Q];
            pull<L, 0, WALLS, 0>(P, fin, gx, gy, 0, g, vi);
            collide<L, FORCED>(P, omega, F0, F1, F2, [&](auto qc, double out) {
                constexpr int q = decltype(qc)::value;
                lds[q][tid] = out;
            });
        }
    }
    __syncthreads();
    if (tid >= T::TX * T::TY) return;
    const int ox = tid % T::TX, oy = tid / T::TX;
    const int x = X0 + ox, y = Y0 + oy;
    if (x >= nx || y >= ny) return;
    const int c0 = (ox + 1) + T::EX * (oy + 1);
    double f[L::Q];
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = lds[q][c0 - L::c(q, 0) - T::EX * L::c(q, 1)];
    });
    if constexpr (WALLS) {
        const int p[2] = {x, y};
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        bool edge = false;
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            lo[d] = !g.per[d] && p[d] == 0;
            hi[d] = !g.per[d] && p[d] == g.n[d] - 1;
            edge = edge || lo[d] || hi[d];
        }
        if (edge) {
            for_q<L>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                fix_wall_at<L, q>(f[q], [&](int m) { return lds[m][c0]; }, g, lo, hi);
            });
        }
    }
    collide_store<L, FORCED>(f, fout, x + (long long)y * vo.sy, vo, omega, F0, F1, F2);
}

// ----------------------------------------------------- copy-ceiling probe --
typedef double dbl2 __attribute__((ext_vector_type(2)));

template <bool NT>
__global__ __launch_bounds__(256) void k_copy(const dbl2* __restrict__ a, dbl2* __restrict__ b,
                                              long long n2)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2) return;
    if constexpr (NT) {
        const dbl2 v = __builtin_nontemporal_load(&a[i]);
        __builtin_nontemporal_store(v, &b[i]);
    } else {
        b[i] = a[i];
    }
}

// reduced-precision storage (mfa_lowprec.hip)
int step_dev(const mfa_field* fin, mfa_field* fout, const Geo& g, int lat, double omega,
             const double* force, hipStream_t st);
int moments_dev(const mfa_field* f, mfa_field* rho, mfa_field* vel, const Geo& g, int lat,
                const double* force, int streamed, hipStream_t st);
int init_dev(mfa_field* rho, mfa_field* vel, mfa_field* f, const Geo& g, int lat,
             const double* force, hipStream_t st);
int convert_dev(const mfa_field* src, mfa_field* dst, const Geo& g, int lat, hipStream_t st);
int reduce_dev(const mfa_field* f, const Geo& g, int lat, int what, double* part, int nblocks,
               hipStream_t st);
int step2_dev(const mfa_field* fin, mfa_field* fout, const Geo& g, int lat, double omega,
              const double* force, int TZ, hipStream_t st);

// -------------------------------------------------------------- helpers ---
static int lattice_dim(int lat) { return lat == MFA_D2Q9 ? 2 : 3; }
static int lattice_q(int lat)
{
    switch (lat) {
    case MFA_D2Q9: return 9;
    case MFA_D3Q15: return 15;
    case MFA_D3Q19: return 19;
    case MFA_D3Q27: return 27;
    default: return -1;
    }
}

static int check_field(const mfa_field* f, int ncomp_min, int need_grow, const char* what,
                       bool allow_reduced = false)
{
    if (!f || !f->data) return fail(MFA_ERR_INVALID, std::string(what) + ": null field");
    if (f->n[0] < 1 || f->n[1] < 1 || f->n[2] < 1)
        return fail(MFA_ERR_INVALID, std::string(what) + ": empty box");
    if (f->dtype < MFA_F64 || f->dtype > MFA_BF16DEV)
        return fail(MFA_ERR_INVALID, std::string(what) + ": unknown dtype");
    if (f->dtype != MFA_F64 && !allow_reduced)
        return fail(MFA_ERR_UNSUPPORTED, std::string(what) + ": dtype not supported by this entry point");
    if (f->ncomp < ncomp_min)
        return fail(MFA_ERR_INVALID, std::string(what) + ": too few components");
    if (f->ngrow < need_grow)
        return fail(MFA_ERR_INVALID, std::string(what) + ": needs a ghost layer");
    // The valid cells of every component must not overlap: taken in order of
    // increasing stride, each dimension with more than one entry must step
    // past the whole extent of the previous one.  This admits the
    // component-major and the row-interleaved orders alike.
    struct Dim { long long stride, extent; } d[4] = {
        {1, f->n[0]}, {f->sy, f->n[1]}, {f->sz, f->n[2]}, {f->sq, f->ncomp}};
    for (int i = 1; i < 4; ++i)
        for (int j = i; j > 0 && d[j].stride < d[j - 1].stride; --j) std::swap(d[j], d[j - 1]);
    long long span = 0;
    for (int i = 0; i < 4; ++i) {
        if (d[i].extent <= 1) continue;
        if (d[i].stride < 1 || d[i].stride < span)
            return fail(MFA_ERR_INVALID, std::string(what) + ": strides smaller than the box");
        span = d[i].stride * d[i].extent;
    }
    return MFA_OK;
}

// Byte range [lo, hi] a field's kernels may touch: every component over the
// valid box grown by its ghost layers, in the lattice's D dimensions (a 3-D
// lattice's halo-mode kernels read the z ghost planes even of a box one
// plane thick; a 2-D view has no z ghost layer).
static void byte_range(const mfa_field& f, int D, long long& lo, long long& hi)
{
    const long long es = f.dtype == MFA_F64 ? 8 : (f.dtype == MFA_F32 ? 4 : 2);
    const long long g = f.ngrow, gz = D == 3 ? g : 0;
    const long long base = (long long)(intptr_t)f.data;
    lo = base - es * (g + g * f.sy + gz * f.sz);
    hi = base + es * ((f.n[0] - 1 + g) + (f.n[1] - 1 + g) * f.sy + (f.n[2] - 1 + gz) * f.sz +
                      (long long)(f.ncomp - 1) * f.sq + 1) - 1;
}

// The same for the multi-box level (mfa_level.hip).
void field_range(const mfa_field& f, int D, long long& lo, long long& hi) { byte_range(f, D, lo, hi); }

// A kernel that reads one field while it writes another races if the two
// share memory, so their byte ranges must be disjoint.
static int check_disjoint(const mfa_field* a, const mfa_field* b, int lattice, const char* what)
{
    long long alo, ahi, blo, bhi;
    const int D = lattice == MFA_D2Q9 ? 2 : 3;
    byte_range(*a, D, alo, ahi);
    byte_range(*b, D, blo, bhi);
    if (alo <= bhi && blo <= ahi)
        return fail(MFA_ERR_INVALID, std::string(what) + ": input and output overlap in memory");
    return MFA_OK;
}

static Geo geo_of(const mfa_field& f, const mfa_geom_bc* gb, int lat)
{
    Geo g;
    std::memset(&g, 0, sizeof(g));
    const int D = lattice_dim(lat);
    for (int d = 0; d < 3; ++d) {
        g.n[d] = f.n[d];
        g.per[d] = (d < D) ? (gb ? gb->periodic[d] != 0 : 1) : 1;
    }
    g.walls = 0;
    for (int d = 0; d < D; ++d)
        if (!g.per[d]) g.walls = 1;
    for (int i = 0; i < 27; ++i) {
        g.rtype[i] = gb ? (signed char)gb->resolved_type[i] : MFA_BC_PERIODIC;
        g.rrho[i] = gb ? gb->resolved_rho[i] : 1.0;
        for (int d = 0; d < 3; ++d) g.ru[i][d] = gb ? gb->resolved_u[i][d] : 0.0;
    }
    return g;
}

// The same checks and geometry for the multi-box level (mfa_level.hip).
int field_check(const mfa_field* f, int ncomp_min, int need_grow, const char* what)
{
    return check_field(f, ncomp_min, need_grow, what);
}

Geo geo_for(const int n[3], const mfa_geom_bc* gb, int lat)
{
    mfa_field f{};
    f.n[0] = n[0];
    f.n[1] = n[1];
    f.n[2] = n[2];
    return geo_of(f, gb, lat);
}

static void launch_shape(const int n[3], dim3& grid, dim3& block, int grow)
{
    const int nx = n[0] + 2 * grow;
    const int bx = nx <= 64 ? 64 : (nx <= 128 ? 128 : 256);
    const int by = 256 / bx;
    block = dim3(bx, by, 1);
    grid = dim3((nx + bx - 1) / bx, (n[1] + 2 * grow + by - 1) / by, 1);
}

// Launch configuration of the fused step kernel, per lattice index (D2Q9,
// D3Q15, D3Q19, D3Q27), set by the autotuner through mfa_set_launch.  The
// defaults are the autotuner's picks on MI355X in the row-interleaved order:
// block 128x2, one cell per thread, non-temporal loads for the c_x = 0 rows
// (the x-shifted rows stay cached: neighbouring wavefronts share their
// lines).  D3Q19 256^3 FP64 forced periodic: 0.844 ms against 0.876 ms for
// 256x1 with every load non-temporal; D3Q27 384^3: 4.22 ms, 5.80 TB/s;
// D2Q9 8192^2: 1.49 ms, 6.48 TB/s.  D3Q15 takes D3Q19's.
// z-planes marched by one block of the two-step kernel (its two prologue
// planes are recomputed by the block below: 2/64 extra work).
constexpr int kStep2Planes = 64;

// z-planes per block of the 3-D two-step kernel (one block per CU: its LDS
// fills the CU).  Each block also computes the two planes below its range,
// and blocks run in rounds of one per CU, so the time goes as
// rounds * (planes + 2); the split of the z-range that minimises it, e.g.
// 256^3 in 64 x 8 columns (128 of them on 256 CUs): two blocks of 128
// planes per column, one round.
static int step2_planes(int tiles, int nz, int ncu)
{
    if (ncu < 1) ncu = 1;
    long long best = -1;
    int best_tz = nz;
    for (int zs = 1; zs <= 64 && zs <= nz; ++zs) {
        const int tz = (nz + zs - 1) / zs;
        const long long blocks = (long long)tiles * ((nz + tz - 1) / tz);
        const long long cost = ((blocks + ncu - 1) / ncu) * (long long)(tz + 2);
        if (best < 0 || cost < best) {
            best = cost;
            best_tz = tz;
        }
    }
    return best_tz;
}

static std::mutex g_cfg_mu;
static const mfa_launch k_default_cfg[4] = {
    {128, 2, 1, 2}, {128, 2, 1, 2}, {128, 2, 1, 2}, {128, 2, 1, 2}};
static mfa_launch g_cfg[4] = {
    {128, 2, 1, 2}, {128, 2, 1, 2}, {128, 2, 1, 2}, {128, 2, 1, 2}};

static bool valid_cfg(const mfa_launch& c)
{
    const bool bx_ok = c.block_x == 64 || c.block_x == 128 || c.block_x == 256;
    return bx_ok && c.block_y >= 1 && c.block_x * c.block_y <= 256 && (c.cells_per_thread == 1 ||
           c.cells_per_thread == 2) && (c.nt_loads >= 0 && c.nt_loads <= 2);
}

template <class L, bool FORCED, bool WALLS, int NC, int NT>
static void launch_fused(dim3 grid, dim3 block, hipStream_t st, const double* a, double* b,
                         const Geo& g, const View& vi, const View& vo, double omega, double F0,
                         double F1, double F2)
{
    hipLaunchKernelGGL((k_step<L, FORCED, 0, WALLS, NC, NT>), grid, block, 0, st, a, b, g, vi, vo,
                       omega, F0, F1, F2);
}

template <class L, bool FORCED, bool WALLS>
static void launch_fused_cfg(const mfa_launch& c, dim3 grid, dim3 block, hipStream_t st,
                             const double* a, double* b, const Geo& g, const View& vi,
                             const View& vo, double omega, double F0, double F1, double F2)
{
    auto go = [&](auto nc, auto nt) {
        launch_fused<L, FORCED, WALLS, decltype(nc)::value, decltype(nt)::value>(grid, block, st, a, b, g, vi,
                                                                                 vo, omega, F0, F1, F2);
    };
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I0 = std::integral_constant<int, 0>;
    if (c.cells_per_thread == 2) {
        if (c.nt_loads == 1) go(I2{}, I1{});
        else if (c.nt_loads == 2) go(I2{}, I2{});
        else go(I2{}, I0{});
    } else {
        if (c.nt_loads == 1) go(I1{}, I1{});
        else if (c.nt_loads == 2) go(I1{}, I2{});
        else go(I1{}, I0{});
    }
}

template <class L>
static int launch_step(const mfa_field* fin, mfa_field* fout, const Geo& g, double omega,
                       const double* force, int mode, hipStream_t st, int lat)
{
    const View vi = view_of(*fin), vo = view_of(*fout);
    const double* a = (const double*)fin->data;
    double* b = (double*)fout->data;
    const double F0 = force ? force[0] : 0.0, F1 = force ? force[1] : 0.0,
                 F2 = (force && L::D == 3) ? force[2] : 0.0;
    if (mode == 0) {
        mfa_launch c;
        {
            std::lock_guard<std::mutex> lk(g_cfg_mu);
            c = g_cfg[lat];
        }
        if (L::D == 2) c.cells_per_thread = 1;
        const dim3 block(c.block_x, c.block_y, 1);
        const dim3 grid((g.n[0] + c.block_x - 1) / c.block_x, (g.n[1] + c.block_y - 1) / c.block_y,
                        (g.n[2] + c.cells_per_thread - 1) / c.cells_per_thread);
        if (force) {
            if (g.walls) launch_fused_cfg<L, true, true>(c, grid, block, st, a, b, g, vi, vo, omega, F0, F1, F2);
            else launch_fused_cfg<L, true, false>(c, grid, block, st, a, b, g, vi, vo, omega, F0, F1, F2);
        } else {
            if (g.walls) launch_fused_cfg<L, false, true>(c, grid, block, st, a, b, g, vi, vo, omega, F0, F1, F2);
            else launch_fused_cfg<L, false, false>(c, grid, block, st, a, b, g, vi, vo, omega, F0, F1, F2);
        }
    } else {
        dim3 grid, block;
        launch_shape(g.n, grid, block, 0);
        grid.z = g.n[2];
        if (force) hipLaunchKernelGGL((k_step<L, true, 1, false, 1, 0>), grid, block, 0, st, a, b, g, vi, vo, omega, F0, F1, F2);
        else hipLaunchKernelGGL((k_step<L, false, 1, false, 1, 0>), grid, block, 0, st, a, b, g, vi, vo, omega, F0, F1, F2);
    }
    MFA_HIP_CHECK(hipGetLastError());
    return MFA_OK;
}

template <class Fn>
static int with_lattice(int lat, Fn&& fn)
{
    switch (lat) {
    case MFA_D2Q9: return fn(Lattice<MFA_D2Q9>{});
    case MFA_D3Q15: return fn(Lattice<MFA_D3Q15>{});
    case MFA_D3Q19: return fn(Lattice<MFA_D3Q19>{});
    case MFA_D3Q27: return fn(Lattice<MFA_D3Q27>{});
    default: return fail(MFA_ERR_INVALID, "unknown lattice id");
    }
}

static int check_2d(const mfa_field* f, int lat)
{
    if (lattice_dim(lat) == 2 && f->n[2] != 1)
        return fail(MFA_ERR_INVALID, "2-D lattice needs n[2] == 1");
    return MFA_OK;
}

// Compute units of the current device, read once per device.
static int cu_count(int* out)
{
    static std::mutex mu;
    static int cus[64] = {0};
    int dev = 0;
    MFA_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(MFA_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!cus[dev]) MFA_HIP_CHECK(hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev));
    *out = cus[dev];
    return MFA_OK;
}

// Reduction workspace: one per device and stream, allocated on first use and
// kept until exit, so that reductions queued on different streams do not
// share their partial sums.
static std::mutex g_ws_mu;
static std::map<std::pair<int, hipStream_t>, double*> g_ws;

static int workspace(hipStream_t st, double** out)
{
    int dev = 0;
    MFA_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ws_mu);
    double*& ws = g_ws[std::make_pair(dev, st)];
    if (!ws) {
        void* p = nullptr;
        MFA_HIP_CHECK(hipMalloc(&p, sizeof(double) * kRedBlocks * kRedMaxOut));
        ws = (double*)p;
    }
    *out = ws;
    return MFA_OK;
}

// Workspace of mfa_reduce_noalloc: the partial sums of the reduction kernels,
// kRedBlocks blocks x kRedMaxOut values (216 KiB) in a file-scope device
// array.  Each GPU holds its own copy, at its own address, from load to
// process exit.  The array's name is a host-side stand-in, not a device
// address: host code uses only the address hipGetSymbolAddress returned on
// that device, kept per device.
constexpr int kScratchDoubles = kRedBlocks * kRedMaxOut;
static __device__ __align__(256) double scratch[kScratchDoubles];
static std::mutex g_scratch_mu;
static double* g_scratch[64] = {nullptr};

static int scratch_ptr(double** out)
{
    int dev = 0;
    MFA_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(MFA_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        MFA_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(MFA_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return MFA_OK;
}

static int reduce_with(const mfa_field* f, int lattice, int what, double* out_dev, void* stream,
                       double* ws)
{
    const int Q = lattice_q(lattice);
    if (Q < 0) return fail(MFA_ERR_INVALID, "mfa_reduce: unknown lattice");
    if (!out_dev) return fail(MFA_ERR_INVALID, "mfa_reduce: null output");
    int rc;
    const int need = (what == MFA_RED_COMP_SUM) ? 1 : Q;
    if ((rc = check_field(f, need, 0, "mfa_reduce f", true))) return rc;
    if (what == MFA_RED_COMP_SUM && f->ncomp > kRedMaxOut)
        return fail(MFA_ERR_INVALID, "mfa_reduce: more than 27 components");
    const Geo geo = geo_of(*f, nullptr, lattice);
    const View v = view_of(*f);
    hipStream_t st = (hipStream_t)stream;
    int nv = 0;
    if (f->dtype != MFA_F64) {
        if ((rc = reduce_dev(f, geo, lattice, what, ws, kRedBlocks, st))) return rc;
        nv = (what == MFA_RED_MASS) ? 1 : (what == MFA_RED_HEALTH ? 2 : lattice_dim(lattice));
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, st, ws, kRedBlocks, nv, out_dev);
        MFA_HIP_CHECK(hipGetLastError());
        return MFA_OK;
    }
    rc = with_lattice(lattice, [&](auto L) {
        using LT = decltype(L);
        if (what == MFA_RED_MASS) {
            hipLaunchKernelGGL((k_reduce_f<LT, MFA_RED_MASS>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, (const double*)f->data, geo, v, f->ncomp, ws);
            nv = 1;
        } else if (what == MFA_RED_MOMENTUM) {
            hipLaunchKernelGGL((k_reduce_f<LT, MFA_RED_MOMENTUM>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, (const double*)f->data, geo, v, f->ncomp, ws);
            nv = LT::D;
        } else if (what == MFA_RED_HEALTH) {
            hipLaunchKernelGGL((k_reduce_f<LT, MFA_RED_HEALTH>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, (const double*)f->data, geo, v, f->ncomp, ws);
            nv = 2;
        } else if (what == MFA_RED_COMP_SUM) {
            hipLaunchKernelGGL((k_reduce_f<LT, MFA_RED_COMP_SUM>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, (const double*)f->data, geo, v, f->ncomp, ws);
            nv = f->ncomp;
        } else {
            return fail(MFA_ERR_INVALID, "mfa_reduce: unknown reduction");
        }
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    });
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, st, ws, kRedBlocks, nv, out_dev);
    MFA_HIP_CHECK(hipGetLastError());
    return MFA_OK;
}

}  // namespace mfa

using namespace mfa;

extern "C" {

const char* mfa_version(void) { return "muphfasa-mi355x 0.1 (gfx950)"; }
const char* mfa_last_error(void) { return g_err.c_str(); }

int mfa_resolve_bc(int lattice, const int face_type[6], const double face_rho[6],
                   const double face_u[18], mfa_geom_bc* out)
{
    // src/LBM/BoundaryConditions.cpp:37-156 and BoundaryConditions.H:82-92
    if (!out || !face_type) return fail(MFA_ERR_INVALID, "mfa_resolve_bc: null argument");
    const int D = lattice_q(lattice) < 0 ? -1 : lattice_dim(lattice);
    if (D < 0) return fail(MFA_ERR_INVALID, "mfa_resolve_bc: unknown lattice");
    auto rank = [](int t) {
        switch (t) {
        case MFA_BC_NO_SLIP: return 3;
        case MFA_BC_MOVING_WALL: return 2;
        case MFA_BC_INFLOW: return 1;
        case MFA_BC_OUTFLOW:
        case MFA_BC_PRESSURE_OUTLET: return 0;
        default: return -1;
        }
    };
    out->lattice = lattice;
    // periodicity is taken from the lo face (src/Geometry/GeometryHandler.cpp:40-47)
    for (int d = 0; d < 3; ++d) out->periodic[d] = (d < D) ? (face_type[2 * d] == MFA_BC_PERIODIC) : 1;
    for (int i = 0; i < 27; ++i) {
        out->resolved_type[i] = MFA_BC_PERIODIC;
        out->resolved_rho[i] = 1.0;
        out->resolved_u[i][0] = out->resolved_u[i][1] = out->resolved_u[i][2] = 0.0;
    }
    for (int f = 0; f < 2 * D; ++f) {
        if (face_type[f] < 0 || face_type[f] > MFA_BC_PERIODIC)
            return fail(MFA_ERR_INVALID, "mfa_resolve_bc: unknown boundary type");
        if ((f & 1) && face_type[f] == MFA_BC_PERIODIC && face_type[f ^ 1] != MFA_BC_PERIODIC)
            return fail(MFA_ERR_INVALID,
                        "buildBCManager: bc input requests \"Periodic\" on a face whose Geometry is "
                        "not periodic in that direction");
    }
    const int ncodes = (D == 2) ? 9 : 27;
    for (int idx = 0; idx < ncodes; ++idx) {
        int code[3] = {0, 0, 0}, tmp = idx;
        bool all_zero = true;
        for (int d = 0; d < D; ++d) {
            code[d] = tmp % 3 - 1;
            tmp /= 3;
            if (code[d]) all_zero = false;
        }
        if (all_zero) continue;
        int types[3], nf = 0;
        double rhos[3], us[3][3];
        for (int d = 0; d < D; ++d) {
            if (!code[d]) continue;
            const int fi = 2 * d + (code[d] < 0 ? 0 : 1);
            if (face_type[fi] == MFA_BC_PERIODIC) continue;
            types[nf] = face_type[fi];
            rhos[nf] = face_rho ? face_rho[fi] : 1.0;
            for (int c = 0; c < 3; ++c) us[nf][c] = face_u ? face_u[3 * fi + c] : 0.0;
            ++nf;
        }
        if (!nf) continue;
        int hin = 0, hout = 0, hp = 0, nmov = 0;
        for (int i = 0; i < nf; ++i) {
            hin |= types[i] == MFA_BC_INFLOW;
            hout |= types[i] == MFA_BC_OUTFLOW;
            hp |= types[i] == MFA_BC_PRESSURE_OUTLET;
            nmov += types[i] == MFA_BC_MOVING_WALL;
        }
        if (hin + hout + hp > 1 || nmov > 1)
            return fail(MFA_ERR_CONFLICT,
                        "resolveBC: conflicting boundary conditions at same edge/corner");
        int best = -1;
        for (int i = 0; i < nf; ++i) best = rank(types[i]) > best ? rank(types[i]) : best;
        int winner = 0;
        for (int i = 0; i < nf; ++i)
            if (rank(types[i]) == best) { winner = types[i]; break; }
        double sr = 0, su[3] = {0, 0, 0};
        int cnt = 0;
        for (int i = 0; i < nf; ++i) {
            if (types[i] != winner) continue;
            sr += rhos[i];
            for (int d = 0; d < D; ++d) su[d] += us[i][d];
            ++cnt;
        }
        out->resolved_type[idx] = winner;
        out->resolved_rho[idx] = sr / cnt;
        for (int d = 0; d < D; ++d) out->resolved_u[idx][d] = su[d] / cnt;
    }
    return MFA_OK;
}

int mfa_step2(const mfa_field* fin, mfa_field* fout, const mfa_geom_bc* g, double omega,
              const double* force, void* stream)
{
    if (!g) return fail(MFA_ERR_INVALID, "mfa_step2: null geometry");
    const int lat = g->lattice;
    const int Q = lattice_q(lat);
    if (Q < 0) return fail(MFA_ERR_INVALID, "mfa_step2: unknown lattice");
    int rc;
    if ((rc = check_field(fin, Q, 0, "mfa_step2 fin", true)) ||
        (rc = check_field(fout, Q, 0, "mfa_step2 fout", true)))
        return rc;
    if (std::memcmp(fin->n, fout->n, sizeof(fin->n)))
        return fail(MFA_ERR_INVALID, "mfa_step2: fin/fout boxes differ");
    if ((rc = check_disjoint(fin, fout, lat, "mfa_step2"))) return rc;
    if (fin->dtype != fout->dtype) return fail(MFA_ERR_INVALID, "mfa_step2: fin/fout storage types differ");
    if ((rc = check_2d(fin, lat))) return rc;
    const Geo geo = geo_of(*fin, g, lat);
    if (lattice_dim(lat) == 2 && fin->dtype != MFA_F64)
        return step2_dev(fin, fout, geo, lat, omega, force, 1, (hipStream_t)stream);
    if (lattice_dim(lat) == 2) {
        // D2Q9, FP64, periodic or with walls; 64 x 4 tiles
        using L = Lattice<MFA_D2Q9>;
        using T = Step2Tile<64, 4>;
        const View vi = view_of(*fin), vo = view_of(*fout);
        const double F0 = force ? force[0] : 0.0, F1 = force ? force[1] : 0.0;
        const dim3 grid((geo.n[0] + T::TX - 1) / T::TX, (geo.n[1] + T::TY - 1) / T::TY);
        const double* a = (const double*)fin->data;
        double* b = (double*)fout->data;
        hipStream_t st = (hipStream_t)stream;
        if (force) {
            if (geo.walls) hipLaunchKernelGGL((k_step2_2d<L, true, true, T>), grid, dim3(T::THREADS), 0, st, a, b, geo, vi, vo, omega, F0, F1, 0.0);
            else hipLaunchKernelGGL((k_step2_2d<L, true, false, T>), grid, dim3(T::THREADS), 0, st, a, b, geo, vi, vo, omega, F0, F1, 0.0);
        } else {
            if (geo.walls) hipLaunchKernelGGL((k_step2_2d<L, false, true, T>), grid, dim3(T::THREADS), 0, st, a, b, geo, vi, vo, omega, F0, F1, 0.0);
            else hipLaunchKernelGGL((k_step2_2d<L, false, false, T>), grid, dim3(T::THREADS), 0, st, a, b, geo, vi, vo, omega, F0, F1, 0.0);
        }
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    }
    // in-plane byte offsets are 32-bit in the 3-D pair kernels (k_step2p,
    // k_step2p_dev) and end below kDropOffset; checked before any device query
    {
        const long long es = fin->dtype == MFA_F64 ? 8 : (fin->dtype == MFA_F32 ? 4 : 2);
        for (const mfa_field* f : {fin, (const mfa_field*)fout})
            if (((long long)(f->n[0] - 1) + (long long)(f->n[1] - 1) * f->sy + 1) * es > (long long)kDropOffset)
                return fail(MFA_ERR_UNSUPPORTED, "mfa_step2: a z-plane spans more than 4 GiB (use mfa_step)");
    }
    if (fin->dtype != MFA_F64) {
        if (geo.walls)
            return fail(MFA_ERR_UNSUPPORTED,
                        "mfa_step2: 3-D boxes with walls need FP64 storage (use mfa_step)");
        // FP32 pairs run k_step2p_dev on 64 x 8 columns: one round of blocks
        // per CU, as the FP64 pass
        int ncu = 0;
        if ((rc = cu_count(&ncu))) return rc;
        const int TZ = step2_planes(((geo.n[0] + 63) / 64) * ((geo.n[1] + 7) / 8), geo.n[2], ncu);
        return step2_dev(fin, fout, geo, lat, omega, force, TZ, (hipStream_t)stream);
    }
    const View vi = view_of(*fin), vo = view_of(*fout);
    const double F0 = force ? force[0] : 0.0, F1 = force ? force[1] : 0.0, F2 = force ? force[2] : 0.0;
    int ncu = 0;
    if ((rc = cu_count(&ncu))) return rc;
    const double* a = (const double*)fin->data;
    double* b = (double*)fout->data;
    auto run_geo = [&](const Geo& gg, auto Lc, auto Tc, auto walls) {
        using L = decltype(Lc);
        using T = decltype(Tc);
        constexpr bool W = decltype(walls)::value;
        const int tiles = ((gg.n[0] + T::TX - 1) / T::TX) * ((gg.n[1] + T::TY - 1) / T::TY);
        const int TZ = step2_planes(tiles, gg.n[2], ncu);
        const dim3 grid((gg.n[0] + T::TX - 1) / T::TX, (gg.n[1] + T::TY - 1) / T::TY,
                        (gg.n[2] + TZ - 1) / TZ);
        if (force)
            hipLaunchKernelGGL((k_step2p<L, true, W, T>), grid, dim3(T::THREADS), 0, (hipStream_t)stream, a, b,
                               gg, vi, vo, omega, F0, F1, F2, TZ);
        else
            hipLaunchKernelGGL((k_step2p<L, false, W, T>), grid, dim3(T::THREADS), 0, (hipStream_t)stream, a,
                               b, gg, vi, vo, omega, F0, F1, F2, TZ);
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    };
    auto run = [&](auto Lc, auto Tc, auto walls) { return run_geo(geo, Lc, Tc, walls); };
    // Walled boxes whose wall dimensions hold at least three cells: the
    // periodic pass for the cells two or more cells from every wall, then
    // k_step2_faces for the shell (kernels_step2.hpp).  D3Q19 256^3 lid
    // cavity: 0.88 ms per step with the walled pass (64 x 4 columns, every
    // row of an x-wall tile divergent at its edge lane) against two single
    // steps' 0.895.
    auto run_shell = [&](auto Lc, auto Tc) {
        using L = decltype(Lc);
        // 16 x 4 face tiles (49 KiB of LDS, three blocks per CU): lid cavity
        // 0.587 ms per step, against 0.627 (16 x 8), 0.632 (32 x 4), 0.598
        // (8 x 8, 128 threads) and 0.601 (16 x 4, 128 threads)
        using FT = FaceTile<16, 4>;
        Geo gp = geo;
        for (int d = 0; d < 3; ++d) gp.per[d] = 1;
        gp.walls = 0;
        int rc2 = run_geo(gp, Lc, Tc, std::false_type{});
        if (rc2) return rc2;
        FaceList fl;
        std::memset(&fl, 0, sizeof(fl));
        int blocks = 0;
        for (int d = 0; d < 3; ++d) {
            if (geo.per[d]) continue;
            const int da = d == 0 ? 1 : 0, db = d == 2 ? 1 : 2;
            for (int h = 0; h < 2; ++h) {
                const int k = fl.n++;
                fl.dim[k] = d;
                fl.hi[k] = h;
                fl.ta[k] = (geo.n[da] + FT::TA - 1) / FT::TA;
                fl.tb[k] = (geo.n[db] + FT::TB - 1) / FT::TB;
                fl.first[k] = blocks;
                blocks += fl.ta[k] * fl.tb[k];
            }
        }
        fl.first[fl.n] = blocks;
        if (force)
            hipLaunchKernelGGL((k_step2_faces<L, true, FT>), dim3(blocks), dim3(FT::THREADS), 0, (hipStream_t)stream,
                               a, b, geo, vi, vo, omega, F0, F1, F2, fl);
        else
            hipLaunchKernelGGL((k_step2_faces<L, false, FT>), dim3(blocks), dim3(FT::THREADS), 0,
                               (hipStream_t)stream, a, b, geo, vi, vo, omega, F0, F1, F2, fl);
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    };
    // Columns (D3Q19 256^3 forced, ms per step): periodic 64 x 8, 0.513,
    // against 0.554 for 64 x 4 and 0.577 for the grown-column layout of
    // 64 x 4; D3Q15 64 x 8, 0.379; D3Q27 (27 rings of planes, 143 KB of
    // LDS) 64 x 5, 1.8 % faster than 64 x 4 (2.68 ms) at 384^3.  With walls
    // class 0 needs a second ring slot and the wall values hold the cell's
    // populations in registers: 64 x 4.
    using Yes = std::true_type;
    using No = std::false_type;
    bool shell = geo.walls != 0;
    for (int d = 0; d < 3; ++d)
        if (!geo.per[d] && geo.n[d] < 3) shell = false;
    if (shell) {
        if (lat == MFA_D3Q15) return run_shell(Lattice<MFA_D3Q15>{}, PushTile<8>{});
        if (lat == MFA_D3Q19) return run_shell(Lattice<MFA_D3Q19>{}, PushTile<8>{});
        return run_shell(Lattice<MFA_D3Q27>{}, PushTile<5>{});
    }
    if (geo.walls) {
        if (lat == MFA_D3Q15) return run(Lattice<MFA_D3Q15>{}, PushTile<4>{}, Yes{});
        if (lat == MFA_D3Q19) return run(Lattice<MFA_D3Q19>{}, PushTile<4>{}, Yes{});
        return run(Lattice<MFA_D3Q27>{}, PushTile<4>{}, Yes{});
    }
    if (lat == MFA_D3Q15) return run(Lattice<MFA_D3Q15>{}, PushTile<8>{}, No{});
    if (lat == MFA_D3Q19) return run(Lattice<MFA_D3Q19>{}, PushTile<8>{}, No{});
    return run(Lattice<MFA_D3Q27>{}, PushTile<5>{}, No{});
}

int mfa_step(const mfa_field* fin, mfa_field* fout, const mfa_geom_bc* g, double omega,
             const double* force, void* stream)
{
    if (!g) return fail(MFA_ERR_INVALID, "mfa_step: null geometry");
    const int lat = g->lattice;
    const int Q = lattice_q(lat);
    if (Q < 0) return fail(MFA_ERR_INVALID, "mfa_step: unknown lattice");
    int rc;
    if ((rc = check_field(fin, Q, 0, "mfa_step fin", true)) ||
        (rc = check_field(fout, Q, 0, "mfa_step fout", true)))
        return rc;
    if (std::memcmp(fin->n, fout->n, sizeof(fin->n)))
        return fail(MFA_ERR_INVALID, "mfa_step: fin/fout boxes differ");
    if ((rc = check_disjoint(fin, fout, lat, "mfa_step"))) return rc;
    if (fin->dtype != fout->dtype) return fail(MFA_ERR_INVALID, "mfa_step: fin/fout storage types differ");
    if ((rc = check_2d(fin, lat))) return rc;
    const Geo geo = geo_of(*fin, g, lat);
    if (fin->dtype != MFA_F64) return step_dev(fin, fout, geo, lat, omega, force, (hipStream_t)stream);
    return with_lattice(lat, [&](auto L) {
        return launch_step<decltype(L)>(fin, fout, geo, omega, force, 0, (hipStream_t)stream, lat);
    });
}

int mfa_stream_copy(const void* src, void* dst, size_t bytes, int nontemporal, void* stream)
{
    if (!src || !dst || bytes % 16) return fail(MFA_ERR_INVALID, "mfa_stream_copy: needs 16-byte multiples");
    const long long n2 = (long long)(bytes / 16);
    const dim3 grid((unsigned)((n2 + 255) / 256));
    if (nontemporal)
        hipLaunchKernelGGL(k_copy<true>, grid, dim3(256), 0, (hipStream_t)stream, (const dbl2*)src, (dbl2*)dst, n2);
    else
        hipLaunchKernelGGL(k_copy<false>, grid, dim3(256), 0, (hipStream_t)stream, (const dbl2*)src, (dbl2*)dst, n2);
    MFA_HIP_CHECK(hipGetLastError());
    return MFA_OK;
}

int mfa_convert(const mfa_field* src, mfa_field* dst, int lattice, void* stream)
{
    const int Q = lattice_q(lattice);
    if (Q < 0) return fail(MFA_ERR_INVALID, "mfa_convert: unknown lattice");
    int rc;
    if ((rc = check_field(src, Q, 0, "mfa_convert src", true)) ||
        (rc = check_field(dst, Q, 0, "mfa_convert dst", true)))
        return rc;
    if (std::memcmp(src->n, dst->n, sizeof(src->n))) return fail(MFA_ERR_INVALID, "mfa_convert: boxes differ");
    if ((src->dtype == MFA_F64) == (dst->dtype == MFA_F64))
        return fail(MFA_ERR_INVALID, "mfa_convert: exactly one side must be FP64");
    if ((rc = check_disjoint(src, dst, lattice, "mfa_convert"))) return rc;
    if ((rc = check_2d(src, lattice))) return rc;
    return convert_dev(src, dst, geo_of(*src, nullptr, lattice), lattice, (hipStream_t)stream);
}

int mfa_set_launch(int lattice, const mfa_launch* cfg)
{
    if (lattice_q(lattice) < 0) return fail(MFA_ERR_INVALID, "mfa_set_launch: unknown lattice");
    const mfa_launch c = cfg ? *cfg : k_default_cfg[lattice];
    if (!valid_cfg(c)) return fail(MFA_ERR_INVALID, "mfa_set_launch: unsupported launch configuration");
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    g_cfg[lattice] = c;
    return MFA_OK;
}

int mfa_get_launch(int lattice, mfa_launch* cfg)
{
    if (lattice_q(lattice) < 0 || !cfg) return fail(MFA_ERR_INVALID, "mfa_get_launch: bad argument");
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    *cfg = g_cfg[lattice];
    return MFA_OK;
}

int mfa_pull_collide(const mfa_field* fin, mfa_field* fout, int lattice, double omega,
                     const double* force, void* stream)
{
    const int Q = lattice_q(lattice);
    if (Q < 0) return fail(MFA_ERR_INVALID, "mfa_pull_collide: unknown lattice");
    int rc;
    if ((rc = check_field(fin, Q, 1, "mfa_pull_collide fin")) ||
        (rc = check_field(fout, Q, 0, "mfa_pull_collide fout")))
        return rc;
    if (std::memcmp(fin->n, fout->n, sizeof(fin->n)))
        return fail(MFA_ERR_INVALID, "mfa_pull_collide: fin/fout boxes differ");
    if ((rc = check_disjoint(fin, fout, lattice, "mfa_pull_collide"))) return rc;
    if ((rc = check_2d(fin, lattice))) return rc;
    const Geo geo = geo_of(*fin, nullptr, lattice);
    return with_lattice(lattice, [&](auto L) {
        return launch_step<decltype(L)>(fin, fout, geo, omega, force, 1, (hipStream_t)stream, lattice);
    });
}

int mfa_fill_periodic(mfa_field* f, const mfa_geom_bc* g, void* stream)
{
    if (!g) return fail(MFA_ERR_INVALID, "mfa_fill_periodic: null geometry");
    const int Q = lattice_q(g->lattice);
    if (Q < 0) return fail(MFA_ERR_INVALID, "mfa_fill_periodic: unknown lattice");
    int rc;
    if ((rc = check_field(f, Q, 1, "mfa_fill_periodic f"))) return rc;
    if ((rc = check_2d(f, g->lattice))) return rc;
    const Geo geo = geo_of(*f, g, g->lattice);
    const View v = view_of(*f);
    return with_lattice(g->lattice, [&](auto L) {
        using LT = decltype(L);
        dim3 grid, block;
        launch_shape(geo.n, grid, block, 1);
        grid.z = geo.n[2] + (LT::D == 3 ? 2 : 0);
        hipLaunchKernelGGL(k_fill_periodic<LT>, grid, block, 0, (hipStream_t)stream, (double*)f->data, geo, v);
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_apply_bc(mfa_field* f, const mfa_geom_bc* g, void* stream)
{
    if (!g) return fail(MFA_ERR_INVALID, "mfa_apply_bc: null geometry");
    const int Q = lattice_q(g->lattice);
    if (Q < 0) return fail(MFA_ERR_INVALID, "mfa_apply_bc: unknown lattice");
    int rc;
    if ((rc = check_field(f, Q, 1, "mfa_apply_bc f"))) return rc;
    if ((rc = check_2d(f, g->lattice))) return rc;
    const Geo geo = geo_of(*f, g, g->lattice);
    if (!geo.walls) return MFA_OK;
    const View v = view_of(*f);
    return with_lattice(g->lattice, [&](auto L) {
        using LT = decltype(L);
        dim3 grid, block;
        launch_shape(geo.n, grid, block, 1);
        grid.z = geo.n[2] + (LT::D == 3 ? 2 : 0);
        hipLaunchKernelGGL(k_apply_bc<LT>, grid, block, 0, (hipStream_t)stream, (double*)f->data, geo, v);
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_moments(const mfa_field* f, mfa_field* rho, mfa_field* vel, const mfa_geom_bc* g,
                const double* force, int streamed, void* stream)
{
    if (!g) return fail(MFA_ERR_INVALID, "mfa_moments: null geometry");
    const int lat = g->lattice;
    const int Q = lattice_q(lat);
    if (Q < 0) return fail(MFA_ERR_INVALID, "mfa_moments: unknown lattice");
    const int D = lattice_dim(lat);
    int rc;
    if ((rc = check_field(f, Q, 0, "mfa_moments f", true)) || (rc = check_field(rho, 1, 0, "mfa_moments rho")) ||
        (rc = check_field(vel, D, 0, "mfa_moments vel")))
        return rc;
    if (std::memcmp(f->n, rho->n, sizeof(f->n)) || std::memcmp(f->n, vel->n, sizeof(f->n)))
        return fail(MFA_ERR_INVALID, "mfa_moments: boxes differ");
    if ((rc = check_2d(f, lat))) return rc;
    if (streamed == 2 && (f->dtype != MFA_F64 || f->ngrow < 1))
        return fail(MFA_ERR_INVALID, "mfa_moments: streamed = 2 pulls through ghost cells: FP64 with ngrow 1");
    const Geo geo = geo_of(*f, g, lat);
    if (f->dtype != MFA_F64)
        return moments_dev(f, rho, vel, geo, lat, force, streamed, (hipStream_t)stream);
    const View vi = view_of(*f), vr = view_of(*rho), vv = view_of(*vel);
    const double F0 = force ? force[0] : 0.0, F1 = force ? force[1] : 0.0,
                 F2 = (force && D == 3) ? force[2] : 0.0;
    return with_lattice(lat, [&](auto L) {
        using LT = decltype(L);
        dim3 grid, block;
        launch_shape(geo.n, grid, block, 0);
        grid.z = geo.n[2];
        hipStream_t st = (hipStream_t)stream;
        const double* a = (const double*)f->data;
        double* r = (double*)rho->data;
        double* u = (double*)vel->data;
        if (streamed == 2) {
            if (force) hipLaunchKernelGGL((k_moments<LT, true, 1>), grid, block, 0, st, a, r, u, geo, vi, vr, vv, F0, F1, F2);
            else hipLaunchKernelGGL((k_moments<LT, false, 1>), grid, block, 0, st, a, r, u, geo, vi, vr, vv, F0, F1, F2);
        } else if (streamed) {
            if (force) hipLaunchKernelGGL((k_moments<LT, true, 0>), grid, block, 0, st, a, r, u, geo, vi, vr, vv, F0, F1, F2);
            else hipLaunchKernelGGL((k_moments<LT, false, 0>), grid, block, 0, st, a, r, u, geo, vi, vr, vv, F0, F1, F2);
        } else {
            if (force) hipLaunchKernelGGL((k_moments<LT, true, 2>), grid, block, 0, st, a, r, u, geo, vi, vr, vv, F0, F1, F2);
            else hipLaunchKernelGGL((k_moments<LT, false, 2>), grid, block, 0, st, a, r, u, geo, vi, vr, vv, F0, F1, F2);
        }
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_init_equilibrium(mfa_field* rho, mfa_field* vel, mfa_field* f, int lattice,
                         const double* force, void* stream)
{
    const int Q = lattice_q(lattice);
    if (Q < 0) return fail(MFA_ERR_INVALID, "mfa_init_equilibrium: unknown lattice");
    const int D = lattice_dim(lattice);
    int rc;
    if ((rc = check_field(f, Q, 0, "mfa_init_equilibrium f", true)) ||
        (rc = check_field(rho, 1, 0, "mfa_init_equilibrium rho")) ||
        (rc = check_field(vel, D, 0, "mfa_init_equilibrium vel")))
        return rc;
    if (std::memcmp(f->n, rho->n, sizeof(f->n)) || std::memcmp(f->n, vel->n, sizeof(f->n)))
        return fail(MFA_ERR_INVALID, "mfa_init_equilibrium: boxes differ");
    if ((rc = check_2d(f, lattice))) return rc;
    const Geo geo = geo_of(*f, nullptr, lattice);
    if (f->dtype != MFA_F64) return init_dev(rho, vel, f, geo, lattice, force, (hipStream_t)stream);
    const View vr = view_of(*rho), vv = view_of(*vel), vo = view_of(*f);
    const double F0 = force ? force[0] : 0.0, F1 = force ? force[1] : 0.0,
                 F2 = (force && D == 3) ? force[2] : 0.0;
    return with_lattice(lattice, [&](auto L) {
        using LT = decltype(L);
        dim3 grid, block;
        launch_shape(geo.n, grid, block, 0);
        grid.z = geo.n[2];
        hipStream_t st = (hipStream_t)stream;
        if (force) hipLaunchKernelGGL((k_init<LT, true>), grid, block, 0, st, (double*)rho->data, (double*)vel->data, (double*)f->data, geo, vr, vv, vo, F0, F1, F2);
        else hipLaunchKernelGGL((k_init<LT, false>), grid, block, 0, st, (double*)rho->data, (double*)vel->data, (double*)f->data, geo, vr, vv, vo, F0, F1, F2);
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_reduce(const mfa_field* f, int lattice, int what, double* out_dev, void* stream)
{
    double* ws = nullptr;
    int rc;
    if ((rc = workspace((hipStream_t)stream, &ws))) return rc;
    return reduce_with(f, lattice, what, out_dev, stream, ws);
}

int mfa_reduce_noalloc(const mfa_field* f, int lattice, int what, double* out_dev, void* stream)
{
    // The kernels' partial sums go to the file-scope scratch array: every
    // block writes each partial the final pass reads, so no clearing is
    // needed, and a field of any size fits (the blocks stride over it).
    double* ws = nullptr;
    int rc;
    if ((rc = scratch_ptr(&ws))) return rc;
    return reduce_with(f, lattice, what, out_dev, stream, ws);
}

int mfa_residual(const mfa_field* a, const mfa_field* b, double* out_dev, void* stream)
{
    int rc;
    if ((rc = check_field(a, 1, 0, "mfa_residual a")) || (rc = check_field(b, 1, 0, "mfa_residual b")))
        return rc;
    if (std::memcmp(a->n, b->n, sizeof(a->n)) || a->ncomp != b->ncomp)
        return fail(MFA_ERR_INVALID, "mfa_residual: shapes differ");
    if (!out_dev) return fail(MFA_ERR_INVALID, "mfa_residual: null output");
    double* ws = nullptr;
    if ((rc = workspace((hipStream_t)stream, &ws))) return rc;
    const Geo geo = geo_of(*a, nullptr, MFA_D3Q19);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_residual, dim3(kRedBlocks), dim3(kRedThreads), 0, st, (const double*)a->data,
                       (const double*)b->data, geo, view_of(*a), view_of(*b), a->ncomp, ws);
    MFA_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, st, ws, kRedBlocks, 2, out_dev);
    MFA_HIP_CHECK(hipGetLastError());
    return MFA_OK;
}

}  // extern "C"
