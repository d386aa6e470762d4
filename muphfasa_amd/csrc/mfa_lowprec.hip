// mfa_lowprec.hip -- reduced-precision population storage for the step
// engine.  Populations are stored as the shifted deviation d_q = f_q - w_q in
// FP32, FP16 or BF16 and collided in FP32; global sums stay FP64.
//
// Storing the deviation instead of f keeps the stored magnitude near the
// size of the flow perturbation (|d_q| ~ w_q * 3 |c.u| ~ 1e-2 for u ~ 0.05)
// rather than w_q ~ 0.3, so the same mantissa resolves the dynamics 10-100x
// more finely; the equilibrium is evaluated directly in deviation form,
//   feq_q - w_q = w_q * (drho + rho * (3 c.u + 4.5 (c.u)^2 - 1.5 u.u)),
// so w_q never enters the arithmetic and is never lost to rounding.
// Conversions to FP16/BF16 round to nearest even.  Bytes per lattice update:
// 2*Q*sizeof(T), i.e. 152 B (FP32) and 76 B (FP16/BF16) for D3Q19 versus
// 304 B in FP64.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "kernels_lowprec.hpp"
#include "kernels_step2.hpp"

namespace mfa {

// One cell per thread.  blockDim.x is a multiple of 64 (checked by the
// host), so a wavefront lies in one row and y is the same on all its lanes:
// the pull-source rows and the output row are then wave-uniform addresses.
template <class L, bool FORCED, bool WALLS, class T>
__global__ __launch_bounds__(256) void k_step_dev(const T* __restrict__ fin, T* __restrict__ fout,
                                                  const Geo g, const View vi, const View vo,
                                                  const float omega, const float F0, const float F1,
                                                  const float F2)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = __builtin_amdgcn_readfirstlane(blockIdx.y * blockDim.y + threadIdx.y);
    const int z = blockIdx.z;
    if (x >= g.n[0] || y >= g.n[1]) return;
    float d[L::Q];
    pull_dev_rows<L, WALLS>(d, fin, x, y, z, g, vi);
    collide_store_dev_row<L, FORCED>(d, fout + y * vo.sy + z * vo.sz, x, vo, omega, F0, F1, F2);
}

// Storage word of T, and two of them loaded or stored as one access.
template <class T> struct Word;
template <> struct Word<float> {
    typedef unsigned int U;
    typedef unsigned int U2 __attribute__((ext_vector_type(2)));
};
template <> struct Word<f16_t> {
    typedef unsigned short U;
    typedef unsigned short U2 __attribute__((ext_vector_type(2)));
};
template <> struct Word<bf16_t> {
    typedef unsigned short U;
    typedef unsigned short U2 __attribute__((ext_vector_type(2)));
};

template <class T>
__device__ __forceinline__ float word_to_f(typename Word<T>::U u) { return to_f(__builtin_bit_cast(T, u)); }
template <class T>
__device__ __forceinline__ typename Word<T>::U f_to_word(float v)
{
    return __builtin_bit_cast(typename Word<T>::U, from_f<T>(v));
}

// The same update with two x-adjacent cells per thread: each population row
// is read and written as aligned pairs (4 B for FP16/BF16, 8 B for FP32)
// instead of single 2- or 4-byte words.  A pull with c_x = +1 needs the pair
// shifted right by one element: its first element is the last element of the
// previous lane's pair (cross-lane shift), or, on lane 0 of a wavefront, one
// extra load; c_x = -1 mirrors this.  Per cell the arithmetic is that of
// k_step_dev, so the two kernels give identical bits.  Requirements (checked
// by the host): even n_x, x strides of pairs aligned, blockDim.x a multiple
// of 64 (a wavefront never spans two rows).
template <class L, bool FORCED, bool WALLS, bool NT, class T>
__global__ __launch_bounds__(256) void k_step_dev_x2(const T* __restrict__ fin, T* __restrict__ fout,
                                                     const Geo g, const View vi, const View vo,
                                                     const float omega, const float F0, const float F1,
                                                     const float F2)
{
    using U = typename Word<T>::U;
    using U2 = typename Word<T>::U2;
    const U* __restrict__ in = (const U*)fin;
    const int x0 = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
    // a wavefront lies in one row (blockDim.x is a multiple of 64)
    const int y = __builtin_amdgcn_readfirstlane(blockIdx.y * blockDim.y + threadIdx.y);
    const int z = blockIdx.z;
    const int nx = g.n[0];
    if (y >= g.n[1]) return;                 // whole wavefronts only
    const bool active = x0 < nx;
    const int lane = threadIdx.x & 63;
    // x of the extra loads: the element left of x0 and right of x0+1
    const int xl = x0 == 0 ? (g.per[0] ? nx - 1 : 0) : x0 - 1;
    const int xr = x0 + 2 >= nx ? (g.per[0] ? 0 : nx - 1) : x0 + 2;
    const bool edge_l = active && lane == 0;
    const bool edge_r = active && (lane == 63 || x0 + 2 >= nx);
    const unsigned b0 = (unsigned)x0 * (unsigned)sizeof(U), bl = (unsigned)xl * (unsigned)sizeof(U),
                   br = (unsigned)xr * (unsigned)sizeof(U);
    long long rows[3][3];
    source_rows<L::D>(rows, y, z, g, vi);
    float da[L::Q], db[L::Q];
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = L::c(q, 0), cy = L::c(q, 1), cz = L::D == 3 ? L::c(q, 2) : 0;
        const U* row = in + q * vi.sq + rows[cy + 1][cz + 1];
        U2 p = {0, 0};
        if (active) {
            if constexpr (NT) p = __builtin_nontemporal_load(at_bytes((const U2*)row, b0));
            else p = *at_bytes((const U2*)row, b0);
        }
        if constexpr (cx == 0) {
            da[q] = word_to_f<T>(p.x);
            db[q] = word_to_f<T>(p.y);
        } else if constexpr (cx == 1) {      // sources x0-1, x0
            U left = (U)__shfl_up((unsigned int)p.y, 1, 64);
            if (edge_l) left = *at_bytes(row, bl);
            da[q] = word_to_f<T>(left);
            db[q] = word_to_f<T>(p.x);
        } else {                             // sources x0+1, x0+2
            U right = (U)__shfl_down((unsigned int)p.x, 1, 64);
            if (edge_r) right = *at_bytes(row, br);
            da[q] = word_to_f<T>(p.y);
            db[q] = word_to_f<T>(right);
        }
    });
    if (!active) return;
    if constexpr (WALLS) {
        wall_fix_dev<L>(da, fin, x0, y, z, g, vi);
        wall_fix_dev<L>(db, fin, x0 + 1, y, z, g, vi);
    }
    float oa[L::Q], ob[L::Q];
    collide_dev<L, FORCED>(da, oa, omega, F0, F1, F2);
    collide_dev<L, FORCED>(db, ob, omega, F0, F1, F2);
    U* __restrict__ orow = (U*)fout + y * vo.sy + z * vo.sz;
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const U2 w = {f_to_word<T>(oa[q]), f_to_word<T>(ob[q])};
        __builtin_nontemporal_store(w, at_bytes((U2*)(orow + q * vo.sq), b0));
    });
}

// Tile of the 2-D two-step pass on reduced storage (k_step2_2d_dev): TX x TY
// output cells, the tile grown by one cell in LDS.
template <int TX_, int TY_>
struct Step2DevTile {
    static constexpr int TX = TX_, TY = TY_;
    static constexpr int EX = TX + 2, EY = TY + 2, EC = EX * EY;
    static constexpr int THREADS = (EC + 63) / 64 * 64;
};

// Moments of reduced storage, evaluated in FP64 from the dequantised
// populations w_q + d_q with the reference's arithmetic.
template <class L, bool FORCED, bool STREAMED, class T>
__global__ __launch_bounds__(256) void k_moments_dev(const T* __restrict__ fin, double* __restrict__ rho_out,
                                                     double* __restrict__ vel_out, const Geo g,
                                                     const View vi, const View vr, const View vv,
                                                     const double F0, const double F1, const double F2)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= g.n[0] || y >= g.n[1]) return;
    float d[L::Q];
    if constexpr (STREAMED) {
        pull_dev<L, true>(d, fin, x, y, z, g, vi);
    } else {
        const long long self = x + y * vi.sy + z * vi.sz;
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            d[q] = to_f(fin[q * vi.sq + self]);
        });
    }
    double f[L::Q];
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = L::w(q) + (double)d[q];
    });
    const double F[3] = {F0, F1, F2};
    double rho, u0, u1, u2;
    moments<L, FORCED>(f, F, rho, u0, u1, u2);
    rho_out[x + y * vr.sy + z * vr.sz] = rho;
    const long long sv = x + y * vv.sy + z * vv.sz;
    vel_out[sv] = u0;
    vel_out[sv + vv.sq] = u1;
    if constexpr (L::D == 3) vel_out[sv + 2 * vv.sq] = u2;
}

// Equilibrium initialisation into reduced storage: feq in FP64 with the
// reference's arithmetic, stored as T(feq - w); with a force the reported
// rho/u are recomputed from the stored values.
template <class L, bool FORCED, class T>
__global__ __launch_bounds__(256) void k_init_dev(double* __restrict__ rho_io, double* __restrict__ vel_io,
                                                  T* __restrict__ fout, const Geo g, const View vr,
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
        const double fe = feq<L, q>(rho, u0, u1, u2, usq);
        const T sv = from_f<T>((float)(fe - L::w(q)));
        fout[q * vo.sq + self] = sv;
        f[q] = L::w(q) + (double)to_f(sv);
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

// Conversion between FP64 populations and reduced storage (valid cells).
template <class L, class T, bool TO_DEV>
__global__ __launch_bounds__(256) void k_convert(const void* __restrict__ src, void* __restrict__ dst,
                                                 const Geo g, const View vs, const View vd)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= g.n[0] || y >= g.n[1]) return;
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const long long is = x + y * vs.sy + z * vs.sz + q * vs.sq;
        const long long id = x + y * vd.sy + z * vd.sz + q * vd.sq;
        if constexpr (TO_DEV) ((T*)dst)[id] = from_f<T>((float)(((const double*)src)[is] - L::w(q)));
        else ((double*)dst)[id] = L::w(q) + (double)to_f(((const T*)src)[is]);
    });
}

// FP64 mass / momentum / health counts of reduced storage (per-block
// partials, see k_reduce_final in mfa_lbm.hip for the deterministic final
// pass).  The stored value is the deviation d_q = f_q - w_q.
template <class L, int WHAT, class T>
__global__ __launch_bounds__(256) void k_reduce_dev(const T* __restrict__ f, const Geo g, const View v,
                                                    double* __restrict__ part)
{
    double acc[3] = {0.0, 0.0, 0.0};
    const long long ncell = (long long)g.n[0] * g.n[1] * g.n[2];
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(c % g.n[0]);
        const long long r = c / g.n[0];
        const int y = (int)(r % g.n[1]);
        const int z = (int)(r / g.n[1]);
        const long long self = x + y * v.sy + z * v.sz;
        if constexpr (WHAT == MFA_RED_HEALTH) {
            double rho = 0.0;
            bool finite = true;
            for_q<L>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const double d = (double)to_f(f[q * v.sq + self]);
                finite = finite && __builtin_isfinite(d);
                rho += L::w(q) + d;
            });
            acc[0] += finite ? 0.0 : 1.0;
            acc[1] += (finite && rho <= 0.0) ? 1.0 : 0.0;
            continue;
        }
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const double d = (double)to_f(f[q * v.sq + self]);
            if constexpr (WHAT == MFA_RED_MASS) acc[0] += L::w(q) + d;
            else {
                if constexpr (L::c(q, 0) != 0) acc[0] += L::c(q, 0) * d;
                if constexpr (L::c(q, 1) != 0) acc[1] += L::c(q, 1) * d;
                if constexpr (L::c(q, 2) != 0) acc[2] += L::c(q, 2) * d;
            }
        });
    }
    // block partial sums (fixed order)
    __shared__ double sh[4][3];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double s = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (lane == 0) sh[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0.0;
        for (int w = 0; w < 4; ++w) s += sh[w][threadIdx.x];
        part[blockIdx.x * 27 + threadIdx.x] = s;
    }
}

// ------------------------------------------------------------ dispatch ----
template <class Fn>
static int with_lattice_dev(int lat, Fn&& fn)
{
    switch (lat) {
    case MFA_D2Q9: return fn(Lattice<MFA_D2Q9>{});
    case MFA_D3Q15: return fn(Lattice<MFA_D3Q15>{});
    case MFA_D3Q19: return fn(Lattice<MFA_D3Q19>{});
    case MFA_D3Q27: return fn(Lattice<MFA_D3Q27>{});
    default: return fail(MFA_ERR_INVALID, "unknown lattice id");
    }
}

template <class Fn>
static int with_store_type(int dtype, Fn&& fn)
{
    switch (dtype) {
    case MFA_F32: return fn(float{});
    case MFA_F16DEV: return fn(f16_t{});
    case MFA_BF16DEV: return fn(bf16_t{});
    default: return fail(MFA_ERR_UNSUPPORTED, "reduced-precision path: unknown storage type");
    }
}

static void shape(const int n[3], dim3& grid, dim3& block)
{
    const int bx = n[0] <= 64 ? 64 : (n[0] <= 128 ? 128 : 256);
    block = dim3(bx, 256 / bx, 1);
    grid = dim3((n[0] + bx - 1) / bx, (n[1] + block.y - 1) / block.y, n[2]);
}

// Launch configuration of the reduced-storage step, per lattice and storage
// type (index dtype - MFA_F32).  cells_per_thread 2 selects k_step_dev_x2
// (block_x then counts threads, i.e. cell pairs); nt_loads applies to it.
// Defaults: the autotuner's picks on MI355X in the row-interleaved order.
// D3Q27 384^3 (also used for D2Q9): FP32 pairs in 64x4 blocks, 5.60 TB/s;
// FP16 / BF16 one cell per thread in 128x1 blocks, 5.70 TB/s.  D3Q19 256^3
// (also used for D3Q15): FP16 / BF16 one cell per thread in 256x1 blocks,
// 0.216 ms against 0.239 ms with 128x1.
static std::mutex g_dev_mu;
static const mfa_launch k_dev_default[4][3] = {
    {{64, 4, 2, 0}, {128, 1, 1, 0}, {128, 1, 1, 0}},
    {{64, 4, 2, 0}, {256, 1, 1, 0}, {256, 1, 1, 0}},
    {{64, 4, 2, 0}, {256, 1, 1, 0}, {256, 1, 1, 0}},
    {{64, 4, 2, 0}, {128, 1, 1, 0}, {128, 1, 1, 0}}};
static mfa_launch g_dev_cfg[4][3] = {
    {k_dev_default[0][0], k_dev_default[0][1], k_dev_default[0][2]},
    {k_dev_default[1][0], k_dev_default[1][1], k_dev_default[1][2]},
    {k_dev_default[2][0], k_dev_default[2][1], k_dev_default[2][2]},
    {k_dev_default[3][0], k_dev_default[3][1], k_dev_default[3][2]}};

static mfa_launch dev_cfg(int lat, int dtype)
{
    std::lock_guard<std::mutex> lk(g_dev_mu);
    return g_dev_cfg[lat][dtype - MFA_F32];
}

// Whether two fields can be walked in aligned x pairs.
template <class T>
static bool pairs_ok(const mfa_field* f)
{
    return f->n[0] % 2 == 0 && (uintptr_t)f->data % (2 * sizeof(T)) == 0 && f->sy % 2 == 0 &&
           f->sz % 2 == 0 && f->sq % 2 == 0;
}

template <class L, bool FORCED, class T>
static void launch_x2(const mfa_launch& c, const Geo& g, hipStream_t st, const T* a, T* b,
                      const View& vi, const View& vo, float om, float F0, float F1, float F2)
{
    const dim3 block(c.block_x, c.block_y, 1);
    const int pairs = g.n[0] / 2;
    const dim3 grid((pairs + c.block_x - 1) / c.block_x, (g.n[1] + c.block_y - 1) / c.block_y, g.n[2]);
    if (g.walls) {
        if (c.nt_loads) hipLaunchKernelGGL((k_step_dev_x2<L, FORCED, true, true, T>), grid, block, 0, st, a, b, g, vi, vo, om, F0, F1, F2);
        else hipLaunchKernelGGL((k_step_dev_x2<L, FORCED, true, false, T>), grid, block, 0, st, a, b, g, vi, vo, om, F0, F1, F2);
    } else {
        if (c.nt_loads) hipLaunchKernelGGL((k_step_dev_x2<L, FORCED, false, true, T>), grid, block, 0, st, a, b, g, vi, vo, om, F0, F1, F2);
        else hipLaunchKernelGGL((k_step_dev_x2<L, FORCED, false, false, T>), grid, block, 0, st, a, b, g, vi, vo, om, F0, F1, F2);
    }
}

int step_dev(const mfa_field* fin, mfa_field* fout, const Geo& g, int lat, double omega,
             const double* force, hipStream_t st)
{
    const View vi{fin->sy, fin->sz, fin->sq}, vo{fout->sy, fout->sz, fout->sq};
    const mfa_launch c = dev_cfg(lat, fin->dtype);
    return with_lattice_dev(lat, [&](auto Lc) {
        using L = decltype(Lc);
        return with_store_type(fin->dtype, [&](auto tv) {
            using T = decltype(tv);
            dim3 grid, block;
            shape(g.n, grid, block);
            const float om = (float)omega;
            const float F0 = force ? (float)force[0] : 0.f, F1 = force ? (float)force[1] : 0.f,
                        F2 = (force && L::D == 3) ? (float)force[2] : 0.f;
            const T* a = (const T*)fin->data;
            T* b = (T*)fout->data;
            if (c.cells_per_thread == 2 && pairs_ok<T>(fin) && pairs_ok<T>(fout)) {
                if (force) launch_x2<L, true, T>(c, g, st, a, b, vi, vo, om, F0, F1, F2);
                else launch_x2<L, false, T>(c, g, st, a, b, vi, vo, om, F0, F1, F2);
                MFA_HIP_CHECK(hipGetLastError());
                return (int)MFA_OK;
            }
            if (c.cells_per_thread == 1) {
                block = dim3(c.block_x, c.block_y, 1);
                grid = dim3((g.n[0] + c.block_x - 1) / c.block_x, (g.n[1] + c.block_y - 1) / c.block_y, g.n[2]);
            }
            if (force) {
                if (g.walls) hipLaunchKernelGGL((k_step_dev<L, true, true, T>), grid, block, 0, st, a, b, g, vi, vo, om, F0, F1, F2);
                else hipLaunchKernelGGL((k_step_dev<L, true, false, T>), grid, block, 0, st, a, b, g, vi, vo, om, F0, F1, F2);
            } else {
                if (g.walls) hipLaunchKernelGGL((k_step_dev<L, false, true, T>), grid, block, 0, st, a, b, g, vi, vo, om, F0, F1, F2);
                else hipLaunchKernelGGL((k_step_dev<L, false, false, T>), grid, block, 0, st, a, b, g, vi, vo, om, F0, F1, F2);
            }
            MFA_HIP_CHECK(hipGetLastError());
            return (int)MFA_OK;
        });
    });
}

// Two steps per pass on a 2-D box of reduced storage, periodic or with
// walls: k_step2_2d (mfa_lbm.hip) for stored deviations.  The intermediate
// state is rounded to the storage type before it enters LDS, and the second
// step's wall values read the cell's own intermediate deviations from LDS
// through the arithmetic of wall_fix_dev, so the pair equals two k_step_dev
// launches bit for bit.
template <class L, bool FORCED, bool WALLS, class T, class TL>
__global__ __launch_bounds__(TL::THREADS) void k_step2_2d_dev(const T* __restrict__ fin, T* __restrict__ fout,
                                                              const Geo g, const View vi, const View vo,
                                                              const float omega, const float F0,
                                                              const float F1)
{
    using U = typename Word<T>::U;
    __shared__ U lds[L::Q][TL::EC];
    const int tid = threadIdx.x;
    const int nx = g.n[0], ny = g.n[1];
    const unsigned nb = gridDim.x * gridDim.y;
    const unsigned b = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned per = nb / kXcds, rem = nb % kXcds, k = b % kXcds;
    const unsigned t = k * per + (k < rem ? k : rem) + b / kXcds;
    const int X0 = (int)(t % gridDim.x) * TL::TX, Y0 = (int)(t / gridDim.x) * TL::TY;
    if (tid < TL::EC) {
        int gx = X0 - 1 + tid % TL::EX, gy = Y0 - 1 + tid / TL::EX;
        bool in = true;
        if (g.per[0]) gx = (gx % nx + nx) % nx;
        else in = in && gx >= 0 && gx < nx;
        if (g.per[1]) gy = (gy % ny + ny) % ny;
        else in = in && gy >= 0 && gy < ny;
        if (in) {
            float d[L::Q], o[L::Q];
            pull_dev<L, WALLS, T, Geo, false>(d, fin, gx, gy, 0, g, vi);
            collide_dev<L, FORCED>(d, o, omega, F0, F1, 0.f);
            for_q<L>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                lds[q][tid] = f_to_word<T>(o[q]);
            });
        }
    }
    __syncthreads();
    if (tid >= TL::TX * TL::TY) return;
    const int ox = tid % TL::TX, oy = tid / TL::TX;
    const int x = X0 + ox, y = Y0 + oy;
    if (x >= nx || y >= ny) return;
    const int c0 = (ox + 1) + TL::EX * (oy + 1);
    float d[L::Q];
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        d[q] = word_to_f<T>(lds[q][c0 - L::c(q, 0) - TL::EX * L::c(q, 1)]);
    });
    if constexpr (WALLS) wall_fix_dev_at<L>(d, [&](int m) { return word_to_f<T>(lds[m][c0]); }, x, y, 0, g);
    collide_store_dev<L, FORCED>(d, fout, x + (long long)y * vo.sy, vo, omega, F0, F1, 0.f);
}

// Two steps per pass on FP32 storage of a periodic 3-D box in the push
// layout of k_step2p (kernels_step2.hpp): a block owns a 64 x TY column of
// output cells; phase A pulls and collides the column grown by one cell and
// pushes each population, rounded to the storage type exactly as a single
// step stores it, into the LDS slot of the tile cell that pulls it in the
// second step (per-c_z rings of 32-bit words); phase B reads its own cell's Q
// words, collides and writes the second step.  Loads address wave-uniform
// plane bases plus fixed 32-bit lane offsets; the march is pair_schedule's,
// with the same branch-free phases as k_step2p.  Per cell the arithmetic is
// k_step_dev's (collide_dev), so a pass equals two single steps bit for bit.
template <class L, bool FORCED, class T, class TL>
__global__ __launch_bounds__(TL::THREADS) __attribute__((amdgpu_waves_per_eu(TL::WAVES_PER_EU)))
void k_step2p_dev(const T* __restrict__ fin, T* __restrict__ fout, const Geo g, const View vi, const View vo,
                  const float omega, const float F0, const float F1, const float F2, const int TZ)
{
    static_assert(sizeof(T) == 4, "FP32 storage only");
    using U = typename Word<T>::U;
    using Rings = PushRings<L, TL::SC, false, U>;
    using Pos = typename Rings::Pos;
    __shared__ Rings lds;
    __shared__ U dummy[TL::SX];
    const int tid = threadIdx.x;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    // one contiguous run of tiles (x fastest) per XCD, as in k_step2p
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned per = nb / kXcds, rem = nb % kXcds, xcd = b % kXcds;
    const unsigned t = xcd * per + (xcd < rem ? xcd : rem) + b / kXcds;
    const int X0 = (int)(t % gridDim.x) * TL::TX, Y0 = (int)((t / gridDim.x) % gridDim.y) * TL::TY;
    const int Z0 = (int)(t / (gridDim.x * gridDim.y)) * TZ;
    const int Z1 = Z0 + TZ < nz ? Z0 + TZ : nz;
    // phase A: one cell of the grown column; its in-plane pull sources as
    // byte offsets (x' + y' sy) * sizeof(T), fixed for the march (the host
    // checks that a plane's offsets fit in 32 bits)
    int ex, ey;
    grown_cell<TL>(tid, ex, ey);
    const bool inA = tid < TL::EC;
    const int gx = ((X0 - 1 + ex) % nx + nx) % nx, gy = ((Y0 - 1 + ey) % ny + ny) % ny;
    unsigned srcA[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int sx = (gx - (a - 1) + nx) % nx, sy = (gy - (c - 1) + ny) % ny;
            srcA[a][c] = (unsigned)(sx + (long long)sy * vi.sy) * (unsigned)sizeof(T);
        }
    // phase B: one output cell
    const int ox = tid % TL::TX, oy = tid / TL::TX;
    const bool inB = tid < TL::TC && X0 + ox < nx && Y0 + oy < ny;
    const int own = ox + 1 + TL::SX * oy;
    const unsigned dstB =
        inB ? (unsigned)((X0 + ox) + (long long)(Y0 + oy) * vo.sy) * (unsigned)sizeof(T) : kDropOffset;
    // every lane loads and collides (a lane beyond the grown column takes
    // the corner cell and pushes to the dummy row), as in k_step2p
    auto loadA = [&](float (&P)[L::Q], int p) {
        const int pw = (p % nz + nz) % nz;
        const T* plane[3];
#pragma unroll
        for (int ci = 0; ci < 3; ++ci) plane[ci] = fin + (long long)(((pw - (ci - 1)) % nz + nz) % nz) * vi.sz;
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = L::c(q, 0), cy = L::c(q, 1), cz = L::c(q, 2);
            P[q] = to_f(*at_bytes(plane[cz + 1] + q * vi.sq, srcA[cx + 1][cy + 1]));
        });
    };
    auto storeA = [&](const float (&P)[L::Q], int p) {
        const Pos pw = Rings::plane(p, Z0);
        float o[L::Q];
        collide_dev<L, FORCED>(P, o, omega, F0, F1, F2);
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int dx = ex - 1 + L::c(q, 0), dy = ey - 1 + L::c(q, 1);
            U* dst = (inA && dx >= 0 && dx < TL::TX && dy >= 0 && dy < TL::TY)
                         ? &lds.template at<q>(pw, dx + 1 + TL::SX * dy)
                         : &dummy[ex];
            *dst = f_to_word<T>(o[q]);
        });
    };
    auto phaseB = [&](int z) {
        const Pos pr = Rings::pulled_by(z, Z0);
        float d[L::Q];
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            d[q] = word_to_f<T>(lds.template at<q>(pr, own));
        });
        float o[L::Q];
        collide_dev<L, FORCED>(d, o, omega, F0, F1, F2);
        T* const orow = fout + (long long)z * vo.sz;
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            store_nt_or_drop(orow + q * vo.sq, dstB, from_f<T>(o[q]));
        });
    };
    float P[L::Q];
    pair_schedule<TL::TC>(tid, Z0, Z1, P, loadA, storeA, phaseB);
}

// Two steps in one pass; the caller has checked a fully periodic 3-D box and
// chosen TZ, the z-planes per block.  FP32 storage only: with FP16 and BF16
// words the pass moves half the bytes but runs the same arithmetic, which
// then bounds it, and it measured slower than two single steps on MI355X
// (D3Q27 384^3 forced: 1.27 against 1.09 ms per step at its best tile).
// k_step2p_dev on 64 x 8 columns (D3Q27: 114 KiB of LDS; D3Q19: 80 KiB):
// D3Q27 384^3 forced 1.33 ms per step and D3Q19 256^3 0.268 ms, against
// 1.56 and 0.305 ms for the grown-column kernel it replaced (D3Q27 32 x 8
// with a ring of four whole planes, D3Q19 64 x 4 with per-c_z rings).
int step2_dev(const mfa_field* fin, mfa_field* fout, const Geo& g, int lat, double omega,
              const double* force, int TZ, hipStream_t st)
{
    if (fin->dtype != MFA_F32)
        return fail(MFA_ERR_UNSUPPORTED, "mfa_step2: reduced storage in FP32 only (use mfa_step)");
    const View vi{fin->sy, fin->sz, fin->sq}, vo{fout->sy, fout->sz, fout->sq};
    if (lat == MFA_D2Q9) {
        using L = Lattice<MFA_D2Q9>;
        using T = float;
        using TL = Step2DevTile<64, 4>;
        const dim3 grid((g.n[0] + TL::TX - 1) / TL::TX, (g.n[1] + TL::TY - 1) / TL::TY);
        const float om = (float)omega;
        const float F0 = force ? (float)force[0] : 0.f, F1 = force ? (float)force[1] : 0.f;
        const T* a = (const T*)fin->data;
        T* b = (T*)fout->data;
        if (force) {
            if (g.walls) hipLaunchKernelGGL((k_step2_2d_dev<L, true, true, T, TL>), grid, dim3(TL::THREADS), 0, st, a, b, g, vi, vo, om, F0, F1);
            else hipLaunchKernelGGL((k_step2_2d_dev<L, true, false, T, TL>), grid, dim3(TL::THREADS), 0, st, a, b, g, vi, vo, om, F0, F1);
        } else {
            if (g.walls) hipLaunchKernelGGL((k_step2_2d_dev<L, false, true, T, TL>), grid, dim3(TL::THREADS), 0, st, a, b, g, vi, vo, om, F0, F1);
            else hipLaunchKernelGGL((k_step2_2d_dev<L, false, false, T, TL>), grid, dim3(TL::THREADS), 0, st, a, b, g, vi, vo, om, F0, F1);
        }
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    }
    // in-plane byte offsets are 32-bit in the kernel
    for (const mfa_field* f : {fin, (const mfa_field*)fout})
        if (((long long)(f->n[0] - 1) + (long long)(f->n[1] - 1) * f->sy + 1) * 4 > (long long)kDropOffset)
            return fail(MFA_ERR_UNSUPPORTED, "mfa_step2: a z-plane spans more than 4 GiB (use mfa_step)");
    return with_lattice_dev(lat, [&](auto Lc) {
        using L = decltype(Lc);
        if constexpr (L::D != 3) {
            return fail(MFA_ERR_UNSUPPORTED, "mfa_step2: 3-D lattices only");
        } else {
            using T = float;
            using TL = PushTile<8>;
            const dim3 grid((g.n[0] + TL::TX - 1) / TL::TX, (g.n[1] + TL::TY - 1) / TL::TY,
                            (g.n[2] + TZ - 1) / TZ);
            const float om = (float)omega;
            const float F0 = force ? (float)force[0] : 0.f, F1 = force ? (float)force[1] : 0.f,
                        F2 = force ? (float)force[2] : 0.f;
            const T* a = (const T*)fin->data;
            T* b = (T*)fout->data;
            if (force)
                hipLaunchKernelGGL((k_step2p_dev<L, true, T, TL>), grid, dim3(TL::THREADS), 0, st, a, b, g,
                                   vi, vo, om, F0, F1, F2, TZ);
            else
                hipLaunchKernelGGL((k_step2p_dev<L, false, T, TL>), grid, dim3(TL::THREADS), 0, st, a, b, g,
                                   vi, vo, om, F0, F1, F2, TZ);
            MFA_HIP_CHECK(hipGetLastError());
            return (int)MFA_OK;
        }
    });
}

int moments_dev(const mfa_field* f, mfa_field* rho, mfa_field* vel, const Geo& g, int lat,
                const double* force, int streamed, hipStream_t st)
{
    const View vi{f->sy, f->sz, f->sq}, vr{rho->sy, rho->sz, rho->sq}, vv{vel->sy, vel->sz, vel->sq};
    return with_lattice_dev(lat, [&](auto Lc) {
        using L = decltype(Lc);
        return with_store_type(f->dtype, [&](auto tv) {
            using T = decltype(tv);
            dim3 grid, block;
            shape(g.n, grid, block);
            const double F0 = force ? force[0] : 0., F1 = force ? force[1] : 0.,
                         F2 = (force && L::D == 3) ? force[2] : 0.;
            const T* a = (const T*)f->data;
            double* r = (double*)rho->data;
            double* u = (double*)vel->data;
            if (force) {
                if (streamed) hipLaunchKernelGGL((k_moments_dev<L, true, true, T>), grid, block, 0, st, a, r, u, g, vi, vr, vv, F0, F1, F2);
                else hipLaunchKernelGGL((k_moments_dev<L, true, false, T>), grid, block, 0, st, a, r, u, g, vi, vr, vv, F0, F1, F2);
            } else {
                if (streamed) hipLaunchKernelGGL((k_moments_dev<L, false, true, T>), grid, block, 0, st, a, r, u, g, vi, vr, vv, F0, F1, F2);
                else hipLaunchKernelGGL((k_moments_dev<L, false, false, T>), grid, block, 0, st, a, r, u, g, vi, vr, vv, F0, F1, F2);
            }
            MFA_HIP_CHECK(hipGetLastError());
            return (int)MFA_OK;
        });
    });
}

int init_dev(mfa_field* rho, mfa_field* vel, mfa_field* f, const Geo& g, int lat, const double* force,
             hipStream_t st)
{
    const View vr{rho->sy, rho->sz, rho->sq}, vv{vel->sy, vel->sz, vel->sq}, vo{f->sy, f->sz, f->sq};
    return with_lattice_dev(lat, [&](auto Lc) {
        using L = decltype(Lc);
        return with_store_type(f->dtype, [&](auto tv) {
            using T = decltype(tv);
            dim3 grid, block;
            shape(g.n, grid, block);
            const double F0 = force ? force[0] : 0., F1 = force ? force[1] : 0.,
                         F2 = (force && L::D == 3) ? force[2] : 0.;
            if (force) hipLaunchKernelGGL((k_init_dev<L, true, T>), grid, block, 0, st, (double*)rho->data, (double*)vel->data, (T*)f->data, g, vr, vv, vo, F0, F1, F2);
            else hipLaunchKernelGGL((k_init_dev<L, false, T>), grid, block, 0, st, (double*)rho->data, (double*)vel->data, (T*)f->data, g, vr, vv, vo, F0, F1, F2);
            MFA_HIP_CHECK(hipGetLastError());
            return (int)MFA_OK;
        });
    });
}

int convert_dev(const mfa_field* src, mfa_field* dst, const Geo& g, int lat, hipStream_t st)
{
    const View vs{src->sy, src->sz, src->sq}, vd{dst->sy, dst->sz, dst->sq};
    const bool to_dev = src->dtype == MFA_F64;
    const int dt = to_dev ? dst->dtype : src->dtype;
    return with_lattice_dev(lat, [&](auto Lc) {
        using L = decltype(Lc);
        return with_store_type(dt, [&](auto tv) {
            using T = decltype(tv);
            dim3 grid, block;
            shape(g.n, grid, block);
            if (to_dev) hipLaunchKernelGGL((k_convert<L, T, true>), grid, block, 0, st, src->data, dst->data, g, vs, vd);
            else hipLaunchKernelGGL((k_convert<L, T, false>), grid, block, 0, st, src->data, dst->data, g, vs, vd);
            MFA_HIP_CHECK(hipGetLastError());
            return (int)MFA_OK;
        });
    });
}

int reduce_dev(const mfa_field* f, const Geo& g, int lat, int what, double* part, int nblocks,
               hipStream_t st)
{
    const View v{f->sy, f->sz, f->sq};
    return with_lattice_dev(lat, [&](auto Lc) {
        using L = decltype(Lc);
        return with_store_type(f->dtype, [&](auto tv) {
            using T = decltype(tv);
            if (what == MFA_RED_MASS)
                hipLaunchKernelGGL((k_reduce_dev<L, MFA_RED_MASS, T>), dim3(nblocks), dim3(256), 0, st, (const T*)f->data, g, v, part);
            else if (what == MFA_RED_MOMENTUM)
                hipLaunchKernelGGL((k_reduce_dev<L, MFA_RED_MOMENTUM, T>), dim3(nblocks), dim3(256), 0, st, (const T*)f->data, g, v, part);
            else if (what == MFA_RED_HEALTH)
                hipLaunchKernelGGL((k_reduce_dev<L, MFA_RED_HEALTH, T>), dim3(nblocks), dim3(256), 0, st, (const T*)f->data, g, v, part);
            else
                return fail(MFA_ERR_UNSUPPORTED, "mfa_reduce: per-component sums need FP64 storage");
            MFA_HIP_CHECK(hipGetLastError());
            return (int)MFA_OK;
        });
    });
}

}  // namespace mfa

using namespace mfa;

int mfa_set_launch_dev(int lattice, int dtype, const mfa_launch* cfg)
{
    if (lattice < MFA_D2Q9 || lattice > MFA_D3Q27 || dtype < MFA_F32 || dtype > MFA_BF16DEV)
        return fail(MFA_ERR_INVALID, "mfa_set_launch_dev: unknown lattice or storage type");
    const mfa_launch c = cfg ? *cfg : k_dev_default[lattice][dtype - MFA_F32];
    const bool bx_ok = c.block_x == 64 || c.block_x == 128 || c.block_x == 256;
    if (!bx_ok || c.block_y < 1 || c.block_x * c.block_y > 256 ||
        (c.cells_per_thread != 1 && c.cells_per_thread != 2) || (c.nt_loads != 0 && c.nt_loads != 1))
        return fail(MFA_ERR_INVALID, "mfa_set_launch_dev: unsupported launch configuration");
    std::lock_guard<std::mutex> lk(g_dev_mu);
    g_dev_cfg[lattice][dtype - MFA_F32] = c;
    return MFA_OK;
}

int mfa_get_launch_dev(int lattice, int dtype, mfa_launch* cfg)
{
    if (lattice < MFA_D2Q9 || lattice > MFA_D3Q27 || dtype < MFA_F32 || dtype > MFA_BF16DEV || !cfg)
        return fail(MFA_ERR_INVALID, "mfa_get_launch_dev: bad argument");
    *cfg = dev_cfg(lattice, dtype);
    return MFA_OK;
}
