// kernels_common.hpp -- device-side building blocks shared by the FP64 and
// reduced-precision kernels of the step engine: kernel argument structs,
// pull-source addressing (periodic wrap / clamp), wall ghost values and the
// compile-time loop over populations.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "lattice.hpp"

namespace mfa {

int fail(int code, const std::string& msg);

// Accelerator complex dies of an MI355X.  The hardware deals consecutive
// workgroups of a launch round-robin to them; each has its own L2.
constexpr unsigned kXcds = 8;

#define MFA_HIP_CHECK(expr)                                                         \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            return fail(MFA_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------- kernel arguments --
struct Geo {
    int n[3];
    int per[3];
    int walls;            // some dimension is non-periodic
    signed char rtype[27];
    double rrho[27];
    double ru[27][3];
};

// Geometry seen through pointers to boundary tables held in device memory
// (used by the batched engine, whose lattices share one boundary table).
struct GeoRef {
    int n[3];
    int per[3];
    int walls;
    const signed char* rtype;
    const double* rrho;
    const double (*ru)[3];
};

struct View {             // one field, addressed relative to cell (0,0,0)
    long long sy, sz, sq;
};

static View view_of(const mfa_field& f) { return View{f.sy, f.sz, f.sq}; }

// ------------------------------------------- two-step LDS plane storage --
// Populations of a lattice by their c_z class k = c_z + 1, and the rank of q
// within its class (in population order).
template <class L>
constexpr int cz_count(int k)
{
    int n = 0;
    for (int q = 0; q < L::Q; ++q) n += (L::c(q, 2) + 1 == k);
    return n;
}
template <class L>
constexpr int cz_rank(int q)
{
    int r = 0;
    for (int p = 0; p < q; ++p) r += (L::c(p, 2) == L::c(q, 2));
    return r;
}

// ------------------------------------------------------------- pull stage --
// Coordinates of the pull source for each lattice velocity component value
// (index c+1): source = x - c.  Fused mode wraps periodic dimensions and
// clamps non-periodic ones (the clamped load is replaced by the wall value).
struct Src {
    long long o[3][3];    // per dimension, per (c+1): element offset
};

template <int D, bool HALO, class G>
__device__ __forceinline__ Src make_src(int x, int y, int z, const G& g, const View& v)
{
    Src s;
    const int p[3] = {x, y, z};
    const long long st[3] = {1, v.sy, v.sz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int ci = 0; ci < 3; ++ci) {
            const int c = ci - 1;
            int q = p[d] - c;
            if (d >= D) q = p[d];
            else if (!HALO) {
                if (q < 0) q = g.per[d] ? g.n[d] - 1 : 0;
                else if (q >= g.n[d]) q = g.per[d] ? 0 : g.n[d] - 1;
            }
            s.o[d][ci] = (long long)q * st[d];
        }
    }
    return s;
}

// Ghost value f(g, q) for g = x - c_q outside a non-periodic face, evaluated
// from the populations of the cell x itself as the reference's BCManager
// would have written it (src/LBM/BoundaryConditions.H:124-189 and the
// src/LBM/BCTypes/ headers).  self_pop(m) returns population m of cell x:
// from HBM in the one-step kernels, from LDS in the two-step kernels.
template <class L, int q, class G, class Ld>
__device__ __forceinline__ double wall_value_at(Ld&& self_pop, const G& g, int idx)
{
    const int t = g.rtype[idx];
    if (t == MFA_BC_MOVING_WALL || t == MFA_BC_NO_SLIP || t == MFA_BC_INFLOW) {
        constexpr int m = L::opp(q);
        return self_pop(m) + bb_contrib<L, m>(g.rrho[idx], g.ru[idx][0], g.ru[idx][1], g.ru[idx][2]);
    }
    if (t == MFA_BC_OUTFLOW) return self_pop(q);
    if (t == MFA_BC_PRESSURE_OUTLET) {
        // src/LBM/BCTypes/PressureOutletKernel.H:29-64
        double rho = 0.0, j0 = 0.0, j1 = 0.0, j2 = 0.0;
#pragma unroll
        for (int m = 0; m < L::Q; ++m) {
            const double fm = self_pop(m);
            rho += fm;
            j0 += L::c(m, 0) * fm;
            j1 += L::c(m, 1) * fm;
            if constexpr (L::D == 3) j2 += L::c(m, 2) * fm;
        }
        const double inv_rho = 1.0 / rho;
        const double cu = cdot<L, q>(j0, j1, j2) * inv_rho;
        double uu;
        if constexpr (L::D == 2) uu = (j0 * j0 + j1 * j1) * inv_rho * inv_rho;
        else uu = (j0 * j0 + j1 * j1 + j2 * j2) * inv_rho * inv_rho;
        const double feq_even =
            L::w(q) * g.rrho[idx] * (1.0 + 0.5 * cu * cu * 3.0 * 3.0 - 0.5 * uu * 3.0);
        constexpr int m = L::opp(q);
        return -self_pop(m) + 2.0 * feq_even;
    }
    return __builtin_nan("");
}

template <class L, int q, class G>
__device__ __forceinline__ double wall_value(const double* __restrict__ fin, long long self,
                                             const View& v, const G& g, int idx)
{
    return wall_value_at<L, q>([&](int m) { return fin[m * v.sq + self]; }, g, idx);
}

// Replace population q, pulled from behind a non-periodic face of the cell
// with edge flags lo/hi, by its wall value.
template <class L, int q, class G, class Ld>
__device__ __forceinline__ void fix_wall_at(double& fq, Ld&& self_pop, const G& g, const int lo[3],
                                            const int hi[3])
{
    int code[3] = {0, 0, 0};
    bool wall = false;
#pragma unroll
    for (int d = 0; d < L::D; ++d) {
        const int c = L::c(q, d);
        if (c == 1 && lo[d]) { code[d] = -1; wall = true; }
        if (c == -1 && hi[d]) { code[d] = 1; wall = true; }
    }
    if (wall) {
        const int idx = (code[0] + 1) + 3 * (code[1] + 1) + 9 * (code[2] + 1) * (L::D == 3);
        fq = wall_value_at<L, q>(self_pop, g, idx);
    }
}

template <class L, int q, class G>
__device__ __forceinline__ void fix_wall(double& fq, const double* __restrict__ fin, long long self,
                                         const View& v, const G& g, const int lo[3],
                                         const int hi[3])
{
    fix_wall_at<L, q>(fq, [&](int m) { return fin[m * v.sq + self]; }, g, lo, hi);
}

// Word of T at a 32-bit byte offset from a wave-uniform base: the load or
// store addresses the base held in scalar registers plus one 32-bit vector
// offset, with no 64-bit address arithmetic per population.
template <class W>
__device__ __forceinline__ const W* at_bytes(const W* base, unsigned off)
{
    return (const W*)((const char*)base + off);
}
template <class W>
__device__ __forceinline__ W* at_bytes(W* base, unsigned off)
{
    return (W*)((char*)base + off);
}

// Non-temporal store of one 8- or 4-byte word at a 32-bit byte offset from a
// wave-uniform base, through a buffer resource whose range ends at
// kDropOffset: a lane that has nothing to write passes kDropOffset and the
// hardware drops its store (out-of-range buffer writes are discarded).  A
// loop of such stores has no branch around them, so every path from a load
// to its first use passes the stores issued after the load, and the
// compiler's wait for the load need not also wait for them.  Valid offsets
// end at or below kDropOffset - sizeof(word).
constexpr unsigned kDropOffset = 0xFFFFFFF0u;
__device__ __forceinline__ void store_nt_or_drop(double* base, unsigned off, double v)
{
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    // dword 3 of the resource: 32-bit data format, the raw-buffer value for
    // gfx9-family targets
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)kDropOffset, 0x00020000);
    // cache policy 2: non-temporal, as __builtin_nontemporal_store
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)off, 0, 2);
}
__device__ __forceinline__ void store_nt_or_drop(float* base, unsigned off, float v)
{
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)kDropOffset, 0x00020000);
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 2);
}

// Element offsets of the pull-source rows of one cell whose y and z are the
// same across its wavefront: rows[cy+1][cz+1] = y' * sy + z' * sz with the
// source coordinates wrapped (periodic) or clamped (walls; the clamped value
// is replaced by the wall value).
template <int D, class G>
__device__ __forceinline__ void source_rows(long long (&rows)[3][3], int y, int z, const G& g,
                                            const View& v)
{
    long long oy[3], oz[3];
#pragma unroll
    for (int ci = 0; ci < 3; ++ci) {
        const int c = ci - 1;
        int sy = y - c, sz = z - c;
        if (sy < 0) sy = g.per[1] ? g.n[1] - 1 : 0;
        else if (sy >= g.n[1]) sy = g.per[1] ? 0 : g.n[1] - 1;
        if (D == 2) sz = z;
        else if (sz < 0) sz = g.per[2] ? g.n[2] - 1 : 0;
        else if (sz >= g.n[2]) sz = g.per[2] ? 0 : g.n[2] - 1;
        oy[ci] = (long long)sy * v.sy;
        oz[ci] = (long long)sz * v.sz;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) rows[a][b] = oy[a] + oz[b];
}

// x of the pull source for c = -1, 0, +1.
template <class G>
__device__ __forceinline__ void source_x(int (&xs)[3], int x, const G& g)
{
#pragma unroll
    for (int ci = 0; ci < 3; ++ci) {
        int sx = x - (ci - 1);
        if (sx < 0) sx = g.per[0] ? g.n[0] - 1 : 0;
        else if (sx >= g.n[0]) sx = g.per[0] ? 0 : g.n[0] - 1;
        xs[ci] = sx;
    }
}

template <class L, int q>
struct Unroll {
    template <class Fn>
    static __device__ __forceinline__ void run(Fn&& fn)
    {
        fn(std::integral_constant<int, q>{});
        if constexpr (q + 1 < L::Q) Unroll<L, q + 1>::run(fn);
    }
};

template <class L, class Fn>
__device__ __forceinline__ void for_q(Fn&& fn) { Unroll<L, 0>::run(fn); }

// Pull the Q populations of cell (x,y,z).  mode: 0 fused (inline wrap and
// walls), 1 halo (reads ghost cells), 2 unstreamed (the cell's own values).
// WALLS: some dimension is non-periodic (compile-time, so the periodic
// kernel carries no boundary code).  NT: 0 cached loads, 1 non-temporal
// loads, 2 non-temporal only where c_x = 0 (rows read x-shifted share cache
// lines with the neighbouring wavefront, which cached loads let it reuse).
template <bool NT>
__device__ __forceinline__ double ld(const double* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <class L, int MODE, bool WALLS = true, int NT = 0, class G = Geo>
__device__ __forceinline__ void pull(double (&f)[L::Q], const double* __restrict__ fin, int x, int y,
                                     int z, const G& g, const View& v)
{
    const long long self = x + y * v.sy + z * v.sz;
    if constexpr (MODE == 2) {
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            f[q] = ld<NT != 0>(&fin[q * v.sq + self]);
        });
        return;
    } else {
        const Src s = make_src<L::D, MODE == 1>(x, y, z, g, v);
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const long long off = s.o[0][L::c(q, 0) + 1] + s.o[1][L::c(q, 1) + 1] +
                                  s.o[2][L::c(q, 2) + 1];
            f[q] = ld<NT == 1 || (NT == 2 && L::c(q, 0) == 0)>(&fin[q * v.sq + off]);
        });
        if constexpr (MODE == 0 && WALLS) {
            const int p[3] = {x, y, z};
            int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            bool edge = false;
#pragma unroll
            for (int d = 0; d < L::D; ++d) {
                lo[d] = !g.per[d] && p[d] == 0;
                hi[d] = !g.per[d] && p[d] == g.n[d] - 1;
                edge = edge || lo[d] || hi[d];
            }
            if (edge) {
                for_q<L>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    fix_wall<L, q>(f[q], fin, self, v, g, lo, hi);
                });
            }
        }
    }
}

// rho and u of Q populations: src/LBM/Macroscopic.H:36-45,115-127,186-201.
template <class L, bool FORCED>
__device__ __forceinline__ void moments(const double (&f)[L::Q], const double F[3], double& rho,
                                        double& u0, double& u1, double& u2)
{
    rho = 0.;
    double j0 = 0., j1 = 0., j2 = 0.;
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        rho += f[q];
    });
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (L::c(q, 0) != 0) j0 = (L::c(q, 0) > 0) ? j0 + f[q] : j0 - f[q];
        if constexpr (L::c(q, 1) != 0) j1 = (L::c(q, 1) > 0) ? j1 + f[q] : j1 - f[q];
        if constexpr (L::D == 3 && L::c(q, 2) != 0) j2 = (L::c(q, 2) > 0) ? j2 + f[q] : j2 - f[q];
    });
    if constexpr (FORCED) {
        u0 = (j0 + 0.5 * F[0]) / rho;
        u1 = (j1 + 0.5 * F[1]) / rho;
        u2 = (L::D == 3) ? (j2 + 0.5 * F[2]) / rho : 0.0;
    } else {
        u0 = j0 / rho;
        u1 = j1 / rho;
        u2 = (L::D == 3) ? j2 / rho : 0.0;
    }
}


// BGK (+Guo) collision of pulled populations: src/LBM/Collision.cpp:24-53
// (plain) and :62-92 (forced).  `emit(q, value)` receives each collided
// population, so the step kernels store them to HBM and the two-step kernel
// to LDS through one copy of the arithmetic.
//
// Populations q and opp(q) are collided together.  Their lattice velocities
// are opposite, so c.u, c.F and every product of two of them change sign or
// not, exactly, in IEEE arithmetic with round-to-nearest: c_opp.u = -(c_q.u),
// 3*(-a) = -(3a), (4.5*(-a))*(-a) = (4.5a)*a, (-a)*(-b) = a*b, and
// x + (-y) is the operation x - y.  Each shared term is computed once, and
// every population gets the bits of the reference's term-by-term formula
// (Equilibrium.H:49-56, Collision.H:42-48).
template <class L, bool FORCED, class Emit>
__device__ __forceinline__ void collide(const double (&f)[L::Q], const double omega, const double F0,
                                        const double F1, const double F2, Emit&& emit)
{
    const double F[3] = {F0, F1, F2};
    double rho, u0, u1, u2;
    moments<L, FORCED>(f, F, rho, u0, u1, u2);
    const double usq = usq_of<L>(u0, u1, u2);
    const double usq15 = 1.5 * usq;
    const double omw = 1 - omega;
    double pref = 0.0, uF = 0.0;
    if constexpr (FORCED) {
        pref = 1.0 - 0.5 * omega;
        uF = (L::D == 2) ? u0 * F0 + u1 * F1 : u0 * F0 + u1 * F1 + u2 * F2;
    }
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int o = L::opp(q);
        if constexpr (q == o) {            // rest population: c = 0
            const double fe = feq<L, q>(rho, u0, u1, u2, usq);
            double out = omw * f[q] + omega * fe;
            if constexpr (FORCED) out = out + guo<L, q>(u0, u1, u2, F0, F1, F2, pref, uF);
            emit(qc, out);
        } else if constexpr (q < o) {      // q and its opposite
            const double cu = cdot<L, q>(u0, u1, u2);
            const double t3 = 3.0 * cu;
            const double t45 = 4.5 * cu * cu;
            const double wr = L::w(q) * rho;
            const double fq = wr * (1.0 + t3 + t45 - usq15);
            const double fo = wr * (1.0 - t3 + t45 - usq15);
            double oq = omw * f[q] + omega * fq;
            double oo = omw * f[o] + omega * fo;
            if constexpr (FORCED) {
                const double cF = cdot<L, q>(F0, F1, F2);
                const double p3 = pref * L::w(q) * 3.0;
                const double tcF = t3 * cF;
                oq = oq + p3 * (cF + tcF - uF);
                oo = oo + p3 * (tcF - cF - uF);
            }
            emit(qc, oq);
            emit(std::integral_constant<int, o>{}, oo);
        }
    });
}

// ... and store to fout with non-temporal stores.
template <class L, bool FORCED>
__device__ __forceinline__ void collide_store(const double (&f)[L::Q], double* __restrict__ fout,
                                              long long self, const View& vo, const double omega,
                                              const double F0, const double F1, const double F2)
{
    collide<L, FORCED>(f, omega, F0, F1, F2, [&](auto qc, double out) {
        constexpr int q = decltype(qc)::value;
        __builtin_nontemporal_store(out, &fout[q * vo.sq + self]);
    });
}

}  // namespace mfa
