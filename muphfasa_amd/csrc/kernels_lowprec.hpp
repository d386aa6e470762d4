// kernels_lowprec.hpp -- device building blocks of the reduced-precision
// storage path: populations stored as the shifted deviation d_q = f_q - w_q
// (FP32 / FP16 / BF16), pulled and collided in FP32.
#pragma once

#include "kernels_common.hpp"

namespace mfa {

typedef __bf16 bf16_t;
typedef _Float16 f16_t;

template <class T>
__device__ __forceinline__ float to_f(T v) { return (float)v; }
template <class T>
__device__ __forceinline__ T from_f(float v) { return (T)v; }

template <class T>
__device__ __forceinline__ T ldnt(const T* p)
{
    if constexpr (sizeof(T) == 4) return __builtin_nontemporal_load(p);
    else return *p;
}

// Wall ghost value in deviation form (see wall_value in kernels_common.hpp):
// bounce-back d(x,opp)+contrib (w_opp = w_q), outflow d(x,q), pressure
// outlet -d(x,opp) + 2 (feq_even - w_q).
template <class L, int q, class T, class G>
__device__ __forceinline__ float wall_value_dev(const T* __restrict__ fin, long long self, const View& v,
                                                const G& g, int idx)
{
    const int t = g.rtype[idx];
    constexpr int m = L::opp(q);
    if (t == MFA_BC_MOVING_WALL || t == MFA_BC_NO_SLIP || t == MFA_BC_INFLOW)
        return to_f(fin[m * v.sq + self]) +
               bb_contrib<L, m, float>((float)g.rrho[idx], (float)g.ru[idx][0], (float)g.ru[idx][1],
                                       (float)g.ru[idx][2]);
    if (t == MFA_BC_OUTFLOW) return to_f(fin[q * v.sq + self]);
    if (t == MFA_BC_PRESSURE_OUTLET) {
        float drho = 0.f, j0 = 0.f, j1 = 0.f, j2 = 0.f;
#pragma unroll
        for (int k = 0; k < L::Q; ++k) {
            const float d = to_f(fin[k * v.sq + self]);
            drho += d;
            j0 += L::c(k, 0) * d;
            j1 += L::c(k, 1) * d;
            if constexpr (L::D == 3) j2 += L::c(k, 2) * d;
        }
        const float inv_rho = 1.0f / (1.0f + drho);
        const float cu = cdot<L, q>(j0, j1, j2) * inv_rho;
        const float uu = (j0 * j0 + j1 * j1 + j2 * j2) * inv_rho * inv_rho;
        const float even_dev =
            (float)L::w(q) * ((float)g.rrho[idx] * (1.0f + 4.5f * cu * cu - 1.5f * uu) - 1.0f);
        return -to_f(fin[m * v.sq + self]) + 2.0f * even_dev;
    }
    return __builtin_nanf("");
}

// Replace the pulled values whose source lies behind a non-periodic face by
// the wall ghost value (cells on the domain edge only).
template <class L, class T, class G>
__device__ __forceinline__ void wall_fix_dev(float (&d)[L::Q], const T* __restrict__ fin, int x, int y,
                                             int z, const G& g, const View& v)
{
    const int p[3] = {x, y, z};
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool edge = false;
#pragma unroll
    for (int dd = 0; dd < L::D; ++dd) {
        lo[dd] = !g.per[dd] && p[dd] == 0;
        hi[dd] = !g.per[dd] && p[dd] == g.n[dd] - 1;
        edge = edge || lo[dd] || hi[dd];
    }
    if (edge) {
        const long long self = x + y * v.sy + z * v.sz;
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            int code[3] = {0, 0, 0};
            bool wall = false;
#pragma unroll
            for (int dd = 0; dd < L::D; ++dd) {
                const int c = L::c(q, dd);
                if (c == 1 && lo[dd]) { code[dd] = -1; wall = true; }
                if (c == -1 && hi[dd]) { code[dd] = 1; wall = true; }
            }
            if (wall) {
                const int idx = (code[0] + 1) + 3 * (code[1] + 1) + (L::D == 3 ? 9 * (code[2] + 1) : 0);
                d[q] = wall_value_dev<L, q>(fin, self, v, g, idx);
            }
        });
    }
}

template <class L, bool WALLS, class T, class G>
__device__ __forceinline__ void pull_dev(float (&d)[L::Q], const T* __restrict__ fin, int x, int y,
                                         int z, const G& g, const View& v)
{
    const Src s = make_src<L::D, false>(x, y, z, g, v);
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const long long off = s.o[0][L::c(q, 0) + 1] + s.o[1][L::c(q, 1) + 1] + s.o[2][L::c(q, 2) + 1];
        d[q] = to_f(ldnt(&fin[q * v.sq + off]));
    });
    if constexpr (WALLS) wall_fix_dev<L>(d, fin, x, y, z, g, v);
}

// FP32 collision in deviation form of one cell; out[q] is the post-collision
// deviation before rounding to the storage type.
template <class L, bool FORCED>
__device__ __forceinline__ void collide_dev(const float (&d)[L::Q], float (&o)[L::Q], const float omega,
                                            const float F0, const float F1, const float F2)
{
    float drho = 0.f, j0 = 0.f, j1 = 0.f, j2 = 0.f;
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        drho += d[q];
        if constexpr (L::c(q, 0) != 0) j0 += (L::c(q, 0) > 0) ? d[q] : -d[q];
        if constexpr (L::c(q, 1) != 0) j1 += (L::c(q, 1) > 0) ? d[q] : -d[q];
        if constexpr (L::D == 3 && L::c(q, 2) != 0) j2 += (L::c(q, 2) > 0) ? d[q] : -d[q];
    });
    const float rho = 1.0f + drho;
    const float inv = 1.0f / rho;
    const float u0 = (FORCED ? j0 + 0.5f * F0 : j0) * inv;
    const float u1 = (FORCED ? j1 + 0.5f * F1 : j1) * inv;
    const float u2 = (L::D == 3) ? (FORCED ? j2 + 0.5f * F2 : j2) * inv : 0.f;
    const float usq = usq_of<L>(u0, u1, u2);
    const float omw = 1.0f - omega;
    const float pref = 1.0f - 0.5f * omega;
    const float uF = u0 * F0 + u1 * F1 + u2 * F2;
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const float cu = cdot<L, q>(u0, u1, u2);
        const float fe = (float)L::w(q) * (drho + rho * (3.0f * cu + 4.5f * cu * cu - 1.5f * usq));
        float out = omw * d[q] + omega * fe;
        if constexpr (FORCED) out += guo<L, q, float>(u0, u1, u2, F0, F1, F2, pref, uF);
        o[q] = out;
    });
}

// Collision and rounded store (RNE) of one cell.
template <class L, bool FORCED, class T>
__device__ __forceinline__ void collide_store_dev(const float (&d)[L::Q], T* __restrict__ fout,
                                                  long long self, const View& vo, const float omega,
                                                  const float F0, const float F1, const float F2)
{
    float o[L::Q];
    collide_dev<L, FORCED>(d, o, omega, F0, F1, F2);
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const T sv = from_f<T>(o[q]);
        if constexpr (sizeof(T) == 4) __builtin_nontemporal_store(sv, &fout[q * vo.sq + self]);
        else fout[q * vo.sq + self] = sv;
    });
}

}  // namespace mfa
