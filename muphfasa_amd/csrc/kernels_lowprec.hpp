// kernels_lowprec.hpp -- device building blocks of the reduced-precision
// storage path: populations stored as the shifted deviation d_q = f_q - w_q
// (FP32 / FP16 / BF16), pulled and collided in FP32.
#pragma once

#include <type_traits>

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

// Wall ghost value in deviation form (see wall_value_at in
// kernels_common.hpp): bounce-back d(x,opp)+contrib (w_opp = w_q), outflow
// d(x,q), pressure outlet -d(x,opp) + 2 (feq_even - w_q).  self_pop(m)
// returns the deviation of population m of the cell itself, as float.
template <class L, int q, class G, class Ld>
__device__ __forceinline__ float wall_value_dev_at(Ld&& self_pop, const G& g, int idx)
{
    const int t = g.rtype[idx];
    constexpr int m = L::opp(q);
    if (t == MFA_BC_MOVING_WALL || t == MFA_BC_NO_SLIP || t == MFA_BC_INFLOW)
        return self_pop(m) +
               bb_contrib<L, m, float>((float)g.rrho[idx], (float)g.ru[idx][0], (float)g.ru[idx][1],
                                       (float)g.ru[idx][2]);
    if (t == MFA_BC_OUTFLOW) return self_pop(q);
    if (t == MFA_BC_PRESSURE_OUTLET) {
        float drho = 0.f, j0 = 0.f, j1 = 0.f, j2 = 0.f;
#pragma unroll
        for (int k = 0; k < L::Q; ++k) {
            const float d = self_pop(k);
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
        return -self_pop(m) + 2.0f * even_dev;
    }
    return __builtin_nanf("");
}

template <class L, int q, class T, class G>
__device__ __forceinline__ float wall_value_dev(const T* __restrict__ fin, long long self, const View& v,
                                                const G& g, int idx)
{
    return wall_value_dev_at<L, q>([&](int k) { return to_f(fin[k * v.sq + self]); }, g, idx);
}

// Replace the pulled values whose source lies behind a non-periodic face by
// the wall value (cells on the domain edge only); self_pop as above.
template <class L, class G, class Ld>
__device__ __forceinline__ void wall_fix_dev_at(float (&d)[L::Q], Ld&& self_pop, int x, int y, int z,
                                                const G& g)
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
                d[q] = wall_value_dev_at<L, q>(self_pop, g, idx);
            }
        });
    }
}

template <class L, class T, class G>
__device__ __forceinline__ void wall_fix_dev(float (&d)[L::Q], const T* __restrict__ fin, int x, int y,
                                             int z, const G& g, const View& v)
{
    const long long self = x + y * v.sy + z * v.sz;
    wall_fix_dev_at<L>(d, [&](int k) { return to_f(fin[k * v.sq + self]); }, x, y, z, g);
}

// NT: non-temporal loads for 4-byte words (2-byte words always load cached).
template <class L, bool WALLS, class T, class G, bool NT = true>
__device__ __forceinline__ void pull_dev(float (&d)[L::Q], const T* __restrict__ fin, int x, int y,
                                         int z, const G& g, const View& v)
{
    const Src s = make_src<L::D, false>(x, y, z, g, v);
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const long long off = s.o[0][L::c(q, 0) + 1] + s.o[1][L::c(q, 1) + 1] + s.o[2][L::c(q, 2) + 1];
        if constexpr (NT) d[q] = to_f(ldnt(&fin[q * v.sq + off]));
        else d[q] = to_f(fin[q * v.sq + off]);
    });
    if constexpr (WALLS) wall_fix_dev<L>(d, fin, x, y, z, g, v);
}

// Compile-time loop over 0 .. N-1.
template <int N, int I = 0, class Fn>
__device__ __forceinline__ void static_for(Fn&& fn)
{
    if constexpr (I < N) {
        fn(std::integral_constant<int, I>{});
        static_for<N, I + 1>(fn);
    }
}

// Index of the population with velocity (cx, cy, cz), or -1.
template <class L>
MFA_HD constexpr int q_of(int cx, int cy, int cz)
{
    for (int q = 0; q < L::Q; ++q)
        if (L::c(q, 0) == cx && L::c(q, 1) == cy && (L::D == 2 ? 0 : L::c(q, 2)) == cz) return q;
    return -1;
}

// Sum over c_z of the column (cx, cy); every column of the four lattices
// holds c_z = 0, or c_z = +1 and -1, or all three.
template <class L, int cx, int cy>
__device__ __forceinline__ float column_sum(const float (&d)[L::Q])
{
    constexpr int qp = q_of<L>(cx, cy, 1), q0 = q_of<L>(cx, cy, 0), qm = q_of<L>(cx, cy, -1);
    static_assert(q0 >= 0 || qp >= 0, "empty velocity column");
    static_assert((qp >= 0) == (qm >= 0), "asymmetric velocity column");
    if constexpr (qp >= 0 && q0 >= 0) return (d[qm] + d[q0]) + d[qp];
    else if constexpr (qp >= 0) return d[qm] + d[qp];
    else return d[q0];
}

// First column (index 3 (cx+1) + (cy+1)) holding c_z = +-1, or 9.
template <class L>
MFA_HD constexpr int first_z_column()
{
    for (int k = 0; k < 9; ++k)
        if (q_of<L>(k / 3 - 1, k % 3 - 1, 1) >= 0) return k;
    return 9;
}

// drho = sum_q d_q and j = sum_q c_q d_q, factored through the velocity
// columns: sums over c_z first, then over c_y, then over c_x, so each
// population enters about twice instead of D + 1 times (D3Q27: 49 additions
// instead of 77).
template <class L>
__device__ __forceinline__ void moments_dev(const float (&d)[L::Q], float& drho, float& j0, float& j1,
                                            float& j2)
{
    float C[3], Dy[3];
    j2 = 0.f;
    static_for<3>([&](auto ic) {
        constexpr int cx = decltype(ic)::value - 1;
        const float am = column_sum<L, cx, -1>(d), a0 = column_sum<L, cx, 0>(d),
                    ap = column_sum<L, cx, 1>(d);
        C[cx + 1] = (am + a0) + ap;
        Dy[cx + 1] = ap - am;
        static_for<3>([&](auto jc) {
            constexpr int cy = decltype(jc)::value - 1;
            constexpr int qp = q_of<L>(cx, cy, 1), qm = q_of<L>(cx, cy, -1);
            if constexpr (L::D == 3 && qp >= 0) {
                const float b = d[qp] - d[qm];
                if constexpr (3 * (cx + 1) + (cy + 1) == first_z_column<L>()) j2 = b;
                else j2 = j2 + b;
            }
        });
    });
    drho = (C[0] + C[1]) + C[2];
    j0 = C[2] - C[0];
    j1 = (Dy[0] + Dy[1]) + Dy[2];
}

// FP32 BGK (+Guo) collision in deviation form of one cell; o[q] is the
// post-collision deviation before rounding to the storage type.
//   feq_q - w_q = w_q (E + O),  feq_opp(q) - w_q = w_q (E - O),
//   E = drho + rho (4.5 (c.u)^2 - 1.5 u.u),  O = 3 rho c.u,
// so each pair q, opp(q) shares c.u, E and O; likewise the Guo term
// (1 - omega/2) 3 w_q (c.F + 3 (c.u)(c.F) - u.F) has the common part
// 3 (c.u)(c.F) - u.F and the odd part c.F.  Products and sums are fused
// (fmaf), which roughly halves the arithmetic of the term-by-term form.
// The grouping keeps that form's accuracy: the small terms of E are summed
// before they are scaled by rho, and omega multiplies the finished
// equilibrium (a premultiplied omega * w_q carries a rounding that is the
// same in every cell and step, and measurably doubles the drift of a
// decaying vortex after 2000 steps).
template <class L, bool FORCED>
__device__ __forceinline__ void collide_dev(const float (&d)[L::Q], float (&o)[L::Q], const float omega,
                                            const float F0, const float F1, const float F2)
{
    float drho, j0, j1, j2;
    moments_dev<L>(d, drho, j0, j1, j2);
    const float rho = 1.0f + drho;
    const float inv = 1.0f / rho;
    const float u0 = (FORCED ? fmaf(0.5f, F0, j0) : j0) * inv;
    const float u1 = (FORCED ? fmaf(0.5f, F1, j1) : j1) * inv;
    const float u2 = (L::D == 3) ? (FORCED ? fmaf(0.5f, F2, j2) : j2) * inv : 0.f;
    const float usq = (L::D == 3) ? fmaf(u2, u2, fmaf(u1, u1, u0 * u0)) : fmaf(u1, u1, u0 * u0);
    const float omw = 1.0f - omega;
    const float r3 = 3.0f * rho;
    const float m15usq = -1.5f * usq;
    const float pref = 1.0f - 0.5f * omega;
    const float uF = (L::D == 3) ? fmaf(u2, F2, fmaf(u1, F1, u0 * F0)) : fmaf(u1, F1, u0 * F0);
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int m = L::opp(q);
        const float w = (float)L::w(q);
        if constexpr (q == m) {                  // rest population: c = 0
            const float fe = w * fmaf(rho, m15usq, drho);
            float out = fmaf(omega, fe, omw * d[q]);
            if constexpr (FORCED) out = fmaf(pref * w * 3.0f, -uF, out);
            o[q] = out;
        } else if constexpr (q < m) {
            const float cu = cdot<L, q>(u0, u1, u2);
            const float E = fmaf(rho, fmaf(4.5f * cu, cu, m15usq), drho);
            const float O = r3 * cu;
            float oq = fmaf(omega, w * (E + O), omw * d[q]);
            float om = fmaf(omega, w * (E - O), omw * d[m]);
            if constexpr (FORCED) {
                const float cF = cdot<L, q>(F0, F1, F2);
                const float kg = pref * w * 3.0f;
                const float ge = fmaf(3.0f * cu, cF, -uF);
                oq = fmaf(kg, ge + cF, oq);
                om = fmaf(kg, ge - cF, om);
            }
            o[q] = oq;
            o[m] = om;
        }
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

// Pull of one cell through wave-uniform source rows (see source_rows).
template <class L, bool WALLS, class T, class G>
__device__ __forceinline__ void pull_dev_rows(float (&d)[L::Q], const T* __restrict__ fin, int x, int y,
                                              int z, const G& g, const View& v)
{
    long long rows[3][3];
    source_rows<L::D>(rows, y, z, g, v);
    int xs[3];
    source_x(xs, x, g);
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = L::c(q, 0), cy = L::c(q, 1), cz = L::D == 3 ? L::c(q, 2) : 0;
        const T* row = fin + q * v.sq + rows[cy + 1][cz + 1];
        d[q] = to_f(ldnt(at_bytes(row, (unsigned)xs[cx + 1] * (unsigned)sizeof(T))));
    });
    if constexpr (WALLS) wall_fix_dev<L>(d, fin, x, y, z, g, v);
}

// Collision and rounded store of one cell whose row (y, z) is wave-uniform:
// orow = fout + y * sy + z * sz.
template <class L, bool FORCED, class T>
__device__ __forceinline__ void collide_store_dev_row(const float (&d)[L::Q], T* __restrict__ orow, int x,
                                                      const View& vo, const float omega, const float F0,
                                                      const float F1, const float F2)
{
    float o[L::Q];
    collide_dev<L, FORCED>(d, o, omega, F0, F1, F2);
    const unsigned xb = (unsigned)x * (unsigned)sizeof(T);
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const T sv = from_f<T>(o[q]);
        T* p = at_bytes(orow + q * vo.sq, xb);
        if constexpr (sizeof(T) == 4) __builtin_nontemporal_store(sv, p);
        else *p = sv;
    });
}

}  // namespace mfa
