// lattice.hpp -- velocity sets and per-cell LBM arithmetic for the gfx950
// step engine (host + device).
//
// Tables follow src/LBM/LatticeDescriptor.H:23-269 of the reference (same
// velocity order, weights written as the same decimal quotients, same
// opposite map), because population indices are part of the storage format.
//
// Floating-point expressions keep the reference's evaluation order
// (Equilibrium.H:28-57, Collision.H:21-76, Macroscopic.H:36-201,
// BCTypes/BounceBack.H:56-63).  Terms whose lattice coefficient is a
// compile-time zero are dropped; adding an exact zero leaves a finite sum
// unchanged (up to the sign of an exact zero), so the values are the ones the
// reference computes.  All code is built with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/mfa_lbm.h"

namespace mfa {

#define MFA_HD __host__ __device__ __forceinline__

template <int ID>
struct Lattice;

template <>
struct Lattice<MFA_D2Q9> {
    static constexpr int D = 2, Q = 9;
    static MFA_HD constexpr int c(int q, int a)
    {
        constexpr int t[9][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},   {-1, 0, 0}, {0, -1, 0},
                                 {1, 1, 0},  {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
        return t[q][a];
    }
    static MFA_HD constexpr double w(int q)
    {
        constexpr double t[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
        return t[q];
    }
    static MFA_HD constexpr int opp(int q)
    {
        constexpr int t[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
        return t[q];
    }
};

template <>
struct Lattice<MFA_D3Q15> {
    static constexpr int D = 3, Q = 15;
    static MFA_HD constexpr int c(int q, int a)
    {
        constexpr int t[15][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},  {0, -1, 0},
                                  {0, 0, 1},   {0, 0, -1},  {1, 1, 1},   {-1, -1, -1}, {1, 1, -1},
                                  {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}};
        return t[q][a];
    }
    static MFA_HD constexpr double w(int q)
    {
        constexpr double t[15] = {2.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                                  1.0 / 9.0,  1.0 / 9.0,  1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
                                  1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
        return t[q];
    }
    static MFA_HD constexpr int opp(int q)
    {
        constexpr int t[15] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13};
        return t[q];
    }
};

template <>
struct Lattice<MFA_D3Q19> {
    static constexpr int D = 3, Q = 19;
    static MFA_HD constexpr int c(int q, int a)
    {
        constexpr int t[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                                  {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
                                  {-1, 0, -1}, {0, 1, 1}, {0, -1, -1}, {1, -1, 0}, {-1, 1, 0},
                                  {1, 0, -1}, {-1, 0, 1}, {0, 1, -1}, {0, -1, 1}};
        return t[q][a];
    }
    static MFA_HD constexpr double w(int q)
    {
        constexpr double t[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                  1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
        return t[q];
    }
    static MFA_HD constexpr int opp(int q)
    {
        constexpr int t[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
        return t[q];
    }
};

template <>
struct Lattice<MFA_D3Q27> {
    static constexpr int D = 3, Q = 27;
    static MFA_HD constexpr int c(int q, int a)
    {
        constexpr int t[27][3] = {
            {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
            {0, 0, -1},  {1, 1, 0},   {-1, -1, 0}, {1, 0, 1},   {-1, 0, -1}, {0, 1, 1},
            {0, -1, -1}, {1, -1, 0},  {-1, 1, 0},  {1, 0, -1},  {-1, 0, 1},  {0, 1, -1},
            {0, -1, 1},  {1, 1, 1},   {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},
            {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};
        return t[q][a];
    }
    static MFA_HD constexpr double w(int q)
    {
        constexpr double t[27] = {8.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,
                                  2.0 / 27.0,  2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
                                  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
                                  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 216.0,
                                  1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
                                  1.0 / 216.0, 1.0 / 216.0};
        return t[q];
    }
    static MFA_HD constexpr int opp(int q)
    {
        constexpr int t[27] = {0,  2,  1,  4,  3,  6,  5,  8,  7,  10, 9,  12, 11, 14,
                               13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};
        return t[q];
    }
};

// c_q . (a0, a1, a2) in the reference's order ((x + y) + z), zero terms dropped.
template <class L, int q, class T>
MFA_HD T cdot(T a0, T a1, T a2)
{
    constexpr int c0 = L::c(q, 0), c1 = L::c(q, 1), c2 = L::c(q, 2);
    T r = T(0);
    bool have = false;
    if constexpr (c0 != 0) { r = (c0 > 0) ? a0 : -a0; have = true; }
    if constexpr (c1 != 0) { const T t = (c1 > 0) ? a1 : -a1; r = have ? r + t : t; have = true; }
    if constexpr (L::D == 3 && c2 != 0) { const T t = (c2 > 0) ? a2 : -a2; r = have ? r + t : t; }
    (void)have;
    return r;
}

// feq_q(rho, u): src/LBM/Equilibrium.H:28-57
template <class L, int q, class T>
MFA_HD T feq(T rho, T u0, T u1, T u2, T usq)
{
    const T cu = cdot<L, q>(u0, u1, u2);
    return T(L::w(q)) * rho * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq);
}

template <class L, class T>
MFA_HD T usq_of(T u0, T u1, T u2)
{
    if constexpr (L::D == 2) return u0 * u0 + u1 * u1;
    else return u0 * u0 + u1 * u1 + u2 * u2;
}

// Guo term: src/LBM/Collision.H:21-49.  pref = (1 - 0.5 omega); uF = u.F
template <class L, int q, class T>
MFA_HD T guo(T u0, T u1, T u2, T F0, T F1, T F2, T pref, T uF)
{
    const T cF = cdot<L, q>(F0, F1, F2);
    const T cu = cdot<L, q>(u0, u1, u2);
    return pref * T(L::w(q)) * T(3) * (cF + T(3) * cu * cF - uF);
}

// bounce-back momentum correction: src/LBM/BCTypes/BounceBack.H:56-63
template <class L, int m, class T>
MFA_HD T bb_contrib(T rho, T U0, T U1, T U2)
{
    const T cu = cdot<L, m>(U0, U1, U2);
    return T(-2.) * rho * T(L::w(m)) * cu * T(3);
}

}  // namespace mfa
