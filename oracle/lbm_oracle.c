/* lbm_oracle.c -- CPU oracle for the MuPhFASa LBM step.
 *
 * TEST INFRASTRUCTURE ONLY: a restatement of the reference's algorithm used
 * as the checker by tests/ and as bench.py's cpu_baseline.  Nothing in the
 * product links or calls this file.
 *
 * Every floating-point expression keeps the reference's evaluation order
 * (C evaluates left to right; compiled with -ffp-contract=off so no FMA is
 * formed), so results are the IEEE-754 values the reference computes on a
 * non-FMA x86-64 build.  Citations are to the reference tree.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stddef.h>
#include <string.h>
#include <stdlib.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- lattice tables: src/LBM/LatticeDescriptor.H:23-269 --------------- */
static const orc_lattice k_lat[4] = {
    /* D2Q9 (:23-62) */
    {2, 9,
     {0, 1, 0, -1, 0, 1, -1, -1, 1},
     {0, 0, 1, 0, -1, 1, 1, -1, -1},
     {0, 0, 0, 0, 0, 0, 0, 0, 0},
     {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0},
     {0, 3, 4, 1, 2, 7, 8, 5, 6}},
    /* D3Q15 (:72-120) */
    {3, 15,
     {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1},
     {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1},
     {0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1, 1, -1},
     {2.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
      1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
      1.0 / 72.0, 1.0 / 72.0},
     {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13}},
    /* D3Q19 (:130-180) */
    {3, 19,
     {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0},
     {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1},
     {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1},
     {1.0 / 3.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
      1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0},
     {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17}},
    /* D3Q27 (:190-263) */
    {3, 27,
     {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0,
      1, -1, 1, -1, 1, -1, -1, 1},
     {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1,
      1, -1, 1, -1, -1, 1, 1, -1},
     {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1,
      1, -1, -1, 1, 1, -1, 1, -1},
     {8.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0,
      2.0 / 27.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
      1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
      1.0 / 54.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
      1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0},
     {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17,
      20, 19, 22, 21, 24, 23, 26, 25}},
};

#define CS2INV 3.0 /* LatticeDescriptor<..>::cs2inv */

const orc_lattice* orc_get_lattice(int lattice)
{
    if (lattice < 0 || lattice > 3) return NULL;
    return &k_lat[lattice];
}

/* ---- cell kernels ----------------------------------------------------- */

/* src/LBM/Equilibrium.H:28-57 (computeEquilibriumCell2D/3D) */
static double feq_cell(const orc_lattice* L, double rho, const double* u, int n)
{
    if (L->D == 2) {
        const double u_sq = u[0] * u[0] + u[1] * u[1];
        const double cdotu = L->cx[n] * u[0] + L->cy[n] * u[1];
        return L->w[n] * rho *
               (1 + CS2INV * cdotu + 0.5 * CS2INV * CS2INV * cdotu * cdotu -
                0.5 * CS2INV * u_sq);
    } else {
        const double u_sq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double cdotu = L->cx[n] * u[0] + L->cy[n] * u[1] + L->cz[n] * u[2];
        return L->w[n] * rho *
               (1 + CS2INV * cdotu + 0.5 * CS2INV * CS2INV * cdotu * cdotu -
                0.5 * CS2INV * u_sq);
    }
}

/* src/LBM/Collision.H:21-49 (guoForceTermCell2D/3D) */
static double guo_cell(const orc_lattice* L, int n, const double* u, const double* F,
                       double omega)
{
    if (L->D == 2) {
        const double cdotF = L->cx[n] * F[0] + L->cy[n] * F[1];
        const double cdotu = L->cx[n] * u[0] + L->cy[n] * u[1];
        const double udotF = u[0] * F[0] + u[1] * F[1];
        return (1.0 - 0.5 * omega) * L->w[n] * CS2INV *
               (cdotF + CS2INV * cdotu * cdotF - udotF);
    } else {
        const double cdotF = L->cx[n] * F[0] + L->cy[n] * F[1] + L->cz[n] * F[2];
        const double cdotu = L->cx[n] * u[0] + L->cy[n] * u[1] + L->cz[n] * u[2];
        const double udotF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
        return (1.0 - 0.5 * omega) * L->w[n] * CS2INV *
               (cdotF + CS2INV * cdotu * cdotF - udotF);
    }
}

/* src/LBM/Collision.H:68-76 (BGK::collide) */
static double bgk(double f, double feq, double omega)
{
    return (1 - omega) * f + omega * feq;
}

/* src/LBM/Macroscopic.H:36-45 (density), :83-93 / :115-127 (plain velocity),
 * :151-162 / :186-201 (half-force velocity).  Accumulators run n ascending. */
static void cell_moments(const orc_lattice* L, const double* f, const double* F,
                         double* rho_out, double* u)
{
    double rho = 0.;
    for (int n = 0; n < L->Q; ++n) rho += f[n];
    double ux = 0., uy = 0., uz = 0.;
    if (L->D == 2) {
        for (int n = 0; n < L->Q; ++n) {
            ux += L->cx[n] * f[n];
            uy += L->cy[n] * f[n];
        }
        if (F) {
            u[0] = (ux + 0.5 * F[0]) / rho;
            u[1] = (uy + 0.5 * F[1]) / rho;
        } else {
            u[0] = ux / rho;
            u[1] = uy / rho;
        }
        u[2] = 0.;
    } else {
        for (int n = 0; n < L->Q; ++n) {
            ux += L->cx[n] * f[n];
            uy += L->cy[n] * f[n];
            uz += L->cz[n] * f[n];
        }
        if (F) {
            u[0] = (ux + 0.5 * F[0]) / rho;
            u[1] = (uy + 0.5 * F[1]) / rho;
            u[2] = (uz + 0.5 * F[2]) / rho;
        } else {
            u[0] = ux / rho;
            u[1] = uy / rho;
            u[2] = uz / rho;
        }
    }
    *rho_out = rho;
}

/* src/LBM/BCTypes/BounceBack.H:56-63 (contrib) */
static double bb_contrib(const orc_lattice* L, int m, double rho, const double* U)
{
    double cdotu;
    if (L->D == 2) cdotu = L->cx[m] * U[0] + L->cy[m] * U[1];
    else cdotu = L->cx[m] * U[0] + L->cy[m] * U[1] + L->cz[m] * U[2];
    return -2. * rho * L->w[m] * cdotu * CS2INV;
}

double orc_feq(int lattice, double rho, const double* u, int n)
{
    return feq_cell(&k_lat[lattice], rho, u, n);
}
double orc_guo(int lattice, int n, const double* u, const double* F, double omega)
{
    return guo_cell(&k_lat[lattice], n, u, F, omega);
}
double orc_bgk(double f, double feq, double omega) { return bgk(f, feq, omega); }
void orc_cell_moments(int lattice, const double* f, const double* F, double* rho,
                      double* u)
{
    cell_moments(&k_lat[lattice], f, F, rho, u);
}
double orc_bb_contrib(int lattice, int m, double rho, const double* U)
{
    return bb_contrib(&k_lat[lattice], m, rho, U);
}

/* ---- boundary resolution: src/LBM/BoundaryConditions.cpp:37-156 ------- */
enum { BC_MOVING = 0, BC_NOSLIP = 1, BC_INFLOW = 2, BC_OUTFLOW = 3, BC_PRESSURE = 4,
       BC_PERIODIC = 5 };

/* src/LBM/BoundaryConditions.H:82-92 */
static int bc_rank(int t)
{
    switch (t) {
    case BC_NOSLIP: return 3;
    case BC_MOVING: return 2;
    case BC_INFLOW: return 1;
    case BC_OUTFLOW: return 0;
    case BC_PRESSURE: return 0;
    default: return -1;
    }
}

/* src/LBM/BoundaryConditions.cpp:37-57 */
int orc_has_conflict(int nfaces, const int* types)
{
    int has_in = 0, has_out = 0, has_p = 0, nmov = 0;
    for (int i = 0; i < nfaces; ++i) {
        if (types[i] == BC_INFLOW) has_in = 1;
        if (types[i] == BC_OUTFLOW) has_out = 1;
        if (types[i] == BC_PRESSURE) has_p = 1;
        if (types[i] == BC_MOVING) nmov++;
    }
    return (has_in + has_out + has_p) > 1 || nmov > 1;
}

/* src/LBM/BoundaryConditions.cpp:59-156 (resolveBC + BCManager::finalize) */
int orc_resolve_bc(int dim, const int* face_type, const double* face_rho,
                   const double* face_u, int* rtype, double* rrho, double* ru)
{
    int ncodes = (dim == 2) ? 9 : 27;
    for (int i = 0; i < 27; ++i) {
        rtype[i] = BC_PERIODIC;
        rrho[i] = 1.0;
        ru[3 * i] = ru[3 * i + 1] = ru[3 * i + 2] = 0.;
    }
    for (int idx = 0; idx < ncodes; ++idx) {
        int code[3] = {0, 0, 0};
        int tmp = idx, all_zero = 1;
        for (int d = 0; d < dim; ++d) {
            code[d] = (tmp % 3) - 1;
            tmp /= 3;
            if (code[d] != 0) all_zero = 0;
        }
        if (all_zero) continue;
        int types[3], nf = 0;
        double rhos[3], us[3][3];
        for (int d = 0; d < dim; ++d) {
            if (code[d] == 0) continue;
            int fi = 2 * d + (code[d] < 0 ? 0 : 1);
            if (face_type[fi] == BC_PERIODIC) continue;
            types[nf] = face_type[fi];
            rhos[nf] = face_rho[fi];
            for (int c = 0; c < 3; ++c) us[nf][c] = face_u[3 * fi + c];
            nf++;
        }
        if (nf == 0) continue;
        if (orc_has_conflict(nf, types)) return -1;
        int max_rank = -1;
        for (int i = 0; i < nf; ++i)
            if (bc_rank(types[i]) > max_rank) max_rank = bc_rank(types[i]);
        int winner = 0;
        for (int i = 0; i < nf; ++i)
            if (bc_rank(types[i]) == max_rank) { winner = types[i]; break; }
        double sum_rho = 0, sum_u[3] = {0, 0, 0};
        int count = 0;
        for (int i = 0; i < nf; ++i) {
            if (types[i] != winner) continue;
            sum_rho += rhos[i];
            for (int d = 0; d < dim; ++d) sum_u[d] += us[i][d];
            ++count;
        }
        rtype[idx] = winner;
        rrho[idx] = sum_rho / count;
        for (int d = 0; d < dim; ++d) ru[3 * idx + d] = sum_u[d] / count;
    }
    return 0;
}

/* ---- fab indexing ----------------------------------------------------- */
typedef struct {
    int n[3], g[3], e[3];
    long sq;
} fab_t;

static fab_t make_fab(const orc_lattice* L, const int* n)
{
    fab_t f;
    for (int d = 0; d < 3; ++d) {
        f.n[d] = n[d];
        f.g[d] = (d < L->D) ? 1 : 0;
        f.e[d] = n[d] + 2 * f.g[d];
    }
    f.sq = (long)f.e[0] * f.e[1] * f.e[2];
    return f;
}

static inline long fidx(const fab_t* f, int i, int j, int k, int q)
{
    return (long)(i + f->g[0]) +
           (long)f->e[0] * ((long)(j + f->g[1]) + (long)f->e[1] * (long)(k + f->g[2])) +
           f->sq * q;
}

/* ---- pull + collide: src/LBM/Collision.cpp:18-93, Streaming.H:20-35 ---- */
void orc_step(int lattice, const double* fin, double* fout, const int* n,
              double omega, const double* F, int nthreads)
{
    const orc_lattice* L = &k_lat[lattice];
    fab_t fb = make_fab(L, n);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) collapse(2) if (nthreads != 1)
#endif
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                double f[27], u[3], rho;
                for (int q = 0; q < L->Q; ++q)
                    f[q] = fin[fidx(&fb, i - L->cx[q], j - L->cy[q], k - L->cz[q], q)];
                cell_moments(L, f, F, &rho, u);
                for (int q = 0; q < L->Q; ++q) {
                    const double feq_n = feq_cell(L, rho, u, q);
                    double out = bgk(f[q], feq_n, omega);
                    if (F) out = out + guo_cell(L, q, u, F, omega);
                    fout[fidx(&fb, i, j, k, q)] = out;
                }
            }
}

/* ---- solid obstacles (no reference counterpart: the reference has no
 *      interior solids).  Half-way bounce-back: a fluid cell pulling from a
 *      solid neighbour takes its own post-collision population pointing into
 *      the solid, f_q(x) = f*_opp(q)(x); solid cells hold w_q.  Pull sources
 *      outside the domain come from the ghost layer as in orc_step.
 *      mask: (n2,n1,n0) uint8, 1 = solid. ---------------------------------- */
void orc_step_masked(int lattice, const double* fin, double* fout, const int* n, double omega,
                     const unsigned char* mask)
{
    const orc_lattice* L = &k_lat[lattice];
    fab_t fb = make_fab(L, n);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const long c = i + (long)n[0] * (j + (long)n[1] * k);
                if (mask[c]) {
                    for (int q = 0; q < L->Q; ++q) fout[fidx(&fb, i, j, k, q)] = L->w[q];
                    continue;
                }
                double f[27], u[3], rho;
                for (int q = 0; q < L->Q; ++q) {
                    const int si = i - L->cx[q], sj = j - L->cy[q], sk = k - L->cz[q];
                    const int inside = si >= 0 && si < n[0] && sj >= 0 && sj < n[1] && sk >= 0 &&
                                       sk < n[2];
                    if (inside && mask[si + (long)n[0] * (sj + (long)n[1] * sk)])
                        f[q] = fin[fidx(&fb, i, j, k, L->opp[q])];
                    else
                        f[q] = fin[fidx(&fb, si, sj, sk, q)];
                }
                cell_moments(L, f, NULL, &rho, u);
                for (int q = 0; q < L->Q; ++q)
                    fout[fidx(&fb, i, j, k, q)] = bgk(f[q], feq_cell(L, rho, u, q), omega);
            }
}

/* ---- periodic ghost fill (FillBoundary(periodicity) on a single box,
 *      src/Simulation/SimulationSetup.cpp:227): a ghost is filled iff every
 *      dimension in which it lies outside the domain is periodic. --------- */
void orc_fill_periodic(int lattice, double* f, const int* n, const int* periodic)
{
    const orc_lattice* L = &k_lat[lattice];
    fab_t fb = make_fab(L, n);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int k = -fb.g[2]; k < n[2] + fb.g[2]; ++k)
        for (int j = -fb.g[1]; j < n[1] + fb.g[1]; ++j)
            for (int i = -fb.g[0]; i < n[0] + fb.g[0]; ++i) {
                const int interior_row = (j >= 0 && j < n[1] && k >= 0 && k < n[2]);
                if (interior_row && i == 0 && n[0] > 1) i = n[0] - 1; /* skip valid run */
                int c[3] = {i, j, k}, s[3], out = 0, ok = 1;
                for (int d = 0; d < 3; ++d) {
                    s[d] = c[d];
                    if (c[d] < 0 || c[d] >= n[d]) {
                        out = 1;
                        if (!periodic[d]) ok = 0;
                        s[d] = (c[d] + n[d]) % n[d];
                    }
                }
                if (!out || !ok) continue;
                for (int q = 0; q < L->Q; ++q)
                    f[fidx(&fb, i, j, k, q)] = f[fidx(&fb, s[0], s[1], s[2], q)];
            }
}

/* ---- physical boundary fill: BCManager::operator()
 *      (src/LBM/BoundaryConditions.H:124-189) over every ghost outside the
 *      periodic-grown domain, dispatching to BounceBack (BounceBack.H:36-53;
 *      NoSlip/MovingWall/Inflow wrappers), Outflow (OutflowKernel.H:26-38) and
 *      PressureOutlet (PressureOutletKernel.H:29-64). ---------------------- */
int orc_apply_bc(int lattice, double* f, const int* n, const int* periodic,
                 const int* rtype, const double* rrho, const double* ru)
{
    const orc_lattice* L = &k_lat[lattice];
    fab_t fb = make_fab(L, n);
    for (int k = -fb.g[2]; k < n[2] + fb.g[2]; ++k)
        for (int j = -fb.g[1]; j < n[1] + fb.g[1]; ++j)
            for (int i = -fb.g[0]; i < n[0] + fb.g[0]; ++i) {
                int c[3] = {i, j, k}, code[3] = {0, 0, 0}, wall = 0;
                for (int d = 0; d < L->D; ++d) {
                    if (c[d] < 0) code[d] = -1;
                    else if (c[d] >= n[d]) code[d] = 1;
                    if (periodic[d]) code[d] = 0;
                    if (code[d]) wall = 1;
                }
                if (!wall) continue;
                int idx = 0, p3 = 1;
                for (int d = 0; d < L->D; ++d) { idx += (code[d] + 1) * p3; p3 *= 3; }
                const int t = rtype[idx];
                if (t == BC_PERIODIC) continue;
                for (int q = 0; q < L->Q; ++q) {
                    int s[3] = {i + L->cx[q], j + L->cy[q], k + L->cz[q]};
                    int inside = 1;
                    for (int d = 0; d < 3; ++d)
                        if (s[d] < 0 || s[d] >= n[d]) inside = 0;
                    if (!inside) continue;
                    const long g = fidx(&fb, i, j, k, q);
                    if (t == BC_MOVING || t == BC_NOSLIP || t == BC_INFLOW) {
                        const int m = L->opp[q];
                        f[g] = f[fidx(&fb, s[0], s[1], s[2], m)] +
                               bb_contrib(L, m, rrho[idx], &ru[3 * idx]);
                    } else if (t == BC_OUTFLOW) {
                        f[g] = f[fidx(&fb, s[0], s[1], s[2], q)];
                    } else if (t == BC_PRESSURE) {
                        double rho = 0.0, jv[3] = {0., 0., 0.};
                        for (int m = 0; m < L->Q; ++m) {
                            const double fm = f[fidx(&fb, s[0], s[1], s[2], m)];
                            rho += fm;
                            jv[0] += L->cx[m] * fm;
                            jv[1] += L->cy[m] * fm;
                            if (L->D == 3) jv[2] += L->cz[m] * fm;
                        }
                        const double inv_rho = 1.0 / rho;
                        double cu, uu;
                        if (L->D == 2) {
                            cu = (L->cx[q] * jv[0] + L->cy[q] * jv[1]) * inv_rho;
                            uu = (jv[0] * jv[0] + jv[1] * jv[1]) * inv_rho * inv_rho;
                        } else {
                            cu = (L->cx[q] * jv[0] + L->cy[q] * jv[1] + L->cz[q] * jv[2]) *
                                 inv_rho;
                            uu = (jv[0] * jv[0] + jv[1] * jv[1] + jv[2] * jv[2]) * inv_rho *
                                 inv_rho;
                        }
                        const double feq_even =
                            L->w[q] * rrho[idx] *
                            (1.0 + 0.5 * cu * cu * CS2INV * CS2INV - 0.5 * uu * CS2INV);
                        f[g] = -f[fidx(&fb, s[0], s[1], s[2], L->opp[q])] + 2.0 * feq_even;
                    } else {
                        return -2;
                    }
                }
            }
    return 0;
}

/* ---- streamMF: src/LBM/Streaming.cpp:15-47 ---------------------------- */
void orc_stream(int lattice, const double* fin, double* fout, const int* n)
{
    const orc_lattice* L = &k_lat[lattice];
    fab_t fb = make_fab(L, n);
    for (int q = 0; q < L->Q; ++q)
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i)
                    fout[fidx(&fb, i, j, k, q)] =
                        fin[fidx(&fb, i - L->cx[q], j - L->cy[q], k - L->cz[q], q)];
}

/* ---- computeMacroscopic (src/LBM/Macroscopic.cpp:115-162), optionally on
 *      the pulled populations (updateMacroscopic,
 *      src/Simulation/SimulationSetup.cpp:141-166).  rho: (n2,n1,n0);
 *      vel: (D,n2,n1,n0). -------------------------------------------------- */
void orc_moments(int lattice, const double* f, const int* n, const double* F,
                 double* rho, double* vel, int streamed)
{
    const orc_lattice* L = &k_lat[lattice];
    fab_t fb = make_fab(L, n);
    const long nc = (long)n[0] * n[1] * n[2];
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                double fl[27], u[3], r;
                for (int q = 0; q < L->Q; ++q)
                    fl[q] = streamed ? f[fidx(&fb, i - L->cx[q], j - L->cy[q], k - L->cz[q], q)]
                                     : f[fidx(&fb, i, j, k, q)];
                cell_moments(L, fl, F, &r, u);
                const long c = i + (long)n[0] * (j + (long)n[1] * k);
                rho[c] = r;
                for (int d = 0; d < L->D; ++d) vel[d * nc + c] = u[d];
            }
}

/* ---- setInitialConditions (src/Simulation/SimulationSetup.cpp:200-219):
 *      unforced: f = feq(rho, u).  Forced: u -= 0.5 F / rho
 *      (Macroscopic.cpp:164-181), f = feq(rho, u), then rho/u are recomputed
 *      from the (unstreamed) f with the half-force shift. ------------------ */
void orc_init_equilibrium(int lattice, double* rho, double* vel, double* f,
                          const int* n, const double* F)
{
    const orc_lattice* L = &k_lat[lattice];
    fab_t fb = make_fab(L, n);
    const long nc = (long)n[0] * n[1] * n[2];
    if (F) {
        for (long c = 0; c < nc; ++c)
            for (int d = 0; d < L->D; ++d) vel[d * nc + c] -= 0.5 * F[d] / rho[c];
    }
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const long c = i + (long)n[0] * (j + (long)n[1] * k);
                double u[3] = {0., 0., 0.};
                for (int d = 0; d < L->D; ++d) u[d] = vel[d * nc + c];
                for (int q = 0; q < L->Q; ++q) f[fidx(&fb, i, j, k, q)] = feq_cell(L, rho[c], u, q);
            }
    if (F) orc_moments(lattice, f, n, F, rho, vel, 0);
}

/* ---- SimulationImpl::advance (src/Simulation/SimulationSetup.cpp:119-139) */
int orc_run(int lattice, double* fa, double* fb, const int* n, double omega,
            const double* F, const int* periodic, const int* rtype,
            const double* rrho, const double* ru, int nsteps, int nthreads)
{
    double* cur = fa;
    double* nxt = fb;
    int walls = 0;
    for (int d = 0; d < k_lat[lattice].D; ++d)
        if (!periodic[d]) walls = 1;
    for (int s = 0; s < nsteps; ++s) {
        orc_step(lattice, cur, nxt, n, omega, F, nthreads);
        double* t = cur;
        cur = nxt;
        nxt = t;
        orc_fill_periodic(lattice, cur, n, periodic);
        /* a fully periodic box has no ghost behind a wall: the boundary
         * fill would visit every ghost and write none */
        if (walls && orc_apply_bc(lattice, cur, n, periodic, rtype, rrho, ru) != 0) return -2;
    }
    return cur == fa ? 0 : 1;
}

double orc_time_run(int lattice, double* fa, double* fb, const int* n, double omega,
                    const double* F, const int* periodic, const int* rtype,
                    const double* rrho, const double* ru, int nsteps, int nthreads)
{
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int r = orc_run(lattice, fa, fb, n, omega, F, periodic, rtype, rrho, ru, nsteps,
                    nthreads);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (r < 0) return -1.0;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ---- the reference's box layout, for the CPU baseline.  The domain of a
 *      fully periodic box is cut into boxes of at most mgs cells per side
 *      (AMReX max_grid_size; 64 by default, src/IO/InputParser.cpp), each
 *      with a one-cell ghost layer.  A step pull-collides every box (OpenMP
 *      over boxes, as MFIter + omp parallel run pullCollideFusedMF,
 *      src/LBM/Collision.cpp:95-143), swaps, and fills every ghost cell of
 *      every box from the valid cell of its periodic image in the box that
 *      owns it (FillBoundary, src/Simulation/SimulationSetup.cpp:221-234).
 *      Starts from the valid cells of the single-box fab f and writes the
 *      final state back into them; returns the seconds taken by the nsteps
 *      steps that follow nwarm untimed ones, or < 0 on bad arguments. ---- */
typedef struct box_t {
    int lo[3], n[3];
    fab_t fb;
    double *a, *b;
} box_t;

static void fill_boxes(const orc_lattice* L, box_t* bx, int nbox, const int* n, int mgs,
                       const int* nb, int use_b)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int bi = 0; bi < nbox; ++bi) {
        box_t* B = &bx[bi];
        double* dst = use_b ? B->b : B->a;
        for (int k = -B->fb.g[2]; k < B->n[2] + B->fb.g[2]; ++k)
            for (int j = -B->fb.g[1]; j < B->n[1] + B->fb.g[1]; ++j)
                for (int i = -B->fb.g[0]; i < B->n[0] + B->fb.g[0]; ++i) {
                    const int c[3] = {i, j, k};
                    int ghost = 0, G[3], ob[3], l[3];
                    for (int d = 0; d < 3; ++d) {
                        if (c[d] < 0 || c[d] >= B->n[d]) ghost = 1;
                        G[d] = ((B->lo[d] + c[d]) % n[d] + n[d]) % n[d];
                        ob[d] = G[d] / mgs;
                        l[d] = G[d] - ob[d] * mgs;
                    }
                    if (!ghost) continue;
                    const box_t* S = &bx[ob[0] + nb[0] * (ob[1] + nb[1] * ob[2])];
                    const double* src = use_b ? S->b : S->a;
                    for (int q = 0; q < L->Q; ++q)
                        dst[fidx(&B->fb, i, j, k, q)] = src[fidx(&S->fb, l[0], l[1], l[2], q)];
                }
    }
}

double orc_time_run_boxes(int lattice, double* f, const int* n, int mgs, double omega,
                          const double* F, int nwarm, int nsteps, int nthreads)
{
    const orc_lattice* L = &k_lat[lattice];
    if (mgs < 1 || nwarm < 0 || nsteps < 1) return -1.0;
    if (L->D == 2 && n[2] != 1) return -1.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    int nb[3];
    for (int d = 0; d < 3; ++d) nb[d] = (d < L->D) ? (n[d] + mgs - 1) / mgs : 1;
    const int nbox = nb[0] * nb[1] * nb[2];
    box_t* bx = (box_t*)calloc((size_t)nbox, sizeof(box_t));
    if (!bx) return -1.0;
    fab_t whole = make_fab(L, n);
    int ok = 1;
    for (int bk = 0; bk < nb[2]; ++bk)
        for (int bj = 0; bj < nb[1]; ++bj)
            for (int bi = 0; bi < nb[0]; ++bi) {
                box_t* B = &bx[bi + nb[0] * (bj + nb[1] * bk)];
                const int bb[3] = {bi, bj, bk};
                for (int d = 0; d < 3; ++d) {
                    B->lo[d] = (d < L->D) ? bb[d] * mgs : 0;
                    B->n[d] = (d < L->D) ? ((n[d] - B->lo[d]) < mgs ? n[d] - B->lo[d] : mgs) : 1;
                }
                B->fb = make_fab(L, B->n);
                B->a = (double*)malloc(sizeof(double) * (size_t)B->fb.sq * L->Q);
                B->b = (double*)malloc(sizeof(double) * (size_t)B->fb.sq * L->Q);
                if (!B->a || !B->b) { ok = 0; continue; }
                for (int q = 0; q < L->Q; ++q)
                    for (int k = 0; k < B->n[2]; ++k)
                        for (int j = 0; j < B->n[1]; ++j)
                            for (int i = 0; i < B->n[0]; ++i)
                                B->a[fidx(&B->fb, i, j, k, q)] =
                                    f[fidx(&whole, B->lo[0] + i, B->lo[1] + j, B->lo[2] + k, q)];
            }
    double secs = -1.0;
    if (ok) {
        fill_boxes(L, bx, nbox, n, mgs, nb, 0);
        int cur_b = 0;   /* the current state lives in a (0) or b (1) */
        struct timespec t0, t1;
        for (int s = 0; s < nwarm + nsteps; ++s) {
            if (s == nwarm) clock_gettime(CLOCK_MONOTONIC, &t0);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
            for (int bi = 0; bi < nbox; ++bi)
                orc_step(lattice, cur_b ? bx[bi].b : bx[bi].a, cur_b ? bx[bi].a : bx[bi].b, bx[bi].n,
                         omega, F, 1);
            cur_b = !cur_b;
            fill_boxes(L, bx, nbox, n, mgs, nb, cur_b);
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        secs = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
        for (int bi = 0; bi < nbox; ++bi) {
            const box_t* B = &bx[bi];
            const double* src = cur_b ? B->b : B->a;
            for (int q = 0; q < L->Q; ++q)
                for (int k = 0; k < B->n[2]; ++k)
                    for (int j = 0; j < B->n[1]; ++j)
                        for (int i = 0; i < B->n[0]; ++i)
                            f[fidx(&whole, B->lo[0] + i, B->lo[1] + j, B->lo[2] + k, q)] =
                                src[fidx(&B->fb, i, j, k, q)];
        }
    }
    for (int bi = 0; bi < nbox; ++bi) {
        free(bx[bi].a);
        free(bx[bi].b);
    }
    free(bx);
    return secs;
}
