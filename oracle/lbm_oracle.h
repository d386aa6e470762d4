/* lbm_oracle.h -- CPU oracle for the MuPhFASa lattice-Boltzmann step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver's per-step arithmetic, used by tests/ and by bench.py's cpu_baseline
 * leg as the checker.  The product path (muphfasa_amd) never links, loads or
 * calls it.
 *
 * Storage convention (mirrors an AMReX FArrayBox with one ghost layer):
 *   valid box [0,n0) x [0,n1) x [0,n2); one ghost layer in each of the first
 *   D dimensions (2-D lattices: n2 == 1 and no ghost in z).  Extents
 *   e_d = n_d + 2 (or 1 for z in 2-D); element (i,j,k,q) with i,j,k measured
 *   from the valid lower corner lives at
 *       (i+gx) + e0*((j+gy) + e1*((k+gz) + e2*q)),
 *   i.e. a C-order array of shape (Q, e2, e1, e0).
 *
 * Lattice ids: 0 D2Q9, 1 D3Q15, 2 D3Q19, 3 D3Q27.
 * Boundary type ids (reference lbmBCtype order): 0 MovingWall, 1 NoSlip,
 *   2 Inflow, 3 Outflow, 4 PressureOutlet, 5 Periodic.
 */
#ifndef MUPHFASA_LBM_ORACLE_H
#define MUPHFASA_LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_lattice {
    int D, Q;
    int cx[27], cy[27], cz[27];
    double w[27];
    int opp[27];
} orc_lattice;

const orc_lattice* orc_get_lattice(int lattice);

/* single-cell kernels */
double orc_feq(int lattice, double rho, const double* u, int n);
double orc_guo(int lattice, int n, const double* u, const double* F, double omega);
double orc_bgk(double f, double feq, double omega);
void   orc_cell_moments(int lattice, const double* f, const double* F /*NULL*/,
                        double* rho, double* u);
double orc_bb_contrib(int lattice, int m, double rho, const double* U);

/* boundary resolution (BCManager::finalize).  face arrays are indexed
 * [xlo,xhi,ylo,yhi,zlo,zhi]; face_u is [6][3].  Outputs: 27 codes.
 * Returns 0, or -1 on a conflicting edge/corner. */
int orc_resolve_bc(int dim, const int* face_type, const double* face_rho,
                   const double* face_u, int* rtype, double* rrho, double* ru);
int orc_has_conflict(int nfaces, const int* types);

/* field-level operations on ghosted fabs */
void orc_step(int lattice, const double* fin, double* fout, const int* n,
              double omega, const double* F, int nthreads);
void orc_step_masked(int lattice, const double* fin, double* fout, const int* n, double omega,
                     const unsigned char* mask);
void orc_fill_periodic(int lattice, double* f, const int* n, const int* periodic);
int  orc_apply_bc(int lattice, double* f, const int* n, const int* periodic,
                  const int* rtype, const double* rrho, const double* ru);
void orc_stream(int lattice, const double* fin, double* fout, const int* n);
void orc_moments(int lattice, const double* f, const int* n, const double* F,
                 double* rho, double* vel, int streamed);
void orc_init_equilibrium(int lattice, double* rho, double* vel, double* f,
                          const int* n, const double* F);

/* whole advance() loop: step, swap, periodic fill, boundary fill.
 * Returns 0 if the final state is in fa, 1 if it is in fb, <0 on error. */
int orc_run(int lattice, double* fa, double* fb, const int* n, double omega,
            const double* F, const int* periodic, const int* rtype,
            const double* rrho, const double* ru, int nsteps, int nthreads);

/* wall-clock of a bounded run (seconds), for the CPU baseline */
double orc_time_run(int lattice, double* fa, double* fb, const int* n,
                    double omega, const double* F, const int* periodic,
                    const int* rtype, const double* rrho, const double* ru,
                    int nsteps, int nthreads);

/* the same loop on a fully periodic domain cut into boxes of at most mgs
 * cells per side, each with its own ghost layer (the reference's
 * max_grid_size layout), OpenMP over boxes; starts from and returns the
 * valid cells of the single-box fab f.  Seconds of the nsteps steps timed
 * after nwarm untimed ones, or < 0. */
double orc_time_run_boxes(int lattice, double* f, const int* n, int mgs, double omega,
                          const double* F, int nwarm, int nsteps, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
