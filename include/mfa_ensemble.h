/* mfa_ensemble.h -- C ABI of the batched engine: hundreds of independent
 * D2Q9 lattices of uneven sizes advanced by one kernel launch per time step
 * (persistent tile loop), optionally captured in a hipGraph, with per-cell
 * solid obstacles and per-lattice convergence retirement.
 *
 * No counterpart in the reference (it advances one lattice per process);
 * every lattice uses the reference's per-cell step (the same kernels as
 * mfa_step) and one shared resolved boundary table (mfa_resolve_bc).
 * Populations of lattice l live at element offset pop_off[l] of one buffer,
 * row-interleaved [y][q][x] (the 9 rows of one y side by side, element
 * (x,y,q) at pop_off + x + y*9*nx + q*nx); per-cell arrays (masks: uint8, 1 = solid; rho;
 * vel as two planes of total_cells) at cell_off[l].  The ensemble object
 * owns only its small metadata; population and field buffers belong to the
 * caller.  Status codes as in mfa_lbm.h; message from mfa_ensemble_error().
 */
#ifndef MFA_ENSEMBLE_H
#define MFA_ENSEMBLE_H

#include "mfa_lbm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfa_ensemble mfa_ensemble;

const char* mfa_ensemble_error(void);

/* lattice: MFA_D2Q9; dtype: MFA_F64 or MFA_F32 (shifted deviations);
 * tile_x*tile_y cells per block (multiple of 64, at most 256); masked: the
 * lattices carry solid masks.  NULL on error. */
mfa_ensemble* mfa_ensemble_create(int lattice, int dtype, int nlat, const int* nx, const int* ny,
                                  const double* omega, const mfa_geom_bc* bc, int tile_x,
                                  int tile_y, int masked);
void mfa_ensemble_destroy(mfa_ensemble* e);
int mfa_ensemble_info(const mfa_ensemble* e, long long* total_populations, long long* total_cells,
                      long long* pop_offsets, long long* cell_offsets, int* ntiles);
/* Active lattices (host array of nlat flags); retired lattices are skipped. */
int mfa_ensemble_set_active(mfa_ensemble* e, const int* active);
/* Number of persistent blocks (0 = one block per tile). */
int mfa_ensemble_set_grid(mfa_ensemble* e, int blocks);
/* f = feq(rho, u) of the given per-cell fields (NULL = rest state rho=1,u=0);
 * solid cells get the rest equilibrium.  With masks it also records which
 * tiles lie next to a solid (the other tiles skip every mask read from then
 * on), which synchronises the stream; the masks must not change afterwards. */
int mfa_ensemble_init(mfa_ensemble* e, void* f, const unsigned char* masks, const double* rho,
                      const double* vel, void* stream);
/* Warm start: fa = feq(rho, u) as mfa_ensemble_init, then one step fa -> fb,
 * then fa = (fa + fb) / 2 (fb is scratch).  A plain equilibrium start from
 * fields that are not the lattice's own steady state carries a part that no
 * run to a residual criterion removes: it converges to a field about 6e-4
 * times the start's error away from the cold start's (measured on a D2Q9
 * lid cavity); the average of two consecutive states does not (2.6e-11). */
int mfa_ensemble_init_warm(mfa_ensemble* e, void* fa, void* fb, const unsigned char* masks,
                           const double* rho, const double* vel, void* stream);
/* Start from per-cell (rho, u) with options (flags, OR-ed):
 *   MFA_START_NEQ      f = feq(rho, u) plus the first-order Chapman-Enskog
 *                      non-equilibrium part of u, -3 (tau - 1) w_q rho Q_q:S
 *                      (Q_q = c_q c_q - I/3, S the strain rate by central
 *                      differences, second-order one-sided through half-way
 *                      walls and solids, first-order towards open faces);
 *                      without it, f = feq(rho, u) as mfa_ensemble_init;
 *   MFA_START_AVERAGE  then one step fa -> fb and fa = (fa + fb) / 2 (fb is
 *                      scratch), as mfa_ensemble_init_warm.
 * mfa_ensemble_init_warm(...) is mfa_ensemble_init_start(..., MFA_START_AVERAGE).
 * The reference initialises feq only (src/Simulation/SimulationSetup.cpp:200-219);
 * the non-equilibrium part is this engine's addition for predicted starts. */
enum { MFA_START_NEQ = 1, MFA_START_AVERAGE = 2 };
int mfa_ensemble_init_start(mfa_ensemble* e, void* fa, void* fb, const unsigned char* masks,
                            const double* rho, const double* vel, int flags, void* stream);
int mfa_ensemble_step(mfa_ensemble* e, const void* fin, void* fout, const unsigned char* masks,
                      void* stream);
/* Two steps fin -> fout in one launch, bit-identical to two
 * mfa_ensemble_step calls (the intermediate state stays in LDS; 32 x 16
 * tiles).  No counterpart in the reference. */
int mfa_ensemble_step2(mfa_ensemble* e, const void* fin, void* fout, const unsigned char* masks,
                       void* stream);
/* nsteps steps alternating fa->fb->fa...: two per launch for whole
 * multiples of four steps, then single steps, so the state ends in fa when
 * nsteps is even; use_graph captures the launches in one hipGraph and
 * replays it. */
int mfa_ensemble_run(mfa_ensemble* e, void* fa, void* fb, const unsigned char* masks, int nsteps,
                     int use_graph, void* stream);
int mfa_ensemble_moments(mfa_ensemble* e, const void* f, const unsigned char* masks, double* rho,
                         double* vel, void* stream);
/* Host output per_lattice[2l] = sum |u - u_prev|^2, [2l+1] = sum |u|^2 over
 * the fluid cells of active lattice l (synchronises the stream). */
int mfa_ensemble_residual(mfa_ensemble* e, const double* vel, const double* vel_prev,
                          const unsigned char* masks, double* per_lattice, void* stream);

#ifdef __cplusplus
}
#endif
#endif
