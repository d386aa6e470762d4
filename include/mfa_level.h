/* mfa_level.h -- several boxes of one lattice on one device: the ghost
 * exchange between them (the role of amrex::MultiFab::FillBoundary in
 * src/Simulation/SimulationSetup.cpp:227 when the domain is chopped into
 * boxes of at most amr.max_grid_size cells) and the physical boundary fill
 * of each box at its place in the domain (PhysBCFunct<BCManager>,
 * SimulationSetup.cpp:230, src/LBM/BoundaryConditions.H:124-189).
 *
 * A level is built once on the host from the box layout, like the
 * BoxArray and the copy metadata AMReX caches for FillBoundary.  The
 * populations of its boxes (FP64, ngrow 1, any strides mfa_field accepts)
 * are bound to numbered slots, typically the two buffers of the A/B pull.
 * A step of the level is then
 *     mfa_level_fill_boundary(lv, a, s); mfa_level_apply_bc(lv, a, g, s);
 *     mfa_level_step(lv, a, b, omega, force, s);
 * (the last is mfa_pull_collide on every box, in one launch) and gives the
 * same valid-cell values, bit for bit, as mfa_step on one box holding the
 * whole domain.  Moments of a box through its ghosts:
 * mfa_moments(..., streamed = 2) after the two fills.
 *
 * Status codes and error messages as in mfa_lbm.h. */
#ifndef MFA_LEVEL_H
#define MFA_LEVEL_H

#include "mfa_lbm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfa_level mfa_level;

/* Build a level: nboxes boxes, box i with valid cells [lo, lo + n) of the
 * domain (box_lo, box_n: nboxes x 3 ints; the third entry 0 / 1 in 2-D).
 * The boxes must tile the domain exactly (no overlap, no gap).  Plans the
 * ghost copies of every box, periodic images included, and keeps them in
 * device memory of the current device (allocated here; freed by
 * mfa_level_destroy).  No counterpart in the reference (AMReX builds this
 * inside FabArrayBase). */
int mfa_level_create(int lattice, const int domain_n[3], const int periodic[3], int nboxes,
                     const int* box_lo, const int* box_n, mfa_level** out);
/* Frees the level's device memory: call it once no work that uses the level
 * is left on any stream. */
int mfa_level_destroy(mfa_level* lv);

/* Number of ghost cells the level's fill_boundary writes (all boxes). */
long long mfa_level_ghost_cells(const mfa_level* lv);

/* Bind the fields of all boxes (boxes[i]: box i, FP64, extents equal to the
 * box's) to slot 0..3.  The exchange, the boundary fill and mfa_level_step
 * need ngrow >= 1; a slot whose boxes have no ghost layer (ngrow 0) serves
 * mfa_level_step_direct only.  Copies their views to device memory
 * with a synchronous copy, so rebind a slot (after reallocating) only while
 * no enqueued level call that uses that slot is still to run. */
int mfa_level_bind(mfa_level* lv, int slot, const mfa_field* boxes);

/* FillBoundary(periodicity): every ghost cell of every box that lies in the
 * periodic-grown domain gets the value (all components) of the valid cell
 * it images, from whichever box owns that cell.  One launch. */
int mfa_level_fill_boundary(const mfa_level* lv, int slot, void* stream);

/* pullCollideFusedMF over the level (src/LBM/Collision.cpp:95-143): every
 * box of slot_in steps through its ghost cells into the same box of
 * slot_out, in one launch; equal, bit for bit, to mfa_pull_collide on each
 * box.  Call after mfa_level_fill_boundary and mfa_level_apply_bc on
 * slot_in.  force: D doubles or NULL. */
int mfa_level_step(const mfa_level* lv, int slot_in, int slot_out, double omega, const double* force,
                   void* stream);

/* The same step without the exchange: what a box pulls from outside itself
 * is read straight from the valid cells of the box that owns it (periodic
 * images included), so slot_in needs no mfa_level_fill_boundary and its
 * ghost cells are never read.  Equal, bit for bit, to
 * mfa_level_fill_boundary + mfa_level_step.  Serves a domain periodic in
 * every dimension whose boxes form a tensor-product grid (the cut positions
 * along each dimension are shared by all boxes, as AMReX's chopping by
 * max_grid_size gives); MFA_ERR_UNSUPPORTED otherwise
 * (mfa_level_direct_available tells beforehand).  No counterpart in the
 * reference, where FillBoundary always precedes the step. */
int mfa_level_step_direct(const mfa_level* lv, int slot_in, int slot_out, double omega,
                          const double* force, void* stream);
int mfa_level_direct_available(const mfa_level* lv);
/* Launch shape of the direct step: cells per thread along z (1, 2 or 4;
 * default 2) and the load kind for a box's own cells (0 plain, default; 1
 * non-temporal; 2 non-temporal for the c_x = 0 rows only).  Results do not
 * depend on it.  Synchronises the device. */
int mfa_level_set_direct_launch(mfa_level* lv, int cells_per_thread, int nontemporal);

/* PhysBCFunct on every box: each ghost cell outside the periodic-grown
 * domain gets, for every population whose source g + c_q is a cell of the
 * box's grown region that lies in the domain, the wall value of its
 * resolved boundary class.  Call after mfa_level_fill_boundary (sources
 * may be ghost cells it filled).  One launch per box. */
int mfa_level_apply_bc(const mfa_level* lv, int slot, const mfa_geom_bc* g, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MFA_LEVEL_H */
