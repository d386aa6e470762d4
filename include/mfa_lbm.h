/* mfa_lbm.h -- C ABI of the MI355X (gfx950) lattice-Boltzmann step engine.
 *
 * Every entry point replaces one MultiFab-level operator of the reference
 * solver (citations are to the reference tree) or one call site of AMReX's
 * runtime on its step path.  All functions:
 *   - borrow device memory owned by the caller (no allocation except where a
 *     function says so), and enqueue work on the given HIP stream
 *     (NULL = the default stream) without synchronising it;
 *   - return MFA_OK (0) or a negative mfa_status, never abort; the message of
 *     the last failure on the calling thread is available from
 *     mfa_last_error().
 *
 * Storage (the "MultiFab population storage" of one box): structure of
 * arrays.  Element (x,y,z,q) of an mfa_field lives at
 *     data + x + y*sy + z*sz + q*sq            (elements, not bytes)
 * with x,y,z measured from the valid lower corner, so ghost cells are
 * addressed with coordinates -1 and n_d.  ngrow says how many ghost layers
 * the allocation holds around the valid box (0 or 1).  Any strides whose
 * valid cells do not overlap are accepted: the MultiFab's component-major
 * order (sq = e0*e1*e2, one plane per population) and the row-interleaved
 * order (sq = e0, sy = ncomp*e0: the Q rows of one (y,z) side by side, the
 * engine's own allocation, about 5 % faster for the step on MI355X) run the
 * same kernels and give identical results.  The input and output of
 * mfa_step, mfa_step2, mfa_pull_collide and mfa_convert must occupy disjoint
 * byte ranges (all components, ghost layers included); overlapping views are
 * refused with MFA_ERR_INVALID before any device work.
 */
#ifndef MFA_LBM_H
#define MFA_LBM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mfa_status {
    MFA_OK = 0,
    MFA_ERR_INVALID = -1,     /* bad argument / shape mismatch */
    MFA_ERR_UNSUPPORTED = -2, /* combination not built (e.g. dtype) */
    MFA_ERR_HIP = -3,         /* HIP runtime error */
    MFA_ERR_CONFLICT = -4,    /* conflicting boundary conditions */
    MFA_ERR_NOMEM = -5
} mfa_status;

/* lattice ids: src/LBM/LatticeDescriptor.H:266-269 */
enum { MFA_D2Q9 = 0, MFA_D3Q15 = 1, MFA_D3Q19 = 2, MFA_D3Q27 = 3 };

/* storage types.  F64: the reference's amrex::Real (exact reference
 * arithmetic).  F32 / F16DEV / BF16DEV: FP32 / IEEE half / bfloat16 of the
 * shifted deviation d_q = f_q - w_q, collided in FP32 (mfa_step,
 * mfa_moments, mfa_init_equilibrium, mfa_reduce mass/momentum and
 * mfa_convert accept them; the halo entry points are FP64 only). */
enum { MFA_F64 = 0, MFA_F32 = 1, MFA_F16DEV = 2, MFA_BF16DEV = 3 };

/* boundary types, numbered as the reference's lbmBCtype
 * (src/LBM/BoundaryConditions.H:41) */
enum { MFA_BC_MOVING_WALL = 0, MFA_BC_NO_SLIP = 1, MFA_BC_INFLOW = 2,
       MFA_BC_OUTFLOW = 3, MFA_BC_PRESSURE_OUTLET = 4, MFA_BC_PERIODIC = 5 };

/* reductions */
enum { MFA_RED_MASS = 0,      /* sum over valid cells and q of f      -> 1 value  */
       MFA_RED_MOMENTUM = 1,  /* sum over valid cells of sum_q c_q f -> D values */
       MFA_RED_COMP_SUM = 2,  /* per-component sum over valid cells  -> ncomp    */
       MFA_RED_HEALTH = 3     /* populations: [cells with a non-finite f_q,
                                 cells with finite rho = sum_q f_q <= 0]  -> 2 */
};

/* Non-owning view of one box of a field (replaces amrex::Array4). */
typedef struct mfa_field {
    void* data;          /* address of element (0,0,0) of component 0 */
    int n[3];            /* valid cells per dimension (n[2] == 1 in 2-D) */
    int ngrow;           /* ghost layers present around the valid box (0 or 1) */
    long long sy, sz, sq;/* strides in elements; x stride is 1 */
    int ncomp;           /* number of components */
    int dtype;           /* MFA_F64, MFA_F32, MFA_F16DEV, MFA_BF16DEV */
} mfa_field;

/* Domain + resolved boundary table of one lattice (replaces the Geometry
 * periodicity and the finalized BCManager<L>).  Index of a ghost class:
 * sum_d (code_d + 1) * 3^d with code_d in {-1,0,1} (0 in periodic dims). */
typedef struct mfa_geom_bc {
    int lattice;
    int periodic[3];
    int resolved_type[27];
    double resolved_rho[27];
    double resolved_u[27][3];
} mfa_geom_bc;

/* Launch configuration of mfa_step's kernel (per lattice, process-wide):
 * block_x in {64,128,256} cells along x, block_y rows per block
 * (block_x*block_y <= 256), cells_per_thread in {1,2} along z (3-D only),
 * nt_loads: 0 cached population loads, 1 non-temporal, 2 non-temporal only
 * for populations with c_x = 0.  Chosen by the autotuner
 * (muphfasa_amd/autotune.py); no counterpart in the reference, whose launch
 * shapes are AMReX's. */
typedef struct mfa_launch {
    int block_x, block_y, cells_per_thread, nt_loads;
} mfa_launch;

/* Library version and last error message (thread-local). */
const char* mfa_version(void);
const char* mfa_last_error(void);

/* Host: build the resolved 3^D boundary table from per-face settings
 * (faces ordered xlo,xhi,ylo,yhi,zlo,zhi; face_u is [6][3]; velocities and
 * densities in lattice units).  Replaces BCManager<L>::setFace* + finalize
 * (src/LBM/BoundaryConditions.cpp:59-156, src/LBM/BoundaryConditions.H:191-199).
 * Returns MFA_ERR_CONFLICT for the combinations resolveBC aborts on. */
int mfa_resolve_bc(int lattice, const int face_type[6], const double face_rho[6],
                   const double face_u[18], mfa_geom_bc* out);

/* THE HOT PATH.  One full time step on valid cells: pull from fin with the
 * periodic wrap and the wall ghost values (bounce-back / moving wall / inflow
 * / outflow / pressure outlet) evaluated inline from fin's valid cells,
 * moments, feq, BGK and, when force != NULL, the Guo term.  Writes fout's
 * valid cells only; reads only fin's valid cells, so ngrow may be 0.
 * Equivalent to SimulationImpl::advance()'s pullCollideFusedMF + swap +
 * fillGhostCells (src/Simulation/SimulationSetup.cpp:119-139,221-234;
 * src/LBM/Collision.cpp:95-143) on the valid cells.  force: D doubles. */
int mfa_step(const mfa_field* fin, mfa_field* fout, const mfa_geom_bc* g,
             double omega, const double* force, void* stream);

/* Two steps in one pass: fout = the state after mfa_step(fin -> tmp) and
 * mfa_step(tmp -> fout), bit for bit, without tmp.  The intermediate state
 * stays in LDS (temporal blocking), so a pair of steps reads fin about 1.1
 * to 1.3 times and writes fout once instead of reading and writing twice.
 * For callers that advance two steps with nothing in between, e.g. two
 * SimulationImpl::advance() calls (src/Simulation/SimulationSetup.cpp:119-139)
 * between outputs.  FP64: every lattice, periodic or with walls of every
 * kind (the second step's ghost values come from the intermediate state).
 * MFA_F32 storage: D2Q9 boxes with or without walls and fully periodic 3-D
 * boxes (the intermediate state rounded to FP32 as a single step stores it).
 * Anything else returns MFA_ERR_UNSUPPORTED and the caller uses mfa_step:
 * 3-D boxes with walls in FP32, 3-D boxes whose z-plane (all populations of
 * one z, from the first cell to the last) spans more than 4 GiB less 16
 * bytes (the pass addresses a plane through 32-bit offsets, and stores of
 * lanes outside the box go to an offset beyond that range, where the
 * hardware drops them), and FP16 / BF16 storage, where the pass
 * measured slower than two single steps (the words halve the bytes but not
 * the arithmetic, which then bounds the pass).  No counterpart in the
 * reference. */
int mfa_step2(const mfa_field* fin, mfa_field* fout, const mfa_geom_bc* g,
              double omega, const double* force, void* stream);

/* Set (cfg != NULL) or reset to the built-in default (cfg == NULL) the
 * launch configuration mfa_step uses for one lattice; query it. */
int mfa_set_launch(int lattice, const mfa_launch* cfg);
int mfa_get_launch(int lattice, mfa_launch* cfg);

/* The same for reduced storage (dtype MFA_F32, MFA_F16DEV or MFA_BF16DEV).
 * Here cells_per_thread 2 means two x-adjacent cells per thread with paired
 * loads and stores (block_x then counts threads); it is used when n_x is even
 * and the fields' pointers and strides keep the pairs aligned, otherwise the
 * one-cell kernel runs.  nt_loads applies to the paired kernel.  No
 * counterpart in the reference (its populations are FP64 only). */
int mfa_set_launch_dev(int lattice, int dtype, const mfa_launch* cfg);
int mfa_get_launch_dev(int lattice, int dtype, mfa_launch* cfg);

/* Literal pullCollideFusedMF (src/LBM/Collision.cpp:95-143): fin must have
 * ngrow >= 1 with ghosts filled (mfa_fill_periodic + mfa_apply_bc). */
int mfa_pull_collide(const mfa_field* fin, mfa_field* fout, int lattice,
                     double omega, const double* force, void* stream);

/* FillBoundary(periodicity) on one box (src/Simulation/SimulationSetup.cpp:227):
 * every ghost whose out-of-range dimensions are all periodic gets its
 * periodic image. */
int mfa_fill_periodic(mfa_field* f, const mfa_geom_bc* g, void* stream);

/* PhysBCFunct<GpuBndryFuncFab<BCManager<L>>> (src/Simulation/SimulationSetup.cpp:230,
 * src/LBM/BoundaryConditions.H:124-189, src/LBM/BCTypes/ headers): every ghost
 * outside the periodic-grown domain, for every population whose source
 * g + c_q is a valid cell. */
int mfa_apply_bc(mfa_field* f, const mfa_geom_bc* g, void* stream);

/* rho/u moments (updateMacroscopic, src/Simulation/SimulationSetup.cpp:141-166:
 * streamMF + computeMacroscopic).  streamed == 1: moments of the pulled
 * populations (inline wrap/boundary, like mfa_step); streamed == 2: of the
 * populations pulled through f's filled ghost cells (FP64, ngrow 1; for the
 * boxes of an mfa_level); streamed == 0: of f at the cell itself
 * (computeMacroscopic on unstreamed f).  rho: 1 comp, vel: D comps, both
 * F64 with ngrow 0.  force may be NULL. */
int mfa_moments(const mfa_field* f, mfa_field* rho, mfa_field* vel,
                const mfa_geom_bc* g, const double* force, int streamed, void* stream);

/* setInitialConditions (src/Simulation/SimulationSetup.cpp:200-219): f =
 * feq(rho, u) on valid cells; with force != NULL u is first shifted by
 * -F/(2 rho) in place (initVelCorrection, src/LBM/Macroscopic.cpp:164-181)
 * and rho/vel are then recomputed from f with the half-force shift. */
int mfa_init_equilibrium(mfa_field* rho, mfa_field* vel, mfa_field* f, int lattice,
                         const double* force, void* stream);

/* FP64 reductions over valid cells, deterministic order (fixed tree), written
 * to out_dev (device memory).  Uses a library-owned device workspace, one per
 * device and stream, allocated on first use (hipMalloc) and kept until exit,
 * so calls queued on different streams do not share partial sums. */
int mfa_reduce(const mfa_field* f, int lattice, int what, double* out_dev, void* stream);

/* mfa_reduce with no allocation, for callers that cannot call hipMalloc
 * (code under HIP graph capture, serving loops, harnesses that forbid it).
 * No counterpart in the reference.  The partial sums go to a file-scope
 * device array of 1024 x 27 doubles (216 KiB) that each GPU the library runs
 * on holds until the process ends; the caller passes only f and out_dev.
 * Same kernels and results as mfa_reduce, bit for bit; any box size fits
 * (the blocks stride over it); per-component sums of more than 27 components
 * are refused.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two streams, two threads) race on it. */
int mfa_reduce_noalloc(const mfa_field* f, int lattice, int what, double* out_dev, void* stream);

/* Steady-state residual: out_dev[0] = sum (a-b)^2, out_dev[1] = sum a^2 over
 * the valid cells of all components of two same-shaped F64 fields. */
int mfa_residual(const mfa_field* a, const mfa_field* b, double* out_dev, void* stream);

/* Convert populations between FP64 and a reduced storage type (valid
 * cells): FP64 -> d = T(f - w_q) with round-to-nearest-even, or back
 * f = w_q + d.  Exactly one of src/dst is FP64.  No reference counterpart
 * (the reference stores FP64 only). */
int mfa_convert(const mfa_field* src, mfa_field* dst, int lattice, void* stream);

/* Measurement utility (no reference counterpart): a streaming device copy of
 * `bytes` (multiple of 16) with 16-byte loads/stores, one per thread.  Its
 * bandwidth is the measured HBM ceiling the step kernel is compared with,
 * and its known size calibrates the FETCH_SIZE / WRITE_SIZE counters. */
int mfa_stream_copy(const void* src, void* dst, size_t bytes, int nontemporal, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MFA_LBM_H */
