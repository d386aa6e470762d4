/* mfa_host.h -- C ABI of the C++ host (the solver's simulation API).
 *
 * Wraps ISimulation / makeSimulation (src/Simulation/ISimulation.H:16-44,
 * src/Simulation/SimulationFactory.cpp:36-60 of the reference) and the run
 * loop of src/main.cpp:14-38 for callers that are not C++ (Python drivers,
 * the test-suite).  Functions return 0 on success or a negative code; the
 * message is available from mfa_sim_error() (thread-local).  Every call on one
 * mfa_sim must come from one thread at a time.
 */
#ifndef MFA_HOST_H
#define MFA_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfa_sim mfa_sim;

const char* mfa_sim_error(void);

/* inputs_text: the text of an inputs file; overrides: "key=value" strings
 * applied after it (like command-line arguments).  NULL on error. */
mfa_sim* mfa_sim_create(const char* inputs_text, int noverrides, const char* const* overrides);
void mfa_sim_destroy(mfa_sim* s);

int mfa_sim_initialize(mfa_sim* s);                 /* ISimulation::initialize */
int mfa_sim_advance(mfa_sim* s, int nsteps);        /* nsteps x ISimulation::advance (replayed
                                                       from captured HIP graphs of 64 steps) */
int mfa_sim_update_macroscopic(mfa_sim* s, int step);
int mfa_sim_write_output(mfa_sim* s, int step);
int mfa_sim_sync(mfa_sim* s);
/* Time per stage since creation, in this order: advance (device time of the
 * steps; calls = steps), updateMacroscopic (device time), writePlotfile and
 * writeCheckpoint (host time, device-to-host copies included), initialize
 * (host time).  Synchronises the stream.  Stands for the reference's
 * BL_PROFILE regions and its declared reportStageTiming
 * (src/Simulation/SimulationSetup.H:110, SimulationSetup.cpp:122,171). */
int mfa_sim_stage_times(mfa_sim* s, double ms[5], long long calls[5]);

/* lattice id, cells n[3], omega, lattice force F[3], time, start step */
int mfa_sim_info(mfa_sim* s, int* lattice, int n[3], double* omega, double force[3],
                 double* time, int* start_step);
/* host copies of the valid cells, component-major [comp][z][y][x] */
int mfa_sim_copy_f(mfa_sim* s, double* f);
int mfa_sim_copy_macro(mfa_sim* s, double* rho, double* vel);
/* Health of the current populations (mfa_reduce MFA_RED_HEALTH): out[0] =
 * cells with a non-finite population, out[1] = cells whose density is finite
 * and <= 0.  mfa_run_main checks it after every run of steps and stops with
 * an error, before writing output, when either is non-zero.  No counterpart
 * in the reference (its solver has no divergence check). */
int mfa_sim_health(mfa_sim* s, double out[2]);

/* Host-only (no GPU): parse inputs and run the UnitConverter chain.
 * out[0..11] = dx, dt, u_lat, nu_lat, tau, omega, Re, nu_phys, Ma,
 *              F_lat[0], F_lat[1], F_lat[2]. */
int mfa_inputs_units(const char* inputs_text, int noverrides, const char* const* overrides,
                     double out[12]);

/* Host-only (no GPU): the populations of a muphfasa_checkpoint_v1 directory
 * (src/IO/Checkpoint.cpp:105-141, readCheckpoint), in whatever box layout and
 * ghost width its VisMF "f" was written with, as ncomp planes of the n[0..dim)
 * domain's valid cells [comp][z][y][x]; the step and time of its Header.
 * The restart of mfa_sim_initialize reads checkpoints the same way. */
int mfa_checkpoint_read(const char* dir, int dim, const int n[3], int ncomp, double* f, int* step,
                        double* time);

/* src/main.cpp: argv[1] = inputs file, then key=value overrides. */
int mfa_run_main(int argc, const char* const* argv);

#ifdef __cplusplus
}
#endif
#endif
