/*
 * jams_host.h -- C ABI of the JAMS-compatible host layer (config parsing,
 * lattice, neighbour list, CSR builder, Solver/Hamiltonian/Thermostat
 * plugins and the run loop) that sits above the engine of jams_hip.h.
 *
 * It replaces the reference driver's jams::initialize_simulation /
 * run_simulation / cleanup_simulation (core/jams++.h:20-23,
 * core/jams++.cc:206-377) for configs that select the GPU Heun or RK4 LLG
 * solver (solver.module = "llg-heun-gpu" or "llg-rk4-gpu").  The command-line driver `jams` is a
 * thin wrapper over these calls.  Return codes are 0 on success, non-zero
 * on failure with the message in jams_sim_last_error.
 */
#ifndef JAMS_HOST_H
#define JAMS_HOST_H

#include <stdint.h>

#include "jams_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jams_sim jams_sim;

/* Build a simulation from config sources merged left to right (each an
 * existing file path or inline config text, core/jams++.cc:48-84).
 * device < 0: set-up only (lattice, neighbour list, CSR; no GPU needed).
 * output_dir and name may be NULL ("." and "jams"). */
int jams_sim_create(const char *const *sources, int nsources, const char *output_dir,
                    const char *name, int device, jams_sim **out);
int jams_sim_destroy(jams_sim *sim);
const char *jams_sim_last_error(const jams_sim *sim);

int64_t jams_sim_num_spins(const jams_sim *sim);
int64_t jams_sim_max_steps(const jams_sim *sim);
/* The engine context (NULL in set-up-only mode). */
jh_ctx *jams_sim_engine(jams_sim *sim);

/* run_simulation: run until max_steps (or `limit` more steps if limit >= 0),
 * notifying monitors on their output_steps cadence. */
int jams_sim_run(jams_sim *sim, int64_t limit);

/* Per-spin arrays after set-up (N values / N x 3 values). */
int jams_sim_sites(const jams_sim *sim, double *mus, double *gyro, double *alpha, double *spins);

/* Exchange term `index` (0-based among exchange Hamiltonians): the number of
 * pairs in the reference neighbour list ("computed interactions"), the
 * scaled neighbour list and the reference-layout CSR.  Pass NULL buffers to
 * query sizes (npairs / nnz written through the count pointers). */
int jams_sim_exchange_counts(const jams_sim *sim, int index, int64_t *computed, int64_t *npairs,
                             int64_t *nnz);
int jams_sim_exchange_pairs(const jams_sim *sim, int index, int32_t *pi, int32_t *pj, double *J9);
int jams_sim_exchange_csr(const jams_sim *sim, int index, int64_t *rowptr, int32_t *col, double *val);
/* The same list as the engine receives it (jh_add_exchange_list): per pair
 * the index of its tensor, and the table of distinct tensors (nvalues x 9). */
int jams_sim_exchange_values(const jams_sim *sim, int index, int32_t *vidx, double *values9,
                             int64_t *nvalues);

/* Iteration and time (ps) of the run (Solver::iteration / Solver::time). */
int jams_sim_iteration(const jams_sim *sim, int64_t *iteration, double *time_ps);

/* Write the current spins as a checkpoint: the reference's TSV spin format
 * (helpers/load.h:22-51, loadable as lattice.spins) in lossless %.17g, headed
 * by '#' lines with the iteration, time and seed.  sim.resume = <file>
 * continues the run from it, noise stream included; the reference has no
 * counterpart (it restarts at t = 0).  The "checkpoint" monitor writes the
 * same file every output_steps and at the end of the run. */
int jams_sim_write_checkpoint(jams_sim *sim, const char *path);

/* Time step (ps), temperature (K) and thermostat seed chosen by the config. */
int jams_sim_params(const jams_sim *sim, double *dt_ps, double *temperature, uint64_t *seed);

#ifdef __cplusplus
}
#endif
#endif
