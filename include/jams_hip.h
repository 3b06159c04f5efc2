/*
 * jams_hip.h -- C ABI of the MI355X (gfx950) atomistic Heun-LLG engine.
 *
 * This is the drop-in boundary for JAMS's GPU Heun solver.  A JAMS build
 * replaces CUDAHeunLLGSolver (solvers/cuda_llg_heun.cu:21-122), its
 * Hamiltonians' calculate_fields (exchange -> SparseMatrix::multiply_gpu,
 * containers/sparse_matrix.h:346-447; zeeman, applied-field, uniaxial,
 * dipole-fft) and CudaThermostatClassical (thermostats/
 * cuda_thermostat_classical.cc:22-56) by calls into this library.  Paths are
 * relative to src/jams/ of stonerlab/jams.
 *
 * Conventions
 *  - Every function returns 0 (JH_OK) on success or a JH_ERR_* code; the
 *    message of the last failure is jh_last_error(ctx) (jh_last_error(NULL)
 *    for failures of jh_create).  The JAMS-side plugin turns a non-zero
 *    return into std::runtime_error, the reference's error convention
 *    (core/jams++.cc:306-323).
 *  - Host arrays are the reference's layout: N x 3 row-major doubles for
 *    spins, fields and noise (core/lattice.cc:688-694), N doubles for
 *    per-spin scalars.  The engine keeps its own device layout (SoA) and
 *    converts at this boundary; nothing is shared implicitly.
 *  - Units are the reference's internal units: time ps, energy meV, field T,
 *    mus meV/T, gyro rad/(ps T) (helpers/consts.h:18-38).
 *  - One context drives one GPU on one HIP stream; a context is not
 *    thread-safe.  All work is stream-ordered; calls that return host data
 *    synchronise the stream.
 */
#ifndef JAMS_HIP_H
#define JAMS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JH_OK 0
#define JH_ERR_INVALID 1      /* bad argument or size */
#define JH_ERR_HIP 2          /* HIP runtime failure (message has the HIP error) */
#define JH_ERR_STATE 3        /* call made in the wrong order */
#define JH_ERR_UNSUPPORTED 4  /* combination not implemented */
#define JH_ERR_NOMEM 5        /* device or host allocation failed */

/* precision forms */
#define JH_PREC_FP64 0        /* FP64 everywhere; exchange sums in CSR order   */
#define JH_PREC_FP32_FIELD 1  /* FP32 spins in the exchange matvec (and FP32
                                 per-entry values in the BLOCK format), tensor
                                 tables exact, products accumulated in FP64;
                                 FP32 dipole transforms; FP64 Heun state,
                                 renormalisation, noise and reductions         */

/* integrators (the solver module that selects them) */
#define JH_INTEGRATOR_HEUN 0  /* llg-heun-gpu: predictor + corrector, two field
                                 evaluations per step (the default)            */
#define JH_INTEGRATOR_RK4 1   /* llg-rk4-gpu: classical RK4, four field
                                 evaluations per step, one noise draw per step,
                                 renormalised once at the end of the step      */

/* exchange storage formats on the device */
#define JH_XFMT_AUTO 0        /* PATTERN, else TYPED, else BLOCK, whichever fits  */
#define JH_XFMT_TYPED 1       /* per neighbour: column | tensor id in 32 bits    */
#define JH_XFMT_BLOCK 2       /* per neighbour: column + explicit 3x3 tensor     */
#define JH_XFMT_PATTERN 3     /* per spin: offset of its row pattern, a shared
                                 sequence of (column - row, tensor id) pairs     */
#define JH_XFMT_PATTERN_GLOBAL 4 /* PATTERN with the pattern table read from
                                 global memory (what PATTERN falls back to when
                                 the table does not fit in LDS)                  */
#define JH_XFMT_STENCIL 5     /* PATTERN, and the Heun step kernels stage each
                                 workgroup's neighbourhood in LDS: a crystal
                                 given by jh_set_lattice whose rows all
                                 translate one stencil per basis site (taken
                                 when asked for or chosen by jh_autotune)        */
#define JH_XFMT_LINE 6        /* PATTERN, and the Heun step kernels hold one
                                 whole z line per wavefront, loading each
                                 neighbour line once and reading neighbours
                                 along z by lane rotation: a periodic crystal
                                 given by jh_set_lattice, Lz = 64 or 128, at
                                 most two basis sites (taken when asked for or
                                 chosen by jh_autotune)                          */

/* term kinds reported by jh_term_kind */
#define JH_TERM_EXCHANGE 0
#define JH_TERM_ZEEMAN 1
#define JH_TERM_UNIAXIAL 2
#define JH_TERM_DIPOLE 3

typedef struct jh_ctx jh_ctx;

/* Library version (major*10000 + minor*100 + patch). */
int jh_version(void);

/* Context on HIP device `device` (replaces Jams::init_device_handles,
 * common.cc:63-80).  The context owns one HIP stream. */
int jh_create(int device, jh_ctx **out);
int jh_destroy(jh_ctx *ctx);
const char *jh_last_error(const jh_ctx *ctx);

/* Per-spin mus (meV/T), gyro (rad/(ps T), already divided by 1+alpha^2 when
 * solver.gilbert_prefactor is set) and alpha: globals::mus/gyro/alpha
 * (core/globals.h:24-37, core/lattice.cc:696-713).  Must come first; calling
 * it again resets every term. */
int jh_set_sites(jh_ctx *ctx, int64_t n, const double *mus, const double *gyro,
                 const double *alpha);

/* globals::s, N x 3.  jh_get_spins synchronises. */
/* Lattice shape of the sites (after jh_set_sites): spin index
 * ((x*size[1] + y)*size[2] + z)*nbasis + b, periodic[a] != 0 on periodic
 * axes (Lattice::generate_supercell, core/lattice.cc:577-749, :622-649;
 * boundaries :987-1007).  Optional: it lets a crystal's exchange run as a
 * stencil over LDS tiles (JH_XFMT_STENCIL) or over whole z lines held in
 * registers (JH_XFMT_LINE). */
int jh_set_lattice(jh_ctx *ctx, const int32_t size[3], const int32_t periodic[3], int nbasis);
int jh_set_spins(jh_ctx *ctx, const double *s);
int jh_get_spins(jh_ctx *ctx, double *s);

/* Time step in ps (CUDAHeunLLGSolver::initialize, cuda_llg_heun.cu:24). */
int jh_set_time_step(jh_ctx *ctx, double dt_ps);

/* ---- Hamiltonian terms, added in config order (the field sum
 *      h = f0 + f1 + ... keeps this order, core/solver.cc:43-57) ---------- */

/* Exchange as the reference's scalar 3N x 3N CSR (SparseInteractionHamiltonian
 * interaction_matrix_, hamiltonian/sparse_interaction.cc:102-138; row 3i+m is
 * h_i,m).  Row pointers are int64 so nnz may exceed 2^31. */
int jh_add_exchange_csr(jh_ctx *ctx, int64_t nrows, int64_t nnz, const int64_t *rowptr,
                        const int32_t *col, const double *val, int *term_id);

/* Exchange as the reference's neighbour list with a table of distinct tensors
 * (jams::InteractionList<Mat3,2>, containers/interaction_list.h:15-92, after
 * the insertion scaling of hamiltonian/exchange.cc:162-169): pairs sorted by
 * (i, j), tensor of pair p = values9[9*value_index[p] .. +9] (row-major).
 * Equivalent to the CSR built from the same list; used where the scalar CSR
 * would not fit in memory. */
int jh_add_exchange_list(jh_ctx *ctx, int64_t npairs, const int32_t *pi, const int32_t *pj,
                         const int32_t *value_index, int64_t nvalues, const double *values9,
                         int *term_id);

/* Zeeman / applied field: per-spin dc field already multiplied by mus (meV,
 * N x 3), optional ac field (N x 3, meV) and angular frequency (N, rad/ps):
 * h = dc + ac cos(omega t)  (hamiltonian/zeeman.cc:121-132,
 * applied_field.cc:146-148).  ac and omega may be NULL. */
int jh_add_zeeman(jh_ctx *ctx, const double *dc, const double *ac, const double *omega,
                  int *term_id);

/* Uniaxial anisotropy: per-spin magnitude K (meV), per-spin unit axis (N x 3),
 * power p (2, 4 or 6 for K1..K3): h = K p (s.e)^(p-1) e
 * (hamiltonian/uniaxial_anisotropy.cc:155-163). */
int jh_add_uniaxial(jh_ctx *ctx, const double *K, const double *axis, int power, int *term_id);

/* Dipole field by zero-padded FFT convolution (hamiltonian/dipole_fft.cc,
 * cuda_dipole_fft.cu).  Lattice of size[3] cells x nbasis sites, site index
 * ((i*Ly + j)*Lz + k)*nbasis + m; cell: unit-cell matrix (row-major, columns
 * a1 a2 a3, lattice constants); basis_frac: nbasis x 3; a_m: lattice
 * parameter in metres; r_cutoff in lattice constants.  Open dimensions are
 * padded to twice their size and the real-space kernel is two-sided. */
int jh_add_dipole_fft(jh_ctx *ctx, const int32_t size[3], const int32_t periodic[3], int nbasis,
                      const double cell[9], const double *basis_frac, double a_m, double r_cutoff,
                      int *term_id);

int jh_num_terms(const jh_ctx *ctx);
int jh_term_kind(const jh_ctx *ctx, int term_id);

/* ---- thermostat (classical white noise, langevin-white-gpu) ------------ */
/* sigma_i = sqrt(2 kB alpha_i / (mus_i gyro_i dt denom)), denom = 1+alpha^2 if
 * gilbert_prefactor; noise w = (xi sigma_i) sqrt(T) with xi from the
 * counter-based Philox4x32-10 stream keyed by seed, counter (spin, step)
 * (thermostats/cuda_thermostat_classical.cc:33-56, solvers/cpu_llg_heun.cc:
 * 36-63).  Noise is generated inside the step kernels; T = 0 disables it. */
int jh_set_thermostat(jh_ctx *ctx, double temperature, uint64_t seed, int gilbert_prefactor);

/* Inject already-scaled noise w (N x 3, Tesla) used by every following step
 * until jh_set_thermostat is called again (parity with an external stream). */
int jh_set_noise(jh_ctx *ctx, const double *w);

/* The generated, scaled noise of iteration `step` (N x 3). */
int jh_get_noise(jh_ctx *ctx, int64_t step, double *w);

/* ---- engine options ----------------------------------------------------- */
int jh_set_precision(jh_ctx *ctx, int mode);
int jh_set_exchange_format(jh_ctx *ctx, int fmt);
/* Solver::create's choice between CudaHeunLLGSolver and CUDALLGRK4Solver
 * (core/solver.cc:61-73; solvers/cuda_llg_rk4.cu:17-35 over
 * solvers/cuda_rk4_base.cu:57-108).  JH_INTEGRATOR_HEUN by default. */
int jh_set_integrator(jh_ctx *ctx, int integrator);

/* ---- stepping (Solver::run, one call = nsteps steps of the integrator) --- */
int jh_step(jh_ctx *ctx, int64_t nsteps);
int jh_get_iteration(const jh_ctx *ctx, int64_t *iteration, double *time_ps);
/* Resume: iteration counter (the thermostat stream position) and time. */
int jh_set_iteration(jh_ctx *ctx, int64_t iteration);

/* ---- fields, energies, reductions --------------------------------------- */
/* Hamiltonian::calculate_fields of one term at time t into out (N x 3). */
int jh_compute_term_field(jh_ctx *ctx, int term_id, double t, double *out);
/* Solver::compute_fields: sum of all terms in order (N x 3). */
int jh_compute_fields(jh_ctx *ctx, double t, double *out);
/* Hamiltonian::calculate_total_energy of one term (meV). */
int jh_term_energy(jh_ctx *ctx, int term_id, double t, double *energy);
/* sum_i mus_i s_i (3) and sum_i mus_i (1): the magnetisation monitor's
 * reductions (helpers/spinops.cc:42-51, monitors/magnetisation.cc:77-104). */
int jh_reduce_magnetisation(jh_ctx *ctx, double out[4]);
/* jh_reduce_magnetisation without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (tight serving or
 * monitoring loops, harnesses that forbid hipMalloc).  The same reduction
 * kernels and result, with their partial sums in a static device array of the
 * library (2 MiB + 32 bytes per GPU, kept until the process ends) instead of
 * the context's buffers; at most 16,777,216 spins (JH_ERR_INVALID beyond).
 * The context must be laid out already (one jh_step, jh_compute_fields or
 * jh_reduce_magnetisation since the last set-up change; JH_ERR_STATE
 * otherwise), so that nothing is allocated.  All calls on one GPU share the
 * array: calls that overlap on one GPU (contexts on several host threads or
 * streams) race. */
int jh_reduce_magnetisation_noalloc(jh_ctx *ctx, double out[4]);

/* ---- batching form: independent replicas in one context ------------------
 * No counterpart in the reference, which runs one system per process (the
 * north-star batching form).  The context's N spins hold nrep independent
 * systems, replica r owning spins [offsets[r], offsets[r+1]); its sites,
 * exchange (pairs may not cross replicas), Zeeman and anisotropy terms are
 * given for all N spins as usual.  Each replica has its own temperature and
 * noise seed: spin offsets[r] + i draws the thermostat stream (seeds[r],
 * step, i), so a replica evolves bit for bit as a context holding it alone.
 * One launch per Heun phase covers all replicas.  nrep = 0 removes them. */
int jh_set_replicas(jh_ctx *ctx, int nrep, const int64_t *offsets, const double *temperature,
                    const uint64_t *seeds);
/* Per replica: sum_i mus_i s_i (3) and sum_i mus_i (1), out is nrep x 4. */
int jh_reduce_magnetisation_replicas(jh_ctx *ctx, double *out);

/* ---- launch shape -------------------------------------------------------- */
/* Run jh_step as replays of one captured hipGraph of `steps` Heun steps
 * (whatever is left over runs as plain launches); 0 disables graphs. */
int jh_set_graph_steps(jh_ctx *ctx, int steps);
/* Threads per workgroup of the step kernels: 64 .. 512, a multiple of 64;
 * 0 (the default) picks 256 or 512 by exchange format. */
int jh_set_block_size(jh_ctx *ctx, int threads);
/* How jh_step launches a Heun step.  JH_STEP_LAUNCH_TWO (the default): a
 * predictor and a corrector launch, as the reference's step launches
 * kernelA and kernelB (solvers/cuda_llg_heun.cu:84-115).
 * JH_STEP_LAUNCH_PERSISTENT: one launch per step whose workgroups, as many as
 * are resident at once, take the predictor over every spin (every replica),
 * meet at a grid barrier and take the corrector; the same results bit for
 * bit.  It needs the Heun integrator, no dipole term and a gathering exchange
 * format (PATTERN, TYPED or BLOCK; AUTO then does not take LINE); the next
 * jh_step reports a context that does not fit.  A barrier that does not
 * complete within seconds is given up and reported by the next
 * jh_synchronize or jh_get_spins; two contexts' persistent launches running
 * on one GPU at the same time can starve each other that way.  No
 * counterpart in the reference: this is the batching form's one launch per
 * step. */
#define JH_STEP_LAUNCH_TWO 0
#define JH_STEP_LAUNCH_PERSISTENT 1
int jh_set_step_launch(jh_ctx *ctx, int mode);
/* Time `steps` steps of every exchange format that fits (PATTERN, TYPED,
 * BLOCK, and on a lattice STENCIL and LINE) at 128, 256 and 512 threads per
 * workgroup (STENCIL also 1024, LINE its one shape), each gathering format
 * also as one persistent launch per step, from the current state; keep the
 * fastest, restore the state; report (may be NULL) gets one line per
 * candidate and a last "chosen ..." line. */
int jh_autotune(jh_ctx *ctx, int steps, char *report, int64_t len);

/* ---- measurement -------------------------------------------------------- */
/* Run nsteps and return the HIP-event time on the context stream (ms). */
int jh_time_steps(jh_ctx *ctx, int64_t nsteps, double *ms);
/* Average duration of one launch of the step's dominant kernel (phase 0 =
 * predictor, 1 = corrector), timed with HIP events over `reps` launches
 * (the two-launch form's phase kernels, whatever jh_set_step_launch says). */
int jh_time_phase(jh_ctx *ctx, int phase, int reps, double *ms_per_launch);
/* Algorithmic bytes per launch of the phase kernel in the format actually
 * used, and the same quantity for the reference's FP64 scalar CSR. */
int jh_phase_bytes(jh_ctx *ctx, int phase, double *fmt_bytes, double *csr_bytes);
/* Average time (ms) of one evaluation of term `term`'s field on the context
 * stream (no host copy), and its algorithmic HBM bytes. */
int jh_time_term(jh_ctx *ctx, int term, int reps, double *ms_per_eval);
int jh_term_bytes(jh_ctx *ctx, int term, double *bytes);
/* Text summary of the device formats chosen (for logs). */
int jh_describe(jh_ctx *ctx, char *buf, int64_t buflen);
int jh_synchronize(jh_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
