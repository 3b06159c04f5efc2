/*
 * oracle.h -- CPU restatement of the JAMS atomistic Heun-LLG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded only by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py.  The product (jams_amd/) never calls it and must fail
 * loudly when its HIP library is missing.
 *
 * Every function restates a piece of stonerlab/jams (paths relative to
 * src/jams/ of the reference) and cites the file:line it follows.  How each
 * function is pinned (golden vectors, closed forms, brute force) is listed
 * in the header comment of oracle.c; functions without such a pin say
 * "parity unpinned" there.
 *
 * Layout conventions follow the reference: spins, fields and noise are
 * row-major N x 3 doubles (AoS, core/lattice.cc:688-694); the exchange matrix
 * is a scalar 3N x 3N CSR whose row 3i+m holds h_i,m (sparse_interaction.cc).
 * Row pointers are int64 here (the reference's int32 overflows at 2^31 nnz).
 */
#ifndef JAMS_ORACLE_H
#define JAMS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- physical constants, helpers/consts.h:18-38 ------------------------ */
#define OR_PI 3.14159265358979323846264338327950288
#define OR_BOHR_MAGNETON 0.0578838181           /* meV / T */
#define OR_HBAR 0.6582119569                    /* meV ps */
#define OR_G_ELECTRON 2.0023193043625
#define OR_KB 0.0861733326                      /* meV / K */
#define OR_JOULE2MEV 6.24150907e21

double or_gyromagnetic_ratio(void);   /* kElectronGFactor*kBohrMagnetonIU/kHBarIU */
double or_vacuum_permeability(void);  /* 4*pi*1e-7 / kJoule2meV */

/* ---- Philox4x32-10 counter RNG (Salmon et al., SC'11) ------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* uniform in (0,1): (2k+1) * 2^-53 with k the top 52 bits of (hi:lo) */
double or_u01(uint32_t hi, uint32_t lo);
/* log(x) and (cos, sin)(2 pi u) as used by the noise stream: correctly
 * rounded IEEE operations (fma included) in a fixed order, reproduced bit for
 * bit by the device code. */
double or_noise_log(double x);
void or_noise_sincos2pi(double u, double *c, double *s);
/* Three standard normals for (seed, step, spin): the thermostat stream. */
void or_normals3(uint64_t seed, uint64_t step, uint64_t spin, double out[3]);
/* Fill xi (N x 3) with the stream of step `step`. */
void or_normals_fill(uint64_t seed, uint64_t step, int64_t n, double *xi);
/* Uniform doubles in [0,1) from the "setup" stream (key seed, stream id). */
double or_setup_uniform(uint64_t seed, uint64_t stream, uint64_t index);
/* Marsaglia point on S^2 (helpers/random.h:12-26) driven by the setup stream;
 * writes n spins into s (N x 3).  */
void or_random_spins(uint64_t seed, int64_t n, double *s);

/* ---- lattice + neighbour list (core/lattice.cc, core/interactions.cc) --- */
typedef struct or_lattice_s {
  double cell[9];        /* unit-cell matrix M, row-major; columns are a1,a2,a3 */
  int nbasis;
  const double *basis_frac;   /* nbasis x 3 fractional positions */
  const int *basis_material;  /* nbasis material ids */
  int size[3];
  int periodic[3];
} or_lattice_t;

typedef struct {
  int type_i, type_j;     /* material ids (JAMS-format template) */
  double r[3];            /* interaction vector, cartesian, lattice constants */
  double J[9];            /* 3x3 tensor, row-major (input energy units) */
} or_template_t;

/* Site index = ((i*Ly + j)*Lz + k)*nb + m  (core/lattice.cc:622-649) */
int64_t or_site_index(const or_lattice_t *lat, int i, int j, int k, int m);

/* Reference neighbour list: post_process_interactions (interactions.cc:292-347)
 * + neighbour_list_from_interactions (interactions.cc:350-395).  Output sorted
 * by (i, j) with the tensor each pair carries.  Returns the number of pairs,
 * -1 on a duplicate pair (the reference throws), -2 if max_out is too small,
 * -3 for an interaction vector that is not a lattice translation. */
int64_t or_neighbour_list(const or_lattice_t *lat, const or_template_t *t, int nt,
                          double tol, double radius_cutoff,
                          int32_t *out_i, int32_t *out_j, double *out_J, int64_t max_out);

/* Independent pin: brute-force geometric enumeration over all ordered site
 * pairs (i, j) with minimum-image displacements matched against the template.
 * Same output contract as or_neighbour_list (sorted by (i, j)). */
int64_t or_neighbour_bruteforce(const or_lattice_t *lat, const or_template_t *t, int nt,
                                double tol, int32_t *out_i, int32_t *out_j, double *out_J,
                                int64_t max_out);

/* Pairs for the distance-based exchange modules (exchange-neartree,
 * exchange-functional): every ordered (i, image of j) of the supercell with
 * distance d = |r_i - r_j'| in the annulus of InteractionNearTree
 * (containers/neartree.h:359-391): !(d >> outer) && (d >> inner) with the
 * relative test of helpers/maths.h:31-34 and tolerance eps.  Images of j:
 * r_j + ((h A1 + k A2) + l A3) for all |h|, |k|, |l| up to ceil(outer / height)
 * + 1 along periodic directions (brute force over all j).  Positions are
 * M (frac + cell).  Self pairs (j = i) are included; the caller drops them.
 * Output sorted by (i, j, d); returns the count, or -2 if max_out is too
 * small. */
int64_t or_annulus_pairs(const or_lattice_t *lat, double inner, double outer, double eps,
                         int32_t *out_i, int32_t *out_j, double *out_rij, double *out_d,
                         int64_t max_out);

/* ---- sparse matrix builder (containers/sparse_matrix_builder.h) -------- */
/* Exchange insertion (exchange.cc:162-169 -> sparse_interaction.cc:25-33) of
 * the neighbour list into COO, then sort (:109-143), merge (:146-164, with
 * its skip-after-erase quirk) and build_csr (:187-248, with its empty-leading-
 * row quirk).  scale = interaction_prefactor * unit_conversion;
 * cutoff = energy_cutoff * unit_conversion.  rowptr has 3N+1 entries; col/val
 * must hold max_nnz.  Returns nnz or -1 if max_nnz is too small. */
int64_t or_exchange_csr(int64_t n, int64_t npairs, const int32_t *pi, const int32_t *pj,
                        const double *pJ, double scale, double cutoff,
                        int64_t *rowptr, int32_t *col, double *val, int64_t max_nnz);

/* Generic COO -> CSR with the reference semantics (for quirk tests). */
int64_t or_coo_to_csr(int64_t nrows, int64_t nnz, int64_t *row, int64_t *col, double *val,
                      int64_t *rowptr, int32_t *out_col, double *out_val);

/* Builder::is_symmetric (sparse_matrix_builder.h:320-361), exact compare. */
int or_csr_is_symmetric(int64_t nrows, const int64_t *rowptr, const int32_t *col,
                        const double *val);

/* y = A x, sequential row sums from 0.0 (interface/sparse_blas.h:15-27, 59-81) */
void or_csrmv(int64_t nrows, const int64_t *rowptr, const int32_t *col, const double *val,
              const double *x, double *y, int nthreads);

/* ---- Hamiltonian terms and the Heun step -------------------------------- */
enum { OR_TERM_EXCHANGE = 0, OR_TERM_ZEEMAN = 1, OR_TERM_UNIAXIAL = 2, OR_TERM_FIELD = 3,
       OR_TERM_DIPOLE = 4 };

typedef struct {
  int64_t n;
  const double *mus, *gyro, *alpha;     /* per spin */
  /* exchange CSR */
  const int64_t *rowptr; const int32_t *col; const double *val;
  /* Zeeman: per-spin dc field (already mus*B, meV) + optional ac part */
  const double *zee_dc; const double *zee_ac; const double *zee_omega;
  /* uniaxial: per-spin magnitude (meV), per-spin unit axis, power p */
  const double *ani_K; const double *ani_axis; int ani_power;
  /* precomputed external field term (e.g. a dipole field), N x 3, meV */
  const double *ext_field;
  int nterms; int kind[8];              /* terms in Hamiltonian config order */
  int nthreads;
  /* dipole term, evaluated by brute force (or_dipole_bruteforce) */
  const struct or_lattice_s *dip_lat; const double *dip_pos; double dip_rcut; double dip_a;
} or_system_t;

/* Field of one term at time t (meV, N x 3). */
void or_term_field(const or_system_t *sys, int term, const double *s, double t, double *h);
/* Solver::compute_fields (core/solver.cc:43-57): h = f0; h += f_k in order. */
void or_compute_fields(const or_system_t *sys, const double *s, double t, double *h,
                       double *scratch);
/* One HeunLLGSolver::run step (solvers/cpu_llg_heun.cc:45-148).  w (N x 3,
 * Tesla) is the scaled noise or NULL at T = 0.  work needs 4*3N doubles. */
void or_heun_step(const or_system_t *sys, double *s, const double *w, double dt, double t0,
                  double *work);

/* One CUDALLGRK4Solver::run step (solvers/cuda_rk4_base.cu:57-108,
 * cuda_llg_rk4_kernel.cuh:8-58, cuda/cuda_spin_ops.cu:4-17); w as above, the
 * same draw for all four stages.  work needs 5*3N doubles. */
void or_rk4_step(const or_system_t *sys, double *s, const double *w, double dt, double t0,
                 double *work);

/* sigma_i (solvers/cpu_llg_heun.cc:36-43) */
void or_thermostat_sigma(int64_t n, const double *mus, const double *gyro,
                         const double *alpha, double dt, int gilbert, double *sigma);
/* w_ij = (xi_ij * sigma_i) * sqrt(T)  (solvers/cpu_llg_heun.cc:55-63) */
void or_scale_noise(int64_t n, const double *xi, const double *sigma, double T, double *w);

/* ---- reductions --------------------------------------------------------- */
/* sum_spins_moments (helpers/spinops.cc:42-51), plain sum. */
void or_sum_moments(int64_t n, const double *s, const double *mus, double out[3]);
/* Kahan scalar sum (helpers/array_ops.h:59-72) */
double or_kahan_sum(int64_t n, const double *x);
/* Kahan N x 3 sum (helpers/array_ops.h:10-30) */
void or_kahan_sum3(int64_t n, const double *x, double out[3]);
/* Per-term total energy (sparse_interaction.cc:86-100; zeeman/uniaxial). */
double or_term_energy(const or_system_t *sys, int term, const double *s, double t);

/* ---- dipole (hamiltonian/dipole_bruteforce.cc, dipole_fft.cc) ---------- */
/* Brute-force dipole field h_i = mus_i sum_j mus_j W(r_ij) s_j (meV), with
 * minimum image over periodic dimensions and r_cutoff (lattice constants).
 * positions: N x 3 cartesian in lattice constants.  a_m: lattice parameter. */
void or_dipole_bruteforce(const or_lattice_t *lat, int64_t n, const double *pos,
                          const double *mus, const double *s, double r_cutoff,
                          double a_m, double *h, int nthreads);
/* The same sum at the listed sites only (h: nsites x 3), for lattices whose
 * full O(N^2) sum is out of reach. */
void or_dipole_bruteforce_sites(const or_lattice_t *lat, int64_t n, const double *pos,
                                const double *mus, const double *s, double r_cutoff, double a_m,
                                int64_t nsites, const int64_t *sites, double *h, int nthreads);

/* The FFT dipole's convolution as a direct sum over cell offsets
 * (dipole_fft.cc:195-350).  one_sided_open = 1 keeps the reference's offsets
 * n in [0, L) on open dimensions; 0 takes -L < n < L (= brute force). */
void or_dipole_fft_direct(const or_lattice_t *lat, const double *mus, const double *s,
                          double r_cutoff, double a_m, int one_sided_open, double *h,
                          int nthreads);

#ifdef __cplusplus
}
#endif
#endif
