/*
 * oracle.c -- CPU restatement of the JAMS Heun-LLG hot path (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product.
 *
 * Compiled with -std=c11 -ffp-contract=off (no FMA contraction), so every
 * +, -, *, / and sqrt is one correctly rounded IEEE-754 double operation in
 * the order written, as in the reference's C++17 ISO build.
 *
 * How each function is pinned (tests/test_oracle.py, tests/test_host.py):
 *   or_philox4x32_10     Random123 known-answer vectors (published KAT).
 *   or_normals3 / fill   integer stream pinned by the KAT; normals checked by
 *                        moments and a Kolmogorov-Smirnov test.
 *   or_random_spins      uniform-on-S^2 distribution tests (z and phi uniform).
 *   or_site_index        closed form of core/lattice.cc:622-649.
 *   or_neighbour_list    equality with or_neighbour_bruteforce (independent
 *                        geometric enumeration) and closed-form counts
 *                        (SC 6/site, bcc 8+6, bcc 5-shell 58, open film pairs).
 *   or_exchange_csr      nnz closed forms, structure rule, dense brute force,
 *                        quirk cases built by hand.
 *   or_csrmv             dense matvec on tiny lattices; uniform-state field
 *                        (sum_j J_ij) s; plane-wave eigenvector J(k).
 *   or_heun_step         macrospin closed form (precession at gamma*B, damping
 *                        tan(theta/2) = tan(theta0/2) exp(-alpha*gamma*B*t)),
 *                        |s| = 1 invariant, energy non-increase at T = 0.
 *   or_annulus_pairs     shell counts in closed form (sc 6/12/8/6, bcc 8/6/12,
 *                        fcc 12/6) on periodic lattices, open-boundary counts;
 *                        the exchange-neartree list built on it equals the
 *                        template enumerator or_neighbour_list on the same
 *                        shells.
 *   or_rk4_step          the same macrospin closed form with fourth-order
 *                        convergence (error ratio ~16 per halving of dt;
 *                        a wrong stage weight or time drops the order);
 *                        |s| = 1; agreement with or_heun_step as dt -> 0.
 *   or_thermostat_sigma  closed form; Langevin-function test of <s_z>.
 *   or_kahan_*           exact sums on integer-valued data.
 *   or_term_energy       FM closed form -1/2 N sum_j J_ij; Zeeman/uniaxial
 *                        closed forms.
 *   or_dipole_bruteforce reference known answers (test/hamiltonian/
 *                        test_dipole.h:256-404) on periodic chains; the
 *                        OR_TERM_DIPOLE term of or_term_field evaluates it.
 *   or_dipole_fft_direct (oracle_dipole_fft.c) equals or_dipole_bruteforce
 *                        to 1e-12 in its two-sided form on open, mixed and
 *                        periodic lattices.
 *   or_noise_log /       libm and long-double references over 1e5 arguments.
 *   or_noise_sincos2pi
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

double or_gyromagnetic_ratio(void) { return OR_G_ELECTRON * OR_BOHR_MAGNETON / OR_HBAR; }
double or_vacuum_permeability(void) { return 4 * OR_PI * 1E-7 / OR_JOULE2MEV; }

/* ======================================================================= */
/* Philox4x32-10.  Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11: round function with multipliers 0xD2511F53 and
 * 0xCD9E8D57, Weyl key schedule 0x9E3779B9 / 0xBB67AE85, 10 rounds.        */
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    philox_round(c, k);
  }
  memcpy(out, c, sizeof c);
}

double or_u01(uint32_t hi, uint32_t lo) {
  const uint64_t k = ((uint64_t)hi << 20) | (uint64_t)(lo >> 12);   /* 52 bits */
  return (double)(2 * k + 1) * 0x1.0p-53;                           /* exact */
}

/* Natural log of a positive normal double with IEEE operations in a fixed
 * order (fma() is the correctly rounded fused multiply-add of C99), so the
 * device engine produces the same bits.  x = m 2^e with m in [sqrt(1/2),
 * sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1), |s| < 0.1716, summed as the
 * odd series 2 s (1 + s^2/3 + ... + s^20/21) by Horner steps. */
double or_noise_log(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
  double m;
  memcpy(&m, &b, 8);
  if (m > 1.4142135623730951) { m = m * 0.5; e = e + 1; }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, z, 1.0 / 19.0); p = fma(p, z, 1.0 / 17.0); p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0); p = fma(p, z, 1.0 / 11.0); p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);  p = fma(p, z, 1.0 / 5.0);  p = fma(p, z, 1.0 / 3.0);
  const double s2 = 2.0 * s;
  const double lm = fma(s2, z * p, s2);
  return fma((double)e, 0.6931471805599453, lm);
}

/* cos(2 pi u) and sin(2 pi u) for u in [0, 1): quadrant reduction u*4 = q + r
 * (exact), x = r pi/2 in [-pi/4, pi/4], Taylor series to x^18 / x^19 by fma
 * Horner steps, then an exact quadrant rotation. */
void or_noise_sincos2pi(double u, double *c, double *s) {
  const double v = u * 4.0;
  const double q = nearbyint(v);
  const double r = v - q;
  const double x = r * 1.5707963267948966;
  const double z = x * x;
  double ps = -1.0 / 121645100408832000.0;               /* -1/19! */
  ps = fma(ps, z, 1.0 / 355687428096000.0);               /*  1/17! */
  ps = fma(ps, z, -1.0 / 1307674368000.0);                /* -1/15! */
  ps = fma(ps, z, 1.0 / 6227020800.0);                    /*  1/13! */
  ps = fma(ps, z, -1.0 / 39916800.0);                     /* -1/11! */
  ps = fma(ps, z, 1.0 / 362880.0);                        /*  1/9!  */
  ps = fma(ps, z, -1.0 / 5040.0);                         /* -1/7!  */
  ps = fma(ps, z, 1.0 / 120.0);                           /*  1/5!  */
  ps = fma(ps, z, -1.0 / 6.0);                            /* -1/3!  */
  const double sn = fma(x, z * ps, x);
  double pc = 1.0 / 6402373705728000.0;                   /*  1/18! */
  pc = fma(pc, z, -1.0 / 20922789888000.0);               /* -1/16! */
  pc = fma(pc, z, 1.0 / 87178291200.0);                   /*  1/14! */
  pc = fma(pc, z, -1.0 / 479001600.0);                    /* -1/12! */
  pc = fma(pc, z, 1.0 / 3628800.0);                       /*  1/10! */
  pc = fma(pc, z, -1.0 / 40320.0);                        /* -1/8!  */
  pc = fma(pc, z, 1.0 / 720.0);                           /*  1/6!  */
  pc = fma(pc, z, -1.0 / 24.0);                           /* -1/4!  */
  pc = fma(pc, z, 0.5);                                   /*  1/2!  */
  const double cs = fma(-z, pc, 1.0);
  const int quad = ((int)q) & 3;
  switch (quad) {
    case 0: *c = cs;  *s = sn;  break;
    case 1: *c = -sn; *s = cs;  break;
    case 2: *c = -cs; *s = -sn; break;
    default: *c = sn; *s = -cs; break;
  }
}

/* Box-Muller pair from two uniforms (z0 = r cos t, z1 = r sin t, t = 2 pi u2). */
static inline void box_muller(double u1, double u2, double *z0, double *z1) {
  const double r = sqrt(-2.0 * or_noise_log(u1));
  double c, s;
  or_noise_sincos2pi(u2, &c, &s);
  *z0 = r * c;
  *z1 = r * s;
}

void or_normals3(uint64_t seed, uint64_t step, uint64_t spin, double out[3]) {
  /* counter = (spin lo, call | spin hi << 1, step lo, step hi); key = seed */
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t c[4], r[4];
  c[0] = (uint32_t)spin; c[2] = (uint32_t)step; c[3] = (uint32_t)(step >> 32);
  c[1] = 0u | ((uint32_t)(spin >> 32) << 1);
  or_philox4x32_10(c, key, r);
  double z0, z1, z2, z3;
  box_muller(or_u01(r[0], r[1]), or_u01(r[2], r[3]), &z0, &z1);
  c[1] = 1u | ((uint32_t)(spin >> 32) << 1);
  or_philox4x32_10(c, key, r);
  box_muller(or_u01(r[0], r[1]), or_u01(r[2], r[3]), &z2, &z3);
  out[0] = z0; out[1] = z1; out[2] = z2;
}

void or_normals_fill(uint64_t seed, uint64_t step, int64_t n, double *xi) {
  for (int64_t i = 0; i < n; ++i) or_normals3(seed, step, (uint64_t)i, xi + 3 * i);
}

double or_setup_uniform(uint64_t seed, uint64_t stream, uint64_t index) {
  /* set-up stream: counter = (index lo, index hi, stream lo, 0xFFFFFFFF) */
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), (uint32_t)stream, 0xFFFFFFFFu};
  uint32_t r[4];
  or_philox4x32_10(c, key, r);
  const uint64_t k = ((uint64_t)r[0] << 21) | (uint64_t)(r[1] >> 11);  /* 53 bits */
  return (double)k * 0x1.0p-53;
}

void or_random_spins(uint64_t seed, int64_t n, double *s) {
  /* helpers/random.h:12-26 (Marsaglia 1972) with the set-up stream as the
   * uniform source; draws are consumed sequentially across all spins. */
  uint64_t idx = 0;
  for (int64_t i = 0; i < n; ++i) {
    double v1, v2, q;
    do {
      v1 = fma(2.0, or_setup_uniform(seed, 0, idx++), -1.0);
      v2 = fma(2.0, or_setup_uniform(seed, 0, idx++), -1.0);
      q = (v1 * v1) + (v2 * v2);
    } while (q > 1.0);
    const double ss = 2.0 * sqrt(1.0 - q);
    s[3 * i + 0] = v1 * ss;
    s[3 * i + 1] = v2 * ss;
    s[3 * i + 2] = fma(-2.0, q, 1.0);
  }
}

/* ======================================================================= */
/* Lattice geometry                                                          */

int64_t or_site_index(const or_lattice_t *lat, int i, int j, int k, int m) {
  return (((int64_t)i * lat->size[1] + j) * lat->size[2] + k) * lat->nbasis + m;
}

static void mat3_mul_vec(const double M[9], const double v[3], double out[3]) {
  for (int r = 0; r < 3; ++r) out[r] = M[3 * r + 0] * v[0] + M[3 * r + 1] * v[1] + M[3 * r + 2] * v[2];
}

static void mat3_inverse(const double M[9], double inv[9]) {
  const double a = M[0], b = M[1], c = M[2], d = M[3], e = M[4], f = M[5], g = M[6], h = M[7], i = M[8];
  const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  inv[0] = (e * i - f * h) / det; inv[1] = (c * h - b * i) / det; inv[2] = (b * f - c * e) / det;
  inv[3] = (f * g - d * i) / det; inv[4] = (a * i - c * g) / det; inv[5] = (c * d - a * f) / det;
  inv[6] = (d * h - e * g) / det; inv[7] = (b * g - a * h) / det; inv[8] = (a * e - b * d) / det;
}

/* lattice_translation_vector (core/interactions.cc:58-75) */
static void lattice_translation(const double r[3], double tol, double T[3]) {
  for (int n = 0; n < 3; ++n) {
    const double ni = nearbyint(r[n]);
    const double fl = floor(r[n]);
    T[n] = (fabs(r[n] - ni) <= tol) ? ni : fl;
  }
}

static int approx_equal3(const double a[3], const double b[3], double tol) {
  for (int n = 0; n < 3; ++n) if (fabs(a[n] - b[n]) > tol) return 0;
  return 1;
}

typedef struct {
  int bi, bj;            /* basis sites */
  int Tn[3];             /* lattice translation */
  int type_i, type_j;
  int src;               /* template row */
} or_tmpl_full_t;

/* approximately_equal (helpers/maths.h:14-22): absolute, then relative */
static int approx_eq(double a, double b, double eps) {
  const double d = fabs(a - b);
  if (d <= eps) return 1;
  const double mx = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
  return d <= mx * eps;
}

static int approx_equal3v(const double a[3], const double b[3], double eps) {
  for (int n = 0; n < 3; ++n) if (!approx_eq(a[n], b[n], eps)) return 0;
  return 1;
}

/* post_process_interactions (core/interactions.cc:292-347) for a JAMS-format
 * template with cartesian vectors and symops disabled:
 * complete_interaction_unitcell_positions (:98-123, partner search :78-87),
 * radius predicate (:324-328), lattice translation vectors (:331-344). */
static int expand_template(const or_lattice_t *lat, const or_template_t *t, int nt, double tol,
                           double radius_cutoff, or_tmpl_full_t **out) {
  double inv[9];
  mat3_inverse(lat->cell, inv);
  or_tmpl_full_t *v = (or_tmpl_full_t *)malloc(sizeof(or_tmpl_full_t) * (size_t)(nt * lat->nbasis + 1));
  int count = 0;
  for (int q = 0; q < nt; ++q) {
    double rf[3];
    mat3_mul_vec(inv, t[q].r, rf);
    for (int bi = 0; bi < lat->nbasis; ++bi) {
      if (lat->basis_material[bi] != t[q].type_i) continue;
      double qq[3], T[3], off[3];
      for (int n = 0; n < 3; ++n) qq[n] = rf[n] + lat->basis_frac[3 * bi + n];
      lattice_translation(qq, tol, T);
      for (int n = 0; n < 3; ++n) off[n] = qq[n] - T[n];
      int bj = -1;
      for (int k = 0; k < lat->nbasis; ++k) {
        if (approx_equal3v(lat->basis_frac + 3 * k, off, tol)) { bj = k; break; }
      }
      if (bj < 0) continue;
      if (lat->basis_material[bj] != t[q].type_j) continue;
      v[count].bi = bi; v[count].bj = bj;
      v[count].type_i = t[q].type_i; v[count].type_j = t[q].type_j;
      v[count].src = q;
      count++;
    }
  }
  int w = 0;
  for (int q = 0; q < count; ++q) {
    const or_template_t *src = &t[v[q].src];
    /* radius predicate: drop definately_greater_than(|r|, r_cut, 1e-4) */
    if (radius_cutoff > 0.0) {
      const double rn = sqrt(src->r[0] * src->r[0] + src->r[1] * src->r[1] + src->r[2] * src->r[2]);
      const double mx = fabs(rn) > fabs(radius_cutoff) ? fabs(rn) : fabs(radius_cutoff);
      if ((rn - radius_cutoff) > mx * 1e-4) continue;
    }
    /* lattice translation T = ltv(r_frac + p_i - p_j) must be integral */
    double rf[3], x[3], T[3];
    mat3_mul_vec(inv, src->r, rf);
    for (int n = 0; n < 3; ++n)
      x[n] = rf[n] + lat->basis_frac[3 * v[q].bi + n] - lat->basis_frac[3 * v[q].bj + n];
    lattice_translation(x, tol, T);
    for (int n = 0; n < 3; ++n) {
      if (fabs(T[n] - x[n]) > tol) { free(v); return -3; }
      v[q].Tn[n] = (int)T[n];
    }
    v[w++] = v[q];
  }
  *out = v;
  return w;
}

typedef struct { int32_t i, j; int64_t src; } or_pair_t;

static int pair_cmp(const void *a, const void *b) {
  const or_pair_t *x = (const or_pair_t *)a, *y = (const or_pair_t *)b;
  if (x->i != y->i) return x->i < y->i ? -1 : 1;
  if (x->j != y->j) return x->j < y->j ? -1 : 1;
  return 0;
}

/* apply_boundary_conditions (core/lattice.cc:987-1007) */
static int apply_bc(const or_lattice_t *lat, int d[3]) {
  for (int n = 0; n < 3; ++n) {
    if (!lat->periodic[n] && (d[n] < 0 || d[n] >= lat->size[n])) return 0;
    d[n] = (d[n] + lat->size[n]) % lat->size[n];
  }
  return 1;
}

int64_t or_neighbour_list(const or_lattice_t *lat, const or_template_t *t, int nt, double tol,
                          double radius_cutoff, int32_t *out_i, int32_t *out_j, double *out_J,
                          int64_t max_out) {
  or_tmpl_full_t *v = NULL;
  const int nv = expand_template(lat, t, nt, tol, radius_cutoff, &v);
  if (nv < 0) return nv;
  const int64_t ncells = (int64_t)lat->size[0] * lat->size[1] * lat->size[2];
  const int64_t cap = ncells * (nv > 0 ? nv : 1);
  or_pair_t *p = (or_pair_t *)malloc(sizeof(or_pair_t) * (size_t)(cap > 0 ? cap : 1));
  int64_t np = 0;
  /* neighbour_list_from_interactions (core/interactions.cc:350-395) */
  for (int i = 0; i < lat->size[0]; ++i)
    for (int j = 0; j < lat->size[1]; ++j)
      for (int k = 0; k < lat->size[2]; ++k)
        for (int q = 0; q < nv; ++q) {
          const int m = v[q].bi;
          const int64_t local = or_site_index(lat, i, j, k, m);
          int d[3] = {i + v[q].Tn[0], j + v[q].Tn[1], k + v[q].Tn[2]};
          if (!apply_bc(lat, d)) continue;
          const int64_t nbr = or_site_index(lat, d[0], d[1], d[2], v[q].bj);
          /* material check against the site's material (no impurities here) */
          if (lat->basis_material[m] != v[q].type_i || lat->basis_material[v[q].bj] != v[q].type_j) continue;
          p[np].i = (int32_t)local; p[np].j = (int32_t)nbr; p[np].src = q;
          np++;
        }
  qsort(p, (size_t)np, sizeof(or_pair_t), pair_cmp);
  for (int64_t x = 1; x < np; ++x) {
    if (p[x].i == p[x - 1].i && p[x].j == p[x - 1].j) { free(p); free(v); return -1; }
  }
  if (np > max_out) { free(p); free(v); return -2; }
  for (int64_t x = 0; x < np; ++x) {
    out_i[x] = p[x].i; out_j[x] = p[x].j;
    memcpy(out_J + 9 * x, t[v[p[x].src].src].J, 9 * sizeof(double));
  }
  free(p); free(v);
  return np;
}

int64_t or_neighbour_bruteforce(const or_lattice_t *lat, const or_template_t *t, int nt, double tol,
                                int32_t *out_i, int32_t *out_j, double *out_J, int64_t max_out) {
  const int64_t ncell = (int64_t)lat->size[0] * lat->size[1] * lat->size[2];
  const int64_t n = ncell * lat->nbasis;
  double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  int *mat = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < lat->size[0]; ++i)
    for (int j = 0; j < lat->size[1]; ++j)
      for (int k = 0; k < lat->size[2]; ++k)
        for (int m = 0; m < lat->nbasis; ++m) {
          const int64_t s = or_site_index(lat, i, j, k, m);
          const double f[3] = {lat->basis_frac[3 * m] + i, lat->basis_frac[3 * m + 1] + j,
                               lat->basis_frac[3 * m + 2] + k};
          mat3_mul_vec(lat->cell, f, pos + 3 * s);
          mat[s] = lat->basis_material[m];
        }
  int64_t np = 0;
  for (int64_t a = 0; a < n; ++a) {
    for (int64_t b = 0; b < n; ++b) {
      /* r_ab = pos_b - pos_a minus any supercell translation along periodic
       * axes: enumerate images within +-2 supercells and test each against
       * every template vector. */
      int hit = -1, nhits = 0;
      for (int ia = -2; ia <= 2; ++ia) {
        if (!lat->periodic[0] && ia != 0) continue;
        for (int ib = -2; ib <= 2; ++ib) {
          if (!lat->periodic[1] && ib != 0) continue;
          for (int ic = -2; ic <= 2; ++ic) {
            if (!lat->periodic[2] && ic != 0) continue;
            const double sh[3] = {(double)ia * lat->size[0], (double)ib * lat->size[1], (double)ic * lat->size[2]};
            double shc[3];
            mat3_mul_vec(lat->cell, sh, shc);
            double r[3];
            for (int q = 0; q < 3; ++q) r[q] = pos[3 * b + q] + shc[q] - pos[3 * a + q];
            for (int q = 0; q < nt; ++q) {
              if (t[q].type_i != mat[a] || t[q].type_j != mat[b]) continue;
              if (approx_equal3(r, t[q].r, tol)) { hit = q; nhits++; }
            }
          }
        }
      }
      if (nhits > 1) { free(pos); free(mat); return -1; }
      if (hit >= 0) {
        if (np >= max_out) { free(pos); free(mat); return -2; }
        out_i[np] = (int32_t)a; out_j[np] = (int32_t)b;
        memcpy(out_J + 9 * np, t[hit].J, 9 * sizeof(double));
        np++;
      }
    }
  }
  free(pos); free(mat);
  return np;
}

/* ======================================================================= */
/* Sparse matrix builder semantics                                           */

typedef struct { int64_t r, c; double v; } or_coo_t;

static int coo_cmp(const void *a, const void *b) {
  const or_coo_t *x = (const or_coo_t *)a, *y = (const or_coo_t *)b;
  if (x->r != y->r) return x->r < y->r ? -1 : 1;
  if (x->c != y->c) return x->c < y->c ? -1 : 1;
  return 0;
}

int64_t or_coo_to_csr(int64_t nrows, int64_t nnz, int64_t *row, int64_t *col, double *val,
                      int64_t *rowptr, int32_t *out_col, double *out_val) {
  or_coo_t *e = (or_coo_t *)malloc(sizeof(or_coo_t) * (size_t)(nnz > 0 ? nnz : 1));
  for (int64_t m = 0; m < nnz; ++m) { e[m].r = row[m]; e[m].c = col[m]; e[m].v = val[m]; }
  /* sort by (row, col)  (sparse_matrix_builder.h:109-143).  qsort is not
   * stable, like std::sort; equal keys are only merged below. */
  qsort(e, (size_t)nnz, sizeof(or_coo_t), coo_cmp);
  /* merge (:146-164): sums adjacent duplicates; after an erase the loop index
   * still advances, so the element that slid into position m is not compared
   * with m-1 (a third consecutive duplicate survives). */
  int64_t len = nnz;
  for (int64_t m = 1; m < len; ++m) {
    if (e[m].r == e[m - 1].r && e[m].c == e[m - 1].c) {
      e[m - 1].v += e[m].v;
      memmove(e + m, e + m + 1, sizeof(or_coo_t) * (size_t)(len - m - 1));
      len--;
    }
  }
  /* build_csr (:187-248): starts from row_[1] with previous_row = 0 */
  rowptr[0] = 0;
  int64_t previous_row = 0;
  for (int64_t m = 1; m < len; ++m) {
    const int64_t current_row = e[m].r;
    if (current_row == previous_row) continue;
    for (int64_t i = previous_row + 1; i < current_row + 1; ++i) rowptr[i] = m;
    previous_row = current_row;
  }
  for (int64_t i = previous_row + 1; i < nrows + 1; ++i) rowptr[i] = len;
  for (int64_t m = 0; m < len; ++m) { out_col[m] = (int32_t)e[m].c; out_val[m] = e[m].v; }
  free(e);
  return len;
}

int64_t or_exchange_csr(int64_t n, int64_t npairs, const int32_t *pi, const int32_t *pj,
                        const double *pJ, double scale, double cutoff,
                        int64_t *rowptr, int32_t *col, double *val, int64_t max_nnz) {
  /* exchange.cc:162-169: Jij = (prefactor * conversion) * J; inserted if
   * max|Jij| > energy_cutoff * conversion; insert_interaction_tensor
   * (sparse_interaction.cc:25-33) skips exactly-zero components. */
  int64_t cap = npairs * 9;
  int64_t *r = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cap > 0 ? cap : 1));
  int64_t *c = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cap > 0 ? cap : 1));
  double *v = (double *)malloc(sizeof(double) * (size_t)(cap > 0 ? cap : 1));
  int64_t nnz = 0;
  for (int64_t q = 0; q < npairs; ++q) {
    double Jij[9];
    double mx = 0.0;
    for (int x = 0; x < 9; ++x) {
      Jij[x] = scale * pJ[9 * q + x];
      if (fabs(Jij[x]) > mx) mx = fabs(Jij[x]);
    }
    if (!(mx > cutoff)) continue;
    for (int m = 0; m < 3; ++m)
      for (int k = 0; k < 3; ++k)
        if (Jij[3 * m + k] != 0.0) {
          r[nnz] = 3 * (int64_t)pi[q] + m; c[nnz] = 3 * (int64_t)pj[q] + k; v[nnz] = Jij[3 * m + k];
          nnz++;
        }
  }
  if (nnz > max_nnz) { free(r); free(c); free(v); return -1; }
  const int64_t len = or_coo_to_csr(3 * n, nnz, r, c, v, rowptr, col, val);
  free(r); free(c); free(v);
  return len;
}

int or_csr_is_symmetric(int64_t nrows, const int64_t *rowptr, const int32_t *col, const double *val) {
  for (int64_t i = 0; i < nrows; ++i) {
    for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const int64_t j = col[p];
      /* find (j, i) */
      int64_t lo = rowptr[j], hi = rowptr[j + 1], found = -1;
      while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (col[mid] < i) lo = mid + 1; else hi = mid;
      }
      if (lo < rowptr[j + 1] && col[lo] == i) found = lo;
      if (found < 0) return 0;
      if (val[found] != val[p]) return 0;
    }
  }
  return 1;
}

void or_csrmv(int64_t nrows, const int64_t *rowptr, const int32_t *col, const double *val,
              const double *x, double *y, int nthreads) {
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t i = 0; i < nrows; ++i) {
    double sum = 0.0;
    for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) sum += x[col[p]] * val[p];
    y[i] = sum;
  }
}

/* ======================================================================= */
/* Field terms                                                               */

void or_term_field(const or_system_t *sys, int term, const double *s, double t, double *h) {
  const int64_t n = sys->n;
  const int nth = sys->nthreads > 0 ? sys->nthreads : 1;
  (void)nth;
  switch (sys->kind[term]) {
    case OR_TERM_EXCHANGE:
      or_csrmv(3 * n, sys->rowptr, sys->col, sys->val, s, h, sys->nthreads);
      break;
    case OR_TERM_ZEEMAN:
      /* zeeman.cc:121-132: field = dc; field += ac * cos(omega t) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nth)
#endif
      for (int64_t i = 0; i < n; ++i) {
        for (int j = 0; j < 3; ++j) h[3 * i + j] = sys->zee_dc[3 * i + j];
        if (sys->zee_ac) {
          for (int j = 0; j < 3; ++j) h[3 * i + j] += sys->zee_ac[3 * i + j] * cos(sys->zee_omega[i] * t);
        }
      }
      break;
    case OR_TERM_UNIAXIAL:
      /* uniaxial_anisotropy.cc:155-163 */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nth)
#endif
      for (int64_t i = 0; i < n; ++i) {
        const double *e = sys->ani_axis + 3 * i;
        const double d = (e[0] * s[3 * i] + e[1] * s[3 * i + 1] + e[2] * s[3 * i + 2]);
        for (int j = 0; j < 3; ++j)
          h[3 * i + j] = sys->ani_K[i] * (double)sys->ani_power * pow(d, sys->ani_power - 1) * e[j];
      }
      break;
    case OR_TERM_FIELD:
      memcpy(h, sys->ext_field, sizeof(double) * 3 * (size_t)n);
      break;
    case OR_TERM_DIPOLE:
      or_dipole_bruteforce(sys->dip_lat, n, sys->dip_pos, sys->mus, s, sys->dip_rcut, sys->dip_a,
                           h, sys->nthreads);
      break;
  }
}

void or_compute_fields(const or_system_t *sys, const double *s, double t, double *h, double *scratch) {
  const int64_t n3 = 3 * sys->n;
  if (sys->nterms == 0) return;
  or_term_field(sys, 0, s, t, h);
  for (int k = 1; k < sys->nterms; ++k) {
    or_term_field(sys, k, s, t, scratch);
    /* cblas_daxpy with alpha = 1: y = y + 1.0 * x */
    for (int64_t x = 0; x < n3; ++x) h[x] = h[x] + 1.0 * scratch[x];
  }
}

/* unit_vector (containers/vec3.h:276-283) */
static inline void unit_vector(double v[3]) {
  const double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (fabs(len) <= DBL_EPSILON) return;
  v[0] = v[0] / len; v[1] = v[1] / len; v[2] = v[2] / len;
}

/* rhs = -gyro * (s x h + alpha * s x (s x h))  (cpu_llg_heun.cc:89, :130) */
static inline void llg_rhs(const double s[3], const double h[3], double gyro, double alpha,
                           double rhs[3]) {
  const double c0 = s[1] * h[2] - s[2] * h[1];
  const double c1 = s[2] * h[0] - s[0] * h[2];
  const double c2 = s[0] * h[1] - s[1] * h[0];
  const double d0 = s[1] * c2 - s[2] * c1;
  const double d1 = s[2] * c0 - s[0] * c2;
  const double d2 = s[0] * c1 - s[1] * c0;
  const double g = -gyro;
  rhs[0] = g * (c0 + alpha * d0);
  rhs[1] = g * (c1 + alpha * d1);
  rhs[2] = g * (c2 + alpha * d2);
}

void or_heun_step(const or_system_t *sys, double *s, const double *w, double dt, double t0,
                  double *work) {
  const int64_t n = sys->n;
  const int nth = sys->nthreads > 0 ? sys->nthreads : 1;
  (void)nth;
  double *s_old = work, *ds = work + 3 * n, *h = work + 6 * n, *scratch = work + 9 * n;
  memcpy(s_old, s, sizeof(double) * 3 * (size_t)n);

  for (int phase = 0; phase < 2; ++phase) {
    const double t = phase == 0 ? t0 : t0 + dt;
    or_compute_fields(sys, s, t, h, scratch);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nth)
#endif
    for (int64_t i = 0; i < n; ++i) {
      double hh[3], si[3], rhs[3];
      for (int j = 0; j < 3; ++j) {
        /* h = w + h / mus  (cpu_llg_heun.cc:68-81, :106-122) */
        hh[j] = w ? (w[3 * i + j] + h[3 * i + j] / sys->mus[i]) : h[3 * i + j] / sys->mus[i];
        si[j] = s[3 * i + j];
      }
      llg_rhs(si, hh, sys->gyro[i], sys->alpha[i], rhs);
      double v[3];
      if (phase == 0) {
        for (int j = 0; j < 3; ++j) {
          ds[3 * i + j] = 0.5 * rhs[j];
          v[j] = si[j] + dt * rhs[j];
        }
      } else {
        for (int j = 0; j < 3; ++j) ds[3 * i + j] = ds[3 * i + j] + 0.5 * rhs[j];
        for (int j = 0; j < 3; ++j) v[j] = s_old[3 * i + j] + dt * ds[3 * i + j];
      }
      unit_vector(v);
      for (int j = 0; j < 3; ++j) s[3 * i + j] = v[j];
    }
  }
}

/* One CUDALLGRK4Solver::run step (solvers/cuda_rk4_base.cu:57-108 with
 * function_kernel = cuda_llg_rk4.cu:17-30 / cuda_llg_rk4_kernel.cuh:8-58 and
 * post_step = cuda/cuda_spin_ops.cu:4-17).  Stage k evaluates the fields at
 * its input y_k (y_0 = s) and time t0, t0 + dt/2, t0 + dt/2, t0 + dt;
 * k = rhs(y_k, h/mus + w) with the step's one noise draw; y_{k+1} =
 * s_old + c k (c = dt/2, dt/2, dt; the reference's copy + axpy, unnormalised);
 * s = v / sqrt(v.v) with v = s_old + dt (k1 + 2 k2 + 2 k3 + k4) / 6 summed
 * left to right.  Each + - * / sqrt is one IEEE operation in that order (the
 * reference's axpy and rsqrt are library and CUDA-math calls whose rounding
 * this restatement does not reproduce).  work needs 5*3N doubles. */
void or_rk4_step(const or_system_t *sys, double *s, const double *w, double dt, double t0,
                 double *work) {
  const int64_t n = sys->n;
  const int nth = sys->nthreads > 0 ? sys->nthreads : 1;
  (void)nth;
  double *s_old = work, *acc = work + 3 * n, *h = work + 6 * n, *y = work + 9 * n,
         *scratch = work + 12 * n;
  memcpy(s_old, s, sizeof(double) * 3 * (size_t)n);
  memcpy(y, s, sizeof(double) * 3 * (size_t)n);
  const double half = 0.5 * dt;
  for (int stage = 0; stage < 4; ++stage) {
    const double t = stage == 0 ? t0 : stage == 3 ? t0 + dt : t0 + half;
    or_compute_fields(sys, y, t, h, scratch);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nth)
#endif
    for (int64_t i = 0; i < n; ++i) {
      double hh[3], yi[3], k[3];
      for (int j = 0; j < 3; ++j) {
        hh[j] = w ? (h[3 * i + j] / sys->mus[i] + w[3 * i + j]) : h[3 * i + j] / sys->mus[i];
        yi[j] = y[3 * i + j];
      }
      llg_rhs(yi, hh, sys->gyro[i], sys->alpha[i], k);
      if (stage < 3) {
        const double c = stage == 2 ? dt : half;
        for (int j = 0; j < 3; ++j) {
          acc[3 * i + j] = stage == 0 ? k[j] : acc[3 * i + j] + 2.0 * k[j];
          y[3 * i + j] = s_old[3 * i + j] + c * k[j];
        }
      } else {
        double v[3];
        for (int j = 0; j < 3; ++j)
          v[j] = s_old[3 * i + j] + (dt * (acc[3 * i + j] + k[j])) / 6.0;
        const double rn = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int j = 0; j < 3; ++j) s[3 * i + j] = v[j] * rn;
      }
    }
  }
}

static void site_pos(const or_lattice_t *lat, int64_t s, double out[3]) {
  const int nb = lat->nbasis;
  const int m = (int)(s % nb);
  int64_t c = s / nb;
  const int k = (int)(c % lat->size[2]);
  c /= lat->size[2];
  const int j = (int)(c % lat->size[1]);
  const int i = (int)(c / lat->size[1]);
  const double f[3] = {lat->basis_frac[3 * m] + i, lat->basis_frac[3 * m + 1] + j,
                       lat->basis_frac[3 * m + 2] + k};
  for (int r = 0; r < 3; ++r)
    out[r] = lat->cell[3 * r] * f[0] + lat->cell[3 * r + 1] * f[1] + lat->cell[3 * r + 2] * f[2];
}

static int dgt(double a, double b, double eps) {
  return (a - b) > (fmax(fabs(a), fabs(b)) * eps);
}

typedef struct { int32_t i, j; double rij[3], d; } ann_pair_t;

static int ann_cmp(const void *pa, const void *pb) {
  const ann_pair_t *a = (const ann_pair_t *)pa, *b = (const ann_pair_t *)pb;
  if (a->i != b->i) return a->i < b->i ? -1 : 1;
  if (a->j != b->j) return a->j < b->j ? -1 : 1;
  return a->d < b->d ? -1 : a->d > b->d ? 1 : 0;
}

int64_t or_annulus_pairs(const or_lattice_t *lat, double inner, double outer, double eps,
                         int32_t *out_i, int32_t *out_j, double *out_rij, double *out_d,
                         int64_t max_out) {
  const int64_t n = (int64_t)lat->nbasis * lat->size[0] * lat->size[1] * lat->size[2];
  double A[3][3];
  for (int d = 0; d < 3; ++d)
    for (int r = 0; r < 3; ++r) A[d][r] = (double)lat->size[d] * lat->cell[3 * r + d];
  int nimg[3];
  for (int d = 0; d < 3; ++d) {
    const double *a = A[(d + 1) % 3], *b = A[(d + 2) % 3], *c = A[d];
    const double x[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double h = fabs(x[0] * c[0] + x[1] * c[1] + x[2] * c[2]) / sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    nimg[d] = lat->periodic[d] ? (int)ceil(outer / h) + 1 : 0;
  }
  ann_pair_t *buf = NULL;
  int64_t cnt = 0, cap = 0;
  for (int64_t i = 0; i < n; ++i) {
    double ri[3];
    site_pos(lat, i, ri);
    for (int64_t j = 0; j < n; ++j) {
      double rj[3];
      site_pos(lat, j, rj);
      for (int h = -nimg[0]; h <= nimg[0]; ++h)
        for (int k = -nimg[1]; k <= nimg[1]; ++k)
          for (int l = -nimg[2]; l <= nimg[2]; ++l) {
            double p[3], rij[3], o[3];
            for (int r = 0; r < 3; ++r) {
              const double t = (h * A[0][r] + k * A[1][r]) + l * A[2][r];
              p[r] = rj[r] + t;
              rij[r] = p[r] - ri[r];
              o[r] = ri[r] - p[r];
            }
            const double d = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
            if (dgt(d, outer, eps) || !dgt(d, inner, eps)) continue;
            if (cnt == cap) {
              cap = cap ? 2 * cap : 1024;
              buf = (ann_pair_t *)realloc(buf, sizeof(ann_pair_t) * (size_t)cap);
            }
            buf[cnt].i = (int32_t)i;
            buf[cnt].j = (int32_t)j;
            memcpy(buf[cnt].rij, rij, sizeof rij);
            buf[cnt].d = d;
            ++cnt;
          }
    }
  }
  if (cnt > max_out) {
    free(buf);
    return -2;
  }
  if (cnt) qsort(buf, (size_t)cnt, sizeof(ann_pair_t), ann_cmp);
  for (int64_t q = 0; q < cnt; ++q) {
    out_i[q] = buf[q].i;
    out_j[q] = buf[q].j;
    memcpy(out_rij + 3 * q, buf[q].rij, 3 * sizeof(double));
    out_d[q] = buf[q].d;
  }
  free(buf);
  return cnt;
}

void or_thermostat_sigma(int64_t n, const double *mus, const double *gyro, const double *alpha,
                         double dt, int gilbert, double *sigma) {
  for (int64_t i = 0; i < n; ++i) {
    double denominator = 1.0;
    if (gilbert) denominator = 1.0 + alpha[i] * alpha[i];
    sigma[i] = sqrt((2.0 * OR_KB * alpha[i]) / (mus[i] * gyro[i] * dt * denominator));
  }
}

void or_scale_noise(int64_t n, const double *xi, const double *sigma, double T, double *w) {
  const double sqrt_t = sqrt(T);
  for (int64_t i = 0; i < n; ++i)
    for (int j = 0; j < 3; ++j) w[3 * i + j] = xi[3 * i + j] * sigma[i] * sqrt_t;
}

/* ======================================================================= */
/* Reductions                                                                */

void or_sum_moments(int64_t n, const double *s, const double *mus, double out[3]) {
  double sum[3] = {0.0, 0.0, 0.0};
  for (int64_t i = 0; i < n; ++i)
    for (int j = 0; j < 3; ++j) sum[j] += mus[i] * s[3 * i + j];
  out[0] = sum[0]; out[1] = sum[1]; out[2] = sum[2];
}

double or_kahan_sum(int64_t n, const double *x) {
  if (n <= 0) return 0.0;
  double sum = x[0], c = 0.0;
  for (int64_t i = 1; i < n; ++i) {
    const double y = x[i] - c;
    const double t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

void or_kahan_sum3(int64_t n, const double *x, double out[3]) {
  if (n <= 0) { out[0] = out[1] = out[2] = 0.0; return; }
  double sum[3] = {x[0], x[1], x[2]}, c[3] = {0, 0, 0};
  for (int64_t i = 1; i < n; ++i)
    for (int j = 0; j < 3; ++j) {
      const double y = x[3 * i + j] - c[j];
      const double t = sum[j] + y;
      c[j] = (t - sum[j]) - y;
      sum[j] = t;
    }
  out[0] = sum[0]; out[1] = sum[1]; out[2] = sum[2];
}

double or_term_energy(const or_system_t *sys, int term, const double *s, double t) {
  const int64_t n = sys->n;
  double *h = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  or_term_field(sys, term, s, t, h);
  double e = 0.0;
  switch (sys->kind[term]) {
    case OR_TERM_EXCHANGE:
    case OR_TERM_DIPOLE:
      /* sparse_interaction.cc:86-100, dipole_bruteforce.cc:55-58:
       * 0.5 * sum_i (-s_i . h_i) */
      for (int64_t i = 0; i < n; ++i)
        e += -(s[3 * i] * h[3 * i] + s[3 * i + 1] * h[3 * i + 1] + s[3 * i + 2] * h[3 * i + 2]);
      e = 0.5 * e;
      break;
    case OR_TERM_ZEEMAN:
    case OR_TERM_FIELD:
      /* zeeman.cc / applied_field.cc: e_i = -s_i . h_i */
      for (int64_t i = 0; i < n; ++i)
        e += -(s[3 * i] * h[3 * i] + s[3 * i + 1] * h[3 * i + 1] + s[3 * i + 2] * h[3 * i + 2]);
      break;
    case OR_TERM_UNIAXIAL:
      /* uniaxial_anisotropy.cc:116-123: e_i = -K (s.e)^p */
      for (int64_t i = 0; i < n; ++i) {
        const double *ax = sys->ani_axis + 3 * i;
        const double d = (ax[0] * s[3 * i] + ax[1] * s[3 * i + 1] + ax[2] * s[3 * i + 2]);
        e += (-sys->ani_K[i] * pow(d, sys->ani_power));
      }
      break;
  }
  free(h);
  return e;
}

/* ======================================================================= */
/* Dipole brute force (hamiltonian/dipole_bruteforce.cc:75-120)              */

/* dipole_bruteforce.cc:75-120: the field at site i from every other site
 * within r_cutoff (Smith's minimum image on periodic axes) */
static void dipole_site(const or_lattice_t *lat, const double S[9], const double Si[9], int64_t n,
                        const double *pos, const double *mus, const double *s, double rc2,
                        double mu0, double a3, int64_t i, double out[3]) {
  const double w0 = mus[i] * mu0 / (4.0 * OR_PI * a3);
  double hx = 0, hy = 0, hz = 0;
  for (int64_t j = 0; j < n; ++j) {
    if (j == i) continue;
    double r[3], f[3];
    for (int q = 0; q < 3; ++q) r[q] = pos[3 * i + q] - pos[3 * j + q];
    mat3_mul_vec(Si, r, f);
    for (int q = 0; q < 3; ++q) if (lat->periodic[q]) f[q] = f[q] - nearbyint(f[q]);
    mat3_mul_vec(S, f, r);
    const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const double mx = r2 > rc2 ? r2 : rc2;
    if ((r2 - rc2) > mx * 1e-4) continue;
    const double sj[3] = {s[3 * j], s[3 * j + 1], s[3 * j + 2]};
    const double sr = sj[0] * r[0] + sj[1] * r[1] + sj[2] * r[2];
    const double nr = sqrt(r2);
    const double r5 = nr * nr * nr * nr * nr;
    hx += w0 * mus[j] * (3.0 * r[0] * sr - r2 * sj[0]) / r5;
    hy += w0 * mus[j] * (3.0 * r[1] * sr - r2 * sj[1]) / r5;
    hz += w0 * mus[j] * (3.0 * r[2] * sr - r2 * sj[2]) / r5;
  }
  out[0] = hx; out[1] = hy; out[2] = hz;
}

/* supercell matrix and its inverse for Smith's minimum image */
static void dipole_setup(const or_lattice_t *lat, double S[9], double Si[9]) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) S[3 * r + c] = lat->cell[3 * r + c] * lat->size[c];
  mat3_inverse(S, Si);
}

void or_dipole_bruteforce(const or_lattice_t *lat, int64_t n, const double *pos, const double *mus,
                          const double *s, double r_cutoff, double a_m, double *h, int nthreads) {
  double S[9], Si[9];
  dipole_setup(lat, S, Si);
  const double rc2 = r_cutoff * r_cutoff;
  const double mu0 = or_vacuum_permeability();
  const double a3 = a_m * a_m * a_m;
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t i = 0; i < n; ++i) dipole_site(lat, S, Si, n, pos, mus, s, rc2, mu0, a3, i, h + 3 * i);
}

void or_dipole_bruteforce_sites(const or_lattice_t *lat, int64_t n, const double *pos,
                                const double *mus, const double *s, double r_cutoff, double a_m,
                                int64_t nsites, const int64_t *sites, double *h, int nthreads) {
  double S[9], Si[9];
  dipole_setup(lat, S, Si);
  const double rc2 = r_cutoff * r_cutoff;
  const double mu0 = or_vacuum_permeability();
  const double a3 = a_m * a_m * a_m;
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t q = 0; q < nsites; ++q)
    dipole_site(lat, S, Si, n, pos, mus, s, rc2, mu0, a3, sites[q], h + 3 * q);
}
