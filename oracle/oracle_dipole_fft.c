/*
 * oracle_dipole_fft.c -- the FFT dipole field's convolution, restated as a
 * direct sum over cell offsets (TEST INFRASTRUCTURE ONLY).
 *
 * hamiltonian/dipole_fft.cc:195-251 fills a tensor W_ij(n) on the (padded)
 * cell grid for offsets n, then convolves it with mus_j s_j by FFT
 * (:254-350):  h_i(x) = mus_i sum_j sum_n W_ij(n) mus_j s_j(x - n).  The FFT
 * only changes how that sum is evaluated, so this direct form pins which
 * offsets enter and with which tensor:
 *   - periodic dimension: n in [0, L), r = minimum image of
 *     pos_i + n - pos_j (lattice/minimum_image.cc via Lattice::displacement);
 *   - open dimension: the reference's loop covers n in [0, L) only
 *     (dipole_fft.cc:227-229 with the padded half left zero), so the sum is
 *     one-sided; one_sided_open = 1 reproduces that, 0 takes -L < n < L,
 *     which equals the brute-force sum (dipole_bruteforce.cc:75-120);
 *   - |r| <= r_cutoff + 1e-4 (helpers/defaults.h:44), no n = 0 self term.
 * W_ab(r) = mu0 / (4 pi a^3) (3 r_a r_b - r^2 delta_ab) / |r|^5.
 * Pinned by: two-sided == or_dipole_bruteforce, which the reference's known
 * answers pin (tests/test_oracle.py).
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

static void inv3(const double *S, double *Si) {
  const double det = S[0] * (S[4] * S[8] - S[5] * S[7]) - S[1] * (S[3] * S[8] - S[5] * S[6]) +
                     S[2] * (S[3] * S[7] - S[4] * S[6]);
  Si[0] = (S[4] * S[8] - S[5] * S[7]) / det; Si[1] = (S[2] * S[7] - S[1] * S[8]) / det;
  Si[2] = (S[1] * S[5] - S[2] * S[4]) / det; Si[3] = (S[5] * S[6] - S[3] * S[8]) / det;
  Si[4] = (S[0] * S[8] - S[2] * S[6]) / det; Si[5] = (S[2] * S[3] - S[0] * S[5]) / det;
  Si[6] = (S[3] * S[7] - S[4] * S[6]) / det; Si[7] = (S[1] * S[6] - S[0] * S[7]) / det;
  Si[8] = (S[0] * S[4] - S[1] * S[3]) / det;
}

void or_dipole_fft_direct(const or_lattice_t *lat, const double *mus, const double *s,
                          double r_cutoff, double a_m, int one_sided_open, double *h,
                          int nthreads) {
  const int nb = lat->nbasis;
  const int L0 = lat->size[0], L1 = lat->size[1], L2 = lat->size[2];
  const int64_t n = (int64_t)L0 * L1 * L2 * nb;
  double S[9], Si[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) S[3 * r + c] = lat->cell[3 * r + c] * lat->size[c];
  inv3(S, Si);
  const double w = or_vacuum_permeability() / (4.0 * OR_PI * a_m * a_m * a_m);
  const double rc2 = (r_cutoff + 1e-4) * (r_cutoff + 1e-4);
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = (lat->periodic[a] || one_sided_open) ? 0 : -(lat->size[a] - 1);
    hi[a] = lat->size[a] - 1;
  }
  const int any_per = lat->periodic[0] || lat->periodic[1] || lat->periodic[2];
  for (int64_t q = 0; q < 3 * n; ++q) h[q] = 0.0;
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int x0 = 0; x0 < L0; ++x0) {
    for (int x1 = 0; x1 < L1; ++x1)
      for (int x2 = 0; x2 < L2; ++x2)
        for (int pi = 0; pi < nb; ++pi) {
          const int64_t i = (((int64_t)x0 * L1 + x1) * L2 + x2) * nb + pi;
          double hx = 0.0, hy = 0.0, hz = 0.0;
          for (int n0 = lo[0]; n0 <= hi[0]; ++n0)
            for (int n1 = lo[1]; n1 <= hi[1]; ++n1)
              for (int n2 = lo[2]; n2 <= hi[2]; ++n2) {
                int y[3] = {x0 - n0, x1 - n1, x2 - n2};
                int ok = 1;
                for (int a = 0; a < 3; ++a) {
                  if (lat->periodic[a]) y[a] = ((y[a] % lat->size[a]) + lat->size[a]) % lat->size[a];
                  else if (y[a] < 0 || y[a] >= lat->size[a]) ok = 0;
                }
                if (!ok) continue;
                for (int pj = 0; pj < nb; ++pj) {
                  if (n0 == 0 && n1 == 0 && n2 == 0 && pi == pj) continue;
                  const double f[3] = {lat->basis_frac[3 * pi] + n0 - lat->basis_frac[3 * pj],
                                       lat->basis_frac[3 * pi + 1] + n1 - lat->basis_frac[3 * pj + 1],
                                       lat->basis_frac[3 * pi + 2] + n2 - lat->basis_frac[3 * pj + 2]};
                  double r[3];
                  for (int a = 0; a < 3; ++a)
                    r[a] = lat->cell[3 * a] * f[0] + lat->cell[3 * a + 1] * f[1] + lat->cell[3 * a + 2] * f[2];
                  if (any_per) {
                    double g[3];
                    for (int a = 0; a < 3; ++a) g[a] = Si[3 * a] * r[0] + Si[3 * a + 1] * r[1] + Si[3 * a + 2] * r[2];
                    for (int a = 0; a < 3; ++a) if (lat->periodic[a]) g[a] -= nearbyint(g[a]);
                    for (int a = 0; a < 3; ++a) r[a] = S[3 * a] * g[0] + S[3 * a + 1] * g[1] + S[3 * a + 2] * g[2];
                  }
                  const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                  if (r2 > rc2 || r2 == 0.0) continue;
                  const int64_t j = (((int64_t)y[0] * L1 + y[1]) * L2 + y[2]) * nb + pj;
                  const double nr = sqrt(r2), r5 = nr * nr * nr * nr * nr;
                  const double sj[3] = {mus[j] * s[3 * j], mus[j] * s[3 * j + 1], mus[j] * s[3 * j + 2]};
                  const double sr = sj[0] * r[0] + sj[1] * r[1] + sj[2] * r[2];
                  hx += w * (3.0 * r[0] * sr - r2 * sj[0]) / r5;
                  hy += w * (3.0 * r[1] * sr - r2 * sj[1]) / r5;
                  hz += w * (3.0 * r[2] * sr - r2 * sj[2]) / r5;
                }
              }
          h[3 * i] = mus[i] * hx;
          h[3 * i + 1] = mus[i] * hy;
          h[3 * i + 2] = mus[i] * hz;
        }
  }
}
