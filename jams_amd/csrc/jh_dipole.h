// jh_dipole.h -- zero-padded FFT dipole field on the device
// (hamiltonian/dipole_fft.cc, hamiltonian/cuda_dipole_fft.cu).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace jh {

class DipoleFFT {
 public:
  // Builds the k-space interaction tensors (real-space tensor on the host,
  // transformed on the device).  Returns nullptr (with *err set) on invalid
  // input.  mus: per-spin moments (meV/T), reference order.
  static DipoleFFT *create(const int32_t size[3], const int32_t periodic[3], int nbasis,
                           const double cell[9], const double *basis_frac, double a_m,
                           double r_cutoff, const double *mus, int64_t n, hipStream_t st,
                           std::string *err);
  virtual ~DipoleFFT() = default;
  // h (SoA, meV) = mus_i sum_j mus_j W(r_ij) s_j for spins s (SoA, reference
  // order); fp32: transforms and tensors in FP32 (the precision form).
  virtual hipError_t field(const double *s, double *h, bool fp32, hipStream_t st) = 0;
  // algorithmic HBM bytes of one field evaluation
  virtual double bytes(bool fp32) const = 0;
  // the kernel variants an evaluation launches: the x axis (fused x forward +
  // multiply + x inverse, or separate passes), the z + y forward and the
  // y + z inverse (fused, or separate passes); "512" marks the compile-time
  // 512-point stages, the number the lines per workgroup
  virtual std::string describe(bool fp32) const = 0;
};

}  // namespace jh
