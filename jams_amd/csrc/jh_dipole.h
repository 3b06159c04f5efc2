// jh_dipole.h -- zero-padded FFT dipole field on the device
// (hamiltonian/dipole_fft.cc, hamiltonian/cuda_dipole_fft.cu).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace jh {

class DipoleFFT {
 public:
  // Builds the k-space interaction tensors on the host and uploads them.
  // Returns nullptr (with *err set) on invalid input.
  static DipoleFFT *create(const int32_t size[3], const int32_t periodic[3], int nbasis,
                           const double cell[9], const double *basis_frac, double a_m,
                           double r_cutoff, const double *mus, hipStream_t st, std::string *err);
  virtual ~DipoleFFT() = default;
  // h (SoA, meV) = mus_i sum_j mus_j W(r_ij) s_j for spins s (SoA).
  virtual hipError_t field(const double *s, double *h, hipStream_t st) = 0;
};

}  // namespace jh
