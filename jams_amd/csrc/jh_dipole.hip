// jh_dipole.hip -- zero-padded FFT dipole field (see jh_dipole.h).
#include "jh_dipole.h"

namespace jh {

DipoleFFT *DipoleFFT::create(const int32_t *, const int32_t *, int, const double *,
                             const double *, double, double, const double *, hipStream_t,
                             std::string *err) {
  if (err) *err = "not available in this build";
  return nullptr;
}

}  // namespace jh
