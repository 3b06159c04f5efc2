// jh_layout.h -- the device layout of FP64 spin arrays (s, the predictor's
// s*, the RK4 stage inputs): one allocation of 3N doubles holding (x, y) of
// spin i at [2i], [2i + 1] and z at [2N + i].  A gathered neighbour is then
// one 16-byte and one 8-byte load; the texture address unit costs an 8-byte
// load about as much as a 16-byte one, so three 8-byte loads (plain SoA) keep
// it about 35 % longer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jh {

// index of component c of spin i in an FP64 spin array of n spins
__host__ __device__ inline int64_t spin_idx(int64_t n, int64_t i, int c) {
  return c < 2 ? 2 * i + c : 2 * n + i;
}
__device__ __forceinline__ void load_spin(const double *__restrict__ s, int64_t n, int64_t i,
                                          double &x, double &y, double &z) {
  const double2 v = reinterpret_cast<const double2 *>(s)[i];
  x = v.x;
  y = v.y;
  z = s[2 * n + i];
}
__device__ __forceinline__ void store_spin(double *s, int64_t n, int64_t i, double x, double y,
                                           double z) {
  reinterpret_cast<double2 *>(s)[i] = make_double2(x, y);
  s[2 * n + i] = z;
}
// the same as non-temporal stores (streamed past the caches)
__device__ __forceinline__ void store_spin_nt(double *s, int64_t n, int64_t i, double x, double y,
                                              double z) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store((d2){x, y}, reinterpret_cast<d2 *>(s) + i);
  __builtin_nontemporal_store(z, s + 2 * n + i);
}

}  // namespace jh
