// jh_device.h -- device-side building blocks of the gfx950 Heun-LLG engine:
// the Philox4x32-10 counter RNG, the thermostat's normal deviates, the LLG
// right-hand side and the spin renormalisation.
//
// Everything here is compiled with -ffp-contract=off: each +, -, *, / and
// sqrt is one correctly rounded IEEE-754 operation in the order written, so
// the FP64 path reproduces the reference CPU solver's arithmetic bit for bit
// (solvers/cpu_llg_heun.cc:84-144, containers/vec3.h:184-188, :276-283).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jh {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Multipliers 0xD2511F53,
// 0xCD9E8D57; Weyl key increments 0x9E3779B9, 0xBB67AE85; 10 rounds.
struct u32x4 { uint32_t x, y, z, w; };

// a ^ b ^ k as one v_bitop3_b32 (truth table 0x96, odd parity, which is
// symmetric in its three inputs); the compiler issues two v_xor_b32.  UK: k
// is wave-uniform and is read from a scalar register; otherwise from a
// vector register (a divergent value must never take the scalar form).
template <bool UK>
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t k) {
  uint32_t r;
  if constexpr (UK)
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "s"(k));
  else
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(k));
  return r;
}

// Each round's two 32 x 32 -> 64-bit products are formed whole (one
// v_mad_u64_u32 each, not a low and a high multiply).  UK: the key (k0, k1)
// is the same in every lane.
template <bool UK>
__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  if constexpr (UK) {
    // each call forms its own ten round keys as it goes (scalar adds):
    // shared between calls, the twenty would stay live and spill
    asm volatile("" : "+s"(k0), "+s"(k1));
  }
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    c = u32x4{xor3<UK>(hi1, c.y, k0), lo1, xor3<UK>(hi0, c.w, k1), lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Uniform in (0, 1): (2k + 1) 2^-53 with k the top 52 bits of hi:lo (exact).
__device__ __forceinline__ double u01(uint32_t hi, uint32_t lo) {
  const uint64_t k = ((uint64_t)hi << 20) | (uint64_t)(lo >> 12);
  return (double)(2 * k + 1) * 0x1.0p-53;
}

// log(x) for a positive normal double: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s) with s = (m-1)/(m+1) summed as the odd series to s^21 by
// fused multiply-add Horner steps.  Every operation is one correctly rounded
// IEEE-754 operation (fma included), so the CPU oracle reproduces the bits.
__device__ __forceinline__ double noise_log(double x) {
  uint64_t b = (uint64_t)__double_as_longlong(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
  double m = __longlong_as_double((long long)b);
  if (m > 1.4142135623730951) { m = m * 0.5; e = e + 1; }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 1.0 / 21.0;
  p = __builtin_fma(p, z, 1.0 / 19.0); p = __builtin_fma(p, z, 1.0 / 17.0);
  p = __builtin_fma(p, z, 1.0 / 15.0); p = __builtin_fma(p, z, 1.0 / 13.0);
  p = __builtin_fma(p, z, 1.0 / 11.0); p = __builtin_fma(p, z, 1.0 / 9.0);
  p = __builtin_fma(p, z, 1.0 / 7.0);  p = __builtin_fma(p, z, 1.0 / 5.0);
  p = __builtin_fma(p, z, 1.0 / 3.0);
  const double s2 = 2.0 * s;
  const double lm = __builtin_fma(s2, z * p, s2);
  return __builtin_fma((double)e, 0.6931471805599453, lm);
}

// cos(2 pi u), sin(2 pi u) for u in (0, 1): exact quadrant reduction, Taylor
// series on [-pi/4, pi/4] by fma Horner steps, exact quadrant rotation.
__device__ __forceinline__ void noise_sincos2pi(double u, double &c, double &s) {
  const double v = u * 4.0;
  const double q = rint(v);
  const double r = v - q;
  const double x = r * 1.5707963267948966;
  const double z = x * x;
  double ps = -1.0 / 121645100408832000.0;
  ps = __builtin_fma(ps, z, 1.0 / 355687428096000.0);
  ps = __builtin_fma(ps, z, -1.0 / 1307674368000.0);
  ps = __builtin_fma(ps, z, 1.0 / 6227020800.0);
  ps = __builtin_fma(ps, z, -1.0 / 39916800.0);
  ps = __builtin_fma(ps, z, 1.0 / 362880.0);
  ps = __builtin_fma(ps, z, -1.0 / 5040.0);
  ps = __builtin_fma(ps, z, 1.0 / 120.0);
  ps = __builtin_fma(ps, z, -1.0 / 6.0);
  const double sn = __builtin_fma(x, z * ps, x);
  double pc = 1.0 / 6402373705728000.0;
  pc = __builtin_fma(pc, z, -1.0 / 20922789888000.0);
  pc = __builtin_fma(pc, z, 1.0 / 87178291200.0);
  pc = __builtin_fma(pc, z, -1.0 / 479001600.0);
  pc = __builtin_fma(pc, z, 1.0 / 3628800.0);
  pc = __builtin_fma(pc, z, -1.0 / 40320.0);
  pc = __builtin_fma(pc, z, 1.0 / 720.0);
  pc = __builtin_fma(pc, z, -1.0 / 24.0);
  pc = __builtin_fma(pc, z, 0.5);
  const double cs = __builtin_fma(-z, pc, 1.0);
  switch (((int)q) & 3) {
    case 0: c = cs;  s = sn;  break;
    case 1: c = -sn; s = cs;  break;
    case 2: c = -cs; s = -sn; break;
    default: c = sn; s = -cs; break;
  }
}

// Three standard normal deviates of the thermostat stream for (seed, step,
// spin): two Philox blocks with counter (spin lo, call | spin hi << 1,
// step lo, step hi) and key (seed lo, seed hi), Box-Muller on each block.
template <bool UK>
__device__ __forceinline__ void normals3(uint64_t seed, uint64_t step, uint64_t spin,
                                         double &z0, double &z1, double &z2) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const uint32_t shi = (uint32_t)(spin >> 32) << 1;
  double c, s;
  {
    const u32x4 a = philox4x32_10<UK>(
        u32x4{(uint32_t)spin, 0u | shi, (uint32_t)step, (uint32_t)(step >> 32)}, k0, k1);
    const double r0 = sqrt(-2.0 * noise_log(u01(a.x, a.y)));
    noise_sincos2pi(u01(a.z, a.w), c, s);
    z0 = r0 * c;
    z1 = r0 * s;
  }
  {
    const u32x4 b = philox4x32_10<UK>(
        u32x4{(uint32_t)spin, 1u | shi, (uint32_t)step, (uint32_t)(step >> 32)}, k0, k1);
    const double r1 = sqrt(-2.0 * noise_log(u01(b.x, b.y)));
    noise_sincos2pi(u01(b.z, b.w), c, s);
    z2 = r1 * c;
  }
}

// rhs = -gyro (s x h + alpha s x (s x h)) with the reference's grouping
// (solvers/cpu_llg_heun.cc:89; containers/vec3.h:184-188).
__device__ __forceinline__ void llg_rhs(double s0, double s1, double s2, double h0, double h1,
                                        double h2, double gyro, double alpha, double &r0,
                                        double &r1, double &r2) {
  const double c0 = s1 * h2 - s2 * h1;
  const double c1 = s2 * h0 - s0 * h2;
  const double c2 = s0 * h1 - s1 * h0;
  const double d0 = s1 * c2 - s2 * c1;
  const double d1 = s2 * c0 - s0 * c2;
  const double d2 = s0 * c1 - s1 * c0;
  const double g = -gyro;
  r0 = g * (c0 + alpha * d0);
  r1 = g * (c1 + alpha * d1);
  r2 = g * (c2 + alpha * d2);
}

// v / |v| as one reciprocal and three products (the FP32-field form, whose
// arithmetic need not reproduce the reference's operation for operation):
// within 1.5 ulp of unit_vector, two FP64 divisions fewer.
__device__ __forceinline__ void unit_vector_rcp(double &v0, double &v1, double &v2) {
  const double len = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
  if (fabs(len) <= 2.220446049250313e-16) return;
  const double r = 1.0 / len;
  v0 = v0 * r;
  v1 = v1 * r;
  v2 = v2 * r;
}

// unit_vector (containers/vec3.h:276-283): v / |v| unless |v| <= DBL_EPSILON.
__device__ __forceinline__ void unit_vector(double &v0, double &v1, double &v2) {
  const double len = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
  if (fabs(len) <= 2.220446049250313e-16) return;
  v0 = v0 / len;
  v1 = v1 / len;
  v2 = v2 / len;
}

// XCD-aware block order: hardware hands consecutive workgroups to the 8
// XCDs round robin; remap so each XCD walks one contiguous range of spins and
// neighbour gathers hit its own L2.
__device__ __forceinline__ int64_t xcd_block(int64_t bid, int64_t nb) {
  const int64_t q = nb >> 3, r = nb & 7;
  const int64_t x = bid & 7, k = bid >> 3;
  const int64_t start = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return start + k;
}

}  // namespace jh
