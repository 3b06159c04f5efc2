// jh_kernels.hip -- gfx950 kernels of the Heun-LLG step.
//
// One Heun step is two launches (solvers/cuda_llg_heun.cu:64-122 does ~10-12):
//   phase 0 (predictor): h(s) = sum of terms in config order, w from the
//     in-register Philox stream, rhs, ds = rhs/2, s* = unit(s + dt rhs);
//   phase 1 (corrector): h(s*) at t+dt, same w, ds += rhs/2,
//     s = unit(s_old + dt ds).
// Each thread owns one spin.  The exchange field is a row-sequential sum in
// the reference CSR order (neighbour column ascending, tensor column
// ascending), which is what makes the FP64 path bit-identical to the CPU
// solver.  The FP32 form gathers FP32 spins, multiplies them by the FP64
// tensors and accumulates in FP64 (one rounding per product and per sum); it
// may order a row by tensor id so that tensors become wave-uniform.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "jh_device.h"
#include "jh_internal.h"
#include "jh_launch.h"

#ifndef JH_GATHER_UNROLL
#define JH_GATHER_UNROLL 4
#endif
#ifndef JH_WAVES_PER_EU
#define JH_WAVES_PER_EU 1
#endif

namespace jh {

constexpr int kGatherUnroll = JH_GATHER_UNROLL;
// Tensor-table rows are indexed with 24-bit multiplies (__mul24, full rate;
// a 32-bit multiply is a quarter-rate instruction): table ids are < 2^24.
// full-tensor rows (9 products per neighbour) gain from more loads in flight
template <int XVk>
constexpr int kUnrollFor = xv_diag(XVk) ? kGatherUnroll : 2 * kGatherUnroll;

// FP32 gather copy of the spins: kG32 floats per spin, (x, y, z[, 0])
__device__ __forceinline__ void load_g32(const float *__restrict__ src, uint32_t j, double &x,
                                         double &y, double &z) {
  if constexpr (kG32 == 4) {
    const float4 v = reinterpret_cast<const float4 *>(src)[j];
    x = v.x; y = v.y; z = v.z;
  } else {
    const float *q = src + (size_t)3 * j;
    x = q[0]; y = q[1]; z = q[2];
  }
}
__device__ __forceinline__ void store_g32(float *dst, int64_t i, double x, double y, double z) {
  if constexpr (kG32 == 4) {
    reinterpret_cast<float4 *>(dst)[i] = make_float4((float)x, (float)y, (float)z, 0.f);
  } else {
    float *q = dst + 3 * i;
    q[0] = (float)x; q[1] = (float)y; q[2] = (float)z;
  }
}

// ---------------------------------------------------------------------------
// exchange field of spin i, FP64, reference order
template <int XVk>
__device__ __forceinline__ void xfield64(const XView &x, const double *__restrict__ src, int64_t n,
                                         int64_t i, const double *tab, double &h0, double &h1,
                                         double &h2) {
  const int64_t sl = i >> 6;
  const int lane = (int)(i & 63);
  const int64_t b0 = x.slice_off[sl], b1 = x.slice_off[sl + 1];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll 4
  for (int64_t e = b0 + lane; e < b1; e += kSlice) {
    const uint32_t w = __builtin_nontemporal_load(x.ent + e);
    const uint32_t j = w & x.colmask;
    double sx, sy, sz;
    load_spin(src, n, j, sx, sy, sz);
    if constexpr (XVk == XV_TDIAG) {
      const double *T = tab + 3 * (w >> x.colbits);
      a0 = a0 + sx * T[0];
      a1 = a1 + sy * T[1];
      a2 = a2 + sz * T[2];
    } else if constexpr (XVk == XV_TFULL) {
      const double *T = tab + 9 * (w >> x.colbits);
      a0 = a0 + sx * T[0]; a0 = a0 + sy * T[1]; a0 = a0 + sz * T[2];
      a1 = a1 + sx * T[3]; a1 = a1 + sy * T[4]; a1 = a1 + sz * T[5];
      a2 = a2 + sx * T[6]; a2 = a2 + sy * T[7]; a2 = a2 + sz * T[8];
    } else if constexpr (XVk == XV_BDIAG) {
      const double *V = x.val64 + (e - lane) * 3 + lane;
      a0 = a0 + sx * __builtin_nontemporal_load(V);
      a1 = a1 + sy * __builtin_nontemporal_load(V + 64);
      a2 = a2 + sz * __builtin_nontemporal_load(V + 128);
    } else if constexpr (XVk == XV_BFULL) {
      const double *V = x.val64 + (e - lane) * 9 + lane;
      double v[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) v[c] = __builtin_nontemporal_load(V + 64 * c);
      a0 = a0 + sx * v[0]; a0 = a0 + sy * v[1]; a0 = a0 + sz * v[2];
      a1 = a1 + sx * v[3]; a1 = a1 + sy * v[4]; a1 = a1 + sz * v[5];
      a2 = a2 + sx * v[6]; a2 = a2 + sy * v[7]; a2 = a2 + sz * v[8];
    }
  }
  h0 = a0; h1 = a1; h2 = a2;
}

// exchange field of spin i from FP32 storage: spins gathered as FP32 (the
// per-spin bytes), tensors from the FP64 table, one fused multiply-add per
// product into an FP64 sum; the row order is the reference's.  BLOCK values
// are stored per entry and therefore in FP32.
template <int XVk>
__device__ __forceinline__ void xfield32(const XView &x, const float *__restrict__ src, int64_t n,
                                         int64_t i, const double *tab, double &h0, double &h1,
                                         double &h2) {
  const int64_t sl = i >> 6;
  const int lane = (int)(i & 63);
  const int64_t b0 = x.slice_off[sl], b1 = x.slice_off[sl + 1];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll 4
  for (int64_t e = b0 + lane; e < b1; e += kSlice) {
    const uint32_t w = __builtin_nontemporal_load(x.ent + e);
    const uint32_t j = w & x.colmask;
    double sx, sy, sz;
    load_g32(src, j, sx, sy, sz);
    if constexpr (XVk == XV_TDIAG) {
      const double *T = tab + __mul24((int)(w >> x.colbits), 3);
      a0 = __builtin_fma(sx, T[0], a0);
      a1 = __builtin_fma(sy, T[1], a1);
      a2 = __builtin_fma(sz, T[2], a2);
    } else if constexpr (XVk == XV_TFULL) {
      const double *T = tab + __mul24((int)(w >> x.colbits), 9);
      a0 = __builtin_fma(sx, T[0], a0); a0 = __builtin_fma(sy, T[1], a0); a0 = __builtin_fma(sz, T[2], a0);
      a1 = __builtin_fma(sx, T[3], a1); a1 = __builtin_fma(sy, T[4], a1); a1 = __builtin_fma(sz, T[5], a1);
      a2 = __builtin_fma(sx, T[6], a2); a2 = __builtin_fma(sy, T[7], a2); a2 = __builtin_fma(sz, T[8], a2);
    } else if constexpr (XVk == XV_BDIAG) {
      const float *V = x.val32 + (e - lane) * 3 + lane;
      a0 = __builtin_fma(sx, (double)__builtin_nontemporal_load(V), a0);
      a1 = __builtin_fma(sy, (double)__builtin_nontemporal_load(V + 64), a1);
      a2 = __builtin_fma(sz, (double)__builtin_nontemporal_load(V + 128), a2);
    } else if constexpr (XVk == XV_BFULL) {
      const float *V = x.val32 + (e - lane) * 9 + lane;
      double v[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) v[c] = (double)__builtin_nontemporal_load(V + 64 * c);
      a0 = __builtin_fma(sx, v[0], a0); a0 = __builtin_fma(sy, v[1], a0); a0 = __builtin_fma(sz, v[2], a0);
      a1 = __builtin_fma(sx, v[3], a1); a1 = __builtin_fma(sy, v[4], a1); a1 = __builtin_fma(sz, v[5], a1);
      a2 = __builtin_fma(sx, v[6], a2); a2 = __builtin_fma(sy, v[7], a2); a2 = __builtin_fma(sz, v[8], a2);
    }
  }
  h0 = a0; h1 = a1; h2 = a2;
}

// exchange field of spin i from its row pattern, FP64, reference order
template <int XVk>
__device__ __forceinline__ void pfield64(const XView &x, const double *__restrict__ src, int64_t n,
                                         int64_t i, const double *tab, const uint32_t *pat,
                                         double &h0, double &h1, double &h2) {
  const uint32_t o = x.pofs[i];
  const int len = (int)pat[o];
  const uint32_t tmask = (1u << x.tbits) - 1u;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  // the products of one neighbour, in the reference order
  auto accum = [&](uint32_t t, double sx, double sy, double sz) {
    if constexpr (xv_diag(XVk)) {
      const double *T = tab + __mul24((int)t, 3);
      a0 = a0 + sx * T[0];
      a1 = a1 + sy * T[1];
      a2 = a2 + sz * T[2];
    } else {
      const double *T = tab + __mul24((int)t, 9);
      a0 = a0 + sx * T[0]; a0 = a0 + sy * T[1]; a0 = a0 + sz * T[2];
      a1 = a1 + sx * T[3]; a1 = a1 + sy * T[4]; a1 = a1 + sz * T[5];
      a2 = a2 + sx * T[6]; a2 = a2 + sy * T[7]; a2 = a2 + sz * T[8];
    }
  };
  if (n < ((int64_t)1 << 28)) {
    // (x, y) pairs and z addressed by 32-bit byte offsets from scalar bases
    const char *__restrict__ xy = reinterpret_cast<const char *>(src);
    const char *__restrict__ zs = reinterpret_cast<const char *>(src + 2 * n);
    const uint32_t ii = (uint32_t)i;
#pragma unroll kUnrollFor<XVk>
    for (int k = 1; k <= len; ++k) {
      const uint32_t w = pat[o + k];
      const uint32_t j = ii + (uint32_t)(((int32_t)w) >> x.tbits);
      const double2 v = *reinterpret_cast<const double2 *>(xy + (j << 4));
      const double sz = *reinterpret_cast<const double *>(zs + (j << 3));
      accum(w & tmask, v.x, v.y, sz);
    }
  } else {
#pragma unroll kUnrollFor<XVk>
    for (int k = 1; k <= len; ++k) {
      const uint32_t w = pat[o + k];
      const int64_t j = i + (int64_t)(((int32_t)w) >> x.tbits);
      double sx, sy, sz;
      load_spin(src, n, j, sx, sy, sz);
      accum(w & tmask, sx, sy, sz);
    }
  }
  h0 = a0; h1 = a1; h2 = a2;
}

// the same from FP32 storage with exact FP64 products (see xfield32)
template <int XVk>
__device__ __forceinline__ void pfield32(const XView &x, const float *__restrict__ src, int64_t n,
                                         int64_t i, const double *tab, const uint32_t *pat,
                                         double &h0, double &h1, double &h2) {
  const uint32_t o = x.pofs[i];
  const int len = (int)pat[o];
  const uint32_t tmask = (1u << x.tbits) - 1u;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll kUnrollFor<XVk>
  for (int k = 1; k <= len; ++k) {
    const uint32_t w = pat[o + k];
    const int64_t j = i + (int64_t)(((int32_t)w) >> x.tbits);
    const uint32_t t = w & tmask;
    double sx, sy, sz;
    load_g32(src, (uint32_t)j, sx, sy, sz);
    if constexpr (xv_diag(XVk)) {
      const double *T = tab + __mul24((int)t, 3);
      a0 = __builtin_fma(sx, T[0], a0);
      a1 = __builtin_fma(sy, T[1], a1);
      a2 = __builtin_fma(sz, T[2], a2);
    } else {
      const double *T = tab + __mul24((int)t, 9);
      a0 = __builtin_fma(sx, T[0], a0); a0 = __builtin_fma(sy, T[1], a0); a0 = __builtin_fma(sz, T[2], a0);
      a1 = __builtin_fma(sx, T[3], a1); a1 = __builtin_fma(sy, T[4], a1); a1 = __builtin_fma(sz, T[5], a1);
      a2 = __builtin_fma(sx, T[6], a2); a2 = __builtin_fma(sy, T[7], a2); a2 = __builtin_fma(sz, T[8], a2);
    }
  }
  h0 = a0; h1 = a1; h2 = a2;
}

// FP32 form with rows ordered by tensor id: in a slice whose rows share one
// tensor sequence the tensor of slot k is wave-uniform and is read through
// scalar loads into SGPRs (no LDS traffic for it); only the neighbour
// offsets come from each lane's pattern.  Other slices take pfield32.
// upat: the pattern words as byte offsets into the gather copy (LDS);
// gpat: the (offset, tensor id) words for the non-uniform slices (global).
template <int XVk>
__device__ __forceinline__ void ufield32(const XView &x, const float *__restrict__ src, int64_t n,
                                         int64_t i, const double *tab, const int32_t *upat,
                                         const uint32_t *gpat, double &h0, double &h1,
                                         double &h2) {
  const int sl = __builtin_amdgcn_readfirstlane((int)(i >> 6));
  const int2 info = x.sseq[sl];
  const int soff = __builtin_amdgcn_readfirstlane(info.x);
  if (soff < 0) {
    pfield32<XVk>(x, src, n, i, tab, gpat, h0, h1, h2);
    return;
  }
  const int len = __builtin_amdgcn_readfirstlane(info.y);
  const double *__restrict__ T = x.seqtab + soff;
  const uint32_t o = x.pofs[i];
  // byte offset of this spin in the gather copy (the copy is < 2 GiB)
  const uint32_t ib = (uint32_t)i * (uint32_t)(kG32 * sizeof(float));
  const char *__restrict__ base = reinterpret_cast<const char *>(src);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll kUnrollFor<XVk>
  for (int k = 0; k < len; ++k) {
    const uint32_t jb = ib + (uint32_t)upat[o + 1 + k];
    const float *q = reinterpret_cast<const float *>(base + jb);
    const double sx = q[0], sy = q[1], sz = q[2];
    if constexpr (xv_diag(XVk)) {
      const double *Tk = T + 3 * k;
      a0 = __builtin_fma(sx, Tk[0], a0);
      a1 = __builtin_fma(sy, Tk[1], a1);
      a2 = __builtin_fma(sz, Tk[2], a2);
    } else {
      const double *Tk = T + 9 * k;
      a0 = __builtin_fma(sx, Tk[0], a0); a0 = __builtin_fma(sy, Tk[1], a0); a0 = __builtin_fma(sz, Tk[2], a0);
      a1 = __builtin_fma(sx, Tk[3], a1); a1 = __builtin_fma(sy, Tk[4], a1); a1 = __builtin_fma(sz, Tk[5], a1);
      a2 = __builtin_fma(sx, Tk[6], a2); a2 = __builtin_fma(sy, Tk[7], a2); a2 = __builtin_fma(sz, Tk[8], a2);
    }
  }
  h0 = a0; h1 = a1; h2 = a2;
}

// D x b into (a0, a1, a2), b = s1 - s2 from FP32 spins, six FP64 multiply-adds
__device__ __forceinline__ void dmi_add(const double *__restrict__ D, float bx, float by, float bz,
                                        double &a0, double &a1, double &a2) {
  const double X = bx, Y = by, Z = bz;
  a0 = __builtin_fma(D[1], Z, a0); a0 = __builtin_fma(-D[2], Y, a0);
  a1 = __builtin_fma(D[2], X, a1); a1 = __builtin_fma(-D[0], Z, a1);
  a2 = __builtin_fma(D[0], Y, a2); a2 = __builtin_fma(-D[1], X, a2);
}

// FP32 form, isotropic (+ DMI) tensors in slices with a pair program
// (XV_QISO / XV_QDMI): per group of one or two bond pairs (T, T^T) sharing J,
// J (s1 + s2 [+ s3 + s4]) with the spin sums in FP32 and J exact in FP64,
// plus D x (s1 - s2) per pair in FP64 (QDMI).  Other slices take ufield32.
template <int XVk>
__device__ __forceinline__ void pairfield32(const XView &x, const float *__restrict__ src,
                                            int64_t n, int64_t i, const double *tab,
                                            const int32_t *upat, const uint32_t *gpat,
                                            double &h0, double &h1, double &h2) {
  const int sl = __builtin_amdgcn_readfirstlane((int)(i >> 6));
  const int2 info = x.spair[sl];
  const int g0 = __builtin_amdgcn_readfirstlane(info.x);
  if (g0 < 0) {
    ufield32<XVk>(x, src, n, i, tab, upat, gpat, h0, h1, h2);
    return;
  }
  const int cnt = __builtin_amdgcn_readfirstlane(info.y);
  const int n2 = cnt & 0xFFFF, n1 = cnt >> 16;
  const int4 *__restrict__ G = x.pgrp + g0;
  const double *__restrict__ V = x.pval + (size_t)8 * g0;
  const uint32_t o = x.pofs[i] + 1;
  // byte offset of this spin in the gather copy (the copy is < 2 GiB)
  const uint32_t ib = (uint32_t)i * (uint32_t)(kG32 * sizeof(float));
  const char *__restrict__ base = reinterpret_cast<const char *>(src);
  auto spin = [&](int k, float &sx, float &sy, float &sz) {
    const float *q = reinterpret_cast<const float *>(base + (ib + (uint32_t)upat[o + k]));
    sx = q[0]; sy = q[1]; sz = q[2];
  };
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  // groups of two pairs: four gathers each, no branch
#pragma unroll 2
  for (int g = 0; g < n2; ++g) {
    const int4 k = G[g];
    const double *__restrict__ Vg = V + 8 * g;
    float x1, y1, z1, x2, y2, z2, x3, y3, z3, x4, y4, z4;
    spin(k.x, x1, y1, z1);
    spin(k.y, x2, y2, z2);
    spin(k.z, x3, y3, z3);
    spin(k.w, x4, y4, z4);
    const float p0 = (x1 + x2) + (x3 + x4);
    const float p1 = (y1 + y2) + (y3 + y4);
    const float p2 = (z1 + z2) + (z3 + z4);
    if constexpr (XVk == XV_QDMI) {
      dmi_add(Vg + 1, x1 - x2, y1 - y2, z1 - z2, a0, a1, a2);
      dmi_add(Vg + 4, x3 - x4, y3 - y4, z3 - z4, a0, a1, a2);
    }
    const double J = Vg[0];
    a0 = __builtin_fma(J, (double)p0, a0);
    a1 = __builtin_fma(J, (double)p1, a1);
    a2 = __builtin_fma(J, (double)p2, a2);
  }
  // single pairs
  for (int g = n2; g < n2 + n1; ++g) {
    const int4 k = G[g];
    const double *__restrict__ Vg = V + 8 * g;
    float x1, y1, z1, x2, y2, z2;
    spin(k.x, x1, y1, z1);
    spin(k.y, x2, y2, z2);
    if constexpr (XVk == XV_QDMI) dmi_add(Vg + 1, x1 - x2, y1 - y2, z1 - z2, a0, a1, a2);
    const double J = Vg[0];
    a0 = __builtin_fma(J, (double)(x1 + x2), a0);
    a1 = __builtin_fma(J, (double)(y1 + y2), a1);
    a2 = __builtin_fma(J, (double)(z1 + z2), a2);
  }
  h0 = a0; h1 = a1; h2 = a2;
}

// bytes of LDS used by the tensor table (the pattern table follows it)
__host__ __device__ inline size_t table_bytes(int prec, int xv, int ntab) {
  if (!xv_typed(xv)) return 0;
  const size_t b = (size_t)ntab * (xv_diag(xv) ? 3 : 9) * 8;
  return (b + 15) & ~(size_t)15;
}

// Copy `words` 32-bit words global -> LDS (both 16-byte aligned) with
// 16-byte loads: every workgroup stages the tables, so their load count
// matters (the five-shell bcc pattern table is 7552 words).
__device__ __forceinline__ void copy_words(uint32_t *dst, const uint32_t *__restrict__ src,
                                           int words) {
  const int w4 = words >> 2;
  uint4 *d4 = reinterpret_cast<uint4 *>(dst);
  const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
  for (int q = threadIdx.x; q < w4; q += blockDim.x) d4[q] = s4[q];
  for (int q = (w4 << 2) + (int)threadIdx.x; q < words; q += blockDim.x) dst[q] = src[q];
}

template <int PREC, int XVk>
__device__ __forceinline__ void load_table(const XView &x, unsigned char *lds) {
  if constexpr (xv_typed(XVk)) {
    const int cnt = x.ntab * (xv_diag(XVk) ? 3 : 9);
    copy_words(reinterpret_cast<uint32_t *>(lds), reinterpret_cast<const uint32_t *>(x.tab64),
               2 * cnt);
  }
  if constexpr (xv_pattern_lds(XVk)) {
    uint32_t *p = reinterpret_cast<uint32_t *>(lds + table_bytes(PREC, XVk, x.ntab));
    const uint32_t *from = xv_uniform(XVk) ? reinterpret_cast<const uint32_t *>(x.upat) : x.pat;
    copy_words(p, from, x.pat_words);
  }
}

template <int PREC, int XVk>
__device__ __forceinline__ void exchange_at(const KParams &p, const double *src64,
                                            const float *src32, int64_t i,
                                            const unsigned char *lds, double &f0, double &f1,
                                            double &f2) {
  if constexpr (XVk == XV_NONE) {
    f0 = f1 = f2 = 0.0;
  } else if constexpr (xv_pattern(XVk)) {
    const uint32_t *pat;
    if constexpr (xv_pattern_lds(XVk))
      pat = reinterpret_cast<const uint32_t *>(lds + table_bytes(PREC, XVk, p.x.ntab));
    else
      pat = p.x.pat;
    if constexpr (PREC == PR_FP64)
      pfield64<XVk>(p.x, src64, p.n, i, reinterpret_cast<const double *>(lds), pat, f0, f1, f2);
    else if constexpr (xv_pair(XVk))
      pairfield32<XVk>(p.x, src32, p.n, i, reinterpret_cast<const double *>(lds),
                       reinterpret_cast<const int32_t *>(pat), p.x.pat, f0, f1, f2);
    else if constexpr (xv_uniform(XVk))
      ufield32<XVk>(p.x, src32, p.n, i, reinterpret_cast<const double *>(lds),
                    reinterpret_cast<const int32_t *>(pat), p.x.pat, f0, f1, f2);
    else
      pfield32<XVk>(p.x, src32, p.n, i, reinterpret_cast<const double *>(lds), pat, f0, f1, f2);
  } else if constexpr (PREC == PR_FP64) {
    xfield64<XVk>(p.x, src64, p.n, i, reinterpret_cast<const double *>(lds), f0, f1, f2);
  } else {
    xfield32<XVk>(p.x, src32, p.n, i, reinterpret_cast<const double *>(lds), f0, f1, f2);
  }
}

__device__ __forceinline__ double cparam(const KParams &p, uint32_t c, int f) {
  return p.ctab[c * CF_STRIDE + f];
}

// Noise stream of a spin: the seed of its replica and its index within it
// (reference order), so a replica draws the same numbers batched or alone.
__device__ __forceinline__ uint64_t noise_seed(const KParams &p, uint32_t c) {
  return (uint64_t)__double_as_longlong(cparam(p, c, CF_SEED));
}
__device__ __forceinline__ uint64_t noise_index_of_ref(const KParams &p, uint32_t c, int64_t ri) {
  return (uint64_t)ri - (uint64_t)__double_as_longlong(cparam(p, c, CF_ROFF));
}
__device__ __forceinline__ uint64_t noise_index(const KParams &p, uint32_t c, int64_t i) {
  return noise_index_of_ref(p, c, ref_index(p, i));
}

// Local (non-exchange) term q of spin i with spin vector (s0, s1, s2).
// GEN = false compiles only the common forms (static Zeeman, K1 anisotropy);
// GEN = true adds the ac Zeeman field and the K2/K3 powers.
template <bool GEN>
__device__ __forceinline__ void local_term(const KParams &p, int kind, uint32_t c, int64_t i,
                                           double s0, double s1, double s2, double t, double &f0,
                                           double &f1, double &f2) {
  if (kind == TK_ZEEMAN) {
    // zeeman.cc:121-132: field = dc; field += ac * cos(omega t)
    f0 = cparam(p, c, CF_ZDC + 0);
    f1 = cparam(p, c, CF_ZDC + 1);
    f2 = cparam(p, c, CF_ZDC + 2);
    if (GEN && p.has_ac) {
      const double cw = cos(cparam(p, c, CF_ZOMEGA) * t);
      f0 = f0 + cparam(p, c, CF_ZAC + 0) * cw;
      f1 = f1 + cparam(p, c, CF_ZAC + 1) * cw;
      f2 = f2 + cparam(p, c, CF_ZAC + 2) * cw;
    }
  } else if (kind == TK_UNIAXIAL) {
    // uniaxial_anisotropy.cc:155-163: K p (s.e)^(p-1) e
    const double e0 = cparam(p, c, CF_AXIS + 0), e1 = cparam(p, c, CF_AXIS + 1),
                 e2 = cparam(p, c, CF_AXIS + 2);
    const double d = (e0 * s0 + e1 * s1 + e2 * s2);
    // pow(d, 1) == d exactly, so K1 needs no library call
    const double pw = (!GEN || p.ani_power == 2) ? d : pow(d, (double)(p.ani_power - 1));
    const double a = cparam(p, c, CF_K) * (double)p.ani_power * pw;
    f0 = a * e0;
    f1 = a * e1;
    f2 = a * e2;
  } else {  // TK_DIPOLE: precomputed field array
    f0 = p.hext[i];
    f1 = p.hext[p.n + i];
    f2 = p.hext[2 * p.n + i];
  }
}

// Total field (meV) of spin i, terms summed in config order
// (core/solver.cc:43-57: h = f0; h = h + 1.0 * f_k), with the exchange term's
// field (x0, x1, x2) already computed.
template <bool GEN>
__device__ __forceinline__ void sum_terms(const KParams &p, uint32_t c, int64_t i, double s0,
                                          double s1, double s2, double t, double x0, double x1,
                                          double x2, double &h0, double &h1, double &h2) {
  for (int q = 0; q < p.nterms; ++q) {
    double f0, f1, f2;
    const int k = p.kind[q];
    if (k == TK_EXCHANGE) {
      f0 = x0; f1 = x1; f2 = x2;
    } else {
      local_term<GEN>(p, k, c, i, s0, s1, s2, t, f0, f1, f2);
    }
    if (q == 0) {
      h0 = f0; h1 = f1; h2 = f2;
    } else {
      h0 = h0 + f0; h1 = h1 + f1; h2 = h2 + f2;
    }
  }
  if (p.nterms == 0) { h0 = h1 = h2 = 0.0; }
}

template <int PREC, int XVk, bool GEN>
__device__ __forceinline__ void total_field(const KParams &p, const double *src64,
                                            const float *src32, int64_t i, uint32_t c,
                                            const unsigned char *lds, double s0, double s1,
                                            double s2, double t, double &h0, double &h1,
                                            double &h2) {
  double x0, x1, x2;
  exchange_at<PREC, XVk>(p, src64, src32, i, lds, x0, x1, x2);
  sum_terms<GEN>(p, c, i, s0, s1, s2, t, x0, x1, x2, h0, h1, h2);
}

// step counter and time of this launch: from the host, or in a graph replay
// the replay's first step (on the device) plus this step's place in the
// graph (t = step dt, the host's own expression)
__device__ __forceinline__ uint64_t launch_step(const KParams &p) {
  return p.dstep ? *p.dstep + p.step : p.step;
}
__device__ __forceinline__ double launch_time(const KParams &p, uint64_t step) {
  return p.dstep ? (double)step * p.dt : p.t;
}

// Renormalisation: the reference's unit_vector in FP64; one reciprocal and
// three products in the FP32-field form (within 1.5 ulp; the predictor and
// the corrector's recomputation of s* use the same one)
template <int PREC>
__device__ __forceinline__ void renorm(double &v0, double &v1, double &v2) {
  if constexpr (PREC == PR_FP64)
    unit_vector(v0, v1, v2);
  else
    unit_vector_rcp(v0, v1, v2);
}

// What a corrector reads of its own spin besides s*: s (the step's start)
// and ds.  The FP32-field corrector loads them before its field (it forms s*
// from them) and keeps them; the FP64 one loads them after its field (fewer
// live registers across the neighbour loop).
struct OwnState {
  double s0, s1, s2, d0, d1, d2;
};
template <int PREC>
constexpr bool kEarlyOwn = PREC == PR_FP32;
__device__ __forceinline__ void load_own(const KParams &p, int64_t i, OwnState &own) {
  const int64_t n = p.n;
  load_spin(p.s, n, i, own.s0, own.s1, own.s2);
  own.d0 = p.ds[i]; own.d1 = p.ds[n + i]; own.d2 = p.ds[2 * n + i];
}

// The spin a Heun phase evaluates the field at: s (predictor) or s*
// (corrector).
template <int PREC, int PHASE>
__device__ __forceinline__ void phase_spin(const KParams &p, int64_t i, double &s0, double &s1,
                                           double &s2, OwnState &own) {
  const int64_t n = p.n;
  if constexpr (PHASE == 1 && kEarlyOwn<PREC>) load_own(p, i, own);
  if (PREC == PR_FP32 && PHASE == 1 && p.spfree) {
    // s* = unit(s + dt rhs) as the predictor formed it: 2 ds is its rhs
    // exactly (ds = rhs / 2; only a subnormal rhs would lose its last bit)
    s0 = own.s0 + p.dt * (2.0 * own.d0);
    s1 = own.s1 + p.dt * (2.0 * own.d1);
    s2 = own.s2 + p.dt * (2.0 * own.d2);
    renorm<PREC>(s0, s1, s2);
  } else {
    load_spin(PHASE == 0 ? p.s : p.sp, n, i, s0, s1, s2);
  }
}

// Everything of a Heun phase after the exchange field (x0, x1, x2) of spin i
// at (s0, s1, s2): the ordered field sum, the thermal field, h / mus, the LLG
// right-hand side and the phase's update and stores.
// The arithmetic of a Heun phase for spin i (reference index ri) after its
// exchange field (x0, x1, x2) at (s0, s1, s2): the ordered field sum, the
// thermal field, h / mus, the LLG right-hand side and the update.  w is the
// step's thermal
// field: read by the caller in the corrector (or injected), drawn here in the
// predictor from the Philox stream.  Out: the predictor's ds = rhs / 2 in
// (d0, d1, d2) and s* in (v0, v1, v2); the corrector's new s in (v0, v1, v2)
// from its own s and ds in `own`.
template <int PREC, int PHASE, bool GEN>
__device__ __forceinline__ void heun_math(const KParams &p, int64_t i, int64_t ri, uint32_t c,
                                          uint64_t step,
                                          double t, double s0, double s1, double s2, double x0,
                                          double x1, double x2, const OwnState &own, double &w0,
                                          double &w1, double &w2, double &d0, double &d1,
                                          double &d2, double &v0, double &v1, double &v2) {
  double h0, h1, h2;
  sum_terms<GEN>(p, c, i, s0, s1, s2, t, x0, x1, x2, h0, h1, h2);

  // thermal field w of this step (solvers/cpu_llg_heun.cc:53-62): one draw
  // per step serves both phases, so the predictor generates it and leaves it
  // for the corrector
  if (p.noise == NM_PHILOX && PHASE == 0) {
    double z0, z1, z2, sT;
    if (p.cls) {
      normals3<false>(noise_seed(p, c), step, noise_index_of_ref(p, c, ri), z0, z1, z2);
      sT = cparam(p, c, CF_SQRTT);
    } else {
      // one site class: the key schedule stays in scalar registers
      normals3<true>(p.seed, step, (uint64_t)ri - p.roff, z0, z1, z2);
      sT = p.sqrtT;
    }
    const double sg = cparam(p, c, CF_SIGMA);
    w0 = z0 * sg * sT; w1 = z1 * sg * sT; w2 = z2 * sg * sT;
  }

  // h = w + h / mus  (solvers/cpu_llg_heun.cc:68-81); the FP32-field form
  // multiplies by one reciprocal
  const double mus = cparam(p, c, CF_MUS);
  if constexpr (PREC == PR_FP32) {
    const double rm = 1.0 / mus;
    h0 = h0 * rm; h1 = h1 * rm; h2 = h2 * rm;
    if (p.noise != NM_NONE) { h0 = w0 + h0; h1 = w1 + h1; h2 = w2 + h2; }
  } else if (p.noise == NM_NONE) {
    h0 = h0 / mus; h1 = h1 / mus; h2 = h2 / mus;
  } else {
    h0 = w0 + h0 / mus; h1 = w1 + h1 / mus; h2 = w2 + h2 / mus;
  }

  double r0, r1, r2;
  llg_rhs(s0, s1, s2, h0, h1, h2, cparam(p, c, CF_GYRO), cparam(p, c, CF_ALPHA), r0, r1, r2);

  if constexpr (PHASE == 0) {
    d0 = 0.5 * r0; d1 = 0.5 * r1; d2 = 0.5 * r2;
    v0 = s0 + p.dt * r0; v1 = s1 + p.dt * r1; v2 = s2 + p.dt * r2;
    renorm<PREC>(v0, v1, v2);
  } else {
    d0 = own.d0 + 0.5 * r0;
    d1 = own.d1 + 0.5 * r1;
    d2 = own.d2 + 0.5 * r2;
    v0 = own.s0; v1 = own.s1; v2 = own.s2;
    v0 = v0 + p.dt * d0; v1 = v1 + p.dt * d1; v2 = v2 + p.dt * d2;
    renorm<PREC>(v0, v1, v2);
  }
}

// The thermal field a phase reads: the predictor's draw (corrector) or the
// injected stream; zero otherwise (the predictor draws its own)
template <int PHASE>
__device__ __forceinline__ void read_noise(const KParams &p, int64_t i, double &w0, double &w1,
                                           double &w2) {
  const int64_t n = p.n;
  w0 = w1 = w2 = 0.0;
  if (p.noise == NM_PHILOX && PHASE == 1) {
    w0 = p.wbuf[i]; w1 = p.wbuf[n + i]; w2 = p.wbuf[2 * n + i];
  } else if (p.noise == NM_INJECTED) {
    w0 = p.winj[i]; w1 = p.winj[n + i]; w2 = p.winj[2 * n + i];
  }
}

// Everything of a Heun phase after the exchange field (x0, x1, x2) of spin i
// at (s0, s1, s2): heun_math and the phase's loads and stores.
template <int PREC, int PHASE, bool GEN>
__device__ __forceinline__ void heun_spin(const KParams &p, int64_t i, uint32_t c, uint64_t step,
                                          double t, double s0, double s1, double s2, double x0,
                                          double x1, double x2, const OwnState &own) {
  const int64_t n = p.n;
  double w0, w1, w2;
  read_noise<PHASE>(p, i, w0, w1, w2);
  OwnState o = own;
  if constexpr (PHASE == 1 && !kEarlyOwn<PREC>) load_own(p, i, o);
  double d0, d1, d2, v0, v1, v2;
  heun_math<PREC, PHASE, GEN>(p, i, ref_index(p, i), c, step, t, s0, s1, s2, x0, x1, x2, o, w0,
                              w1, w2, d0, d1,
                              d2, v0, v1, v2);
  if (p.ntstore) {
    if constexpr (PHASE == 0) {
      if (p.noise == NM_PHILOX) {
        __builtin_nontemporal_store(w0, p.wbuf + i);
        __builtin_nontemporal_store(w1, p.wbuf + n + i);
        __builtin_nontemporal_store(w2, p.wbuf + 2 * n + i);
      }
      __builtin_nontemporal_store(d0, p.ds + i);
      __builtin_nontemporal_store(d1, p.ds + n + i);
      __builtin_nontemporal_store(d2, p.ds + 2 * n + i);
      if (!(PREC == PR_FP32 && p.spfree)) store_spin_nt(p.sp, n, i, v0, v1, v2);
    } else {
      store_spin_nt(p.snew, n, i, v0, v1, v2);
    }
  } else if constexpr (PHASE == 0) {
    if (p.noise == NM_PHILOX) { p.wbuf[i] = w0; p.wbuf[n + i] = w1; p.wbuf[2 * n + i] = w2; }
    p.ds[i] = d0; p.ds[n + i] = d1; p.ds[2 * n + i] = d2;
    if (!(PREC == PR_FP32 && p.spfree)) store_spin(p.sp, n, i, v0, v1, v2);
  } else {
    store_spin(p.snew, n, i, v0, v1, v2);
  }
  if constexpr (PREC == PR_FP32) store_g32(p.w32, i, v0, v1, v2);
}

// Minimum waves per SIMD the compiler keeps room for (VGPR budget): the
// wave-uniform FP32 variants are latency-bound and gain from occupancy.
#ifndef JH_UNIFORM_WAVES
#define JH_UNIFORM_WAVES 1
#endif
#ifndef JH_DIAG_WAVES
#define JH_DIAG_WAVES 1
#endif
template <int PREC, int XVk>
constexpr int kMinWaves = (PREC == PR_FP32 && xv_uniform(XVk)) ? JH_UNIFORM_WAVES
                          : xv_diag(XVk) ? JH_DIAG_WAVES : JH_WAVES_PER_EU;

// The spin block a gathering step kernel's workgroup takes for its XCD-
// ordered block number blk.  With basis_ilv the workgroups alternate between
// the basis sites' index ranges, so the workgroups of one XCD hold the same
// cells of every basis site and the other sites' spins their rows gather stay
// in that XCD's L2.
__device__ __forceinline__ int64_t step_block(const KParams &p, int64_t blk) {
  if (p.basis_ilv && !p.rowlist && p.order_nb > 1 && p.order_cells % blockDim.x == 0) {
    const int64_t per = p.order_cells / blockDim.x;
    blk = (blk % p.order_nb) * per + blk / p.order_nb;
  }
  return blk;
}

// How the gathering step kernels find a spin's site class (its row of the
// class table: mus, gyro, alpha, sigma, the noise seed and offset, the local
// terms).  A per-lane class index makes every parameter one vector load per
// spin; a wave-uniform one lets the compiler read the table through the
// scalar unit.
//   CM_LANE:  per lane (the GEN forms).
//   CM_SPLIT: contexts with a class table (replica batches, several
//             materials): when every lane of the wave has the same class,
//             which holds for all but the waves that straddle a replica
//             boundary, the update runs with the first lane's class; the
//             other waves run it per lane.
//   CM_ONE:   no class table: every spin has class 0, known at compile time.
enum { CM_LANE = 0, CM_SPLIT = 1, CM_ONE = 2 };

// One Heun phase of spin i with the gathering exchange formats: the field at
// the phase's spins and the phase's update and stores.
template <int PREC, int XVk, int PHASE, bool GEN, int CM = CM_LANE>
__device__ __forceinline__ void heun_gather_spin(const KParams &p, int64_t i,
                                                 const unsigned char *lds) {
  const uint32_t c = CM == CM_ONE ? 0u : p.cls ? (uint32_t)p.cls[i] : 0u;
  const uint64_t step = launch_step(p);
  const double t0 = launch_time(p, step);
  double s0, s1, s2;
  OwnState own;
  phase_spin<PREC, PHASE>(p, i, s0, s1, s2, own);
  const double t = PHASE == 0 ? t0 : t0 + p.dt;
  double x0, x1, x2;
  exchange_at<PREC, XVk>(p, PHASE == 0 ? p.s : p.sp, p.g32, i, lds, x0, x1, x2);
  if constexpr (CM == CM_SPLIT) {
    const uint32_t c0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)c);
    if (__builtin_amdgcn_ballot_w64(c != c0) == 0)
      heun_spin<PREC, PHASE, GEN>(p, i, c0, step, t, s0, s1, s2, x0, x1, x2, own);
    else
      heun_spin<PREC, PHASE, GEN>(p, i, c, step, t, s0, s1, s2, x0, x1, x2, own);
  } else {
    heun_spin<PREC, PHASE, GEN>(p, i, c, step, t, s0, s1, s2, x0, x1, x2, own);
  }
}

template <int PREC, int XVk, int PHASE, bool GEN, int CM>
__global__ void __launch_bounds__(kMaxBlock)
__attribute__((amdgpu_waves_per_eu(kMinWaves<PREC, XVk>))) k_heun(const KParams p) {
  extern __shared__ __align__(16) unsigned char lds[];
  load_table<PREC, XVk>(p.x, lds);
  if constexpr (xv_typed(XVk)) __syncthreads();
  int64_t i = step_block(p, xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
  if (p.rowlist) {
    if (i >= p.nrowlist) return;
    i = p.rowlist[i];
  } else if (i >= p.n) {
    return;
  }
  heun_gather_spin<PREC, XVk, PHASE, GEN, CM>(p, i, lds);
}

// Grid barrier of the persistent step kernel over sync = {arrived count,
// generation, failed flag} in device memory (zero before the first launch;
// each barrier leaves the count at zero again).  Every work-item publishes
// its stores at agent scope first (the release fence writes back the XCD's
// L2, which other XCDs do not share), then one work-item per workgroup counts
// the workgroup in: the last to arrive resets the count and advances the
// generation, the others poll the generation.  The polling is bounded: after
// `limit` polls a workgroup sets the failed flag and leaves, so every wave
// ends even if the grid were not resident at once (the host launches at most
// as many workgroups as fit, and reports the flag).  The acquire fence after
// the barrier drops stale lines before the next phase's loads.
__device__ __forceinline__ bool grid_barrier(unsigned *sync, unsigned limit) {
  __shared__ int passed;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned gen = __hip_atomic_load(sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned before =
        __hip_atomic_fetch_add(sync, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (before == gridDim.x - 1) {
      __hip_atomic_store(sync, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(sync + 1, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned polls = 0;
      while (__hip_atomic_load(sync + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        if (++polls > limit) {
          __hip_atomic_store(sync + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(4);
      }
    }
    passed = ok;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return passed != 0;
}

// A whole Heun step in one persistent launch (the batching form: every
// replica's predictor, one grid barrier, every replica's corrector).  The
// grid holds at most as many workgroups as are resident at once; workgroup w
// takes the spin blocks w, w + G, w + 2G, ... of the XCD-ordered numbering
// (with G a multiple of 8 each XCD keeps a contiguous range, as in k_heun).
// The tables are staged in LDS once for both phases.  The corrector reads the
// FP32 copy the predictor wrote and writes the other (g32 and w32 swap).
template <int PREC, int XVk, bool GEN>
__global__ void __launch_bounds__(kMaxBlock)
__attribute__((amdgpu_waves_per_eu(kMinWaves<PREC, XVk>)))
k_heun_step(const KParams p, unsigned *sync, unsigned limit) {
  extern __shared__ __align__(16) unsigned char lds[];
  load_table<PREC, XVk>(p.x, lds);
  if constexpr (xv_typed(XVk)) __syncthreads();
  const int64_t nv = (p.n + blockDim.x - 1) / blockDim.x;
  for (int64_t v = blockIdx.x; v < nv; v += gridDim.x) {
    const int64_t i = step_block(p, xcd_block(v, nv)) * blockDim.x + threadIdx.x;
    if (i < p.n) heun_gather_spin<PREC, XVk, 0, GEN>(p, i, lds);
  }
  if (!grid_barrier(sync, limit)) return;
  KParams q = p;
  q.g32 = p.w32;
  q.w32 = const_cast<float *>(p.g32);
  for (int64_t v = blockIdx.x; v < nv; v += gridDim.x) {
    const int64_t i = step_block(q, xcd_block(v, nv)) * blockDim.x + threadIdx.x;
    if (i < q.n) heun_gather_spin<PREC, XVk, 1, GEN>(q, i, lds);
  }
}

// ---------------------------------------------------------------------------
// STENCIL step kernels (StView in jh_internal.h).
//
// LDS tile of E elements, one spin each: FP64 as a z plane (8 E bytes, z of
// element e at 8 e) followed by an (x, y) plane ((x, y) at 8 E + 16 e), so
// lanes read consecutive 8- and 16-byte words (no bank conflicts); FP32 as
// (x, y, z, 0) at 16 e.  Cells outside an open boundary hold zero spins:
// their products add +-0 to a sum that starts at +0, which leaves it
// unchanged.
template <int PREC>
constexpr int kStElem = PREC == PR_FP64 ? 24 : 16;

__device__ __forceinline__ bool st_wrap(int &a, int l, int periodic) {
  if (a < 0) {
    if (!periodic) return false;
    a += l;
  } else if (a >= l) {
    if (!periodic) return false;
    a -= l;
  }
  return true;
}

// Stage the workgroup's tile: wave w loads tile rows w, w + nw, ... (a row is
// one (basis site, x, y) line of tz cells along z, lanes along z).
template <int PREC>
__device__ __forceinline__ void st_load_tile(const KParams &p, const StView &st,
                                             const double *__restrict__ src64,
                                             const float *__restrict__ src32, int b, int X0,
                                             int Y0, int Z0, unsigned char *lds) {
  const int TX = st.tx[b], TY = st.ty[b], TZ = st.tz[b];
  const int X1 = X0 + st.ox[b], Y1 = Y0 + st.oy[b], Z1 = Z0 + st.oz[b];
  const int rows = st.nb * TX * TY;
  const int E = rows * TZ;
  const int nw = (int)(blockDim.x >> 6);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int64_t C = (int64_t)st.lx * st.ly * st.lz;
  const int64_t n = p.n;
  for (int row = wave; row < rows; row += nw) {
    const int bb = row / (TX * TY);
    const int rem = row - bb * (TX * TY);
    int xx = X1 + rem / TY, yy = Y1 + rem % TY;
    const bool okxy = st_wrap(xx, st.lx, st.per[0]) && st_wrap(yy, st.ly, st.per[1]);
    const int64_t rowbase = (int64_t)bb * C + ((int64_t)xx * st.ly + yy) * st.lz;
    for (int tz = lane; tz < TZ; tz += 64) {
      int zz = Z1 + tz;
      const bool ok = okxy && st_wrap(zz, st.lz, st.per[2]);
      const int e = row * TZ + tz;
      if constexpr (PREC == PR_FP64) {
        double x = 0.0, y = 0.0, z = 0.0;
        if (ok) load_spin(src64, n, rowbase + zz, x, y, z);
        reinterpret_cast<double *>(lds)[e] = z;
        reinterpret_cast<double2 *>(lds + (size_t)8 * E)[e] = make_double2(x, y);
      } else {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok) {
          const float *q = src32 + (size_t)kG32 * (rowbase + zz);
          v = make_float4(q[0], q[1], q[2], 0.f);
        }
        reinterpret_cast<float4 *>(lds)[e] = v;
      }
    }
  }
}

// Exchange field of the lane whose tile element starts at byte `base`, over
// the entries of sequence sid.  FP64: the reference order with one rounding
// per product and per sum; FP32 form: one fused multiply-add per product into
// FP64 sums.
template <int PREC, bool DIAG>
__device__ __forceinline__ void st_field(const StView &st, int sid, const unsigned char *lds,
                                         uint32_t xybase, uint32_t base, double &h0, double &h1,
                                         double &h2) {
  const int2 sq = st.seq[sid];
  const int first = __builtin_amdgcn_readfirstlane(sq.x);
  const int len = __builtin_amdgcn_readfirstlane(sq.y);
  const int32_t *__restrict__ off = st.off + first;
  constexpr int V = DIAG ? 3 : 9;
  const double *__restrict__ T = st.val + (size_t)first * V;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll 8
  for (int k = 0; k < len; ++k) {
    const uint32_t a = base + (uint32_t)off[k];
    const unsigned char *q = lds + a;
    double sx, sy, sz;
    if constexpr (PREC == PR_FP64) {
      sz = *reinterpret_cast<const double *>(q);
      const double2 v = *reinterpret_cast<const double2 *>(lds + (xybase + 2 * a));
      sx = v.x; sy = v.y;
    } else {
      const float4 v = *reinterpret_cast<const float4 *>(q);
      sx = v.x; sy = v.y; sz = v.z;
    }
    const double *Tk = T + V * k;
    if constexpr (PREC == PR_FP64) {
      if constexpr (DIAG) {
        a0 = a0 + sx * Tk[0];
        a1 = a1 + sy * Tk[1];
        a2 = a2 + sz * Tk[2];
      } else {
        a0 = a0 + sx * Tk[0]; a0 = a0 + sy * Tk[1]; a0 = a0 + sz * Tk[2];
        a1 = a1 + sx * Tk[3]; a1 = a1 + sy * Tk[4]; a1 = a1 + sz * Tk[5];
        a2 = a2 + sx * Tk[6]; a2 = a2 + sy * Tk[7]; a2 = a2 + sz * Tk[8];
      }
    } else {
      if constexpr (DIAG) {
        a0 = __builtin_fma(sx, Tk[0], a0);
        a1 = __builtin_fma(sy, Tk[1], a1);
        a2 = __builtin_fma(sz, Tk[2], a2);
      } else {
        a0 = __builtin_fma(sx, Tk[0], a0); a0 = __builtin_fma(sy, Tk[1], a0); a0 = __builtin_fma(sz, Tk[2], a0);
        a1 = __builtin_fma(sx, Tk[3], a1); a1 = __builtin_fma(sy, Tk[4], a1); a1 = __builtin_fma(sz, Tk[5], a1);
        a2 = __builtin_fma(sx, Tk[6], a2); a2 = __builtin_fma(sy, Tk[7], a2); a2 = __builtin_fma(sz, Tk[8], a2);
      }
    }
  }
  h0 = a0; h1 = a1; h2 = a2;
}

// One Heun phase over the STENCIL layout: workgroup -> (basis site b, block
// of wx x wy lines, 64 z), wave -> line, lane -> z.
template <int PREC, bool DIAG, int PHASE, bool GEN>
__global__ void __launch_bounds__(kStMaxBlock) k_heun_st(const KParams p, const StView st) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int bz = (int)(lb % st.nbz);
  int64_t r = lb / st.nbz;
  const int by = (int)(r % st.nby);
  r /= st.nby;
  const int bx = (int)(r % st.nbx);
  const int b = (int)(r / st.nbx);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int xi = wave / st.wy, yi = wave - (wave / st.wy) * st.wy;
  const int X0 = bx * st.wx, Y0 = by * st.wy, Z0 = bz * 64;
  const int x = X0 + xi, y = Y0 + yi;
  const int64_t C = (int64_t)st.lx * st.ly * st.lz;
  const int64_t i = (int64_t)b * C + ((int64_t)x * st.ly + y) * st.lz + Z0 + lane;

  st_load_tile<PREC>(p, st, PHASE == 0 ? p.s : p.sp, p.g32, b, X0, Y0, Z0, lds);
  __syncthreads();

  const uint32_t c = p.cls ? (uint32_t)p.cls[i] : 0u;
  const uint64_t step = launch_step(p);
  const double t0 = launch_time(p, step);
  double s0, s1, s2;
  OwnState own;
  phase_spin<PREC, PHASE>(p, i, s0, s1, s2, own);
  const double t = PHASE == 0 ? t0 : t0 + p.dt;
  const int sid = st.wseq ? st.wseq[((int64_t)b * st.lx + x) * st.ly + y] : b;
  const uint32_t e0 = (uint32_t)((xi * st.ty[b] + yi) * st.tz[b] + lane);
  const uint32_t E = (uint32_t)(st.nb * st.tx[b] * st.ty[b] * st.tz[b]);
  double x0, x1, x2;
  st_field<PREC, DIAG>(st, sid, lds, 8 * E, e0 * (PREC == PR_FP64 ? 8u : 16u), x0, x1, x2);
  if (st.fixmask) {
    const uint64_t fm = st.fixmask[lb * (int64_t)(blockDim.x >> 6) + wave];
    if ((fm >> lane) & 1ull) return;  // this row is the row-list launch's
  }
  heun_spin<PREC, PHASE, GEN>(p, i, c, step, t, s0, s1, s2, x0, x1, x2, own);
}

// ---------------------------------------------------------------------------
// LINE step kernels (LnView in jh_internal.h): wave -> one (basis site, x, y)
// line of Lz = 64 P cells along z, cell z = P lane + h in lane `lane`.

// Wave-wide lane rotation by one (DPP): lane l reads lane l + 1 (wave_rol)
// or lane l - 1 (wave_ror), mod 64.  Every lane of the wave is active
// wherever these run, so each reads a live value.
constexpr int kDppRol = 0x134;
constexpr int kDppRor = 0x13C;
template <int CTRL>
__device__ __forceinline__ float lane_rot(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double lane_rot(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// the value lane l + S holds (|S| <= 2)
template <int S, typename T>
__device__ __forceinline__ T lane_shift(T v) {
  if constexpr (S > 0) return lane_shift<S - 1>(lane_rot<kDppRol>(v));
  else if constexpr (S < 0) return lane_shift<S + 1>(lane_rot<kDppRor>(v));
  else return v;
}

// Two adjacent spins i0, i0 + 1 (i0 even, so every access is 16-byte
// aligned) of an FP64 spin array ((x, y) pairs, then z) or SoA vector, and
// of the FP32 gather copy (24 bytes, 8-byte aligned).
__device__ __forceinline__ void ld_spin2(const double *__restrict__ a, int64_t n, int64_t i0,
                                         double (&v)[2][3]) {
  const double2 p0 = reinterpret_cast<const double2 *>(a)[i0];
  const double2 p1 = reinterpret_cast<const double2 *>(a)[i0 + 1];
  const double2 z = *reinterpret_cast<const double2 *>(a + 2 * n + i0);
  v[0][0] = p0.x; v[0][1] = p0.y; v[0][2] = z.x;
  v[1][0] = p1.x; v[1][1] = p1.y; v[1][2] = z.y;
}
// Two adjacent doubles at p (16-byte aligned), non-temporal (nt) or not.
// The step kernels store non-temporally when p.ntstore says so: a phase's
// outputs are read only by the next launch, and at millions of spins they
// outgrow the caches anyway, so streaming them past the L2 leaves it to the
// neighbour spins that the waves in flight re-read.
typedef double nt_double2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_d2(double *p, double x, double y, bool nt) {
  if (nt)
    __builtin_nontemporal_store((nt_double2){x, y}, reinterpret_cast<nt_double2 *>(p));
  else
    *reinterpret_cast<double2 *>(p) = make_double2(x, y);
}
__device__ __forceinline__ void st_spin2(double *a, int64_t n, int64_t i0, const double (&v)[2][3],
                                         bool nt) {
  st_d2(a + 2 * i0, v[0][0], v[0][1], nt);
  st_d2(a + 2 * (i0 + 1), v[1][0], v[1][1], nt);
  st_d2(a + 2 * n + i0, v[0][2], v[1][2], nt);
}
__device__ __forceinline__ void ld_soa2(const double *__restrict__ a, int64_t n, int64_t i0,
                                        double (&v)[2][3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double2 q = *reinterpret_cast<const double2 *>(a + c * n + i0);
    v[0][c] = q.x;
    v[1][c] = q.y;
  }
}
__device__ __forceinline__ void st_soa2(double *a, int64_t n, int64_t i0, const double (&v)[2][3],
                                        bool nt) {
#pragma unroll
  for (int c = 0; c < 3; ++c) st_d2(a + c * n + i0, v[0][c], v[1][c], nt);
}
__device__ __forceinline__ void st_g32_2(float *dst, int64_t i0, const double (&v)[2][3]) {
  static_assert(kG32 == 3, "the two-spin FP32 store assumes 12-byte spins");
  typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
  typedef float f2a8 __attribute__((ext_vector_type(2), aligned(8)));
  float *q = dst + (size_t)3 * i0;
  f4a8 u;
  u.x = (float)v[0][0]; u.y = (float)v[0][1]; u.z = (float)v[0][2]; u.w = (float)v[1][0];
  f2a8 w;
  w.x = (float)v[1][1]; w.y = (float)v[1][2];
  *reinterpret_cast<f4a8 *>(q) = u;
  *reinterpret_cast<f2a8 *>(q + 4) = w;
}

template <int PREC>
using LnVal = typename std::conditional<PREC == PR_FP64, double, float>::type;
template <int PREC, int P>
struct LnLine {
  LnVal<PREC> v[P][3];  // cell P lane + h: v[h][component]
};

// Load line (bb, xx, yy) (wrapped coordinates): lane `lane` gets its P cells,
// coalesced (FP64: the (x, y) pairs as P 16-byte loads and the z values as
// one P * 8-byte load; FP32: the P * 12 bytes of the gather copy).
template <int PREC, int P>
__device__ __forceinline__ void ln_load(const LnView &ln, const double *__restrict__ src64,
                                        const float *__restrict__ src32, int64_t n, int bb,
                                        int xx, int yy, int lane, LnLine<PREC, P> &L) {
  // the line's first cell (wave-uniform, 32-bit: the host keeps N below
  // 2^31 / 3) and the lane's cell within it
  const uint32_t cell0 = (((uint32_t)bb * (uint32_t)ln.lx + (uint32_t)xx) * (uint32_t)ln.ly +
                          (uint32_t)yy) * (uint32_t)ln.lz;
  const uint32_t lc = (uint32_t)(P * lane);
  if constexpr (PREC == PR_FP64) {
    const double2 *xy = reinterpret_cast<const double2 *>(src64) + cell0;
#pragma unroll
    for (int h = 0; h < P; ++h) {
      const double2 v = xy[lc + h];
      L.v[h][0] = v.x;
      L.v[h][1] = v.y;
    }
    const double *zl = src64 + 2 * n + cell0;
    if constexpr (P == 2) {
      const double2 z = *reinterpret_cast<const double2 *>(zl + lc);
      L.v[0][2] = z.x;
      L.v[1][2] = z.y;
    } else {
      L.v[0][2] = zl[lc];
    }
  } else if constexpr (P == 2 && kG32 == 3) {
    // the lane's 24 bytes (8-byte aligned) as one 16- and one 8-byte load,
    // taken apart in registers
    typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
    typedef float f2a8 __attribute__((ext_vector_type(2), aligned(8)));
    const float *q = src32 + (size_t)kG32 * cell0 + kG32 * lc;
    const f4a8 u = *reinterpret_cast<const f4a8 *>(q);
    const f2a8 w = *reinterpret_cast<const f2a8 *>(q + 4);
    L.v[0][0] = u.x; L.v[0][1] = u.y; L.v[0][2] = u.z;
    L.v[1][0] = u.w; L.v[1][1] = w.x; L.v[1][2] = w.y;
  } else {
    const float *q = src32 + (size_t)kG32 * cell0 + kG32 * lc;
#pragma unroll
    for (int h = 0; h < P; ++h) {
      L.v[h][0] = q[kG32 * h];
      L.v[h][1] = q[kG32 * h + 1];
      L.v[h][2] = q[kG32 * h + 2];
    }
  }
}

// one neighbour's product into (a0, a1, a2): FP64 in the reference's order
// with one rounding per product and per sum; FP32 form as FP64 FMAs
template <int PREC, bool DIAG>
__device__ __forceinline__ void ln_acc(const double *__restrict__ T, double sx, double sy,
                                       double sz, double &a0, double &a1, double &a2) {
  if constexpr (PREC == PR_FP64) {
    if constexpr (DIAG) {
      a0 = a0 + sx * T[0];
      a1 = a1 + sy * T[1];
      a2 = a2 + sz * T[2];
    } else {
      a0 = a0 + sx * T[0]; a0 = a0 + sy * T[1]; a0 = a0 + sz * T[2];
      a1 = a1 + sx * T[3]; a1 = a1 + sy * T[4]; a1 = a1 + sz * T[5];
      a2 = a2 + sx * T[6]; a2 = a2 + sy * T[7]; a2 = a2 + sz * T[8];
    }
  } else {
    if constexpr (DIAG) {
      a0 = __builtin_fma(sx, T[0], a0);
      a1 = __builtin_fma(sy, T[1], a1);
      a2 = __builtin_fma(sz, T[2], a2);
    } else {
      a0 = __builtin_fma(sx, T[0], a0); a0 = __builtin_fma(sy, T[1], a0); a0 = __builtin_fma(sz, T[2], a0);
      a1 = __builtin_fma(sx, T[3], a1); a1 = __builtin_fma(sy, T[4], a1); a1 = __builtin_fma(sz, T[5], a1);
      a2 = __builtin_fma(sx, T[6], a2); a2 = __builtin_fma(sy, T[7], a2); a2 = __builtin_fma(sz, T[8], a2);
    }
  }
}

// The neighbours at z + DZ of the lane's P cells from line L, as read (FP32
// or FP64).
template <int P, int DZ, typename Ln, typename T>
__device__ __forceinline__ void ln_gather(const Ln &L, T (&v)[P][3]) {
#pragma unroll
  for (int h = 0; h < P; ++h) {
    const int t = h + DZ;
    const int R = ((t % P) + P) % P;
    const int Q = (t - R) / P;
    // Q is -1, 0 or 1 (P = 2) or DZ (P = 1): resolved at compile time
    if (Q == 1) {
      v[h][0] = lane_shift<1>(L.v[R][0]); v[h][1] = lane_shift<1>(L.v[R][1]);
      v[h][2] = lane_shift<1>(L.v[R][2]);
    } else if (Q == -1) {
      v[h][0] = lane_shift<-1>(L.v[R][0]); v[h][1] = lane_shift<-1>(L.v[R][1]);
      v[h][2] = lane_shift<-1>(L.v[R][2]);
    } else if (Q == 2) {
      v[h][0] = lane_shift<2>(L.v[R][0]); v[h][1] = lane_shift<2>(L.v[R][1]);
      v[h][2] = lane_shift<2>(L.v[R][2]);
    } else if (Q == -2) {
      v[h][0] = lane_shift<-2>(L.v[R][0]); v[h][1] = lane_shift<-2>(L.v[R][1]);
      v[h][2] = lane_shift<-2>(L.v[R][2]);
    } else {
      v[h][0] = L.v[R][0]; v[h][1] = L.v[R][1]; v[h][2] = L.v[R][2];
    }
  }
}

struct LnGroup {
  uint32_t mask;  // entry codes present
  int vbase;      // values of the group's first entry
};

// Entry code CODE = slot | (dz + 2) << 1 of a group, if present: its
// neighbours come from slot CODE & 1 by the rotations of dz, its values are
// the next in the group's run (T advances).
template <int PREC, bool DIAG, int P, int CODE>
__device__ __forceinline__ void ln_code(const LnLine<PREC, P> (&S)[kLnSlots], uint32_t mask,
                                        const double *__restrict__ &T, double (&a)[P][3]) {
  constexpr int V = DIAG ? 3 : 9;
  if (mask & (1u << CODE)) {
    LnVal<PREC> v[P][3];
    ln_gather<P, (CODE >> 1) - 2>(S[CODE & 1], v);
#pragma unroll
    for (int h = 0; h < P; ++h)
      ln_acc<PREC, DIAG>(T, (double)v[h][0], (double)v[h][1], (double)v[h][2], a[h][0], a[h][1],
                         a[h][2]);
    T += V;
  }
}

// Pair code PC = (dz + 2) * 3 + delta + 1 of an FP32-form pair group, if
// present: slot 0's neighbours at dz and slot 1's at -dz + delta, whose
// tensors are J I + [D]x and its transpose, add J (s + s') + D x (s - s')
// (the FP32 spins' sums and differences, exact in FP64, and FP64 products
// and sums); (J, D) are the next 4 values (Tp advances).
template <bool DIAG, int P, int PC>
__device__ __forceinline__ void ln_pair(const LnLine<PR_FP32, P> (&S)[kLnSlots], uint32_t mask,
                                        const double *__restrict__ &Tp, double (&a)[P][3]) {
  if (mask & (1u << PC)) {
    constexpr int dz = PC / 3 - 2;
    constexpr int dz2 = -dz + (PC % 3 - 1);
    float u[P][3], w[P][3];
    ln_gather<P, dz>(S[0], u);
    ln_gather<P, dz2>(S[1], w);
    const double J = Tp[0];
#pragma unroll
    for (int h = 0; h < P; ++h) {
      // the sum and difference of two FP32 values are exact in FP64
      const double u0 = u[h][0], u1 = u[h][1], u2 = u[h][2];
      const double w0 = w[h][0], w1 = w[h][1], w2 = w[h][2];
      a[h][0] = __builtin_fma(J, u0 + w0, a[h][0]);
      a[h][1] = __builtin_fma(J, u1 + w1, a[h][1]);
      a[h][2] = __builtin_fma(J, u2 + w2, a[h][2]);
      if constexpr (!DIAG) {
        const double *D = Tp + 1;
        const double X = u0 - w0, Y = u1 - w1, Z = u2 - w2;
        a[h][0] = __builtin_fma(D[1], Z, a[h][0]); a[h][0] = __builtin_fma(-D[2], Y, a[h][0]);
        a[h][1] = __builtin_fma(D[2], X, a[h][1]); a[h][1] = __builtin_fma(-D[0], Z, a[h][1]);
        a[h][2] = __builtin_fma(D[0], Y, a[h][2]); a[h][2] = __builtin_fma(-D[1], X, a[h][2]);
      }
    }
    Tp += 4;
  }
}

// Exchange field of the wave's line (b, xw, yw) over its program sid.  The
// two lines of group g + 1 are loaded (buffer B, or A) while group g's
// entries sum, so a group's loads are in flight during the previous group's
// arithmetic; the steady-state loop issues its loads unconditionally and the
// last one or two groups are peeled, which keeps the compiler's memory waits
// exact.  A group's entries are its present codes in ascending order, each a
// wave-uniform branch over straight-line code.
template <int PREC, bool DIAG, int P>
__device__ __forceinline__ void ln_field(const LnView &ln, int sid, int xw, int yw,
                                         const double *__restrict__ src64,
                                         const float *__restrict__ src32, int64_t n, int lane,
                                         double (&a)[P][3]) {
  constexpr int V = DIAG ? 3 : 9;
#pragma unroll
  for (int h = 0; h < P; ++h) a[h][0] = a[h][1] = a[h][2] = 0.0;
  const int2 sq = ln.seq[sid];
  const int g0 = __builtin_amdgcn_readfirstlane(sq.x);
  const int G = __builtin_amdgcn_readfirstlane(sq.y);
  const int32_t *__restrict__ prog = ln.prog + (size_t)kLnGroupWords * g0;
  LnLine<PREC, P> A[kLnSlots], B[kLnSlots];
  LnGroup hA{}, hB{};
  auto load_group = [&](LnLine<PREC, P>(&S)[kLnSlots], LnGroup &hg, int g) {
    const int32_t *w = prog + (size_t)kLnGroupWords * g;
    hg.mask = (uint32_t)__builtin_amdgcn_readfirstlane(w[0]);
    hg.vbase = __builtin_amdgcn_readfirstlane(w[1]);
    const int used = __builtin_amdgcn_readfirstlane(w[2]);
#pragma unroll
    for (int k = 0; k < kLnSlots; ++k) {
#if JH_LN_SLOT_MASK
      if (!((used >> k) & 1)) continue;  // no entry reads this slot
#endif
      const int lw = __builtin_amdgcn_readfirstlane(w[4 + k]);
      int xx = xw + ((lw >> 4) & 0xFFF) - 2048;
      int yy = yw + ((lw >> 16) & 0xFFF) - 2048;
      xx = xx < 0 ? xx + ln.lx : xx >= ln.lx ? xx - ln.lx : xx;
      yy = yy < 0 ? yy + ln.ly : yy >= ln.ly ? yy - ln.ly : yy;
      ln_load<PREC, P>(ln, src64, src32, n, lw & 15, xx, yy, lane, S[k]);
    }
  };
  auto run_group = [&](const LnLine<PREC, P>(&S)[kLnSlots], const LnGroup &hg) {
    const uint32_t m = hg.mask;
    if constexpr (PREC == PR_FP32) {
      if (m & 0x80000000u) {
        const double *__restrict__ Tp = ln.pval + (size_t)hg.vbase * 4;
        ln_pair<DIAG, P, 0>(S, m, Tp, a);
        ln_pair<DIAG, P, 1>(S, m, Tp, a);
        ln_pair<DIAG, P, 2>(S, m, Tp, a);
        ln_pair<DIAG, P, 3>(S, m, Tp, a);
        ln_pair<DIAG, P, 4>(S, m, Tp, a);
        ln_pair<DIAG, P, 5>(S, m, Tp, a);
        ln_pair<DIAG, P, 6>(S, m, Tp, a);
        ln_pair<DIAG, P, 7>(S, m, Tp, a);
        ln_pair<DIAG, P, 8>(S, m, Tp, a);
        ln_pair<DIAG, P, 9>(S, m, Tp, a);
        ln_pair<DIAG, P, 10>(S, m, Tp, a);
        ln_pair<DIAG, P, 11>(S, m, Tp, a);
        ln_pair<DIAG, P, 12>(S, m, Tp, a);
        ln_pair<DIAG, P, 13>(S, m, Tp, a);
        ln_pair<DIAG, P, 14>(S, m, Tp, a);
        return;
      }
    }
    const double *__restrict__ T = ln.val + (size_t)hg.vbase * V;
    ln_code<PREC, DIAG, P, 0>(S, m, T, a);
    ln_code<PREC, DIAG, P, 1>(S, m, T, a);
    ln_code<PREC, DIAG, P, 2>(S, m, T, a);
    ln_code<PREC, DIAG, P, 3>(S, m, T, a);
    ln_code<PREC, DIAG, P, 4>(S, m, T, a);
    ln_code<PREC, DIAG, P, 5>(S, m, T, a);
    ln_code<PREC, DIAG, P, 6>(S, m, T, a);
    ln_code<PREC, DIAG, P, 7>(S, m, T, a);
    ln_code<PREC, DIAG, P, 8>(S, m, T, a);
    ln_code<PREC, DIAG, P, 9>(S, m, T, a);
  };
  if (G <= 0) return;
  load_group(A, hA, 0);
  int g = 0;
  for (; g + 2 < G; g += 2) {
    load_group(B, hB, g + 1);
    run_group(A, hA);
    load_group(A, hA, g + 2);
    run_group(B, hB);
  }
  if (g + 1 < G) {
    load_group(B, hB, g + 1);
    run_group(A, hA);
    run_group(B, hB);
  } else {
    run_group(A, hA);
  }
}

// One Heun phase over the LINE layout: workgroup -> 4 consecutive lines,
// wave -> line (b, x, y), lane -> cells P lane .. P lane + P - 1.
// Waves per SIMD the compiler keeps VGPR room for: the predictors fit 5
// (96 VGPRs or fewer) without spilling, the correctors would spill.
#ifndef JH_LN_SLOT_MASK
#define JH_LN_SLOT_MASK 1
#endif
#ifndef JH_LN_PRED_WAVES
#define JH_LN_PRED_WAVES 5
#endif
template <int PHASE>
constexpr int kLnWaves = PHASE == 0 ? JH_LN_PRED_WAVES : 1;
// The corrector's own s, ds and noise: read before the field (FP32-field
// form: it forms s* from them) or after it (FP64, fewer live registers)
#ifndef JH_LN_EARLY64
#define JH_LN_EARLY64 0
#endif
template <int PREC>
constexpr bool kLnEarlyOwn = kEarlyOwn<PREC> || JH_LN_EARLY64;
// ONE: no class table (one site class), so every spin's class is the
// constant 0 and its parameters are read through the scalar unit.
template <int PREC, bool DIAG, int PHASE, bool GEN, int P, bool ONE>
__global__ void __launch_bounds__(kLnBlock)
__attribute__((amdgpu_waves_per_eu(kLnWaves<PHASE>))) k_heun_ln(const KParams p, const LnView ln) {
  const int rlb = ln.rlblocks;
  if ((int)blockIdx.x < rlb) {
    // FP64 row list: the rows whose z neighbours wrap, each summed in its own
    // column order from its row pattern (pattern table in global memory,
    // tensor table in LDS)
    if constexpr (PREC == PR_FP64) {
      extern __shared__ __align__(16) unsigned char lds[];
      constexpr int XG = DIAG ? XV_GDIAG : XV_GFULL;
      load_table<PREC, XG>(p.x, lds);
      __syncthreads();
      const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
      if (j >= ln.nrows) return;
      const int64_t i = ln.rows[j];
      const uint32_t c = ONE ? 0u : p.cls ? (uint32_t)p.cls[i] : 0u;
      const uint64_t step = launch_step(p);
      const double t0 = launch_time(p, step);
      double s0, s1, s2;
      OwnState own;
      phase_spin<PREC, PHASE>(p, i, s0, s1, s2, own);
      const double t = PHASE == 0 ? t0 : t0 + p.dt;
      double x0, x1, x2;
      exchange_at<PREC, XG>(p, PHASE == 0 ? p.s : p.sp, p.g32, i, lds, x0, x1, x2);
      heun_spin<PREC, PHASE, GEN>(p, i, c, step, t, s0, s1, s2, x0, x1, x2, own);
    }
    return;
  }
  const int64_t blk = xcd_block(blockIdx.x - rlb, gridDim.x - rlb);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  // waves take the lines basis site fastest, then y within a band of ty
  // rows, then x, then band: the lines of one (x, y) column, which read each
  // other, run side by side on one XCD, and the lines in flight on an XCD
  // stay within a few x of each other in one band, so their neighbour lines
  // stay in its L2
  const int64_t q = blk * (int64_t)(blockDim.x >> 6) + wave;
  const int lane = (int)(threadIdx.x & 63);
  const int b = (int)(q % ln.nb);
  int64_t r = q / ln.nb;
  const int yin = (int)(r % ln.ty);
  r /= ln.ty;
  const int xw = (int)(r % ln.lx);
  const int yw = (int)(r / ln.lx) * ln.ty + yin;
  const int64_t line = ((int64_t)b * ln.lx + xw) * ln.ly + yw;
  const int64_t i0 = line * ln.lz + (int64_t)P * lane;
  const int64_t n = p.n;
  const uint64_t step = launch_step(p);
  const double t0 = launch_time(p, step);
  const double t = PHASE == 0 ? t0 : t0 + p.dt;
  const int sid = ln.wseq ? __builtin_amdgcn_readfirstlane(ln.wseq[line]) : b;
  // a lane whose cells the row list owns (FP64) stores nothing
  const bool fixed = ln.fixmask && ((ln.fixmask[line] >> lane) & 1ull);
  double s[P][3];
  OwnState own[P];
  double a[P][3];
  if constexpr (P == 2) {
    // the lane's two cells are adjacent: each array is one 16-byte access
    // per (x, y) pair and per SoA component
    double w[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    if constexpr (PHASE == 0) {
      ld_spin2(p.s, n, i0, s);
    } else if constexpr (kLnEarlyOwn<PREC>) {
      if (p.noise == NM_PHILOX) ld_soa2(p.wbuf, n, i0, w);
      double os[2][3], od[2][3];
      ld_spin2(p.s, n, i0, os);
      ld_soa2(p.ds, n, i0, od);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        own[h] = OwnState{os[h][0], os[h][1], os[h][2], od[h][0], od[h][1], od[h][2]};
        if (p.spfree) {
          // s* = unit(s + dt rhs) as the predictor formed it (2 ds is its rhs)
          s[h][0] = os[h][0] + p.dt * (2.0 * od[h][0]);
          s[h][1] = os[h][1] + p.dt * (2.0 * od[h][1]);
          s[h][2] = os[h][2] + p.dt * (2.0 * od[h][2]);
          renorm<PREC>(s[h][0], s[h][1], s[h][2]);
        }
      }
      if (!p.spfree) ld_spin2(p.sp, n, i0, s);
    } else {
      ld_spin2(p.sp, n, i0, s);
    }
    ln_field<PREC, DIAG, P>(ln, sid, xw, yw, PHASE == 0 ? p.s : p.sp, p.g32, n, lane, a);
    if (fixed) return;
    if (p.noise == NM_PHILOX && PHASE == 1 && !kLnEarlyOwn<PREC>) ld_soa2(p.wbuf, n, i0, w);
    else if (p.noise == NM_INJECTED) ld_soa2(p.winj, n, i0, w);
    if constexpr (PHASE == 1 && !kLnEarlyOwn<PREC>) {
      double os[2][3], od[2][3];
      ld_spin2(p.s, n, i0, os);
      ld_soa2(p.ds, n, i0, od);
#pragma unroll
      for (int h = 0; h < 2; ++h)
        own[h] = OwnState{os[h][0], os[h][1], os[h][2], od[h][0], od[h][1], od[h][2]};
    }
    double d[2][3], v[2][3];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t c = ONE ? 0u : p.cls ? (uint32_t)p.cls[i0 + h] : 0u;
      // (formed here from i0: a reference index computed from the line's
      // coordinates up front stays live across the field and spills)
      heun_math<PREC, PHASE, GEN>(p, i0 + h, ref_index(p, i0 + h), c, step, t, s[h][0], s[h][1],
                                  s[h][2], a[h][0],
                                  a[h][1], a[h][2], own[h], w[h][0], w[h][1], w[h][2], d[h][0],
                                  d[h][1], d[h][2], v[h][0], v[h][1], v[h][2]);
    }
    if constexpr (PHASE == 0) {
      if (p.noise == NM_PHILOX) st_soa2(p.wbuf, n, i0, w, p.ntstore);
      st_soa2(p.ds, n, i0, d, p.ntstore);
      if (!(PREC == PR_FP32 && p.spfree)) st_spin2(p.sp, n, i0, v, p.ntstore);
    } else {
      st_spin2(p.snew, n, i0, v, p.ntstore);
    }
    if constexpr (PREC == PR_FP32) st_g32_2(p.w32, i0, v);
  } else {
#pragma unroll
    for (int h = 0; h < P; ++h) phase_spin<PREC, PHASE>(p, i0 + h, s[h][0], s[h][1], s[h][2], own[h]);
    ln_field<PREC, DIAG, P>(ln, sid, xw, yw, PHASE == 0 ? p.s : p.sp, p.g32, n, lane, a);
    if (fixed) return;
#pragma unroll
    for (int h = 0; h < P; ++h) {
      const uint32_t c = ONE ? 0u : p.cls ? (uint32_t)p.cls[i0 + h] : 0u;
      heun_spin<PREC, PHASE, GEN>(p, i0 + h, c, step, t, s[h][0], s[h][1], s[h][2], a[h][0],
                                  a[h][1], a[h][2], own[h]);
    }
  }
}

// One stage of the classical RK4 step of llg-rk4-gpu
// (solvers/cuda_rk4_base.cu:57-108, cuda_llg_rk4_kernel.cuh:8-58,
// cuda_rk4_base_kernel.cuh:1-18, cuda/cuda_spin_ops.cu:4-17):
//   k = rhs(y, h(y, t) / mus + w) at the stage input y (y = s at stage 0,
//   else p.sp), t = t0, t0 + dt/2, t0 + dt/2, t0 + dt, with the step's noise
//   w drawn at stage 0 and reused;
//   stages 0-2 write the next input y' = s + c k, c = dt/2, dt/2, dt, to p.sy
//   (the reference's copy + axpy, without renormalisation);
//   the weighted sum k1 + 2 k2 + 2 k3 + k4 accumulates in p.ds left to right,
//   as the reference's combination is written;
//   stage 3 writes s = v / sqrt(v.v), v = s + (dt sum) / 6 (the reference
//   multiplies by CUDA's rsqrt; here 1 / sqrt, each correctly rounded).
// ONE: no class table, the class is the constant 0 (its parameters then
// come through the scalar unit).
template <int PREC, int XVk, bool GEN, bool ONE>
__global__ void __launch_bounds__(kMaxBlock)
__attribute__((amdgpu_waves_per_eu(kMinWaves<PREC, XVk>))) k_rk4(const KParams p) {
  extern __shared__ __align__(16) unsigned char lds[];
  load_table<PREC, XVk>(p.x, lds);
  if constexpr (xv_typed(XVk)) __syncthreads();
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t i = blk * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const int64_t n = p.n;
  const uint32_t c = ONE ? 0u : p.cls ? (uint32_t)p.cls[i] : 0u;
  const int stage = p.stage;
  const uint64_t step = p.dstep ? *p.dstep + p.step : p.step;
  const double t0 = p.dstep ? (double)step * p.dt : p.t;
  const double half = 0.5 * p.dt;
  const double t = stage == 0 ? t0 : stage == 3 ? t0 + p.dt : t0 + half;

  const double *src = stage == 0 ? p.s : p.sp;
  double y0, y1, y2;
  load_spin(src, n, i, y0, y1, y2);
  double h0, h1, h2;
  total_field<PREC, XVk, GEN>(p, src, p.g32, i, c, lds, y0, y1, y2, t, h0, h1, h2);

  double w0 = 0.0, w1 = 0.0, w2 = 0.0;
  if (p.noise == NM_PHILOX && stage == 0) {
    double z0, z1, z2, sT;
    if (p.cls) {
      normals3<false>(noise_seed(p, c), step, noise_index(p, c, i), z0, z1, z2);
      sT = cparam(p, c, CF_SQRTT);
    } else {
      normals3<true>(p.seed, step, (uint64_t)ref_index(p, i) - p.roff, z0, z1, z2);
      sT = p.sqrtT;
    }
    const double sg = cparam(p, c, CF_SIGMA);
    w0 = z0 * sg * sT; w1 = z1 * sg * sT; w2 = z2 * sg * sT;
    p.wbuf[i] = w0; p.wbuf[n + i] = w1; p.wbuf[2 * n + i] = w2;
  } else if (p.noise == NM_PHILOX) {
    w0 = p.wbuf[i]; w1 = p.wbuf[n + i]; w2 = p.wbuf[2 * n + i];
  } else if (p.noise == NM_INJECTED) {
    w0 = p.winj[i]; w1 = p.winj[n + i]; w2 = p.winj[2 * n + i];
  }
  const double mus = cparam(p, c, CF_MUS);
  if (p.noise == NM_NONE) {
    h0 = h0 / mus; h1 = h1 / mus; h2 = h2 / mus;
  } else {
    h0 = h0 / mus + w0; h1 = h1 / mus + w1; h2 = h2 / mus + w2;
  }
  double r0, r1, r2;
  llg_rhs(y0, y1, y2, h0, h1, h2, cparam(p, c, CF_GYRO), cparam(p, c, CF_ALPHA), r0, r1, r2);

  double s0, s1, s2;
  load_spin(p.s, n, i, s0, s1, s2);
  if (stage < 3) {
    if (stage == 0) {
      p.ds[i] = r0; p.ds[n + i] = r1; p.ds[2 * n + i] = r2;
    } else {
      p.ds[i] = p.ds[i] + 2.0 * r0;
      p.ds[n + i] = p.ds[n + i] + 2.0 * r1;
      p.ds[2 * n + i] = p.ds[2 * n + i] + 2.0 * r2;
    }
    const double cst = stage == 2 ? p.dt : half;
    const double v0 = s0 + cst * r0, v1 = s1 + cst * r1, v2 = s2 + cst * r2;
    store_spin(p.sy, n, i, v0, v1, v2);
    if constexpr (PREC == PR_FP32) store_g32(p.w32, i, v0, v1, v2);
  } else {
    double v0 = s0 + (p.dt * (p.ds[i] + r0)) / 6.0;
    double v1 = s1 + (p.dt * (p.ds[n + i] + r1)) / 6.0;
    double v2 = s2 + (p.dt * (p.ds[2 * n + i] + r2)) / 6.0;
    const double rn = 1.0 / sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    v0 = v0 * rn; v1 = v1 * rn; v2 = v2 * rn;
    store_spin(p.snew, n, i, v0, v1, v2);
    if constexpr (PREC == PR_FP32) store_g32(p.w32, i, v0, v1, v2);
  }
}

// Field of one term (term >= 0) or of the ordered sum (term < 0) at spins p.s,
// time p.t, into out (SoA).
template <int PREC, int XVk>
__global__ void __launch_bounds__(kBlock) k_field(const KParams p, int term, double *out) {
  extern __shared__ __align__(16) unsigned char lds[];
  load_table<PREC, XVk>(p.x, lds);
  if constexpr (xv_typed(XVk)) __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= p.n) return;
  const int64_t n = p.n;
  const uint32_t c = p.cls ? (uint32_t)p.cls[i] : 0u;
  double s0, s1, s2;
  load_spin(p.s, n, i, s0, s1, s2);
  double h0, h1, h2;
  if (term < 0) {
    total_field<PREC, XVk, true>(p, p.s, p.g32, i, c, lds, s0, s1, s2, p.t, h0, h1, h2);
  } else if (p.kind[term] == TK_EXCHANGE) {
    exchange_at<PREC, XVk>(p, p.s, p.g32, i, lds, h0, h1, h2);
  } else {
    local_term<true>(p, p.kind[term], c, i, s0, s1, s2, p.t, h0, h1, h2);
  }
  out[i] = h0; out[n + i] = h1; out[2 * n + i] = h2;
}

// Generated, scaled noise of one step (for inspection and parity tests).
__global__ void __launch_bounds__(kBlock) k_noise(const KParams p, double *out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= p.n) return;
  const uint32_t c = p.cls ? (uint32_t)p.cls[i] : 0u;
  double z0, z1, z2;
  normals3<false>(noise_seed(p, c), p.step, noise_index(p, c, i), z0, z1, z2);
  const double sg = cparam(p, c, CF_SIGMA), sT = cparam(p, c, CF_SQRTT);
  out[i] = z0 * sg * sT;
  out[p.n + i] = z1 * sg * sT;
  out[2 * p.n + i] = z2 * sg * sT;
}

// FP32 gather copy of FP64 spins (kG32 floats per spin)
__global__ void __launch_bounds__(kBlock) k_to_f32(const double *a, float *b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    double x, y, z;
    load_spin(a, n, i, x, y, z);
    store_g32(b, i, x, y, z);
  }
}

// Deterministic pairwise block reduction of 4 FP64 quantities per spin:
//   mode 0: (mus s_x, mus s_y, mus s_z, mus)
//   mode 1: (-s.h, 0, 0, 0) with h = field array (energy density)
//   mode 2: (-K (s.e)^p, 0, 0, 0) (uniaxial energy)
__global__ void __launch_bounds__(kBlock) k_reduce(const KParams p, int mode, const double *h,
                                                   double *partial) {
  __shared__ double red[4][kBlock];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < p.n) {
    const int64_t n = p.n;
    const uint32_t c = p.cls ? (uint32_t)p.cls[i] : 0u;
    double s0, s1, s2;
    load_spin(p.s, n, i, s0, s1, s2);
    if (mode == 0) {
      const double mu = cparam(p, c, CF_MUS);
      v[0] = mu * s0; v[1] = mu * s1; v[2] = mu * s2; v[3] = mu;
    } else if (mode == 1) {
      v[0] = -(s0 * h[i] + s1 * h[n + i] + s2 * h[2 * n + i]);
    } else {
      const double d = (cparam(p, c, CF_AXIS) * s0 + cparam(p, c, CF_AXIS + 1) * s1 +
                        cparam(p, c, CF_AXIS + 2) * s2);
      v[0] = (-cparam(p, c, CF_K) * pow(d, (double)p.ani_power));
    }
  }
  for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = red[q][threadIdx.x] + red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[(int64_t)blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// Final pass: pairwise sum of nb partial 4-vectors in one workgroup.
__global__ void __launch_bounds__(kBlock) k_reduce_final(const double *partial, int64_t nb,
                                                         double *out) {
  __shared__ double red[4][kBlock];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t b = threadIdx.x; b < nb; b += kBlock)
    for (int q = 0; q < 4; ++q) v[q] = v[q] + partial[b * 4 + q];
  for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = red[q][threadIdx.x] + red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = red[threadIdx.x][0];
}

// Per-replica magnetisation: workgroup r sums (mus s, mus) over the spins
// [off[r], off[r+1]) in a fixed pairwise order.
__global__ void __launch_bounds__(kBlock) k_reduce_replicas(const KParams p, const int64_t *off,
                                                            double *out) {
  __shared__ double red[4][kBlock];
  const int r = blockIdx.x;
  const int64_t a = off[r], b = off[r + 1], n = p.n;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = a + threadIdx.x; i < b; i += kBlock) {
    const uint32_t c = p.cls ? (uint32_t)p.cls[i] : 0u;
    const double mu = cparam(p, c, CF_MUS);
    double s0, s1, s2;
    load_spin(p.s, n, i, s0, s1, s2);
    v[0] = v[0] + mu * s0;
    v[1] = v[1] + mu * s1;
    v[2] = v[2] + mu * s2;
    v[3] = v[3] + mu;
  }
  for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = red[q][threadIdx.x] + red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[(int64_t)r * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// One thread sets the device step counter of a graph replay.
__global__ void k_set_step(uint64_t *step, uint64_t value) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *step = value;
}

// ---------------------------------------------------------------------------
// launchers

static size_t table_lds(const KParams &p, int prec, int xv) {
  size_t b = table_bytes(prec, xv, p.x.ntab);
  if (xv_pattern_lds(xv)) b += (size_t)p.x.pat_words * 4;
  return b;
}

template <int PREC, int XVk>
static hipError_t launch_heun_t(const KParams &p, int phase, int bs, hipStream_t st) {
  const size_t lds = table_lds(p, PREC, XVk);
  const bool gen = p.has_ac || p.ani_power != 2;
  const int64_t rows = p.rowlist ? p.nrowlist : p.n;
  const int64_t nb = (rows + bs - 1) / bs;
  // the common forms find the site class wave-uniformly (CM_SPLIT with a
  // class table, CM_ONE without); the GEN forms per lane
  const bool uc = p.cls != nullptr;
  if (phase == 0 && !gen && uc)
    hipLaunchKernelGGL((k_heun<PREC, XVk, 0, false, CM_SPLIT>), dim3((unsigned)nb), dim3(bs), lds, st, p);
  else if (phase == 0 && !gen)
    hipLaunchKernelGGL((k_heun<PREC, XVk, 0, false, CM_ONE>), dim3((unsigned)nb), dim3(bs), lds, st, p);
  else if (phase == 0)
    hipLaunchKernelGGL((k_heun<PREC, XVk, 0, true, CM_LANE>), dim3((unsigned)nb), dim3(bs), lds, st, p);
  else if (!gen && uc)
    hipLaunchKernelGGL((k_heun<PREC, XVk, 1, false, CM_SPLIT>), dim3((unsigned)nb), dim3(bs), lds, st, p);
  else if (!gen)
    hipLaunchKernelGGL((k_heun<PREC, XVk, 1, false, CM_ONE>), dim3((unsigned)nb), dim3(bs), lds, st, p);
  else
    hipLaunchKernelGGL((k_heun<PREC, XVk, 1, true, CM_LANE>), dim3((unsigned)nb), dim3(bs), lds, st, p);
  return hipGetLastError();
}

template <int PREC>
static hipError_t launch_heun_p(const KParams &p, int xv, int phase, int bs, hipStream_t st) {
  switch (xv) {
    case XV_NONE: return launch_heun_t<PREC, XV_NONE>(p, phase, bs, st);
    case XV_TDIAG: return launch_heun_t<PREC, XV_TDIAG>(p, phase, bs, st);
    case XV_TFULL: return launch_heun_t<PREC, XV_TFULL>(p, phase, bs, st);
    case XV_BDIAG: return launch_heun_t<PREC, XV_BDIAG>(p, phase, bs, st);
    case XV_BFULL: return launch_heun_t<PREC, XV_BFULL>(p, phase, bs, st);
    case XV_PDIAG: return launch_heun_t<PREC, XV_PDIAG>(p, phase, bs, st);
    case XV_PFULL: return launch_heun_t<PREC, XV_PFULL>(p, phase, bs, st);
    case XV_GDIAG: return launch_heun_t<PREC, XV_GDIAG>(p, phase, bs, st);
    case XV_GFULL: return launch_heun_t<PREC, XV_GFULL>(p, phase, bs, st);
    case XV_UDIAG:
      if constexpr (PREC == PR_FP32) return launch_heun_t<PREC, XV_UDIAG>(p, phase, bs, st);
      break;
    case XV_UFULL:
      if constexpr (PREC == PR_FP32) return launch_heun_t<PREC, XV_UFULL>(p, phase, bs, st);
      break;
    case XV_QISO:
      if constexpr (PREC == PR_FP32) return launch_heun_t<PREC, XV_QISO>(p, phase, bs, st);
      break;
    case XV_QDMI:
      if constexpr (PREC == PR_FP32) return launch_heun_t<PREC, XV_QDMI>(p, phase, bs, st);
      break;
  }
  return hipErrorInvalidValue;
}

// Workgroups of one persistent step launch: as many as are resident at once
// (the occupancy of this instantiation at bs threads and its LDS, times the
// CUs), at most one per spin block, a multiple of 8 when there are 8 or more.
// The occupancy query is cached per (kernel, bs, LDS bytes).
template <int PREC, int XVk, bool GEN>
static hipError_t launch_heun_step_k(const KParams &p, int bs, int cus, unsigned *sync,
                                     unsigned limit, hipStream_t st) {
  const auto fn = k_heun_step<PREC, XVk, GEN>;
  const size_t lds = table_lds(p, PREC, XVk);
  static std::mutex mu;
  static std::map<std::tuple<const void *, int, size_t>, int> occ;
  int per = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_tuple(reinterpret_cast<const void *>(fn), bs, lds);
    const auto it = occ.find(key);
    if (it != occ.end()) {
      per = it->second;
    } else {
      const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, bs, lds);
      if (e != hipSuccess) return e;
      occ[key] = per;
    }
  }
  if (per < 1 || cus < 1) return hipErrorInvalidConfiguration;
  const int64_t nv = (p.n + bs - 1) / bs;
  int64_t g = std::min<int64_t>(nv, (int64_t)per * cus);
  if (g >= 8) g -= g % 8;
  hipLaunchKernelGGL(fn, dim3((unsigned)g), dim3(bs), lds, st, p, sync, limit);
  return hipGetLastError();
}

template <int PREC, int XVk>
static hipError_t launch_heun_step_t(const KParams &p, int bs, int cus, unsigned *sync,
                                     unsigned limit, hipStream_t st) {
  const bool gen = p.has_ac || p.ani_power != 2;
  return gen ? launch_heun_step_k<PREC, XVk, true>(p, bs, cus, sync, limit, st)
             : launch_heun_step_k<PREC, XVk, false>(p, bs, cus, sync, limit, st);
}

template <int PREC>
static hipError_t launch_heun_step_p(const KParams &p, int xv, int bs, int cus, unsigned *sync,
                                     unsigned limit, hipStream_t st) {
  switch (xv) {
    case XV_NONE: return launch_heun_step_t<PREC, XV_NONE>(p, bs, cus, sync, limit, st);
    case XV_TDIAG: return launch_heun_step_t<PREC, XV_TDIAG>(p, bs, cus, sync, limit, st);
    case XV_TFULL: return launch_heun_step_t<PREC, XV_TFULL>(p, bs, cus, sync, limit, st);
    case XV_BDIAG: return launch_heun_step_t<PREC, XV_BDIAG>(p, bs, cus, sync, limit, st);
    case XV_BFULL: return launch_heun_step_t<PREC, XV_BFULL>(p, bs, cus, sync, limit, st);
    case XV_PDIAG: return launch_heun_step_t<PREC, XV_PDIAG>(p, bs, cus, sync, limit, st);
    case XV_PFULL: return launch_heun_step_t<PREC, XV_PFULL>(p, bs, cus, sync, limit, st);
    case XV_GDIAG: return launch_heun_step_t<PREC, XV_GDIAG>(p, bs, cus, sync, limit, st);
    case XV_GFULL: return launch_heun_step_t<PREC, XV_GFULL>(p, bs, cus, sync, limit, st);
    case XV_UDIAG:
      if constexpr (PREC == PR_FP32) return launch_heun_step_t<PREC, XV_UDIAG>(p, bs, cus, sync, limit, st);
      break;
    case XV_UFULL:
      if constexpr (PREC == PR_FP32) return launch_heun_step_t<PREC, XV_UFULL>(p, bs, cus, sync, limit, st);
      break;
    case XV_QISO:
      if constexpr (PREC == PR_FP32) return launch_heun_step_t<PREC, XV_QISO>(p, bs, cus, sync, limit, st);
      break;
    case XV_QDMI:
      if constexpr (PREC == PR_FP32) return launch_heun_step_t<PREC, XV_QDMI>(p, bs, cus, sync, limit, st);
      break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_heun_step(const KParams &p, int prec, int xv, int bs, int cus, unsigned *sync,
                            unsigned limit, hipStream_t st) {
  if (bs < 64 || bs > kMaxBlock || bs % 64 != 0 || !sync || p.rowlist) return hipErrorInvalidValue;
  return prec == PR_FP64 ? launch_heun_step_p<PR_FP64>(p, xv, bs, cus, sync, limit, st)
                         : launch_heun_step_p<PR_FP32>(p, xv, bs, cus, sync, limit, st);
}

hipError_t launch_heun(const KParams &p, int prec, int xv, int phase, int bs, hipStream_t st) {
  if (bs < 64 || bs > kMaxBlock || bs % 64 != 0) return hipErrorInvalidValue;
  return prec == PR_FP64 ? launch_heun_p<PR_FP64>(p, xv, phase, bs, st)
                         : launch_heun_p<PR_FP32>(p, xv, phase, bs, st);
}

// LDS bytes of a STENCIL workgroup of source basis site b
size_t st_tile_bytes(const StView &v, int prec, int b) {
  return (size_t)v.nb * v.tx[b] * v.ty[b] * v.tz[b] * (prec == PR_FP64 ? kStElem<PR_FP64>
                                                                       : kStElem<PR_FP32>);
}

template <int PREC, bool DIAG, int PHASE, bool GEN>
static hipError_t launch_heun_st_k(const KParams &p, const StView &v, size_t lds,
                                   hipStream_t st) {
  const auto fn = k_heun_st<PREC, DIAG, PHASE, GEN>;
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const int64_t nwg = (int64_t)v.nb * v.nbx * v.nby * v.nbz;
  hipLaunchKernelGGL(fn, dim3((unsigned)nwg), dim3(64 * v.wx * v.wy), lds, st, p, v);
  return hipGetLastError();
}

template <int PREC, bool DIAG>
static hipError_t launch_heun_st_t(const KParams &p, const StView &v, int phase, size_t lds,
                                   hipStream_t st) {
  const bool gen = p.has_ac || p.ani_power != 2;
  if (phase == 0)
    return gen ? launch_heun_st_k<PREC, DIAG, 0, true>(p, v, lds, st)
               : launch_heun_st_k<PREC, DIAG, 0, false>(p, v, lds, st);
  return gen ? launch_heun_st_k<PREC, DIAG, 1, true>(p, v, lds, st)
             : launch_heun_st_k<PREC, DIAG, 1, false>(p, v, lds, st);
}

hipError_t launch_heun_st(const KParams &p, const StView &v, int prec, bool diag, int phase,
                          hipStream_t st) {
  if (v.nb < 1 || v.nb > kStMaxBasis || v.wx * v.wy * 64 > kStMaxBlock) return hipErrorInvalidValue;
  size_t lds = 0;
  for (int b = 0; b < v.nb; ++b) lds = std::max(lds, st_tile_bytes(v, prec, b));
  if (prec == PR_FP64)
    return diag ? launch_heun_st_t<PR_FP64, true>(p, v, phase, lds, st)
                : launch_heun_st_t<PR_FP64, false>(p, v, phase, lds, st);
  return diag ? launch_heun_st_t<PR_FP32, true>(p, v, phase, lds, st)
              : launch_heun_st_t<PR_FP32, false>(p, v, phase, lds, st);
}

template <int PREC, bool DIAG, int P>
static hipError_t launch_heun_ln_t(const KParams &p, const LnView &v, int phase, int bs,
                                   hipStream_t st) {
  const int64_t lines = (int64_t)v.nb * v.lx * v.ly;
  const dim3 grid((unsigned)(lines / (bs / 64) + v.rlblocks)), block((unsigned)bs);
  const size_t lds = v.rlblocks > 0 ? table_bytes(PREC, DIAG ? XV_GDIAG : XV_GFULL, p.x.ntab) : 0;
  const bool gen = p.has_ac || p.ani_power != 2;
  const bool one = p.cls == nullptr;
  if (phase == 0) {
    if (gen) hipLaunchKernelGGL((k_heun_ln<PREC, DIAG, 0, true, P, false>), grid, block, lds, st, p, v);
    else if (one) hipLaunchKernelGGL((k_heun_ln<PREC, DIAG, 0, false, P, true>), grid, block, lds, st, p, v);
    else hipLaunchKernelGGL((k_heun_ln<PREC, DIAG, 0, false, P, false>), grid, block, lds, st, p, v);
  } else {
    if (gen) hipLaunchKernelGGL((k_heun_ln<PREC, DIAG, 1, true, P, false>), grid, block, lds, st, p, v);
    else if (one) hipLaunchKernelGGL((k_heun_ln<PREC, DIAG, 1, false, P, true>), grid, block, lds, st, p, v);
    else hipLaunchKernelGGL((k_heun_ln<PREC, DIAG, 1, false, P, false>), grid, block, lds, st, p, v);
  }
  return hipGetLastError();
}

template <int PREC, bool DIAG>
static hipError_t launch_heun_ln_p(const KParams &p, const LnView &v, int phase, int bs,
                                   hipStream_t st) {
  if (v.lz == 64) return launch_heun_ln_t<PREC, DIAG, 1>(p, v, phase, bs, st);
  return launch_heun_ln_t<PREC, DIAG, 2>(p, v, phase, bs, st);
}

hipError_t launch_heun_ln(const KParams &p, const LnView &v, int prec, bool diag, int phase,
                          hipStream_t st) {
  // the grid covers the lines exactly: 4 lines per workgroup, else one
  const int64_t lines = (int64_t)v.nb * v.lx * v.ly;
  if (v.nb < 1 || v.nb > kLnSlots || (v.lz != 64 && v.lz != 128) || lines <= 0 ||
      lines * v.lz != p.n)
    return hipErrorInvalidValue;
  const int bs = lines % (kLnBlock / 64) == 0 ? kLnBlock : 64;
  if (v.rlblocks != (v.nrows + bs - 1) / bs || (v.nrows > 0 && (!v.rows || prec != PR_FP64)))
    return hipErrorInvalidValue;
  if (prec == PR_FP64)
    return diag ? launch_heun_ln_p<PR_FP64, true>(p, v, phase, bs, st)
                : launch_heun_ln_p<PR_FP64, false>(p, v, phase, bs, st);
  return diag ? launch_heun_ln_p<PR_FP32, true>(p, v, phase, bs, st)
              : launch_heun_ln_p<PR_FP32, false>(p, v, phase, bs, st);
}

template <int PREC, int XVk>
static hipError_t launch_rk4_t(const KParams &p, int bs, hipStream_t st) {
  const int64_t nb = (p.n + bs - 1) / bs;
  const size_t lds = table_lds(p, PREC, XVk);
  if (p.has_ac || p.ani_power != 2)
    hipLaunchKernelGGL((k_rk4<PREC, XVk, true, false>), dim3((unsigned)nb), dim3(bs), lds, st, p);
  else if (p.cls)
    hipLaunchKernelGGL((k_rk4<PREC, XVk, false, false>), dim3((unsigned)nb), dim3(bs), lds, st, p);
  else
    hipLaunchKernelGGL((k_rk4<PREC, XVk, false, true>), dim3((unsigned)nb), dim3(bs), lds, st, p);
  return hipGetLastError();
}

template <int PREC>
static hipError_t launch_rk4_p(const KParams &p, int xv, int bs, hipStream_t st) {
  switch (xv) {
    case XV_NONE: return launch_rk4_t<PREC, XV_NONE>(p, bs, st);
    case XV_TDIAG: return launch_rk4_t<PREC, XV_TDIAG>(p, bs, st);
    case XV_TFULL: return launch_rk4_t<PREC, XV_TFULL>(p, bs, st);
    case XV_BDIAG: return launch_rk4_t<PREC, XV_BDIAG>(p, bs, st);
    case XV_BFULL: return launch_rk4_t<PREC, XV_BFULL>(p, bs, st);
    case XV_PDIAG: return launch_rk4_t<PREC, XV_PDIAG>(p, bs, st);
    case XV_PFULL: return launch_rk4_t<PREC, XV_PFULL>(p, bs, st);
    case XV_GDIAG: return launch_rk4_t<PREC, XV_GDIAG>(p, bs, st);
    case XV_GFULL: return launch_rk4_t<PREC, XV_GFULL>(p, bs, st);
    case XV_UDIAG:
      if constexpr (PREC == PR_FP32) return launch_rk4_t<PREC, XV_UDIAG>(p, bs, st);
      break;
    case XV_UFULL:
      if constexpr (PREC == PR_FP32) return launch_rk4_t<PREC, XV_UFULL>(p, bs, st);
      break;
    case XV_QISO:
      if constexpr (PREC == PR_FP32) return launch_rk4_t<PREC, XV_QISO>(p, bs, st);
      break;
    case XV_QDMI:
      if constexpr (PREC == PR_FP32) return launch_rk4_t<PREC, XV_QDMI>(p, bs, st);
      break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_rk4(const KParams &p, int prec, int xv, int bs, hipStream_t st) {
  if (bs < 64 || bs > kMaxBlock || bs % 64 != 0 || p.stage < 0 || p.stage > 3)
    return hipErrorInvalidValue;
  return prec == PR_FP64 ? launch_rk4_p<PR_FP64>(p, xv, bs, st)
                         : launch_rk4_p<PR_FP32>(p, xv, bs, st);
}

hipError_t launch_reduce_replicas(const KParams &p, const int64_t *off, int nrep, double *out,
                                  hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_replicas, dim3((unsigned)nrep), dim3(kBlock), 0, st, p, off, out);
  return hipGetLastError();
}

hipError_t launch_set_step(uint64_t *step, uint64_t value, hipStream_t st) {
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, st, step, value);
  return hipGetLastError();
}

template <int PREC, int XVk>
static hipError_t launch_field_t(const KParams &p, int term, double *out, hipStream_t st) {
  const int64_t nb = (p.n + kBlock - 1) / kBlock;
  const size_t lds = table_lds(p, PREC, XVk);
  hipLaunchKernelGGL((k_field<PREC, XVk>), dim3((unsigned)nb), dim3(kBlock), lds, st, p, term, out);
  return hipGetLastError();
}

template <int PREC>
static hipError_t launch_field_p(const KParams &p, int xv, int term, double *out, hipStream_t st) {
  switch (xv) {
    case XV_NONE: return launch_field_t<PREC, XV_NONE>(p, term, out, st);
    case XV_TDIAG: return launch_field_t<PREC, XV_TDIAG>(p, term, out, st);
    case XV_TFULL: return launch_field_t<PREC, XV_TFULL>(p, term, out, st);
    case XV_BDIAG: return launch_field_t<PREC, XV_BDIAG>(p, term, out, st);
    case XV_BFULL: return launch_field_t<PREC, XV_BFULL>(p, term, out, st);
    case XV_PDIAG: return launch_field_t<PREC, XV_PDIAG>(p, term, out, st);
    case XV_PFULL: return launch_field_t<PREC, XV_PFULL>(p, term, out, st);
    case XV_GDIAG: return launch_field_t<PREC, XV_GDIAG>(p, term, out, st);
    case XV_GFULL: return launch_field_t<PREC, XV_GFULL>(p, term, out, st);
    case XV_UDIAG:
      if constexpr (PREC == PR_FP32) return launch_field_t<PREC, XV_UDIAG>(p, term, out, st);
      break;
    case XV_UFULL:
      if constexpr (PREC == PR_FP32) return launch_field_t<PREC, XV_UFULL>(p, term, out, st);
      break;
    case XV_QISO:
      if constexpr (PREC == PR_FP32) return launch_field_t<PREC, XV_QISO>(p, term, out, st);
      break;
    case XV_QDMI:
      if constexpr (PREC == PR_FP32) return launch_field_t<PREC, XV_QDMI>(p, term, out, st);
      break;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_field(const KParams &p, int prec, int xv, int term, double *out,
                        hipStream_t st) {
  return prec == PR_FP64 ? launch_field_p<PR_FP64>(p, xv, term, out, st)
                         : launch_field_p<PR_FP32>(p, xv, term, out, st);
}

hipError_t launch_noise(const KParams &p, double *out, hipStream_t st) {
  const int64_t nb = (p.n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_noise, dim3((unsigned)nb), dim3(kBlock), 0, st, p, out);
  return hipGetLastError();
}

hipError_t launch_to_f32(const double *a, float *b, int64_t n, hipStream_t st) {
  const int64_t nb = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)nb), dim3(kBlock), 0, st, a, b, n);
  return hipGetLastError();
}

hipError_t launch_reduce(const KParams &p, int mode, const double *h, double *partial,
                         double *out, hipStream_t st) {
  const int64_t nb = (p.n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_reduce, dim3((unsigned)nb), dim3(kBlock), 0, st, p, mode, h, partial);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, st, (const double *)partial, nb, out);
  return hipGetLastError();
}

}  // namespace jh
