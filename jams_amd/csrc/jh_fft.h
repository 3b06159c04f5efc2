// jh_fft.h -- batched mixed-radix FFT passes for gfx950, used by the dipole
// convolution.
//
// A pass transforms B lines of length n per workgroup along one axis of a
// strided array.  The lines are staged in LDS with coalesced loads: adjacent
// lines when the axis is strided, consecutive elements when it is
// contiguous.  The radix stages run in LDS with Stockham autosort
// (Govindaraju et al., "High performance discrete Fourier transforms on
// graphics processors", SC'08): stage s with radix R and sub-length Ns reads
// x[j + r n/R], twiddles it by exp(-+2 pi i r (j mod Ns) / (Ns R)), does an
// R-point DFT and writes y[(j / Ns) Ns R + (j mod Ns) + r Ns].  Radices 2, 3,
// 4, 5 and 8 are written out; any other factor takes a direct DFT.
// Twiddles come from a table exp(-2 pi i t / n), t < n, built on the host.
//
// Loads and stores can be pruned: inputs k >= n_in read as zero (zero
// padding), outputs k >= n_out are not written (the padded half is never
// needed).  Special loads and stores fuse the scatter of spins into the
// padded grid and the gather of the field back to the spins.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jh_layout.h"

// The transforms need no particular rounding sequence (the dipole's parity is
// a tolerance against the direct sum), so multiply-adds may fuse here even
// though the build disables contraction for the step kernels' bit-exact
// arithmetic.
#pragma clang fp contract(fast)

namespace jh {
namespace fft {

constexpr int kMaxStages = 24;
constexpr int kLogThreads = 9;
constexpr int kThreads = 1 << kLogThreads;

template <class T> struct C2;
template <> struct C2<double> { using type = double2; };
template <> struct C2<float> { using type = float2; };
template <class T2> struct Re;
template <> struct Re<double2> { using type = double; };
template <> struct Re<float2> { using type = float; };

template <class T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  T2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <class T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) { T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <class T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) { T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// multiply by -i (forward) or +i (inverse)
template <class T2>
__device__ __forceinline__ T2 crot(T2 a, bool inv) {
  T2 r;
  if (inv) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
  return r;
}

// Line addressing: line L = (a cb + b) cc + c, element k at
// a sa + b sb + c sc + k se (all in complex elements).
struct LineMap {
  int64_t cb, cc;
  int64_t sa, sb, sc, se;
  __host__ __device__ int64_t at(int64_t L, int64_t k) const {
    const int64_t c = L % cc, ab = L / cc;
    const int64_t b = ab % cb, a = ab / cb;
    return a * sa + b * sb + c * sc + k * se;
  }
  // element 0 of line L < 2^31 (32-bit divisions; the host checks the range)
  __device__ int64_t base32(uint32_t L) const {
    const uint32_t c = L % (uint32_t)cc, ab = L / (uint32_t)cc;
    const uint32_t b = ab % (uint32_t)cb, a = ab / (uint32_t)cb;
    return (int64_t)a * sa + (int64_t)b * sb + (int64_t)c * sc;
  }
};

// Spin-grid geometry for the fused scatter / gather: spin index of cell
// (x, y, z), basis m is ((x Ly + y) Lz + z) nb + m (core/lattice.cc:622-649).
struct SpinGrid {
  int64_t n;        // spins
  int L[3];
  int nb;
  const double *mus;    // per spin (meV/T)
};

enum LoadKind { LD_CPLX = 0, LD_SPIN = 1, LD_HERM = 2 };
enum StoreKind { ST_CPLX = 0, ST_FIELD = 1 };

template <class T>
struct Pass {
  using T2 = typename C2<T>::type;
  int n, nst;
  int radix[kMaxStages];
  int n_in, n_out;
  int B, lb;             // lines per workgroup, a power of two: B = 1 << lb
  int inverse;
  int pingpong;          // 1: two LDS buffers (plans with a direct-DFT factor)
  int64_t nlines;
  LineMap src, dst;
  const T2 *in;
  T2 *out;
  const T2 *tw;          // n twiddles exp(-2 pi i t / n)
  // per stage: sub-length Ns before it, 1/Ns rounded to float, n / (Ns R)
  int ns[kMaxStages];
  float ns_rcp[kMaxStages];
  int tws[kMaxStages];
  // LD_SPIN: lines are (g, x, y), g = 3 m + component; value mus * s
  // ST_FIELD: the same lines; h[comp N + i] = mus_i * Re(value)
  // LD_HERM: the line holds the first n_in = n/2 + 1 terms of a Hermitian
  // sequence of length n
  SpinGrid grid;
  const double *spins;   // FP64 spin array (jh_layout.h)
  double *field;         // SoA FP64
};

// R-point DFT of v in place (forward sign exp(-2 pi i qr/R); inverse +).
template <class T2, int R>
__device__ __forceinline__ void dft_fixed(T2 *v, bool inv) {
  using T = typename Re<T2>::type;
  if constexpr (R == 2) {
    const T2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const T2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const T2 b0 = cadd(v[1], v[3]), b1 = crot(csub(v[1], v[3]), inv);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
  } else if constexpr (R == 8) {
    // two radix-4 halves (even, odd) and the 8th roots
    T2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft_fixed<T2, 4>(e, inv);
    dft_fixed<T2, 4>(o, inv);
    const T h = (T)0.70710678118654752440;
    // o[q] *= exp(-+ 2 pi i q / 8)
    T2 w1, w3;
    if (inv) { w1.x = h; w1.y = h; w3.x = -h; w3.y = h; }
    else { w1.x = h; w1.y = -h; w3.x = -h; w3.y = -h; }
    o[1] = cmul(o[1], w1);
    o[2] = crot(o[2], inv);
    o[3] = cmul(o[3], w3);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[q] = cadd(e[q], o[q]);
      v[q + 4] = csub(e[q], o[q]);
    }
  } else if constexpr (R == 3) {
    const T c = (T)-0.5, s = inv ? (T)0.86602540378443864676 : (T)-0.86602540378443864676;
    const T2 a = v[0], b = v[1], d = v[2];
    const T2 t = cadd(b, d), u = csub(b, d);
    T2 m;
    m.x = a.x + c * t.x;
    m.y = a.y + c * t.y;
    v[0] = cadd(a, t);
    // v1 = m + i s u, v2 = m - i s u
    v[1].x = m.x - s * u.y; v[1].y = m.y + s * u.x;
    v[2].x = m.x + s * u.y; v[2].y = m.y - s * u.x;
  } else if constexpr (R == 5) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = inv ? (T)0.95105651629515357212 : (T)-0.95105651629515357212;
    const T s2 = inv ? (T)0.58778525229247312917 : (T)-0.58778525229247312917;
    const T2 a = v[0];
    const T2 t1 = cadd(v[1], v[4]), u1 = csub(v[1], v[4]);
    const T2 t2 = cadd(v[2], v[3]), u2 = csub(v[2], v[3]);
    T2 m1, m2, n1, n2;
    m1.x = a.x + c1 * t1.x + c2 * t2.x; m1.y = a.y + c1 * t1.y + c2 * t2.y;
    m2.x = a.x + c2 * t1.x + c1 * t2.x; m2.y = a.y + c2 * t1.y + c1 * t2.y;
    // n1 = s1 u1 + s2 u2, n2 = s2 u1 - s1 u2 (times i below)
    n1.x = s1 * u1.x + s2 * u2.x; n1.y = s1 * u1.y + s2 * u2.y;
    n2.x = s2 * u1.x - s1 * u2.x; n2.y = s2 * u1.y - s1 * u2.y;
    v[0].x = a.x + t1.x + t2.x; v[0].y = a.y + t1.y + t2.y;
    v[1].x = m1.x - n1.y; v[1].y = m1.y + n1.x;
    v[4].x = m1.x + n1.y; v[4].y = m1.y - n1.x;
    v[2].x = m2.x - n2.y; v[2].y = m2.y + n2.x;
    v[3].x = m2.x + n2.y; v[3].y = m2.y - n2.x;
  }
}

// q = x / d and r = x % d for 0 <= x < 2^22, 1 <= d < 2^22, with rd the
// float nearest 1/d: the float estimate is within one of the quotient, and
// one correction step makes it exact (no integer division instruction).
__device__ __forceinline__ int divmod_small(int x, int d, float rd, int &r) {
  int q = (int)((float)x * rd);
  r = x - __mul24(q, d);
  if (r < 0) {
    q -= 1;
    r += d;
  } else if (r >= d) {
    q += 1;
    r -= d;
  }
  return q;
}

// twiddle t from the workgroup's LDS copy of the table
template <class T, bool INV>
__device__ __forceinline__ typename C2<T>::type twiddle(const typename C2<T>::type *tw, int t) {
  typename C2<T>::type w = tw[t];
  if (INV) w.y = -w.y;
  return w;
}

// One Stockham stage over B lines: in/out are LDS arrays [k][b].  Butterfly
// t is line b = t mod B, index j = t / B < n / R.
template <class T, int R, bool INV>
__device__ __forceinline__ void stage_fixed(const Pass<T> &p, const typename C2<T>::type *tw,
                                            const typename C2<T>::type *in,
                                            typename C2<T>::type *out, int st) {
  using T2 = typename C2<T>::type;
  const int n = p.n, B = p.B, lb = p.lb, m = n / R;
  const int Ns = p.ns[st], tws = p.tws[st];
  const float rns = p.ns_rcp[st];
  for (int t = threadIdx.x; t < (m << lb); t += kThreads) {
    const int b = t & (B - 1), j = t >> lb;
    int jm;
    const int jd = divmod_small(j, Ns, rns, jm);
    T2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[((j + r * m) << lb) + b];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<T, INV>(tw, jm * r * tws));
    dft_fixed<T2, R>(v, INV);
    const int d = jd * Ns * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) out[((d + r * Ns) << lb) + b] = v[r];
  }
}

// Butterflies per thread of an in-place stage: R * kBF <= 8 complex values
// in registers (with 16, the 512-thread workgroups hold 108-188 VGPRs and one
// fits per CU; with 8, two do, and the passes run 25-30 % faster).
template <class T, int R>
constexpr int kBF = (8 / R) > 0 ? 8 / R : 1;

// The same stage in place on one LDS buffer: every thread first reads and
// transforms its butterflies (at most kBF<T, R>, which the host guarantees by
// its choice of B), then all write.
template <class T, int R, bool INV>
__device__ __forceinline__ void stage_inplace(const Pass<T> &p, const typename C2<T>::type *tw,
                                              typename C2<T>::type *A, int st) {
  using T2 = typename C2<T>::type;
  constexpr int NBF = kBF<T, R>;
  const int n = p.n, B = p.B, lb = p.lb, m = n / R, total = m << lb;
  const int Ns = p.ns[st], tws = p.tws[st];
  const float rns = p.ns_rcp[st];
  T2 v[NBF][R];
  int dq[NBF];
#pragma unroll
  for (int q = 0; q < NBF; ++q) {
    const int t = threadIdx.x + q * kThreads;
    if (t < total) {
      const int b = t & (B - 1), j = t >> lb;
      int jm;
      const int jd = divmod_small(j, Ns, rns, jm);
      dq[q] = ((jd * Ns * R + jm) << lb) + b;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = A[((j + r * m) << lb) + b];
#pragma unroll
      for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], twiddle<T, INV>(tw, jm * r * tws));
      dft_fixed<T2, R>(v[q], INV);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NBF; ++q) {
    const int t = threadIdx.x + q * kThreads;
    if (t < total) {
#pragma unroll
      for (int r = 0; r < R; ++r) A[dq[q] + ((r * Ns) << lb)] = v[q][r];
    }
  }
  __syncthreads();
}

// Any other factor R: direct R-point DFT with roots from the n-table.
template <class T, bool INV>
__device__ void stage_generic(const Pass<T> &p, const typename C2<T>::type *tw,
                              const typename C2<T>::type *in, typename C2<T>::type *out, int st,
                              int R) {
  using T2 = typename C2<T>::type;
  const int n = p.n, B = p.B, lb = p.lb, m = n / R;
  const int Ns = p.ns[st], tws = p.tws[st];
  const float rns = p.ns_rcp[st];
  const int root = n / R;  // exp(-2 pi i / R) = tw[n / R]
  for (int t = threadIdx.x; t < (m << lb); t += kThreads) {
    const int b = t & (B - 1), j = t >> lb;
    int jm;
    const int jd = divmod_small(j, Ns, rns, jm);
    const int d = jd * Ns * R + jm;
    for (int q = 0; q < R; ++q) {
      T2 acc;
      acc.x = 0;
      acc.y = 0;
      for (int r = 0; r < R; ++r) {
        T2 x = in[((j + r * m) << lb) + b];
        if (r > 0) x = cmul(x, twiddle<T, INV>(tw, jm * r * tws));
        acc = cadd(acc, cmul(x, twiddle<T, INV>(tw, ((q * r) % R) * root)));
      }
      out[((d + q * Ns) << lb) + b] = acc;
    }
  }
}

// Thread -> (line b, first element k0, element step ks) for staging B lines
// of n elements with the 512 threads of a workgroup: element-fastest when
// the axis is contiguous in memory, line-fastest when it is strided.
struct Stager {
  int b, k0, ks;
  __device__ Stager(bool kfast, int lb) {
    const int tid = threadIdx.x;
    if (kfast) {
      const int sh = kLogThreads - lb;  // 512 / B threads per line
      b = tid >> sh;
      k0 = tid & ((1 << sh) - 1);
      ks = 1 << sh;
    } else {
      b = tid & ((1 << lb) - 1);
      k0 = tid >> lb;
      ks = kThreads >> lb;
    }
  }
};

// Spin-grid line L = (g, x, y), g = 3 m + comp: index of its spin at z = 0
// (a spin at z has index base + z nb) and the component.
template <class T>
__device__ __forceinline__ int64_t spin_line(const Pass<T> &p, uint32_t L, int &comp) {
  const uint32_t cells = (uint32_t)p.grid.L[0] * (uint32_t)p.grid.L[1];
  const uint32_t g = L / cells, xy = L % cells;
  comp = (int)(g % 3u);
  return (int64_t)xy * p.grid.L[2] * p.grid.nb + (int)(g / 3u);
}

template <class T, int LD, int ST, bool INV>
__global__ void __launch_bounds__(kThreads) k_fft_pass(const Pass<T> p) {
  using T2 = typename C2<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  T2 *A = reinterpret_cast<T2 *>(smem);
  T2 *Bf = A + (size_t)p.n * p.B;
  const int n = p.n, lb = p.lb;
  const int64_t L0 = (int64_t)blockIdx.x << lb;
  // the twiddle table follows the line buffer(s); the barrier after the
  // load makes it visible
  T2 *tw = (p.pingpong ? Bf : A) + ((size_t)n << lb);
  for (int t = threadIdx.x; t < n; t += kThreads) tw[t] = p.tw[t];

  // ---- load: each thread keeps one line; its address is computed once ----
  {
    const Stager s(LD == LD_SPIN || p.src.se == 1, lb);
    const int64_t L = L0 + s.b;
    const bool live = L < p.nlines;
    int comp = 0;
    int64_t base = 0;
    if (live) {
      if constexpr (LD == LD_SPIN) base = spin_line(p, (uint32_t)L, comp);
      else base = p.src.base32((uint32_t)L);
    }
    for (int k = s.k0; k < n; k += s.ks) {
      T2 v;
      v.x = 0;
      v.y = 0;
      if (live) {
        if constexpr (LD == LD_SPIN) {
          if (k < p.n_in) {
            const int64_t idx = base + (int64_t)k * p.grid.nb;
            v.x = (T)(p.grid.mus[idx] * p.spins[spin_idx(p.grid.n, idx, comp)]);
          }
        } else if constexpr (LD == LD_HERM) {
          if (k < p.n_in) {
            v = p.in[base + (int64_t)k * p.src.se];
          } else {
            v = p.in[base + (int64_t)(n - k) * p.src.se];
            v.y = -v.y;
          }
        } else {
          if (k < p.n_in) v = p.in[base + (int64_t)k * p.src.se];
        }
      }
      A[(k << lb) + s.b] = v;
    }
  }
  __syncthreads();

  // ---- stages: in place on A, or ping-pong between A and Bf ----
  T2 *in = A, *out = Bf;
  if (!p.pingpong) {
    for (int st = 0; st < p.nst; ++st) {
      switch (p.radix[st]) {
        case 2: stage_inplace<T, 2, INV>(p, tw, A, st); break;
        case 3: stage_inplace<T, 3, INV>(p, tw, A, st); break;
        case 4: stage_inplace<T, 4, INV>(p, tw, A, st); break;
        case 5: stage_inplace<T, 5, INV>(p, tw, A, st); break;
        default: stage_inplace<T, 8, INV>(p, tw, A, st); break;
      }
    }
  } else {
    for (int st = 0; st < p.nst; ++st) {
      const int R = p.radix[st];
      switch (R) {
        case 2: stage_fixed<T, 2, INV>(p, tw, in, out, st); break;
        case 3: stage_fixed<T, 3, INV>(p, tw, in, out, st); break;
        case 4: stage_fixed<T, 4, INV>(p, tw, in, out, st); break;
        case 5: stage_fixed<T, 5, INV>(p, tw, in, out, st); break;
        case 8: stage_fixed<T, 8, INV>(p, tw, in, out, st); break;
        default: stage_generic<T, INV>(p, tw, in, out, st, R); break;
      }
      __syncthreads();
      T2 *tmp = in;
      in = out;
      out = tmp;
    }
  }

  // ---- store ----
  {
    const Stager s(ST == ST_FIELD || p.dst.se == 1, lb);
    const int64_t L = L0 + s.b;
    if (L < p.nlines) {
      int comp = 0;
      int64_t base;
      if constexpr (ST == ST_FIELD) base = spin_line(p, (uint32_t)L, comp);
      else base = p.dst.base32((uint32_t)L);
      for (int k = s.k0; k < p.n_out; k += s.ks) {
        const T2 v = in[(k << lb) + s.b];
        if constexpr (ST == ST_FIELD) {
          const int64_t idx = base + (int64_t)k * p.grid.nb;
          p.field[comp * p.grid.n + idx] = p.grid.mus[idx] * (double)v.x;
        } else {
          p.out[base + (int64_t)k * p.dst.se] = v;
        }
      }
    }
  }
}

}  // namespace fft
}  // namespace jh

#pragma clang fp contract(off)
