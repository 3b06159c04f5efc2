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

namespace jh {
namespace fft {

constexpr int kMaxStages = 24;
constexpr int kThreads = 512;

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
  int B;                 // lines per workgroup
  int inverse;
  int pingpong;          // 1: two LDS buffers (plans with a direct-DFT factor)
  int64_t nlines;
  LineMap src, dst;
  const T2 *in;
  T2 *out;
  const T2 *tw;          // n twiddles exp(-2 pi i t / n)
  // LD_SPIN: lines are (g, x, y), g = 3 m + component; value mus * s
  // ST_FIELD: the same lines; h[comp N + i] = mus_i * Re(value)
  // LD_HERM: the line holds the first n_in = n/2 + 1 terms of a Hermitian
  // sequence of length n
  SpinGrid grid;
  const double *spins;   // SoA FP64
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

template <class T>
__device__ __forceinline__ typename C2<T>::type twiddle(const Pass<T> &p, int64_t t) {
  typename C2<T>::type w = p.tw[t];
  if (p.inverse) w.y = -w.y;
  return w;
}

// One Stockham stage over B lines: in/out are LDS arrays [k][b].
template <class T, int R>
__device__ __forceinline__ void stage_fixed(const Pass<T> &p, const typename C2<T>::type *in,
                                            typename C2<T>::type *out, int Ns) {
  using T2 = typename C2<T>::type;
  const int n = p.n, B = p.B, m = n / R;
  const int64_t stride_t = (int64_t)n / ((int64_t)Ns * R);
  for (int t = threadIdx.x; t < m * B; t += blockDim.x) {
    const int b = t % B, j = t / B;
    const int jm = j % Ns;
    T2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[(int64_t)(j + r * m) * B + b];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle(p, (int64_t)jm * r * stride_t));
    dft_fixed<T2, R>(v, p.inverse != 0);
    const int d = (j / Ns) * Ns * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) out[(int64_t)(d + r * Ns) * B + b] = v[r];
  }
}

// Butterflies per thread of an in-place stage: R * kBF<R> <= 16 values in
// registers.
template <int R>
constexpr int kBF = (16 / R) > 0 ? 16 / R : 1;

// The same stage in place on one LDS buffer: every thread first reads and
// transforms its butterflies (at most kBF<R>, which the host guarantees by
// its choice of B), then all write.
template <class T, int R>
__device__ __forceinline__ void stage_inplace(const Pass<T> &p, typename C2<T>::type *A, int Ns) {
  using T2 = typename C2<T>::type;
  constexpr int NBF = kBF<R>;
  const int n = p.n, B = p.B, m = n / R, total = m * B;
  const int64_t stride_t = (int64_t)n / ((int64_t)Ns * R);
  T2 v[NBF][R];
#pragma unroll
  for (int q = 0; q < NBF; ++q) {
    const int t = threadIdx.x + q * blockDim.x;
    if (t < total) {
      const int b = t % B, j = t / B, jm = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = A[(int64_t)(j + r * m) * B + b];
#pragma unroll
      for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], twiddle(p, (int64_t)jm * r * stride_t));
      dft_fixed<T2, R>(v[q], p.inverse != 0);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NBF; ++q) {
    const int t = threadIdx.x + q * blockDim.x;
    if (t < total) {
      const int b = t % B, j = t / B, jm = j % Ns;
      const int d = (j / Ns) * Ns * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) A[(int64_t)(d + r * Ns) * B + b] = v[q][r];
    }
  }
  __syncthreads();
}

// Any other factor R: direct R-point DFT with roots from the n-table.
template <class T>
__device__ void stage_generic(const Pass<T> &p, const typename C2<T>::type *in,
                              typename C2<T>::type *out, int Ns, int R) {
  using T2 = typename C2<T>::type;
  const int n = p.n, B = p.B, m = n / R;
  const int64_t stride_t = (int64_t)n / ((int64_t)Ns * R);
  const int64_t root = n / R;  // exp(-2 pi i / R) = tw[n / R]
  for (int t = threadIdx.x; t < m * B; t += blockDim.x) {
    const int b = t % B, j = t / B;
    const int jm = j % Ns;
    const int d = (j / Ns) * Ns * R + jm;
    for (int q = 0; q < R; ++q) {
      T2 acc;
      acc.x = 0;
      acc.y = 0;
      for (int r = 0; r < R; ++r) {
        T2 x = in[(int64_t)(j + r * m) * B + b];
        if (r > 0) x = cmul(x, twiddle(p, (int64_t)jm * r * stride_t));
        acc = cadd(acc, cmul(x, twiddle(p, ((int64_t)q * r % R) * root)));
      }
      out[(int64_t)(d + q * Ns) * B + b] = acc;
    }
  }
}

template <class T, int LD, int ST>
__global__ void __launch_bounds__(kThreads) k_fft_pass(const Pass<T> p) {
  using T2 = typename C2<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  T2 *A = reinterpret_cast<T2 *>(smem);
  T2 *Bf = A + (size_t)p.n * p.B;
  const int n = p.n, B = p.B;
  const int64_t L0 = (int64_t)blockIdx.x * B;
  // element-fastest staging when the axis is contiguous, line-fastest else
  const bool kfast_in = LD == LD_SPIN || p.src.se == 1;
  const bool kfast_out = ST == ST_FIELD || p.dst.se == 1;

  // ---- load ----
  for (int q = threadIdx.x; q < n * B; q += blockDim.x) {
    const int b = kfast_in ? q / n : q % B;
    const int k = kfast_in ? q % n : q / B;
    const int64_t L = L0 + b;
    T2 v;
    v.x = 0;
    v.y = 0;
    if (L < p.nlines) {
      if constexpr (LD == LD_SPIN) {
        if (k < p.n_in) {
          const int64_t cells = (int64_t)p.grid.L[0] * p.grid.L[1];
          const int g = (int)(L / cells);
          const int64_t xy = L % cells;
          const int m = g / 3, comp = g % 3;
          const int64_t idx = (xy * p.grid.L[2] + k) * p.grid.nb + m;
          v.x = (T)(p.grid.mus[idx] * p.spins[comp * p.grid.n + idx]);
        }
      } else if constexpr (LD == LD_HERM) {
        if (k < p.n_in) {
          v = p.in[p.src.at(L, k)];
        } else {
          v = p.in[p.src.at(L, n - k)];
          v.y = -v.y;
        }
      } else {
        if (k < p.n_in) v = p.in[p.src.at(L, k)];
      }
    }
    A[(int64_t)k * B + b] = v;
  }
  __syncthreads();

  // ---- stages: in place on A, or ping-pong between A and Bf ----
  T2 *in = A, *out = Bf;
  int Ns = 1;
  if (!p.pingpong) {
    for (int st = 0; st < p.nst; ++st) {
      const int R = p.radix[st];
      switch (R) {
        case 2: stage_inplace<T, 2>(p, A, Ns); break;
        case 3: stage_inplace<T, 3>(p, A, Ns); break;
        case 4: stage_inplace<T, 4>(p, A, Ns); break;
        case 5: stage_inplace<T, 5>(p, A, Ns); break;
        default: stage_inplace<T, 8>(p, A, Ns); break;
      }
      Ns *= R;
    }
  } else
  for (int st = 0; st < p.nst; ++st) {
    const int R = p.radix[st];
    switch (R) {
      case 2: stage_fixed<T, 2>(p, in, out, Ns); break;
      case 3: stage_fixed<T, 3>(p, in, out, Ns); break;
      case 4: stage_fixed<T, 4>(p, in, out, Ns); break;
      case 5: stage_fixed<T, 5>(p, in, out, Ns); break;
      case 8: stage_fixed<T, 8>(p, in, out, Ns); break;
      default: stage_generic<T>(p, in, out, Ns, R); break;
    }
    __syncthreads();
    T2 *tmp = in;
    in = out;
    out = tmp;
    Ns *= R;
  }

  // ---- store ----
  for (int q = threadIdx.x; q < n * B; q += blockDim.x) {
    const int b = kfast_out ? q / n : q % B;
    const int k = kfast_out ? q % n : q / B;
    const int64_t L = L0 + b;
    if (L >= p.nlines || k >= p.n_out) continue;
    const T2 v = in[(int64_t)k * B + b];
    if constexpr (ST == ST_FIELD) {
      const int64_t cells = (int64_t)p.grid.L[0] * p.grid.L[1];
      const int g = (int)(L / cells);
      const int64_t xy = L % cells;
      const int m = g / 3, comp = g % 3;
      const int64_t idx = (xy * p.grid.L[2] + k) * p.grid.nb + m;
      p.field[comp * p.grid.n + idx] = p.grid.mus[idx] * (double)v.x;
    } else {
      p.out[p.dst.at(L, k)] = v;
    }
  }
}

}  // namespace fft
}  // namespace jh
