// jh_dipole.hip -- zero-padded FFT dipole field (hamiltonian/dipole_fft.cc:
// 35-350, hamiltonian/cuda_dipole_fft.cu:62-342), written for gfx950.
//
// h_i = mus_i sum_j mus_j W(r_ij) s_j with
//   W_ab(r) = mu0 / (4 pi a^3) (3 r_a r_b - r^2 delta_ab) / |r|^5
// (r in lattice constants, |r| <= r_cutoff + 1e-4, no self term;
// dipole_fft.cc:225-251), evaluated as a convolution on the padded cell grid:
// a periodic dimension keeps its L cells and takes the minimum image; an open
// dimension is padded to 2L and its kernel is two-sided (offsets -L < n < L),
// which is the brute-force sum (dipole_bruteforce.cc:75-120).  The reference's
// FFT forms are one-sided (CPU) or unpadded (GPU) on open dimensions; this
// build does not copy that.
//
// Per evaluation (spins s in the jh_layout.h layout, reference order):
//   z pass   r2c along the contiguous axis, fused with the scatter of mus s
//            into the zero-padded grid (only the L0 x L1 occupied lines);
//   y pass   forward, inputs pruned to y < L1, lines x < L0 only;
//   x pass   forward, inputs pruned to x < L0;
//   multiply H_i(k) = sum_j W_ij(k) S_j(k), in place;
//   x pass   inverse, outputs x < L0 only;
//   y pass   inverse, lines x < L0, outputs y < L1;
//   z pass   c2r, fused with the gather h_i = mus_i Re(...), outputs z < L2.
// After the z pass the grid is held as [g][q2][x][y] (y contiguous), so the
// y transforms read and write whole contiguous lines and the x transforms
// read rows of adjacent y lines.
// Where they apply, three kernels do this work with fewer trips to memory:
// k_dip_zy_fwd (z + y forward), k_dip_x_fused (x forward, multiply, x
// inverse; one basis site, real tensor) and k_dip_zy_inv (y + z inverse).
// They repeat the separate passes' operations line by line.
// The 1/P normalisation is folded into the tensor.  For one basis site W is
// even in n (except for the half-way image of an even periodic length inside
// r_cutoff), and the transformed tensor is then real: 6 values per k-point.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "jh_device.h"
#include "jh_dipole.h"
#include "jh_fft.h"

namespace jh {

using fft::LineMap;
using fft::Pass;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kMu0IU = 4.0 * kPi * 1e-7 / 6.24150907e21;  // helpers/consts.h:38
constexpr double kDistTol = 1e-4;                              // helpers/defaults.h:44
constexpr size_t kMaxPassLds = 64 * 1024;  // one line of the longest axis per workgroup
constexpr int kMaxBasis = 8;

#define DCHK(x)                              \
  do {                                       \
    hipError_t e_ = (x);                     \
    if (e_ != hipSuccess) return e_;         \
  } while (0)

template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t cnt) {
    release();
    n = cnt;
    return hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(cnt, 1));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
};

// radices of n: 8s first, then 4, 2, 5, 3, then any other prime factor
std::vector<int> factorise(int n) {
  std::vector<int> r;
  while (n % 8 == 0) { r.push_back(8); n /= 8; }
  if (n % 4 == 0) { r.push_back(4); n /= 4; }
  if (n % 2 == 0) { r.push_back(2); n /= 2; }
  for (int f : {5, 3}) while (n % f == 0) { r.push_back(f); n /= f; }
  for (int f = 7; n > 1; f += 2) while (n % f == 0) { r.push_back(f); n /= f; }
  return r;
}

// exp(-2 pi i t / n), t < n.  Entries of the first half are evaluated in long
// double and rounded once; for even n the second half is their exact
// negation, w(t + n/2) = -w(t), so the rounded table keeps that symmetry by
// construction (the 512-point kernels hold only the first half).
template <class T2>
std::vector<T2> twiddles(int n) {
  std::vector<T2> w((size_t)n);
  const int half = (n % 2 == 0) ? n / 2 : n;
  for (int t = 0; t < half; ++t) {
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / n;
    w[t].x = (decltype(w[t].x))cosl(a);
    w[t].y = (decltype(w[t].y))sinl(a);
  }
  for (int t = half; t < n; ++t) {
    w[t].x = -w[t - half].x;
    w[t].y = -w[t - half].y;
  }
  return w;
}

template <class T>
__global__ void __launch_bounds__(256) k_dip_mult_real(typename fft::C2<T>::type *K,
                                                       const T *W, int64_t M) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M) return;
  using T2 = typename fft::C2<T>::type;
  const T2 sx = K[p], sy = K[M + p], sz = K[2 * M + p];
  const T xx = W[p], xy = W[M + p], xz = W[2 * M + p], yy = W[3 * M + p], yz = W[4 * M + p],
          zz = W[5 * M + p];
  T2 hx, hy, hz;
  hx.x = xx * sx.x + xy * sy.x + xz * sz.x; hx.y = xx * sx.y + xy * sy.y + xz * sz.y;
  hy.x = xy * sx.x + yy * sy.x + yz * sz.x; hy.y = xy * sx.y + yy * sy.y + yz * sz.y;
  hz.x = xz * sx.x + yz * sy.x + zz * sz.x; hz.y = xz * sx.y + yz * sy.y + zz * sz.y;
  K[p] = hx;
  K[M + p] = hy;
  K[2 * M + p] = hz;
}

// several basis sites: complex tensors per ordered pair (i, j), 6 per k
template <class T>
__global__ void __launch_bounds__(256) k_dip_mult_cplx(typename fft::C2<T>::type *K,
                                                       const typename fft::C2<T>::type *W,
                                                       int64_t M, int nb) {
  using T2 = typename fft::C2<T>::type;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M) return;
  T2 S[3 * kMaxBasis];
  for (int g = 0; g < 3 * nb; ++g) S[g] = K[(int64_t)g * M + p];
  for (int i = 0; i < nb; ++i) {
    T2 h[3];
    for (int a = 0; a < 3; ++a) { h[a].x = 0; h[a].y = 0; }
    for (int j = 0; j < nb; ++j) {
      const T2 *Wij = W + (int64_t)(i * nb + j) * 6 * M + p;
      // symmetric tensor: xx xy xz yy yz zz
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const int c = a <= b ? (a == 0 ? b : a + b + 1) : (b == 0 ? a : a + b + 1);
          h[a] = fft::cadd(h[a], fft::cmul(Wij[c * M], S[3 * j + b]));
        }
    }
    for (int a = 0; a < 3; ++a) K[(int64_t)(3 * i + a) * M + p] = h[a];
  }
}

// ---------------------------------------------------------------------------
// x transforms and the multiply in one kernel (one basis site, real tensor).
// A workgroup owns `by` adjacent y lines of one q2 plane for the three
// components (lw = 3 by lines): it loads their x < L0 terms, runs the forward
// x transform in LDS, multiplies by W, runs the inverse and stores x < L0.
// Every line goes through the stages, twiddles and operations of the
// separate passes, so the field is the same to the bit; the full-length x
// spectra never leave LDS (for a 256 x 256 x 8 film padded to 512 x 512 x 16
// that removes two thirds of the bytes the x passes and the multiply move).
template <class T>
struct XFused {
  using T2 = typename fft::C2<T>::type;
  int n, nst;
  int radix[fft::kMaxStages];
  int ns[fft::kMaxStages];
  float ns_rcp[fft::kMaxStages];
  int tws[fft::kMaxStages];
  int L0;          // occupied x (inputs and outputs)
  int by, lby;     // y lines per workgroup, by = 1 << lby
  int lw;          // 3 by lines in LDS
  float lw_rcp;
  int P1, nyb;     // y lines per plane, y blocks per plane
  int64_t M, sq;   // component stride, q2-plane stride (P0 P1)
  T2 *K;
  const T *W;      // [y block][6][n][by]
  const T2 *tw;
};

// one in-place stage over the lw lines of A ([k][line]) of n points; thread
// t handles line t mod lw, butterfly t / lw (the host keeps every thread
// within kBF).  PP: XFused or ZYFused (n, lw, lw_rcp, ns, ns_rcp, tws).
template <class T, int R, bool INV, class PP>
__device__ __forceinline__ void xf_stage(const PP &p, const typename fft::C2<T>::type *tw,
                                         typename fft::C2<T>::type *A, int st) {
  using T2 = typename fft::C2<T>::type;
  constexpr int NBF = fft::kBF<T, R>;
  const int n = p.n, lw = p.lw, m = n / R, total = m * lw, nt = blockDim.x;
  const int Ns = p.ns[st], tws = p.tws[st];
  const float rns = p.ns_rcp[st];
  T2 v[NBF][R];
  int dq[NBF];
#pragma unroll
  for (int q = 0; q < NBF; ++q) {
    const int t = threadIdx.x + q * nt;
    if (t < total) {
      int l, jm;
      const int j = fft::divmod_small(t, lw, p.lw_rcp, l);
      const int jd = fft::divmod_small(j, Ns, rns, jm);
      dq[q] = (jd * Ns * R + jm) * lw + l;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = A[(j + r * m) * lw + l];
#pragma unroll
      for (int r = 1; r < R; ++r) v[q][r] = fft::cmul(v[q][r], fft::twiddle<T, INV>(tw, jm * r * tws));
      fft::dft_fixed<T2, R>(v[q], INV);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NBF; ++q) {
    const int t = threadIdx.x + q * nt;
    if (t < total) {
#pragma unroll
      for (int r = 0; r < R; ++r) A[dq[q] + r * Ns * lw] = v[q][r];
    }
  }
  __syncthreads();
}

template <class T, bool INV, class PP>
__device__ __forceinline__ void xf_stages(const PP &p, const typename fft::C2<T>::type *tw,
                                          typename fft::C2<T>::type *A) {
  for (int st = 0; st < p.nst; ++st) {
    switch (p.radix[st]) {
      case 2: xf_stage<T, 2, INV>(p, tw, A, st); break;
      case 3: xf_stage<T, 3, INV>(p, tw, A, st); break;
      case 4: xf_stage<T, 4, INV>(p, tw, A, st); break;
      case 5: xf_stage<T, 5, INV>(p, tw, A, st); break;
      default: xf_stage<T, 8, INV>(p, tw, A, st); break;
    }
  }
}

// Twiddle t < 512 of a 512-point line from the first half of the table:
// exp(-2 pi i (t + 256) / 512) = -exp(-2 pi i t / 512), and twiddles() builds
// the second half as the exact negation of the first, so the value is the
// full table's.  The half table (4 KB in FP64) lets a third fused x
// workgroup fit per CU.
constexpr int kTw512 = 256;
template <class T, bool INV>
__device__ __forceinline__ typename fft::C2<T>::type twiddle512(const typename fft::C2<T>::type *tw,
                                                                 int t) {
  typename fft::C2<T>::type w = tw[t & (kTw512 - 1)];
  if (t & kTw512) {
    w.x = -w.x;
    w.y = -w.y;
  }
  if (INV) w.y = -w.y;
  return w;
}

// 512-point lines (radix 8, 8, 8) and LW lines known at compile time: the
// thread's line l = t mod LW and butterfly j = t / LW are the same in every
// stage, the stage's sub-length is a power of two, and the LDS reads and
// writes of a butterfly sit at constant offsets from one base address.  The
// operations on the data are those of xf_stage.
template <class T, int NS, bool INV, int LW>
__device__ __forceinline__ void xf_stage512(const typename fft::C2<T>::type *tw,
                                            typename fft::C2<T>::type *A, int j, int l) {
  using T2 = typename fft::C2<T>::type;
  constexpr int R = 8, m = 512 / R, tws = 512 / (NS * R);
  const bool live = j < m;
  T2 v[R];
  int d = 0;
  if (live) {
    const int jm = j & (NS - 1), jd = j / NS;
    d = (jd * NS * R + jm) * LW + l;
    const T2 *a = A + j * LW + l;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = a[r * m * LW];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = fft::cmul(v[r], twiddle512<T, INV>(tw, jm * r * tws));
    fft::dft_fixed<T2, R>(v, INV);
  }
  __syncthreads();
  if (live) {
#pragma unroll
    for (int r = 0; r < R; ++r) A[d + r * NS * LW] = v[r];
  }
  __syncthreads();
}

template <class T, bool INV, int LW>
__device__ __forceinline__ void xf_stages512(const typename fft::C2<T>::type *tw,
                                             typename fft::C2<T>::type *A) {
  const int t = threadIdx.x;
  const int j = t / LW, l = t - j * LW;
  xf_stage512<T, 1, INV, LW>(tw, A, j, l);
  xf_stage512<T, 8, INV, LW>(tw, A, j, l);
  xf_stage512<T, 64, INV, LW>(tw, A, j, l);
}

// e = q lw + l with a compile-time lw when LW > 0
template <int LW>
__device__ __forceinline__ int divmod_lw(int e, int lw, float rcp, int &l) {
  if constexpr (LW > 0) {
    const int q = e / LW;
    l = e - q * LW;
    return q;
  } else {
    return fft::divmod_small(e, lw, rcp, l);
  }
}

// the y or x transform of a fused kernel: LW > 0 selects the 512-point stages
template <class T, bool INV, int LW, class PP>
__device__ __forceinline__ void line_stages(const PP &p, const typename fft::C2<T>::type *tw,
                                            typename fft::C2<T>::type *A) {
  if constexpr (LW > 0) xf_stages512<T, INV, LW>(tw, A);
  else xf_stages<T, INV>(p, tw, A);
}

// LWK > 0: 512-point x lines and LWK = 3 by lines (compile-time geometry)
template <class T, int LWK>
__global__ void __launch_bounds__(512) k_dip_x_fused(const XFused<T> p) {
  using T2 = typename fft::C2<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = LWK ? 512 : p.n, lw = LWK ? LWK : p.lw;
  const int by = LWK ? LWK / 3 : p.by, lby = LWK ? (LWK == 6 ? 1 : LWK == 12 ? 2 : 0) : p.lby;
  const int nt = blockDim.x;
  T2 *A = reinterpret_cast<T2 *>(smem);
  T2 *tw = A + (size_t)n * lw;
  for (int t = threadIdx.x; t < (LWK ? kTw512 : n); t += nt) tw[t] = p.tw[t];
  // consecutive blocks on one XCD share the L2 lines of adjacent y
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int q2 = (int)(blk / p.nyb), y0 = (int)(blk % p.nyb) << lby;
  T2 *K = p.K + (int64_t)q2 * p.sq + y0;
  // load x < L0 (zero padding above), element e = x lw + c by + yb
  for (int e = threadIdx.x; e < n * lw; e += nt) {
    int l;
    const int x = divmod_lw<LWK>(e, lw, p.lw_rcp, l);
    const int c = l >> lby, yb = l & (by - 1);
    T2 v;
    v.x = 0;
    v.y = 0;
    if (x < p.L0 && y0 + yb < p.P1) v = K[c * p.M + (int64_t)x * p.P1 + yb];
    A[e] = v;
  }
  __syncthreads();
  line_stages<T, false, LWK>(p, tw, A);
  // H(k) = W(k) S(k) with the arithmetic of k_dip_mult_real
  {
    const int nby = n * by;
    const T *W = p.W + blk * 6 * nby;
    for (int e = threadIdx.x; e < nby; e += nt) {
      T2 *a = A + (e >> lby) * lw + (e & (by - 1));
      const T2 sx = a[0], sy = a[by], sz = a[2 * by];
      const T xx = W[e], xy = W[nby + e], xz = W[2 * nby + e], yy = W[3 * nby + e],
              yz = W[4 * nby + e], zz = W[5 * nby + e];
      T2 hx, hy, hz;
      hx.x = xx * sx.x + xy * sy.x + xz * sz.x; hx.y = xx * sx.y + xy * sy.y + xz * sz.y;
      hy.x = xy * sx.x + yy * sy.x + yz * sz.x; hy.y = xy * sx.y + yy * sy.y + yz * sz.y;
      hz.x = xz * sx.x + yz * sy.x + zz * sz.x; hz.y = xz * sx.y + yz * sy.y + zz * sz.y;
      a[0] = hx;
      a[by] = hy;
      a[2 * by] = hz;
    }
  }
  __syncthreads();
  line_stages<T, true, LWK>(p, tw, A);
  // store x < L0
  for (int e = threadIdx.x; e < p.L0 * lw; e += nt) {
    int l;
    const int x = divmod_lw<LWK>(e, lw, p.lw_rcp, l);
    const int c = l >> lby, yb = l & (by - 1);
    if (y0 + yb < p.P1) K[c * p.M + (int64_t)x * p.P1 + yb] = A[e];
  }
}

// the tensor in the fused kernel's order: [y block][6][x][by] from [6][q2][x][y]
template <class T>
__global__ void __launch_bounds__(256) k_dip_w_fused(const T *Wr, T *Wf, int64_t M, int n, int P1,
                                                     int lby, int nyb, int64_t total) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int by = 1 << lby;
  const int yb = (int)(q & (by - 1));
  int64_t r = q >> lby;
  const int x = (int)(r % n);
  r /= n;
  const int c = (int)(r % 6);
  const int64_t blk = r / 6;
  const int64_t q2 = blk / nyb;
  const int y = (int)(blk % nyb) * by + yb;
  Wf[q] = y < P1 ? Wr[c * M + q2 * n * P1 + (int64_t)x * P1 + y] : (T)0;
}

// Short contiguous axis (P2 <= kDirectP, L2 <= kDirectL): one thread per line
// evaluates the DFT directly in registers, fused with the spin scatter
// (forward, real input z < L2, outputs q < Q2) or the field gather (inverse,
// Hermitian input q < Q2, real outputs z < L2).  A workgroup of T
// consecutive lines stages its inputs and outputs through LDS so that global
// loads and stores run in memory order (consecutive lines are adjacent in
// both the spin arrays and the padded grid); each thread then reads its own
// line from LDS (odd strides: no bank conflicts).
constexpr int kDirectP = 64, kDirectL = 32;
constexpr size_t kDirectLds = 64 * 1024;

__host__ __device__ inline int odd_stride(int len) { return len | 1; }

// LDS bytes of a direct z workgroup of T lines: twiddles, the complex side
// (outputs forward, inputs inverse), the real side and the line tables
template <class T>
__host__ __device__ inline size_t direct_stage(int T_, int n_in, int n_out, bool inverse) {
  using T2 = typename fft::C2<T>::type;
  const int nc = inverse ? n_in : n_out, nr = inverse ? n_out : n_in;
  const size_t c = (size_t)T_ * odd_stride(nc) * sizeof(T2), r = (size_t)T_ * odd_stride(nr) * sizeof(T);
  return ((c > r ? c : r) + 15) & ~(size_t)15;
}
// the input and output staging share one area (they are live one after the
// other, with a barrier between)
template <class T>
size_t direct_lds(int T_, int n, int n_in, int n_out, bool inverse) {
  using T2 = typename fft::C2<T>::type;
  return (size_t)n * sizeof(T2) + direct_stage<T>(T_, n_in, n_out, inverse) +
         (size_t)T_ * (sizeof(int64_t) * 2 + sizeof(int));
}

// LM: register bound of the line (L2 <= LM, P2 <= 2 LM), 8, 16 or 32
template <class T, int LM>
__global__ void __launch_bounds__(256) k_dip_z_direct_fwd(const Pass<T> p) {
  using T2 = typename fft::C2<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int nt = blockDim.x, tid = threadIdx.x;  // nt: a power of two
  const int lnt = 31 - __clz(nt);
  const int n = p.n, n_in = p.n_in, n_out = p.n_out;
  const int SV = odd_stride(n_in), SO = odd_stride(n_out);
  T2 *tw = reinterpret_cast<T2 *>(smem);
  T2 *outs = tw + n;
  T *vals = reinterpret_cast<T *>(outs);  // the same area, used first
  int64_t *lidx = reinterpret_cast<int64_t *>(reinterpret_cast<unsigned char *>(outs) +
                                              direct_stage<T>(nt, n_in, n_out, false));
  int64_t *lout = lidx + nt;
  int *lcomp = reinterpret_cast<int *>(lout + nt);
  const int64_t L0 = (int64_t)blockIdx.x * nt;
  for (int t = tid; t < n; t += nt) tw[t] = p.tw[t];
  {
    const int64_t L = L0 + tid;
    int comp = 0;
    int64_t ib = -1, ob = -1;
    if (L < p.nlines) {
      ib = fft::spin_line(p, (uint32_t)L, comp);
      ob = p.dst.base32((uint32_t)L);
    }
    lidx[tid] = ib;
    lout[tid] = ob;
    lcomp[tid] = comp;
  }
  __syncthreads();
  // spins of the workgroup's lines, in memory order
  const float rin = 1.0f / (float)n_in, rout = 1.0f / (float)n_out;
  for (int e = tid; e < nt * n_in; e += nt) {
    int z;
    const int b = fft::divmod_small(e, n_in, rin, z);
    const int64_t base = lidx[b];
    T v = 0;
    if (base >= 0) {
      const int64_t idx = base + (int64_t)z * p.grid.nb;
      v = (T)(p.grid.mus[idx] * p.spins[spin_idx(p.grid.n, idx, lcomp[b])]);
    }
    vals[b * SV + z] = v;
  }
  __syncthreads();
  {
    T v[LM];
#pragma unroll
    for (int z = 0; z < LM; ++z) v[z] = z < n_in ? vals[tid * SV + z] : (T)0;
    __syncthreads();  // vals is overwritten by outs below
    for (int q = 0; q < n_out; ++q) {
      T2 acc;
      acc.x = 0;
      acc.y = 0;
      int t = 0;  // (q z) mod n
#pragma unroll
      for (int z = 0; z < LM; ++z) {
        if (z < n_in) {
          const T2 w = tw[t];
          acc.x = acc.x + v[z] * w.x;
          acc.y = acc.y + v[z] * w.y;
          t += q;
          if (t >= n) t -= n;
        }
      }
      outs[tid * SO + q] = acc;
    }
  }
  __syncthreads();
  // grid side: element-fastest when q is contiguous, line-fastest otherwise
  const bool qfast = p.dst.se == 1;
  for (int e = tid; e < nt * n_out; e += nt) {
    int q, b;
    if (qfast) b = fft::divmod_small(e, n_out, rout, q);
    else { q = e >> lnt; b = e & (nt - 1); }
    const int64_t base = lout[b];
    if (base >= 0) p.out[base + (int64_t)q * p.dst.se] = outs[b * SO + q];
  }
}

template <class T, int LM>
__global__ void __launch_bounds__(256) k_dip_z_direct_inv(const Pass<T> p) {
  using T2 = typename fft::C2<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int nt = blockDim.x, tid = threadIdx.x;  // nt: a power of two
  const int lnt = 31 - __clz(nt);
  const int n = p.n, n_in = p.n_in, n_out = p.n_out;
  const int SX = odd_stride(n_in), SZ = odd_stride(n_out);
  T2 *tw = reinterpret_cast<T2 *>(smem);
  T2 *xs = tw + n;
  T *outs = reinterpret_cast<T *>(xs);  // the same area, used second
  int64_t *lidx = reinterpret_cast<int64_t *>(reinterpret_cast<unsigned char *>(xs) +
                                              direct_stage<T>(nt, n_in, n_out, true));
  int64_t *lin = lidx + nt;
  int *lcomp = reinterpret_cast<int *>(lin + nt);
  const int64_t L0 = (int64_t)blockIdx.x * nt;
  for (int t = tid; t < n; t += nt) tw[t] = p.tw[t];
  {
    const int64_t L = L0 + tid;
    int comp = 0;
    int64_t fb = -1, ib = -1;
    if (L < p.nlines) {
      fb = fft::spin_line(p, (uint32_t)L, comp);
      ib = p.src.base32((uint32_t)L);
    }
    lidx[tid] = fb;
    lin[tid] = ib;
    lcomp[tid] = comp;
  }
  __syncthreads();
  const float rin = 1.0f / (float)n_in, rout = 1.0f / (float)n_out;
  const bool qfast = p.src.se == 1;
  for (int e = tid; e < nt * n_in; e += nt) {
    int q, b;
    if (qfast) b = fft::divmod_small(e, n_in, rin, q);
    else { q = e >> lnt; b = e & (nt - 1); }
    const int64_t base = lin[b];
    T2 v;
    v.x = 0;
    v.y = 0;
    if (base >= 0) v = p.in[base + (int64_t)q * p.src.se];
    xs[b * SX + q] = v;
  }
  __syncthreads();
  {
    T2 X[LM + 1];
#pragma unroll
    for (int q = 0; q <= LM; ++q)
      if (q < n_in) X[q] = xs[tid * SX + q];
    __syncthreads();  // xs is overwritten by outs below
    for (int z = 0; z < n_out; ++z) {
      // Re sum_k X[k] e^{+2 pi i k z/n} over the Hermitian sequence:
      // X[0] + 2 Re(X[k] w^{kz}) for 0 < k < n/2 (+ X[n/2] (-1)^z, n even)
      T acc = X[0].x;
      int t = z;  // (k z) mod n, k = 1
#pragma unroll
      for (int k = 1; k <= LM; ++k) {
        if (k < n_in) {
          const T2 w = tw[t];  // e^{-2 pi i t/n}; the inverse uses its conjugate
          const T re = X[k].x * w.x + X[k].y * w.y;
          if (2 * k == n) acc = acc + re;
          else acc = acc + 2 * re;
          t += z;
          if (t >= n) t -= n;
        }
      }
      outs[tid * SZ + z] = acc;
    }
  }
  __syncthreads();
  for (int e = tid; e < nt * n_out; e += nt) {
    int z;
    const int b = fft::divmod_small(e, n_out, rout, z);
    const int64_t base = lidx[b];
    if (base >= 0) {
      const int64_t idx = base + (int64_t)z * p.grid.nb;
      p.field[lcomp[b] * p.grid.n + idx] = p.grid.mus[idx] * (double)outs[b * SZ + z];
    }
  }
}

// threads (lines) per direct z workgroup: 256, or fewer so that LDS stays
// within 64 KB
template <class T>
int direct_threads(int n, int n_in, int n_out, bool inverse) {
  int t = 256;
  while (t > 64 && direct_lds<T>(t, n, n_in, n_out, inverse) > kDirectLds) t /= 2;
  return t;
}

template <class T>
hipError_t run_direct_z(const Pass<T> &p, bool inverse, hipStream_t st) {
  if (p.nlines <= 0) return hipSuccess;
  if (p.nlines >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  const int t = direct_threads<T>(p.n, p.n_in, p.n_out, inverse);
  const size_t lds = direct_lds<T>(t, p.n, p.n_in, p.n_out, inverse);
  const unsigned blocks = (unsigned)((p.nlines + t - 1) / t);
  const int lmax = std::max(inverse ? p.n_out : p.n_in, (p.n + 1) / 2);
  if (lmax <= 8) {
    if (inverse) hipLaunchKernelGGL((k_dip_z_direct_inv<T, 8>), dim3(blocks), dim3(t), lds, st, p);
    else hipLaunchKernelGGL((k_dip_z_direct_fwd<T, 8>), dim3(blocks), dim3(t), lds, st, p);
  } else if (lmax <= 16) {
    if (inverse) hipLaunchKernelGGL((k_dip_z_direct_inv<T, 16>), dim3(blocks), dim3(t), lds, st, p);
    else hipLaunchKernelGGL((k_dip_z_direct_fwd<T, 16>), dim3(blocks), dim3(t), lds, st, p);
  } else {
    if (inverse) hipLaunchKernelGGL((k_dip_z_direct_inv<T, 32>), dim3(blocks), dim3(t), lds, st, p);
    else hipLaunchKernelGGL((k_dip_z_direct_fwd<T, 32>), dim3(blocks), dim3(t), lds, st, p);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// z forward and y forward in one kernel (one workgroup per (g, x < L0) and
// q2 range): the direct r2c DFT of the workgroup's L1 spin lines along z,
// written straight into LDS as lw lines along y (one per q2), zero padded to
// P1, then the y transform, then a contiguous store of each q2 line.  The
// occupied planes the separate z pass would write and the y pass read back
// never leave LDS.  The DFT and the stages repeat the separate kernels'
// operations, so the grid is the same to the bit.
// LDS bound of the fused y + z inverse kernel: within 48 KB three workgroups
// fit per CU (with a [y][q2] staging area in LDS instead of register sums
// the FP64 kernel needed 54-80 KB and ran 3-8 % slower than two separate
// kernels at a 256 x 256 x 8 film).
#ifndef JH_ZYI_LDS
#define JH_ZYI_LDS (48 * 1024)
#endif
// workgroups per (g, x) of the fused z + y forward kernel, at least
#ifndef JH_ZY_SPLIT
#define JH_ZY_SPLIT 3
#endif
template <class T>
struct ZYFused {
  using T2 = typename fft::C2<T>::type;
  // y transform (in place, lw lines of n = P1 points)
  int n, nst;
  int radix[fft::kMaxStages];
  int ns[fft::kMaxStages];
  float ns_rcp[fft::kMaxStages];
  int tws[fft::kMaxStages];
  int lw;          // q2 lines per workgroup
  float lw_rcp;
  int nsplit;      // workgroups per (g, x): q2 in [s lw, s lw + lw) ∩ [0, Q2)
  // z transform and the spin grid
  int P2, L2, Q2, L1, L0, nb;
  int64_t nspin;   // spins (component stride of the spin arrays)
  int64_t sa, sq;  // K strides of g and q2 (P0 P1 Q2, P0 P1)
  const double *spins, *mus;
  double *field;   // inverse: SoA field out
  T2 *K;
  const T2 *twy, *twz;
};

// LWK > 0: 512-point y lines and LWK q2 lines per workgroup (compile time)
template <class T, int LM, int LWK>
__global__ void __launch_bounds__(512) k_dip_zy_fwd(const ZYFused<T> p) {
  using T2 = typename fft::C2<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = LWK ? 512 : p.n, lw = LWK ? LWK : p.lw, nt = blockDim.x, tid = threadIdx.x;
  T2 *twy = reinterpret_cast<T2 *>(smem);
  T2 *twz = twy + (LWK ? kTw512 : n);
  T2 *A = twz + p.P2;                   // [y][line], P1 x lw
  T *vals = reinterpret_cast<T *>(A);   // the same area, used first: [y][z]
  for (int t = tid; t < (LWK ? kTw512 : n); t += nt) twy[t] = p.twy[t];
  for (int t = tid; t < p.P2; t += nt) twz[t] = p.twz[t];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int sp = (int)(blk % p.nsplit);
  const int64_t gx = blk / p.nsplit;
  const int x = (int)(gx % p.L0), g = (int)(gx / p.L0);
  const int m = g / 3, comp = g % 3;
  const int qa = sp * lw;
  const int nq = min(lw, p.Q2 - qa);
  const int L1 = p.L1, L2 = p.L2, SV = odd_stride(L2);
  // spins of cells (x, y < L1, z < L2), basis m, in memory order
  const int64_t base = (int64_t)x * L1 * L2 * p.nb + m;
  for (int e = tid; e < L1 * L2; e += nt) {
    const int64_t idx = base + (int64_t)e * p.nb;
    // (y, z) of element e: a shift when L2 fills the register bound LM
    const int y = L2 == LM ? e / LM : e / L2, z = e - y * L2;
    vals[y * SV + z] = (T)(p.mus[idx] * p.spins[spin_idx(p.nspin, idx, comp)]);
  }
  __syncthreads();
  // direct DFT along z of line y (the arithmetic of k_dip_z_direct_fwd)
  T v[LM];
  if (tid < L1) {
#pragma unroll
    for (int z = 0; z < LM; ++z) v[z] = z < L2 ? vals[tid * SV + z] : (T)0;
  }
  __syncthreads();  // vals is overwritten by A below
  if (tid < L1) {
    for (int l = 0; l < lw; ++l) {
      const int q = qa + l;
      T2 acc;
      acc.x = 0;
      acc.y = 0;
      if (l < nq) {
        int t = 0;  // (q z) mod P2
#pragma unroll
        for (int z = 0; z < LM; ++z) {
          if (z < L2) {
            const T2 w = twz[t];
            acc.x = acc.x + v[z] * w.x;
            acc.y = acc.y + v[z] * w.y;
            t += q;
            if (t >= p.P2) t -= p.P2;
          }
        }
      }
      A[tid * lw + l] = acc;
    }
  }
  // zero padding y >= L1
  for (int e = L1 * lw + tid; e < n * lw; e += nt) {
    T2 z0;
    z0.x = 0;
    z0.y = 0;
    A[e] = z0;
  }
  __syncthreads();
  line_stages<T, false, LWK>(p, twy, A);
  // q2 lines, y contiguous
  T2 *K = p.K + (int64_t)g * p.sa + (int64_t)qa * p.sq + (int64_t)x * n;
  for (int e = tid; e < n * nq; e += nt) {
    const int l = e / n, y = e - l * n;
    K[(int64_t)l * p.sq + y] = A[y * lw + l];
  }
}

// y inverse and z inverse in one kernel: one workgroup per (g, x < L0)
// transforms the Q2 q2 lines along y in batches of lw; after each batch the
// thread of line y < L1 adds the batch's terms to its c2r sums along z
// (registers, in the order k = 0, 1, ... of k_dip_z_direct_inv, so the sums
// are the same to the bit); then the field is stored in memory order.
// LWK > 0: 512-point y lines and batches of LWK q2 lines (compile time)
template <class T, int LM, int LWK>
__global__ void __launch_bounds__(512) k_dip_zy_inv(const ZYFused<T> p) {
  using T2 = typename fft::C2<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = LWK ? 512 : p.n, lw = LWK ? LWK : p.lw, nt = blockDim.x, tid = threadIdx.x;
  const int L1 = p.L1, L2 = p.L2, Q2 = p.Q2, P2 = p.P2;
  const int SZ = odd_stride(L2);
  T2 *twy = reinterpret_cast<T2 *>(smem);
  T2 *twz = twy + (LWK ? kTw512 : n);
  T2 *A = twz + P2;                      // [y][line], P1 x lw; later outs
  T *outs = reinterpret_cast<T *>(A);    // [y < L1][SZ]
  for (int t = tid; t < (LWK ? kTw512 : n); t += nt) twy[t] = p.twy[t];
  for (int t = tid; t < P2; t += nt) twz[t] = p.twz[t];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int x = (int)(blk % p.L0), g = (int)(blk / p.L0);
  const int m = g / 3, comp = g % 3;
  const T2 *K = p.K + (int64_t)g * p.sa + (int64_t)x * n;
  // c2r sums of line y = tid: X[0] + 2 Re(X[k] w^{kz}) for 0 < k < P2/2
  // (+ X[P2/2] (-1)^z, P2 even)
  T acc[LM];
  for (int q0 = 0; q0 < Q2; q0 += lw) {
    const int nq = min(lw, Q2 - q0);
    for (int e = tid; e < n * lw; e += nt) {
      const int l = e / n, y = e - l * n;
      T2 v;
      v.x = 0;
      v.y = 0;
      if (l < nq) v = K[(int64_t)(q0 + l) * p.sq + y];
      A[y * lw + l] = v;
    }
    __syncthreads();
    line_stages<T, true, LWK>(p, twy, A);
    if (tid < L1) {
      for (int l = 0; l < nq; ++l) {
        const int k = q0 + l;
        const T2 X = A[tid * lw + l];
        int t = 0;  // (k z) mod P2
#pragma unroll
        for (int z = 0; z < LM; ++z) {
          if (z < L2) {
            if (k == 0) {
              acc[z] = X.x;
            } else {
              const T2 w = twz[t];  // e^{-2 pi i kz/P2}; conjugated below
              const T re = X.x * w.x + X.y * w.y;
              if (2 * k == P2) acc[z] = acc[z] + re;
              else acc[z] = acc[z] + 2 * re;
            }
            t += k;
            if (t >= P2) t -= P2;
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid < L1) {
#pragma unroll
    for (int z = 0; z < LM; ++z)
      if (z < L2) outs[tid * SZ + z] = acc[z];
  }
  __syncthreads();
  const int64_t base = (int64_t)x * L1 * L2 * p.nb + m;
  for (int e = tid; e < L1 * L2; e += nt) {
    const int y = L2 == LM ? e / LM : e / L2, z = e - y * L2;
    const int64_t idx = base + (int64_t)e * p.nb;
    p.field[comp * p.nspin + idx] = p.mus[idx] * (double)outs[y * SZ + z];
  }
}

// tensor set-up: the q2 < Q2 half of the transformed grid into the tensor
// arrays (real part for one basis site)
template <class T>
__global__ void __launch_bounds__(256) k_dip_extract(const double2 *F, int64_t P2, int64_t Q2,
                                                     int64_t rows, int ncomp, bool real_only,
                                                     T *Wr, typename fft::C2<T>::type *Wc) {
  const int64_t M = rows * Q2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= M * ncomp) return;
  // output [g][q2][row], row = x P1 + y; F is [g][row][P2]
  const int64_t g = q / M, r = q % M;
  const double2 v = F[g * rows * P2 + (r % rows) * P2 + r / rows];
  if (real_only) {
    Wr[q] = (T)v.x;
  } else {
    Wc[q].x = (T)v.x;
    Wc[q].y = (T)v.y;
  }
}

template <class T>
size_t pass_lds(int n, int B, bool pingpong) {
  return (pingpong ? 2 : 1) * (size_t)n * B * sizeof(typename fft::C2<T>::type);
}
// the same plus the twiddle table (at most 64 KB + 32 KB: gfx950 gives a
// workgroup up to 160 KB)
template <class T>
size_t pass_lds_total(int n, int B, bool pingpong) {
  return pass_lds<T>(n, B, pingpong) + (size_t)n * sizeof(typename fft::C2<T>::type);
}

// Lines per workgroup: as many as LDS (64 KB) and, for in-place plans, the
// per-thread butterfly budget of every stage allow (at most 64).
template <class T>
int pick_lines(const Pass<T> &p) {
  const int n = p.n;
  int B = (int)std::min<size_t>(64, (64 * 1024) / pass_lds<T>(n, 1, p.pingpong));
  if (!p.pingpong)
    for (int st = 0; st < p.nst; ++st) {
      const int R = p.radix[st];
      const int nbf = R == 2 ? fft::kBF<T, 2> : R == 3 ? fft::kBF<T, 3> : R == 4 ? fft::kBF<T, 4>
                    : R == 5 ? fft::kBF<T, 5> : fft::kBF<T, 8>;
      B = std::min<int>(B, (int)((int64_t)nbf * fft::kThreads / (n / R)));
    }
  B = (int)std::min<int64_t>(std::max(B, 1), std::max<int64_t>(p.nlines, 1));
  // a power of two (lines map to threads by shifts and masks)
  int P2 = 1;
  while (2 * P2 <= B) P2 *= 2;
  return P2;
}

template <class T, int LD, int ST>
hipError_t run_pass(Pass<T> p, hipStream_t st) {
  if (p.nlines <= 0) return hipSuccess;
  p.pingpong = 0;
  for (int s = 0; s < p.nst; ++s)
    if (p.radix[s] != 2 && p.radix[s] != 3 && p.radix[s] != 4 && p.radix[s] != 5 && p.radix[s] != 8)
      p.pingpong = 1;
  if (p.nlines >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  p.B = pick_lines<T>(p);
  p.lb = 0;
  while ((1 << p.lb) < p.B) ++p.lb;
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    p.ns[s] = Ns;
    p.ns_rcp[s] = 1.0f / (float)Ns;
    p.tws[s] = p.n / (Ns * p.radix[s]);
    Ns *= p.radix[s];
  }
  const size_t lds = pass_lds_total<T>(p.n, p.B, p.pingpong);
  const int64_t blocks = (p.nlines + p.B - 1) / p.B;
  // above 64 KB of dynamic LDS the kernel has to be told so
  const void *fn = p.inverse ? reinterpret_cast<const void *>(&fft::k_fft_pass<T, LD, ST, true>)
                             : reinterpret_cast<const void *>(&fft::k_fft_pass<T, LD, ST, false>);
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  if (p.inverse)
    hipLaunchKernelGGL((fft::k_fft_pass<T, LD, ST, true>), dim3((unsigned)blocks),
                       dim3(fft::kThreads), lds, st, p);
  else
    hipLaunchKernelGGL((fft::k_fft_pass<T, LD, ST, false>), dim3((unsigned)blocks),
                       dim3(fft::kThreads), lds, st, p);
  return hipGetLastError();
}

// the fused y + z inverse kernel for a compile-time lwk (3, 5) or none (0)
template <class T, int LM>
hipError_t launch_zy_inv(const ZYFused<T> &q, int lwk, unsigned blocks, int nt, size_t lds,
                         hipStream_t st) {
  const void *fn = lwk == 3   ? reinterpret_cast<const void *>(&k_dip_zy_inv<T, LM, 3>)
                   : lwk == 5 ? reinterpret_cast<const void *>(&k_dip_zy_inv<T, LM, 5>)
                              : reinterpret_cast<const void *>(&k_dip_zy_inv<T, LM, 0>);
  if (lds > 64 * 1024)
    DCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (lwk == 3) hipLaunchKernelGGL((k_dip_zy_inv<T, LM, 3>), dim3(blocks), dim3(nt), lds, st, q);
  else if (lwk == 5) hipLaunchKernelGGL((k_dip_zy_inv<T, LM, 5>), dim3(blocks), dim3(nt), lds, st, q);
  else hipLaunchKernelGGL((k_dip_zy_inv<T, LM, 0>), dim3(blocks), dim3(nt), lds, st, q);
  return hipGetLastError();
}

// per-precision device state
template <class T>
struct Plane {
  using T2 = typename fft::C2<T>::type;
  DBuf<T2> K;            // [3 nb][Q2][P0][P1]
  DBuf<T> Wr;            // nb == 1: [6][Q2 P0 P1]
  DBuf<T2> Wc;           // nb > 1: [nb nb][6][Q2 P0 P1]
  DBuf<T> Wf;            // fused x kernel: [Q2 y blocks][6][P0][by] (Wr released)
  int by = 0;            // y lines per fused workgroup, 0: separate passes
  DBuf<T2> tw[3];
  bool ready = false;
};

class DipoleFFTImpl : public DipoleFFT {
 public:
  int L[3], P[3], per[3], nb;
  int64_t n, Q2, M;  // M = P0 P1 Q2
  bool real_tensor = false;  // W(n) = W(-n): the transformed tensor is real
  std::vector<int> rad[3];
  DBuf<double> mus;
  Plane<double> f64;
  Plane<float> f32;
  DBuf<double2> Fsetup;  // transformed tensor grid (FP64), kept until both planes are built

  template <class T>
  Pass<T> base_pass(Plane<T> &pl, int axis, bool inverse) {
    Pass<T> p;
    std::memset(&p, 0, sizeof p);
    p.n = P[axis];
    p.nst = (int)rad[axis].size();
    for (int s = 0; s < p.nst; ++s) p.radix[s] = rad[axis][s];
    p.inverse = inverse ? 1 : 0;
    p.tw = pl.tw[axis].p;
    p.grid.n = n;
    p.grid.L[0] = L[0];
    p.grid.L[1] = L[1];
    p.grid.L[2] = L[2];
    p.grid.nb = nb;
    p.grid.mus = mus.p;
    return p;
  }

  // threads of the fused x kernel for by y lines: enough that no thread
  // holds more than kBF butterflies of any stage
  template <class T>
  int xf_threads(int by) const {
    const int n = P[0], lw = 3 * by;
    int need = 64;
    for (int R : rad[0]) {
      const int kbf = R == 2 ? fft::kBF<T, 2> : R == 3 ? fft::kBF<T, 3> : R == 4 ? fft::kBF<T, 4>
                    : R == 5 ? fft::kBF<T, 5> : fft::kBF<T, 8>;
      need = std::max(need, (n / R * lw + kbf - 1) / kbf);
    }
    return (need + 63) / 64 * 64;
  }
  // axis of 512 points factored 8 x 8 x 8: the fused kernels' compile-time
  // stage geometry applies
  bool lines512(int axis) const { return P[axis] == 512 && rad[axis] == std::vector<int>{8, 8, 8}; }
  // the compile-time line count of each fused kernel (0: general stages); the
  // LDS sizes and the launches both follow these
  int x_lwk(int lw) const { return lines512(0) && lw == 6 ? 6 : 0; }
  int zy_fwd_lwk(int lw) const { return lines512(1) && lw == 3 ? 3 : 0; }
  int zy_inv_lwk(int lw) const { return lines512(1) && (lw == 3 || lw == 5) ? lw : 0; }

  // y lines per workgroup of the fused x kernel (0: not applicable): one
  // basis site with a real tensor, in-place radices, LDS within 64 KB (two
  // workgroups per CU), at most 512 threads
  template <class T>
  int fused_lines() const {
    if (nb != 1 || !real_tensor) return 0;
    for (int r : rad[0])
      if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) return 0;
    const size_t c = sizeof(typename fft::C2<T>::type);
    for (int by = 8; by >= 1; by /= 2)
      if ((size_t)P[0] * (3 * by + 1) * c <= kMaxPassLds && xf_threads<T>(by) <= 512) return by;
    return 0;
  }

  template <class T>
  hipError_t run_x_fused(Plane<T> &pl, hipStream_t st) {
    using T2 = typename fft::C2<T>::type;
    XFused<T> q;
    std::memset(&q, 0, sizeof q);
    q.n = P[0];
    q.nst = (int)rad[0].size();
    int Ns = 1;
    for (int k = 0; k < q.nst; ++k) {
      q.radix[k] = rad[0][k];
      q.ns[k] = Ns;
      q.ns_rcp[k] = 1.0f / (float)Ns;
      q.tws[k] = q.n / (Ns * q.radix[k]);
      Ns *= q.radix[k];
    }
    q.L0 = L[0];
    q.by = pl.by;
    while ((1 << q.lby) < q.by) ++q.lby;
    q.lw = 3 * q.by;
    q.lw_rcp = 1.0f / (float)q.lw;
    q.P1 = P[1];
    q.nyb = (P[1] + q.by - 1) / q.by;
    q.M = M;
    q.sq = (int64_t)P[0] * P[1];
    q.K = pl.K.p;
    q.W = pl.Wf.p;
    q.tw = pl.tw[0].p;
    const size_t lds = ((size_t)q.n * q.lw + (x_lwk(q.lw) ? kTw512 : q.n)) * sizeof(T2);
    const unsigned blocks = (unsigned)(Q2 * q.nyb);
    if (x_lwk(q.lw) == 6)
      hipLaunchKernelGGL((k_dip_x_fused<T, 6>), dim3(blocks), dim3(xf_threads<T>(q.by)), lds, st, q);
    else
      hipLaunchKernelGGL((k_dip_x_fused<T, 0>), dim3(blocks), dim3(xf_threads<T>(q.by)), lds, st, q);
    return hipGetLastError();
  }

  template <class T>
  hipError_t ensure(Plane<T> &pl, hipStream_t st) {
    if (pl.ready) return hipSuccess;
    using T2 = typename fft::C2<T>::type;
    for (int a = 0; a < 3; ++a) {
      const auto w = twiddles<T2>(P[a]);
      DCHK(pl.tw[a].alloc(w.size()));
      DCHK(hipMemcpyAsync(pl.tw[a].p, w.data(), sizeof(T2) * w.size(), hipMemcpyHostToDevice, st));
      DCHK(hipStreamSynchronize(st));
    }
    DCHK(pl.K.alloc((size_t)(3 * nb) * M));
    const int ncomp = 6 * nb * nb;
    if (real_tensor) {
      DCHK(pl.Wr.alloc((size_t)ncomp * M));
    } else {
      DCHK(pl.Wc.alloc((size_t)ncomp * M));
    }
    const int64_t blocks = ((int64_t)ncomp * M + 255) / 256;
    hipLaunchKernelGGL((k_dip_extract<T>), dim3((unsigned)blocks), dim3(256), 0, st,
                       (const double2 *)Fsetup.p, (int64_t)P[2], Q2, (int64_t)P[0] * P[1], ncomp,
                       real_tensor, pl.Wr.p, pl.Wc.p);
    DCHK(hipGetLastError());
    DCHK(hipStreamSynchronize(st));
    const int by = fused_lines<T>();
    if (by > 0) {
      int lby = 0;
      while ((1 << lby) < by) ++lby;
      const int nyb = (P[1] + by - 1) / by;
      const int64_t total = Q2 * nyb * 6 * (int64_t)P[0] * by;
      DCHK(pl.Wf.alloc((size_t)total));
      hipLaunchKernelGGL((k_dip_w_fused<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                         st, (const T *)pl.Wr.p, pl.Wf.p, M, P[0], P[1], lby, nyb, total);
      DCHK(hipGetLastError());
      DCHK(hipStreamSynchronize(st));
      pl.Wr.release();
      pl.by = by;
    }
    pl.ready = true;
    return hipSuccess;
  }

  // transform the real-space tensor grid (complex, [6 nb nb][P0][P1][P2]) in place
  hipError_t transform_tensor(hipStream_t st) {
    const int ncomp = 6 * nb * nb;
    const int64_t P01 = (int64_t)P[0] * P[1], Pall = P01 * P[2];
    Plane<double> tmp;
    for (int a = 0; a < 3; ++a) {
      const auto w = twiddles<double2>(P[a]);
      DCHK(tmp.tw[a].alloc(w.size()));
      DCHK(hipMemcpyAsync(tmp.tw[a].p, w.data(), sizeof(double2) * w.size(), hipMemcpyHostToDevice, st));
      DCHK(hipStreamSynchronize(st));
    }
    // z: lines (g, x, y), contiguous
    {
      auto p = base_pass<double>(tmp, 2, false);
      p.n_in = p.n_out = P[2];
      p.nlines = ncomp * P01;
      p.src = p.dst = LineMap{P[0], P[1], Pall, (int64_t)P[1] * P[2], P[2], 1};
      p.in = Fsetup.p;
      p.out = Fsetup.p;
      DCHK((run_pass<double, fft::LD_CPLX, fft::ST_CPLX>(p, st)));
    }
    // y: lines (g, x, z)
    {
      auto p = base_pass<double>(tmp, 1, false);
      p.n_in = p.n_out = P[1];
      p.nlines = ncomp * (int64_t)P[0] * P[2];
      p.src = p.dst = LineMap{P[0], P[2], Pall, (int64_t)P[1] * P[2], 1, P[2]};
      p.in = Fsetup.p;
      p.out = Fsetup.p;
      DCHK((run_pass<double, fft::LD_CPLX, fft::ST_CPLX>(p, st)));
    }
    // x: lines (g, y z)
    {
      auto p = base_pass<double>(tmp, 0, false);
      p.n_in = p.n_out = P[0];
      p.nlines = ncomp * (int64_t)P[1] * P[2];
      p.src = p.dst = LineMap{1, (int64_t)P[1] * P[2], Pall, 0, 1, (int64_t)P[1] * P[2]};
      p.in = Fsetup.p;
      p.out = Fsetup.p;
      DCHK((run_pass<double, fft::LD_CPLX, fft::ST_CPLX>(p, st)));
    }
    return hipStreamSynchronize(st);
  }

  bool direct_z() const { return P[2] <= kDirectP && L[2] <= kDirectL; }

  // threads of the fused z + y forward kernel for lw q2 lines: one per y
  // line for the z DFT, and no more than kBF butterflies per thread in any
  // y stage
  template <class T>
  int zy_threads(int lw) const {
    int need = std::max(64, L[1]);
    for (int R : rad[1]) {
      const int kbf = R == 2 ? fft::kBF<T, 2> : R == 3 ? fft::kBF<T, 3> : R == 4 ? fft::kBF<T, 4>
                    : R == 5 ? fft::kBF<T, 5> : fft::kBF<T, 8>;
      need = std::max(need, (P[1] / R * lw + kbf - 1) / kbf);
    }
    return (need + 63) / 64 * 64;
  }
  template <class T>
  size_t zy_lds(int lw) const {
    using T2 = typename fft::C2<T>::type;
    const size_t a = std::max((size_t)P[1] * lw * sizeof(T2),
                              (size_t)L[1] * (size_t)odd_stride(L[2]) * sizeof(T));
    return (size_t)((zy_fwd_lwk(lw) ? kTw512 : P[1]) + P[2]) * sizeof(T2) + a;
  }
  // workgroups per (g, x) of the fused z + y forward kernel, 0 when it does
  // not apply (direct z axis, in-place y radices, LDS within 64 KB, at most
  // 512 threads); JH_ZY_SPLIT or more when Q2 >= 4, for more workgroups in
  // flight (at a 256 x 256 x 8 film, 3 q2 lines per workgroup ran faster
  // than 5 or 2)
  template <class T>
  int zy_split() const {
    if (!direct_z()) return 0;
    for (int r : rad[1])
      if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) return 0;
    for (int ns = Q2 >= 4 ? JH_ZY_SPLIT : 1; ns <= Q2; ++ns) {
      const int lw = (int)((Q2 + ns - 1) / ns);
      if (zy_lds<T>(lw) <= kMaxPassLds && zy_threads<T>(lw) <= 512) return ns;
    }
    return 0;
  }

  template <class T>
  size_t zy_inv_lds(int lw) const {
    using T2 = typename fft::C2<T>::type;
    const size_t a = std::max((size_t)P[1] * lw * sizeof(T2),
                              (size_t)L[1] * (size_t)odd_stride(L[2]) * sizeof(T));
    return (size_t)((zy_inv_lwk(lw) ? kTw512 : P[1]) + P[2]) * sizeof(T2) + a;
  }
  // q2 lines per batch of the fused y + z inverse kernel, 0 when it does not
  // apply: LDS within JH_ZYI_LDS, at most 512 threads; batches of equal size
  template <class T>
  int zy_inv_lines() const {
    if (zy_split<T>() == 0) return 0;
    int lmax = 0;
    for (int lw = (int)Q2; lw >= 1 && !lmax; --lw)
      if (zy_inv_lds<T>(lw) <= (size_t)JH_ZYI_LDS && zy_threads<T>(lw) <= 512) lmax = lw;
    if (!lmax) return 0;
    const int nbatch = (int)((Q2 + lmax - 1) / lmax);
    return (int)((Q2 + nbatch - 1) / nbatch);
  }

  template <class T>
  ZYFused<T> zy_params(Plane<T> &pl, const double *s, double *h, int lw) const {
    ZYFused<T> q;
    std::memset(&q, 0, sizeof q);
    q.n = P[1];
    q.nst = (int)rad[1].size();
    int Ns = 1;
    for (int k = 0; k < q.nst; ++k) {
      q.radix[k] = rad[1][k];
      q.ns[k] = Ns;
      q.ns_rcp[k] = 1.0f / (float)Ns;
      q.tws[k] = q.n / (Ns * q.radix[k]);
      Ns *= q.radix[k];
    }
    q.lw = lw;
    q.lw_rcp = 1.0f / (float)lw;
    q.P2 = P[2];
    q.L2 = L[2];
    q.Q2 = (int)Q2;
    q.L1 = L[1];
    q.L0 = L[0];
    q.nb = nb;
    q.nspin = n;
    q.sa = (int64_t)P[0] * P[1] * Q2;
    q.sq = (int64_t)P[0] * P[1];
    q.spins = s;
    q.mus = mus.p;
    q.field = h;
    q.K = pl.K.p;
    q.twy = pl.tw[1].p;
    q.twz = pl.tw[2].p;
    return q;
  }

  template <class T>
  hipError_t run_zy_inv(Plane<T> &pl, double *h, hipStream_t st) {
    const int lw = zy_inv_lines<T>();
    const ZYFused<T> q = zy_params<T>(pl, nullptr, h, lw);
    const unsigned blocks = (unsigned)((int64_t)3 * nb * L[0]);
    const int nt = zy_threads<T>(lw);
    const size_t lds = zy_inv_lds<T>(lw);
    const int lmax = std::max(L[2], (P[2] + 1) / 2);
    const int lwk = zy_inv_lwk(lw);
    if (lmax <= 8) DCHK((launch_zy_inv<T, 8>(q, lwk, blocks, nt, lds, st)));
    else if (lmax <= 16) DCHK((launch_zy_inv<T, 16>(q, lwk, blocks, nt, lds, st)));
    else DCHK((launch_zy_inv<T, 32>(q, lwk, blocks, nt, lds, st)));
    return hipSuccess;
  }

  template <class T>
  hipError_t run_zy_fwd(Plane<T> &pl, const double *s, hipStream_t st) {
    const int nsplit = zy_split<T>();
    ZYFused<T> q = zy_params<T>(pl, s, nullptr, (int)((Q2 + nsplit - 1) / nsplit));
    q.nsplit = nsplit;
    const unsigned blocks = (unsigned)((int64_t)3 * nb * L[0] * q.nsplit);
    const int nt = zy_threads<T>(q.lw);
    const size_t lds = zy_lds<T>(q.lw);
    const bool k3 = zy_fwd_lwk(q.lw) == 3;
    if (L[2] <= 8) {
      if (k3) hipLaunchKernelGGL((k_dip_zy_fwd<T, 8, 3>), dim3(blocks), dim3(nt), lds, st, q);
      else hipLaunchKernelGGL((k_dip_zy_fwd<T, 8, 0>), dim3(blocks), dim3(nt), lds, st, q);
    } else if (L[2] <= 16) {
      if (k3) hipLaunchKernelGGL((k_dip_zy_fwd<T, 16, 3>), dim3(blocks), dim3(nt), lds, st, q);
      else hipLaunchKernelGGL((k_dip_zy_fwd<T, 16, 0>), dim3(blocks), dim3(nt), lds, st, q);
    } else {
      if (k3) hipLaunchKernelGGL((k_dip_zy_fwd<T, 32, 3>), dim3(blocks), dim3(nt), lds, st, q);
      else hipLaunchKernelGGL((k_dip_zy_fwd<T, 32, 0>), dim3(blocks), dim3(nt), lds, st, q);
    }
    return hipGetLastError();
  }

  template <class T>
  hipError_t eval(Plane<T> &pl, const double *s, double *h, hipStream_t st) {
    using T2 = typename fft::C2<T>::type;
    DCHK(ensure(pl, st));
    const int64_t G = 3 * nb;
    const int64_t sa = (int64_t)P[0] * P[1] * Q2;
    if (zy_split<T>() > 0) {
      // z forward and y forward in one kernel
      DCHK(run_zy_fwd<T>(pl, s, st));
    } else {
      // z forward, scatter
      {
        auto p = base_pass<T>(pl, 2, false);
        p.n_in = L[2];
        p.n_out = (int)Q2;
        p.nlines = G * L[0] * L[1];
        p.dst = LineMap{L[0], L[1], sa, P[1], 1, (int64_t)P[0] * P[1]};
        p.out = pl.K.p;
        p.spins = s;
        if (direct_z()) {
          DCHK(run_direct_z<T>(p, false, st));
        } else {
          DCHK((run_pass<T, fft::LD_SPIN, fft::ST_CPLX>(p, st)));
        }
      }
      // y forward: lines (g, q2, x < L0), contiguous
      {
        auto p = base_pass<T>(pl, 1, false);
        p.n_in = L[1];
        p.n_out = P[1];
        p.nlines = G * L[0] * Q2;
        p.src = p.dst = LineMap{Q2, L[0], sa, (int64_t)P[0] * P[1], P[1], 1};
        p.in = pl.K.p;
        p.out = pl.K.p;
        DCHK((run_pass<T, fft::LD_CPLX, fft::ST_CPLX>(p, st)));
      }
    }
    if (pl.by > 0) {
      // x forward, multiply, x inverse in one kernel
      DCHK(run_x_fused<T>(pl, st));
    } else {
      // x forward: lines (g, q2, y), stride P1
      {
        auto p = base_pass<T>(pl, 0, false);
        p.n_in = L[0];
        p.n_out = P[0];
        p.nlines = G * P[1] * Q2;
        p.src = p.dst = LineMap{Q2, P[1], sa, (int64_t)P[0] * P[1], 1, P[1]};
        p.in = pl.K.p;
        p.out = pl.K.p;
        DCHK((run_pass<T, fft::LD_CPLX, fft::ST_CPLX>(p, st)));
      }
      // multiply
      {
        const int64_t blocks = (M + 255) / 256;
        if (real_tensor)
          hipLaunchKernelGGL((k_dip_mult_real<T>), dim3((unsigned)blocks), dim3(256), 0, st, pl.K.p,
                             (const T *)pl.Wr.p, M);
        else
          hipLaunchKernelGGL((k_dip_mult_cplx<T>), dim3((unsigned)blocks), dim3(256), 0, st, pl.K.p,
                             (const T2 *)pl.Wc.p, M, nb);
        DCHK(hipGetLastError());
      }
      // x inverse, outputs x < L0
      {
        auto p = base_pass<T>(pl, 0, true);
        p.n_in = P[0];
        p.n_out = L[0];
        p.nlines = G * P[1] * Q2;
        p.src = p.dst = LineMap{Q2, P[1], sa, (int64_t)P[0] * P[1], 1, P[1]};
        p.in = pl.K.p;
        p.out = pl.K.p;
        DCHK((run_pass<T, fft::LD_CPLX, fft::ST_CPLX>(p, st)));
      }
    }
    if (zy_inv_lines<T>() > 0) {
      // y inverse and z inverse in one kernel
      DCHK(run_zy_inv<T>(pl, h, st));
    } else {
      // y inverse: lines (g, q2, x < L0), outputs y < L1
      {
        auto p = base_pass<T>(pl, 1, true);
        p.n_in = P[1];
        p.n_out = L[1];
        p.nlines = G * L[0] * Q2;
        p.src = p.dst = LineMap{Q2, L[0], sa, (int64_t)P[0] * P[1], P[1], 1};
        p.in = pl.K.p;
        p.out = pl.K.p;
        DCHK((run_pass<T, fft::LD_CPLX, fft::ST_CPLX>(p, st)));
      }
      // z inverse (c2r), gather
      {
        auto p = base_pass<T>(pl, 2, true);
        p.n_in = (int)Q2;
        p.n_out = L[2];
        p.nlines = G * L[0] * L[1];
        p.src = LineMap{L[0], L[1], sa, P[1], 1, (int64_t)P[0] * P[1]};
        p.in = pl.K.p;
        p.field = h;
        if (direct_z()) {
          DCHK(run_direct_z<T>(p, true, st));
        } else {
          DCHK((run_pass<T, fft::LD_HERM, fft::ST_FIELD>(p, st)));
        }
      }
    }
    return hipSuccess;
  }

  hipError_t field(const double *s, double *h, bool fp32, hipStream_t st) override {
    return fp32 ? eval(f32, s, h, st) : eval(f64, s, h, st);
  }

  template <class T>
  std::string describe_t() const {
    const int by = fused_lines<T>();
    const int ns = zy_split<T>();
    const int lwf = ns ? (int)((Q2 + ns - 1) / ns) : 0;
    const int lwi = zy_inv_lines<T>();
    auto part = [](const char *name, int lw, int lwk) {
      return std::string(name) + "=" +
             (lw ? (lwk ? "fused512/" : "fused/") + std::to_string(lw) : std::string("passes"));
    };
    return part("dip_x", by ? 3 * by : 0, by ? x_lwk(3 * by) : 0) + " " +
           part("dip_zyf", lwf, lwf ? zy_fwd_lwk(lwf) : 0) + " " +
           part("dip_zyi", lwi, lwi ? zy_inv_lwk(lwi) : 0);
  }
  std::string describe(bool fp32) const override {
    return fp32 ? describe_t<float>() : describe_t<double>();
  }

  double bytes(bool fp32) const override {
    const double c = fp32 ? 8.0 : 16.0, G = 3.0 * nb;
    const double occ = (double)L[0] * L[1] * Q2, half = (double)L[0] * P[1] * Q2, full = (double)M;
    const double spin = 24.0 * n + 8.0 * n;  // s + mus (gather side: h + mus)
    double b = 2 * spin;
    const bool fused = fp32 ? fused_lines<float>() > 0 : fused_lines<double>() > 0;
    const double xm = fused ? half + half /*x fused r w*/
                            : half + full /*x fwd*/ + full + full /*mult r w*/ + full + half /*x inv*/;
    const bool zy = fp32 ? zy_split<float>() > 0 : zy_split<double>() > 0;
    const double zf = zy ? half /*z + y fwd w*/ : occ /*z fwd w*/ + occ + half /*y fwd r w*/;
    const bool zyi = fp32 ? zy_inv_lines<float>() > 0 : zy_inv_lines<double>() > 0;
    const double zi = zyi ? half /*y + z inv r*/ : half + occ /*y inv r w*/ + occ /*z inv r*/;
    b += G * c * (zf + xm + zi);
    b += (real_tensor ? 6.0 * (c / 2) : 6.0 * nb * nb * c) * full;  // tensor
    return b;
  }
};

}  // namespace

DipoleFFT *DipoleFFT::create(const int32_t size[3], const int32_t periodic[3], int nbasis,
                             const double cell[9], const double *basis_frac, double a_m,
                             double r_cutoff, const double *mus, int64_t nspins, hipStream_t st,
                             std::string *err) {
  auto fail = [&](const std::string &m) -> DipoleFFT * {
    if (err) *err = m;
    return nullptr;
  };
  if (nbasis < 1 || nbasis > kMaxBasis) return fail("1 to 8 basis sites are supported");
  if (!(a_m > 0) || !(r_cutoff > 0)) return fail("lattice parameter and r_cutoff must be positive");
  std::unique_ptr<DipoleFFTImpl> d(new DipoleFFTImpl);
  d->nb = nbasis;
  d->n = nspins;
  for (int a = 0; a < 3; ++a) {
    if (size[a] < 1) return fail("lattice size must be positive");
    d->L[a] = size[a];
    d->per[a] = periodic[a] ? 1 : 0;
    d->P[a] = periodic[a] ? size[a] : 2 * size[a];
    if (pass_lds<double>(d->P[a], 1, true) > kMaxPassLds)
      return fail("padded dimension too long for one workgroup's LDS (at most 2048)");
    d->rad[a] = factorise(d->P[a]);
    if ((int)d->rad[a].size() > fft::kMaxStages) return fail("too many FFT stages");
    for (int r : d->rad[a])
      if (r > 64) return fail("padded dimension has a prime factor above 64");
  }
  d->Q2 = d->P[2] / 2 + 1;
  d->M = (int64_t)d->P[0] * d->P[1] * d->Q2;
  if (d->mus.alloc((size_t)nspins) != hipSuccess) return fail("device allocation failed");
  if (hipMemcpy(d->mus.p, mus, sizeof(double) * nspins, hipMemcpyHostToDevice) != hipSuccess)
    return fail("upload failed");

  // ---- real-space tensor grid, [pair][6][P0][P1][P2], normalised by 1/P ----
  const int64_t P01 = (int64_t)d->P[0] * d->P[1], Pall = P01 * d->P[2];
  const int ncomp = 6 * nbasis * nbasis;
  std::vector<double2> F((size_t)ncomp * Pall);
  for (auto &v : F) { v.x = 0.0; v.y = 0.0; }
  const double w0 = kMu0IU / (4.0 * kPi * a_m * a_m * a_m) / (double)Pall;
  const double rc2 = (r_cutoff + kDistTol) * (r_cutoff + kDistTol);
  // supercell matrix (columns L_a a_a) and its inverse for the minimum image
  double S[9], Si[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) S[3 * r + c] = cell[3 * r + c] * d->L[c];
  {
    const double det = S[0] * (S[4] * S[8] - S[5] * S[7]) - S[1] * (S[3] * S[8] - S[5] * S[6]) +
                       S[2] * (S[3] * S[7] - S[4] * S[6]);
    if (std::fabs(det) < 1e-300) return fail("singular unit cell");
    Si[0] = (S[4] * S[8] - S[5] * S[7]) / det; Si[1] = (S[2] * S[7] - S[1] * S[8]) / det;
    Si[2] = (S[1] * S[5] - S[2] * S[4]) / det; Si[3] = (S[5] * S[6] - S[3] * S[8]) / det;
    Si[4] = (S[0] * S[8] - S[2] * S[6]) / det; Si[5] = (S[2] * S[3] - S[0] * S[5]) / det;
    Si[6] = (S[3] * S[7] - S[4] * S[6]) / det; Si[7] = (S[1] * S[6] - S[0] * S[7]) / det;
    Si[8] = (S[0] * S[4] - S[1] * S[3]) / det;
  }
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = d->per[a] ? 0 : -(d->L[a] - 1);
    hi[a] = d->L[a] - 1;
  }
  const int64_t n0cnt = hi[0] - lo[0] + 1;
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t) {
    pool.emplace_back([&, t] {
      for (int64_t i0 = t; i0 < n0cnt; i0 += nt) {
        const int n0 = lo[0] + (int)i0;
        for (int n1 = lo[1]; n1 <= hi[1]; ++n1)
          for (int n2 = lo[2]; n2 <= hi[2]; ++n2)
            for (int pi = 0; pi < nbasis; ++pi)
              for (int pj = 0; pj < nbasis; ++pj) {
                if (n0 == 0 && n1 == 0 && n2 == 0 && pi == pj) continue;
                const double f[3] = {basis_frac[3 * pi] + n0 - basis_frac[3 * pj],
                                     basis_frac[3 * pi + 1] + n1 - basis_frac[3 * pj + 1],
                                     basis_frac[3 * pi + 2] + n2 - basis_frac[3 * pj + 2]};
                double r[3];
                for (int a = 0; a < 3; ++a)
                  r[a] = cell[3 * a] * f[0] + cell[3 * a + 1] * f[1] + cell[3 * a + 2] * f[2];
                bool any_per = d->per[0] || d->per[1] || d->per[2];
                if (any_per) {
                  double g[3];
                  for (int a = 0; a < 3; ++a) g[a] = Si[3 * a] * r[0] + Si[3 * a + 1] * r[1] + Si[3 * a + 2] * r[2];
                  for (int a = 0; a < 3; ++a) if (d->per[a]) g[a] -= std::nearbyint(g[a]);
                  for (int a = 0; a < 3; ++a) r[a] = S[3 * a] * g[0] + S[3 * a + 1] * g[1] + S[3 * a + 2] * g[2];
                }
                const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                if (r2 > rc2 || r2 == 0.0) continue;
                const double nr = std::sqrt(r2);
                const double r5 = nr * nr * nr * nr * nr;
                const int64_t g0 = ((int64_t)n0 % d->P[0] + d->P[0]) % d->P[0];
                const int64_t g1 = ((int64_t)n1 % d->P[1] + d->P[1]) % d->P[1];
                const int64_t g2 = ((int64_t)n2 % d->P[2] + d->P[2]) % d->P[2];
                const int64_t at = (g0 * d->P[1] + g1) * d->P[2] + g2;
                static const int ab[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
                for (int q = 0; q < 6; ++q) {
                  const int a = ab[q][0], b = ab[q][1];
                  const double v = w0 * (3.0 * r[a] * r[b] - (a == b ? r2 : 0.0)) / r5;
                  F[((size_t)((pi * nbasis + pj) * 6 + q)) * Pall + at].x = v;
                }
              }
      }
    });
  }
  for (auto &th : pool) th.join();
  // One basis site and W(n) = W(-n mod P): the transform is real and 6 reals
  // per k-point suffice.  A periodic dimension of even length with
  // r_cutoff >= L/2 breaks the symmetry (the minimum image of the half-way
  // offset is not negated), and then the complex form is kept.
  d->real_tensor = false;
  if (nbasis == 1) {
    double wmax = 0.0;
    for (const auto &v : F) wmax = std::max(wmax, std::fabs(v.x));
    bool sym = true;
    for (int64_t g0 = 0; g0 < d->P[0] && sym; ++g0)
      for (int64_t g1 = 0; g1 < d->P[1] && sym; ++g1)
        for (int64_t g2 = 0; g2 < d->P[2] && sym; ++g2) {
          const int64_t at = (g0 * d->P[1] + g1) * d->P[2] + g2;
          const int64_t m0 = (d->P[0] - g0) % d->P[0], m1 = (d->P[1] - g1) % d->P[1],
                        m2 = (d->P[2] - g2) % d->P[2];
          const int64_t mt = (m0 * d->P[1] + m1) * d->P[2] + m2;
          for (int q = 0; q < 6 && sym; ++q)
            sym = std::fabs(F[(size_t)q * Pall + at].x - F[(size_t)q * Pall + mt].x) <= 1e-13 * wmax;
        }
    d->real_tensor = sym;
  }
  if (d->Fsetup.alloc(F.size()) != hipSuccess) return fail("device allocation failed");
  if (hipMemcpy(d->Fsetup.p, F.data(), sizeof(double2) * F.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail("upload failed");
  F.clear();
  F.shrink_to_fit();
  if (d->transform_tensor(st) != hipSuccess) return fail("tensor transform failed");
  if (d->ensure(d->f64, st) != hipSuccess) return fail("tensor set-up failed");
  if (d->ensure(d->f32, st) != hipSuccess) return fail("tensor set-up failed");
  d->Fsetup.release();
  return d.release();
}

}  // namespace jh
