// jh_engine.hip -- host side of the gfx950 Heun-LLG engine and its C ABI
// (include/jams_hip.h).
//
// The context owns one HIP stream and all device buffers.  Terms are added in
// Hamiltonian config order; the first jh_step / jh_compute_* call after a
// change "prepares" the device: it converts the exchange input into the
// sliced-ELL layout, de-duplicates per-spin constants into site classes and
// uploads everything.  Nothing is allocated inside jh_step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <thread>
#include <vector>

#include "../../include/jams_hip.h"
#include "jh_dipole.h"
#include "jh_internal.h"
#include "jh_launch.h"

#ifndef JH_LN_BAND
#define JH_LN_BAND 64
#endif

using namespace jh;

// ---------------------------------------------------------------------------
// Device scratch of jh_reduce_magnetisation_noalloc: the per-workgroup partial
// sums of k_reduce (4 doubles per kBlock spins) followed by the 4-double
// result, for at most kNoAllocMaxSpins spins: 2 MiB + 32 bytes that every GPU
// the library runs on holds until the process ends.  k_reduce writes every
// partial that k_reduce_final reads, so nothing in it has to start at zero.
// In host code `scratch` is only a symbol: its device address comes from
// hipGetSymbolAddress, once per device (bind_scratch, called by jh_create),
// and that address is what kernels and copies are given.
constexpr int64_t kNoAllocMaxSpins = (int64_t)1 << 24;
constexpr int64_t kNoAllocBlocks = (kNoAllocMaxSpins + kBlock - 1) / kBlock;
static __device__ __align__(256) double scratch[4 * kNoAllocBlocks + 4];

namespace {

std::mutex g_scratch_mu;
std::unordered_map<int, double *> g_scratch_addr;  // device -> address of scratch

// Looks up the device address of `scratch` on `device`, which must be the
// current device; once per device.
hipError_t bind_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_addr.count(device)) return hipSuccess;
  double *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void **>(&ptr), HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_scratch_addr[device] = ptr;
  return hipSuccess;
}

double *scratch_on(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  const auto it = g_scratch_addr.find(device);
  return it == g_scratch_addr.end() ? nullptr : it->second;
}


thread_local std::string g_create_error;

constexpr double kBoltzmannIU = 0.0861733326;  // meV/K, helpers/consts.h:35
constexpr size_t kMaxTableLds = 48 * 1024;     // bytes of LDS for the tensor table
constexpr size_t kStMaxTileLds = 160 * 1024;   // bytes of LDS for a STENCIL tile (a CU's LDS)
constexpr int64_t kLnAutoMinRow = 24;           // neighbours per row from which AUTO takes LINE
// polls of the persistent step's grid barrier before a workgroup gives up:
// each is a 256-cycle sleep plus an uncached load, so at least 0.1 s and
// typically about a second, against a phase of under a millisecond for a few
// million spins
constexpr unsigned kStepPollLimit = 1u << 20;
// spins from which the step kernels store their outputs non-temporally
constexpr int64_t kNtStoreMin = (int64_t)1 << 21;
constexpr int kLnBand = JH_LN_BAND;             // y rows per band of the LINE wave order
#ifndef JH_LN_PAIRS
#define JH_LN_PAIRS 1
#endif
constexpr bool kLnPairs = JH_LN_PAIRS;          // FP32-form LINE pair programs
// Tensor tables of the FP32 form (LDS table, wave-uniform sequences) keep
// the exact FP64 values: they are per tensor, not per spin, so FP32 would
// save no bandwidth, and rounding them perturbs the Hamiltonian itself (a
// systematic error that grows along a trajectory).  Per-entry values of the
// BLOCK format are FP32.
#ifndef JH_FP32_TENSOR_TABLES
#define JH_FP32_TENSOR_TABLES 0
#endif
constexpr bool kRoundTables = JH_FP32_TENSOR_TABLES != 0;
// FP32 form: pair programs for isotropic + DMI tensors (build_pairs)
#ifndef JH_PAIR_PROGRAMS
#define JH_PAIR_PROGRAMS 1
#endif
constexpr bool kPairPrograms = JH_PAIR_PROGRAMS != 0 && !kRoundTables;

// Static partition of [0, n) over host threads (set-up work only).
template <class F>
void parallel_for(int64_t n, F &&f) {
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 16u));
  if (n < 65536 || nt == 1) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  const int64_t chunk = (n + nt - 1) / nt;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t a = (int64_t)t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    pool.emplace_back([a, b, &f] {
      for (int64_t i = a; i < b; ++i) f(i);
    });
  }
  for (auto &th : pool) th.join();
}

struct JhError {
  int code;
  std::string msg;
};

#define JH_THROW(code, msg) throw JhError{(code), (msg)}
#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      JH_THROW(JH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      JH_THROW(JH_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                 " bytes failed: " + hipGetErrorString(e));
    }
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

// Exchange operator on the host: per-spin neighbour blocks (j ascending) with
// a tensor id into a table of distinct 3x3 tensors.
struct XHost {
  int64_t n = 0;
  std::vector<int64_t> ptr;   // n + 1
  std::vector<int32_t> col;   // neighbour spin per block
  std::vector<int32_t> typ;   // tensor id per block
  std::vector<double> tab;    // ntypes x 9
  int64_t csr_nnz = 0;        // non-zeros of the equivalent scalar CSR
};

struct TermHost {
  int kind;
  int xidx = -1;                       // exchange: index into xh
  std::vector<double> dc, ac, omega;   // zeeman (N x 3, N x 3, N)
  std::vector<double> K, axis;         // uniaxial
  int power = 2;
  int dipole = -1;                     // dipole: index into dipoles
};

// 72-byte key hashing for tensor de-duplication
struct Key9 {
  uint64_t w[9];
  bool operator==(const Key9 &o) const { return std::memcmp(w, o.w, sizeof w) == 0; }
};
struct Key9Hash {
  size_t operator()(const Key9 &k) const {
    uint64_t h = 1469598103934665603ull;
    for (int q = 0; q < 9; ++q) {
      h ^= k.w[q] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    }
    return (size_t)h;
  }
};
static Key9 key9(const double *t) {
  Key9 k;
  for (int q = 0; q < 9; ++q) {
    // +0 and -0 are the same tensor entry for the matvec (a zero product adds
    // nothing), so canonicalise zeros
    double v = t[q] == 0.0 ? 0.0 : t[q];
    std::memcpy(&k.w[q], &v, 8);
  }
  return k;
}

}  // namespace

struct jh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // sites
  int64_t n = 0;
  std::vector<double> mus, gyro, alpha;
  double dt = 0.0;
  bool have_dt = false;

  // state
  DevBuf<double> s, sp, ds, winj, wbuf, hext, fout;
  DevBuf<float> s32, sp32;
  DevBuf<double> sq;        // RK4: second stage-input buffer (stages alternate sp, sq)
  DevBuf<float> sq32;
  DevBuf<double> red_partial, red_out;
  std::vector<double> h_stage;   // host staging for AoS <-> SoA
  bool spins_set = false;
  int order_nb = 1;              // device order: basis-major with this period
  int64_t order_cells = 0;
  int64_t iteration = 0;
  double time = 0.0;

  // terms
  std::vector<TermHost> terms;
  std::vector<XHost> xh;
  std::vector<std::unique_ptr<DipoleFFT>> dipoles;

  // thermostat
  double temperature = 0.0;
  uint64_t seed = 0;
  int gilbert = 0;
  int noise_mode = NM_NONE;

  // options
  int prec = PR_FP64;
  int integrator = JH_INTEGRATOR_HEUN;
  int xfmt_req = JH_XFMT_AUTO;

  // prepared device state
  bool dirty = true;
  int xv = XV_NONE;
  int xfmt = JH_XFMT_AUTO;
  int colbits = 0;
  int ntab = 0;
  int64_t nent = 0;
  DevBuf<int64_t> slice_off;
  DevBuf<uint32_t> ent;
  DevBuf<double> tab64, val64;
  DevBuf<float> val32;
  DevBuf<uint32_t> pofs, pat;
  int pat_words = 0, pat_lds = 0, tbits = 0, npatterns = 0;
  DevBuf<int32_t> sseq;          // UNIFORM: (offset, length) per slice
  DevBuf<double> seqtab;
  DevBuf<int32_t> upat;          // UNIFORM: byte-offset pattern words
  DevBuf<int32_t> spair, pgrp;   // pair programs (XV_QISO / XV_QDMI)
  DevBuf<double> pval;
  int64_t npaired = 0;
  int64_t nuniform = 0, nslices = 0;
  DevBuf<uint16_t> cls;
  DevBuf<double> ctab;
  std::vector<double> ctab_host;
  int nclass = 0;
  int ani_power = 2;
  int has_ac = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // lattice shape (jh_set_lattice): spin ((x Ly + y) Lz + z) nb + b
  bool have_lattice = false;
  int lsize[3] = {0, 0, 0}, lper[3] = {0, 0, 0}, lnb = 0;
  // STENCIL step kernels (a layer over the PATTERN format, which k_field,
  // RK4 and the FP64 row-list launch keep using)
  bool st_on = false;
  StView st{};
  DevBuf<int32_t> st_seq, st_off, st_wseq, st_rows;
  DevBuf<double> st_val;
  DevBuf<uint64_t> st_fixmask;
  int64_t st_nfix = 0;
  int st_nseq = 0;
  // LINE step kernels (also over the PATTERN format; the FP64 row list is
  // st_rows / st_nfix)
  bool ln_on = false;
  LnView ln{};
  DevBuf<int32_t> ln_seq, ln_prog, ln_wseq;
  DevBuf<double> ln_val, ln_pval;
  DevBuf<uint64_t> ln_fixmask;
  bool ln_pairs = false;
  int ln_nseq = 0, ln_ngroups = 0;

  // replicas: independent systems in consecutive index ranges
  int nrep = 0;
  std::vector<int64_t> rep_off;     // nrep + 1
  std::vector<double> rep_T;
  std::vector<uint64_t> rep_seed;
  DevBuf<int64_t> d_rep_off;
  DevBuf<double> rep_out;

  // launch shape and graph replay
  int block = 0;                    // threads per workgroup of the step kernels
                                    // (0: chosen per exchange variant)
  // measured defaults: diagonal tensors (few neighbours) run best at 256
  // threads, full tensors and the wave-uniform FP32 variants at 512 (their
  // occupancy is bounded by LDS per workgroup, not per thread)
  int launch_block() const {
    if (block) return std::min(block, kMaxBlock);  // 1024 applies to the STENCIL launch
    return (xv_diag(xv) && !xv_uniform(xv)) ? 256 : 512;
  }
  // JH_STEP_LAUNCH_PERSISTENT: one launch per Heun step with a grid barrier
  // between the phases; step_sync holds the barrier's words, cus the
  // device's compute units (the grid is sized to what is resident)
  int step_launch = JH_STEP_LAUNCH_TWO;
  DevBuf<unsigned> step_sync;
  int cus = 0;
  int graph_steps = 0;              // steps per captured graph (0: plain launches)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  DevBuf<uint64_t> dstep;

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }
};

namespace {

void require_sites(jh_ctx *c) {
  if (c->n <= 0) JH_THROW(JH_ERR_STATE, "jh_set_sites must be called first");
}

// device index <-> reference index under the basis-major order
inline int64_t ref_of(const jh_ctx *c, int64_t d) {
  return c->order_nb == 1 ? d : (d % c->order_cells) * c->order_nb + d / c->order_cells;
}
inline int64_t dev_of(const jh_ctx *c, int64_t i) {
  return c->order_nb == 1 ? i : (i % c->order_nb) * c->order_cells + i / c->order_nb;
}

// index of component k of element d of a device vector array: a spin array
// (jh_layout.h) or SoA
inline int64_t vec_idx(int64_t n, int64_t d, int k, bool spin) {
  return spin ? spin_idx(n, d, k) : k * n + d;
}

// reference AoS (N x 3) -> device layout in device order
void upload_vec(jh_ctx *c, const double *aos, double *dst, bool spin) {
  const int64_t n = c->n;
  c->h_stage.resize((size_t)(3 * n));
  double *st = c->h_stage.data();
  parallel_for(n, [&](int64_t d) {
    const int64_t i = ref_of(c, d);
    for (int k = 0; k < 3; ++k) st[vec_idx(n, d, k, spin)] = aos[3 * i + k];
  });
  HIPCHK(hipMemcpyAsync(dst, st, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

void download_vec(jh_ctx *c, const double *src, double *aos, bool spin) {
  const int64_t n = c->n;
  c->h_stage.resize((size_t)(3 * n));
  double *st = c->h_stage.data();
  HIPCHK(hipMemcpyAsync(st, src, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  parallel_for(n, [&](int64_t d) {
    const int64_t i = ref_of(c, d);
    for (int k = 0; k < 3; ++k) aos[3 * i + k] = st[vec_idx(n, d, k, spin)];
  });
}

// ---- exchange input conversion -------------------------------------------

// Merge rows 3i, 3i+1, 3i+2 of the scalar CSR into 3x3 blocks keyed by
// neighbour j = col / 3.  Rows must be strictly column-ascending (the
// reference's build_csr output is).
XHost xhost_from_csr(int64_t n, int64_t nrows, int64_t nnz, const int64_t *rowptr,
                     const int32_t *col, const double *val) {
  if (nrows != 3 * n) JH_THROW(JH_ERR_INVALID, "exchange CSR must have 3N rows");
  if (rowptr[0] != 0 || rowptr[nrows] != nnz)
    JH_THROW(JH_ERR_INVALID, "exchange CSR row pointers do not span [0, nnz]");
  for (int64_t r = 0; r < nrows; ++r) {
    if (rowptr[r + 1] < rowptr[r]) JH_THROW(JH_ERR_INVALID, "exchange CSR row pointers decrease");
    for (int64_t q = rowptr[r]; q < rowptr[r + 1]; ++q) {
      if (col[q] < 0 || col[q] >= 3 * n) JH_THROW(JH_ERR_INVALID, "exchange CSR column out of range");
      if (q > rowptr[r] && col[q] <= col[q - 1])
        JH_THROW(JH_ERR_INVALID, "exchange CSR columns must be strictly ascending in each row");
    }
  }
  XHost x;
  x.n = n;
  x.csr_nnz = nnz;
  x.ptr.assign((size_t)n + 1, 0);
  // pass 1: block counts
  parallel_for(n, [&](int64_t i) {
    int64_t q[3], e[3];
    for (int m = 0; m < 3; ++m) { q[m] = rowptr[3 * i + m]; e[m] = rowptr[3 * i + m + 1]; }
    int64_t cnt = 0;
    for (;;) {
      int64_t jmin = INT64_MAX;
      for (int m = 0; m < 3; ++m) if (q[m] < e[m]) jmin = std::min<int64_t>(jmin, col[q[m]] / 3);
      if (jmin == INT64_MAX) break;
      for (int m = 0; m < 3; ++m) while (q[m] < e[m] && col[q[m]] / 3 == jmin) ++q[m];
      ++cnt;
    }
    x.ptr[i + 1] = cnt;
  });
  for (int64_t i = 0; i < n; ++i) x.ptr[i + 1] += x.ptr[i];
  const int64_t nb = x.ptr[n];
  x.col.resize((size_t)nb);
  std::vector<double> blk((size_t)nb * 9, 0.0);
  parallel_for(n, [&](int64_t i) {
    int64_t q[3], e[3];
    for (int m = 0; m < 3; ++m) { q[m] = rowptr[3 * i + m]; e[m] = rowptr[3 * i + m + 1]; }
    int64_t b = x.ptr[i];
    for (;;) {
      int64_t jmin = INT64_MAX;
      for (int m = 0; m < 3; ++m) if (q[m] < e[m]) jmin = std::min<int64_t>(jmin, col[q[m]] / 3);
      if (jmin == INT64_MAX) break;
      x.col[b] = (int32_t)jmin;
      for (int m = 0; m < 3; ++m)
        while (q[m] < e[m] && col[q[m]] / 3 == jmin) {
          blk[(size_t)b * 9 + 3 * m + (col[q[m]] % 3)] = val[q[m]];
          ++q[m];
        }
      ++b;
    }
  });
  // de-duplicate tensors
  x.typ.resize((size_t)nb);
  std::unordered_map<Key9, int32_t, Key9Hash> map;
  for (int64_t b = 0; b < nb; ++b) {
    const Key9 k = key9(&blk[(size_t)b * 9]);
    auto it = map.find(k);
    if (it == map.end()) {
      const int32_t id = (int32_t)(x.tab.size() / 9);
      map.emplace(k, id);
      for (int q2 = 0; q2 < 9; ++q2) x.tab.push_back(blk[(size_t)b * 9 + q2]);
      x.typ[b] = id;
    } else {
      x.typ[b] = it->second;
    }
  }
  return x;
}

XHost xhost_from_list(int64_t n, int64_t np, const int32_t *pi, const int32_t *pj,
                      const int32_t *vi, int64_t nv, const double *v9) {
  XHost x;
  x.n = n;
  x.ptr.assign((size_t)n + 1, 0);
  for (int64_t q = 0; q < np; ++q) {
    if (pi[q] < 0 || pi[q] >= n || pj[q] < 0 || pj[q] >= n)
      JH_THROW(JH_ERR_INVALID, "neighbour list index out of range");
    if (vi[q] < 0 || vi[q] >= nv) JH_THROW(JH_ERR_INVALID, "neighbour list value index out of range");
    if (q > 0 && (pi[q] < pi[q - 1] || (pi[q] == pi[q - 1] && pj[q] <= pj[q - 1])))
      JH_THROW(JH_ERR_INVALID, "neighbour list must be strictly sorted by (i, j)");
    x.ptr[(size_t)pi[q] + 1]++;
  }
  for (int64_t i = 0; i < n; ++i) x.ptr[i + 1] += x.ptr[i];
  x.col.assign(pj, pj + np);
  // de-duplicate the value table (identical tensors share an id)
  std::vector<int32_t> remap((size_t)nv);
  std::unordered_map<Key9, int32_t, Key9Hash> map;
  for (int64_t v = 0; v < nv; ++v) {
    const Key9 k = key9(v9 + 9 * v);
    auto it = map.find(k);
    if (it == map.end()) {
      const int32_t id = (int32_t)(x.tab.size() / 9);
      map.emplace(k, id);
      for (int q2 = 0; q2 < 9; ++q2) x.tab.push_back(v9[9 * v + q2]);
      remap[v] = id;
    } else {
      remap[v] = it->second;
    }
  }
  x.typ.resize((size_t)np);
  int64_t nnz = 0;
  std::vector<int64_t> nz_per_val((size_t)nv, 0);
  for (int64_t v = 0; v < nv; ++v)
    for (int q2 = 0; q2 < 9; ++q2) nz_per_val[v] += (v9[9 * v + q2] != 0.0);
  for (int64_t q = 0; q < np; ++q) {
    x.typ[q] = remap[vi[q]];
    nnz += nz_per_val[vi[q]];
  }
  x.csr_nnz = nnz;
  return x;
}

// ---- prepare: device formats and site classes ----------------------------

// Distinct tensors for the LDS table (diagonal or full rows), FP64.
void upload_table(jh_ctx *c, const XHost &x, int ntypes, int ntab, int vpt, bool diag) {
  std::vector<double> t64((size_t)ntab * vpt, 0.0);
  for (int t = 0; t < ntypes; ++t)
    for (int q = 0; q < vpt; ++q) {
      const double v = diag ? x.tab[9 * t + 4 * q] : x.tab[9 * t + q];
      t64[(size_t)t * vpt + q] = (c->prec == PR_FP32 && kRoundTables) ? (double)(float)v : v;
    }
  c->tab64.alloc(t64.size());
  HIPCHK(hipMemcpy(c->tab64.p, t64.data(), sizeof(double) * t64.size(), hipMemcpyHostToDevice));
}

// Smallest period nb (N % nb == 0, nb <= 16) with which most rows repeat
// their (column - row, tensor id) sequence: the number of basis sites of a
// crystal.  1 when there is no exchange or no such period.
int choose_order(const XHost &x) {
  const int64_t n = x.n;
  std::vector<uint64_t> h((size_t)n);
  parallel_for(n, [&](int64_t i) {
    uint64_t v = 1469598103934665603ull ^ (uint64_t)(x.ptr[i + 1] - x.ptr[i]);
    for (int64_t q = x.ptr[i]; q < x.ptr[i + 1]; ++q) {
      v = (v ^ (uint64_t)((int64_t)x.col[q] - i)) * 1099511628211ull;
      v = (v ^ (uint64_t)x.typ[q]) * 1099511628211ull;
    }
    h[i] = v;
  });
  for (int nb = 1; nb <= 16; ++nb) {
    if (n % nb != 0 || n < 4 * nb) continue;
    int64_t same = 0;
    for (int64_t i = 0; i + nb < n; ++i) same += h[i] == h[i + nb];
    if (same * 2 > n) return nb;
  }
  return 1;
}

// Group rows by their sequence of (column - row, tensor id).  Entries pack
// the signed offset above `tbits` bits of tensor id.  Fails (returns false)
// when an offset does not fit or the rows are too diverse to share patterns.
// by_tensor: order each row by (tensor id, column) instead of by column (the
// FP32 form's order, which makes tensor sequences shareable across lanes).
bool build_patterns(const jh_ctx *c, const XHost &x, int tbits, bool by_tensor,
                    std::vector<uint32_t> &pofs, std::vector<uint32_t> &pat, int &npatterns) {
  const int64_t n = x.n;
  const int64_t lim = (int64_t)1 << (31 - tbits);
  const uint32_t tmask = (1u << tbits) - 1u;
  pofs.assign((size_t)n, 0);
  pat.clear();
  npatterns = 0;
  std::unordered_map<uint64_t, std::vector<uint32_t>> seen;  // hash -> pattern offsets
  std::vector<uint32_t> row;
  for (int64_t d = 0; d < n; ++d) {
    const int64_t i = ref_of(c, d);
    row.clear();
    for (int64_t q = x.ptr[i]; q < x.ptr[i + 1]; ++q) {
      const int64_t off = dev_of(c, x.col[q]) - d;
      if (off >= lim || off < -lim) return false;
      row.push_back(((uint32_t)(int32_t)off << tbits) | (uint32_t)x.typ[q]);
    }
    if (by_tensor)
      std::sort(row.begin(), row.end(), [&](uint32_t a, uint32_t b) {
        if ((a & tmask) != (b & tmask)) return (a & tmask) < (b & tmask);
        return ((int32_t)a >> tbits) < ((int32_t)b >> tbits);
      });
    uint64_t h = 1469598103934665603ull ^ (uint64_t)row.size();
    for (uint32_t w : row) h = (h ^ w) * 1099511628211ull;
    auto &bucket = seen[h];
    uint32_t found = UINT32_MAX;
    for (uint32_t o : bucket) {
      if (pat[o] == row.size() && std::equal(row.begin(), row.end(), pat.begin() + o + 1)) {
        found = o;
        break;
      }
    }
    if (found == UINT32_MAX) {
      if (pat.size() + row.size() + 1 > ((size_t)1 << 24)) return false;
      found = (uint32_t)pat.size();
      pat.push_back((uint32_t)row.size());
      pat.insert(pat.end(), row.begin(), row.end());
      bucket.push_back(found);
      ++npatterns;
    }
    pofs[d] = found;
  }
  // worth it only if the table is much smaller than per-neighbour storage
  return (int64_t)pat.size() * 8 < x.ptr[n] + 1024;
}

// Slices (64 device rows) whose rows all carry the same tensor-id sequence
// (rows ordered by tensor id): for each, the (offset, length) of that
// sequence's tensors in seqtab; (-1, 0) for the others.
// Returns the number of uniform slices, or -1 when the sequences would not
// fit the table.
int64_t build_uniform(const XHost &x, int64_t n, int tbits, bool diag,
                      const std::vector<uint32_t> &pofs, const std::vector<uint32_t> &pat,
                      std::vector<int32_t> &sseq, std::vector<double> &seqtab,
                      std::vector<std::vector<uint32_t>> &seq_ids_out,
                      std::vector<int32_t> &slice_seq) {
  const uint32_t tmask = (1u << tbits) - 1u;
  const int vpt = diag ? 3 : 9;
  // tensor sequence of each pattern, de-duplicated by its id sequence
  std::unordered_map<uint32_t, int32_t> pat_seq;              // pattern offset -> sequence
  std::unordered_map<uint64_t, std::vector<int32_t>> byhash;  // hash -> sequences
  std::vector<std::pair<int32_t, int32_t>> seqs;              // (seqtab offset, length)
  std::vector<std::vector<uint32_t>> seq_ids;
  std::vector<uint32_t> ids;
  const size_t kMaxSeqTab = (size_t)1 << 20;  // doubles (8 MB)
  for (int64_t d = 0; d < n; ++d) {
    const uint32_t o = pofs[d];
    if (pat_seq.count(o)) continue;
    const uint32_t len = pat[o];
    ids.assign(len, 0);
    uint64_t h = 1469598103934665603ull ^ len;
    for (uint32_t k = 0; k < len; ++k) {
      ids[k] = pat[o + 1 + k] & tmask;
      h = (h ^ ids[k]) * 1099511628211ull;
    }
    int32_t found = -1;
    for (int32_t sq : byhash[h])
      if (seq_ids[sq] == ids) { found = sq; break; }
    if (found < 0) {
      if (seqtab.size() + (size_t)len * vpt > kMaxSeqTab) return -1;
      found = (int32_t)seqs.size();
      seqs.push_back({(int32_t)seqtab.size(), (int32_t)len});
      seq_ids.push_back(ids);
      for (uint32_t k = 0; k < len; ++k) {
        const double *T = &x.tab[(size_t)ids[k] * 9];
        for (int q = 0; q < vpt; ++q) {
          const double v = diag ? T[4 * q] : T[q];
          seqtab.push_back(kRoundTables ? (double)(float)v : v);
        }
      }
      byhash[h].push_back(found);
    }
    pat_seq[o] = found;
  }
  const int64_t nsl = (n + kSlice - 1) / kSlice;
  sseq.assign((size_t)nsl * 2, 0);
  slice_seq.assign((size_t)nsl, -1);
  int64_t nu = 0;
  for (int64_t sl = 0; sl < nsl; ++sl) {
    const int32_t first = pat_seq[pofs[(size_t)sl * kSlice]];
    bool uni = true;
    for (int64_t d = sl * kSlice; d < std::min(n, (sl + 1) * kSlice) && uni; ++d)
      uni = pat_seq[pofs[d]] == first;
    sseq[2 * sl] = uni ? seqs[first].first : -1;
    sseq[2 * sl + 1] = uni ? seqs[first].second : 0;
    slice_seq[sl] = uni ? first : -1;
    nu += uni;
  }
  seq_ids_out = std::move(seq_ids);
  return nu;
}

// Pair programs of the FP32 form (XV_QISO / XV_QDMI).  Where every tensor of
// a uniform sequence is J I + [D]x (isotropic plus antisymmetric: T[m][m] = J,
// T[m][q] = -T[q][m]) and the entries pair up as (T, T^T) (a symmetric
// Hamiltonian's bonds r and -r), a pair's field is
//     T s1 + T^T s2 = J (s1 + s2) + D x (s1 - s2),
// so the pairs of one J are summed two at a time in FP32 (|s1 + s2| <= 2,
// their sum <= 4: rounding of the size of the FP32 spins themselves) and
// multiplied by the exact FP64 J once, and D x (s1 - s2) takes six FP64
// multiply-adds: about 10.75 operations per neighbour against 13 for a full
// tensor.  Per group of one or two pairs sharing J: pgrp = the four slots
// (k1A, k2A, k1B, k2B; k1B = -1 for one pair) of the sequence (the per-lane
// pattern words give each slot's offset) and pval = (J, DA, DB, 0); the
// two-pair groups come first.  Per slice: spair = (first group, two-pair
// groups | one-pair groups << 16), (-1, 0) without a program.
// Returns the number of slices with a program; dmi is set when some D != 0.
int64_t build_pairs(const XHost &x, const std::vector<std::vector<uint32_t>> &seq_ids,
                    const std::vector<int32_t> &slice_seq, std::vector<int32_t> &spair,
                    std::vector<int32_t> &pgrp, std::vector<double> &pval, bool &dmi) {
  dmi = false;
  std::vector<std::pair<int32_t, int32_t>> prog(seq_ids.size(), {-1, 0});
  for (size_t q = 0; q < seq_ids.size(); ++q) {
    const auto &ids = seq_ids[q];
    if (ids.size() >= 0x7FFF) continue;
    struct Iso { double J, D[3]; };
    std::vector<Iso> iso(ids.size());
    bool ok = true;
    for (size_t k = 0; k < ids.size() && ok; ++k) {
      const double *T = &x.tab[(size_t)ids[k] * 9];
      ok = T[0] == T[4] && T[4] == T[8] && T[1] == -T[3] && T[2] == -T[6] && T[5] == -T[7];
      // [D]x = ((0, -Dz, Dy), (Dz, 0, -Dx), (-Dy, Dx, 0))
      iso[k] = {T[0], {T[7], T[2], T[3]}};
    }
    if (!ok) continue;
    // pair each entry with an unpaired one of tensor T^T: same J, -D
    auto key = [](double J, double a, double b, double c) {
      // canonical zeros so that D = 0 pairs with itself
      auto z = [](double v) { return v == 0.0 ? 0.0 : v; };
      return std::array<double, 4>{z(J), z(a), z(b), z(c)};
    };
    std::map<std::array<double, 4>, std::vector<int>> open;
    std::vector<std::pair<int, int>> pairs;
    for (size_t k = 0; k < ids.size(); ++k) {
      const Iso &e = iso[k];
      auto &cands = open[key(e.J, -e.D[0], -e.D[1], -e.D[2])];
      if (!cands.empty()) {
        pairs.push_back({cands.back(), (int)k});
        cands.pop_back();
      } else {
        open[key(e.J, e.D[0], e.D[1], e.D[2])].push_back((int)k);
      }
    }
    if (pairs.size() * 2 != ids.size()) continue;  // an entry without its partner
    // groups of two pairs of one J, then the pairs left over one by one (the
    // kernel runs the two kinds as two loops without branches)
    std::stable_sort(pairs.begin(), pairs.end(), [&](const auto &a, const auto &b) {
      return iso[a.first].J < iso[b.first].J;
    });
    std::vector<std::array<int, 4>> twos, ones;
    for (size_t p = 0; p < pairs.size();) {
      if (p + 1 < pairs.size() && iso[pairs[p + 1].first].J == iso[pairs[p].first].J) {
        twos.push_back({pairs[p].first, pairs[p].second, pairs[p + 1].first, pairs[p + 1].second});
        p += 2;
      } else {
        ones.push_back({pairs[p].first, pairs[p].second, -1, -1});
        p += 1;
      }
    }
    const int32_t g0 = (int32_t)(pgrp.size() / 4);
    for (const auto *list : {&twos, &ones})
      for (const auto &g : *list) {
        for (int k : g) pgrp.push_back(k);
        pval.push_back(iso[g[0]].J);
        for (int a = 0; a < 3; ++a) pval.push_back(iso[g[0]].D[a]);
        for (int a = 0; a < 3; ++a) pval.push_back(g[2] >= 0 ? iso[g[2]].D[a] : 0.0);
        pval.push_back(0.0);
        for (int a = 0; a < 3; ++a)
          dmi = dmi || iso[g[0]].D[a] != 0.0 || (g[2] >= 0 && iso[g[2]].D[a] != 0.0);
      }
    if (twos.size() >= 0x8000 || ones.size() >= 0x8000) continue;
    prog[q] = {g0, (int32_t)(twos.size() | (ones.size() << 16))};
  }
  spair.assign(slice_seq.size() * 2, 0);
  int64_t np = 0;
  for (size_t sl = 0; sl < slice_seq.size(); ++sl) {
    const int32_t q = slice_seq[sl];
    const auto pr = q >= 0 ? prog[q] : std::pair<int32_t, int32_t>{-1, 0};
    spair[2 * sl] = pr.first;
    spair[2 * sl + 1] = pr.second;
    np += pr.first >= 0;
  }
  return np;
}

// ---- crystal stencils (STENCIL and LINE layouts) ---------------------------
//
// Both apply to a context holding one crystal whose shape jh_set_lattice
// gave, in basis-major device order, whose exchange rows are all the
// translates of one stencil per basis site: every neighbour (dx, dy, dz, b',
// tensor) of the row at the middle of the lattice, wrapped on periodic axes
// and dropped outside open ones.

struct StEnt {
  int d[3];
  int bb;
  int t;
};

// The stencil of each basis site (from its middle row) and its extent;
// false if the context is no such crystal.
static bool crystal_stencil(jh_ctx *c, const XHost &x, std::vector<std::vector<StEnt>> &sten,
                            int lo[kStMaxBasis][3], int hi[kStMaxBasis][3]) {
  if (!c->have_lattice || c->nrep > 0 || c->integrator != JH_INTEGRATOR_HEUN) return false;
  const int L[3] = {c->lsize[0], c->lsize[1], c->lsize[2]};
  const int nb = c->lnb;
  const int64_t C = (int64_t)L[0] * L[1] * L[2];
  if (nb < 1 || nb > kStMaxBasis || C * nb != c->n || L[2] % 64 != 0) return false;
  if (c->order_nb != nb && !(nb == 1 && c->order_nb == 1)) return false;
  sten.assign((size_t)nb, {});
  for (int b = 0; b < nb; ++b) {
    const int mid[3] = {L[0] / 2, L[1] / 2, L[2] / 2};
    const int64_t i = (((int64_t)mid[0] * L[1] + mid[1]) * L[2] + mid[2]) * nb + b;
    for (int a = 0; a < 3; ++a) { lo[b][a] = 0; hi[b][a] = 0; }
    for (int64_t q = x.ptr[i]; q < x.ptr[i + 1]; ++q) {
      StEnt e;
      const int64_t j = x.col[q];
      e.bb = (int)(j % nb);
      int64_t r = j / nb;
      int cell[3];
      cell[2] = (int)(r % L[2]);
      r /= L[2];
      cell[1] = (int)(r % L[1]);
      cell[0] = (int)(r / L[1]);
      for (int a = 0; a < 3; ++a) {
        int d = cell[a] - mid[a];
        if (c->lper[a]) {
          if (d > L[a] / 2) d -= L[a];
          if (d < -(L[a] - 1) / 2) d += L[a];
        }
        // the stencil must reach less than half way round (a unique minimum
        // image) and stay clear of the middle row's open boundaries
        if (2 * std::abs(d) >= L[a]) return false;
        e.d[a] = d;
        lo[b][a] = std::min(lo[b][a], d);
        hi[b][a] = std::max(hi[b][a], d);
      }
      e.t = x.typ[q];
      sten[b].push_back(e);
    }
  }
  return true;
}

// reference column of entry e of the row at `cell`, or -1 outside an open
// boundary
static int64_t stencil_col(const jh_ctx *c, const StEnt &e, const int cell[3]) {
  const int *L = c->lsize;
  int cc[3];
  for (int a = 0; a < 3; ++a) {
    cc[a] = cell[a] + e.d[a];
    if (cc[a] < 0 || cc[a] >= L[a]) {
      if (!c->lper[a]) return -1;
      cc[a] = (cc[a] + L[a]) % L[a];
    }
  }
  return (((int64_t)cc[0] * L[1] + cc[1]) * L[2] + cc[2]) * c->lnb + e.bb;
}

// FP64 sequences: per line (b, x, y) the column order of its rows away from
// the z ends (where no z neighbour wraps), entries outside an open boundary
// last; wseq[(b lx + x) ly + y] = sequence id, seq_basis[id] = b.
static bool line_orders(const jh_ctx *c, const std::vector<std::vector<StEnt>> &sten,
                        std::vector<std::vector<int>> &seqs, std::vector<int32_t> &wseq,
                        std::vector<int> &seq_basis) {
  const int *L = c->lsize;
  const int nb = c->lnb;
  wseq.assign((size_t)nb * L[0] * L[1], -1);
  seqs.clear();
  seq_basis.clear();
  std::map<std::vector<int>, int> ids;
  for (int b = 0; b < nb; ++b)
    for (int xx = 0; xx < L[0]; ++xx)
      for (int yy = 0; yy < L[1]; ++yy) {
        const int cell[3] = {xx, yy, L[2] / 2};
        const auto &sb = sten[b];
        std::vector<std::pair<int64_t, int>> keyed(sb.size());
        for (size_t k = 0; k < sb.size(); ++k) {
          const int64_t j = stencil_col(c, sb[k], cell);
          keyed[k] = {j < 0 ? INT64_MAX : j, (int)k};
        }
        std::sort(keyed.begin(), keyed.end());
        std::vector<int> o(sb.size() + 1);
        o[0] = b;  // the key carries the basis site (entries are per basis)
        for (size_t k = 0; k < sb.size(); ++k) o[k + 1] = keyed[k].second;
        auto it = ids.find(o);
        int id;
        if (it == ids.end()) {
          id = (int)seqs.size();
          ids.emplace(o, id);
          seqs.emplace_back(o.begin() + 1, o.end());
          seq_basis.push_back(b);
        } else {
          id = it->second;
        }
        wseq[((size_t)b * L[0] + xx) * L[1] + yy] = id;
      }
  return seqs.size() <= 4096;
}

// Every row must be its stencil's translate (else false); with `fp64`, rows
// whose column order is not their line's sequence (a z neighbour wraps) are
// marked in fix.
static bool check_rows(const jh_ctx *c, const XHost &x, const std::vector<std::vector<StEnt>> &sten,
                       const std::vector<std::vector<int>> &seqs, const std::vector<int32_t> &wseq,
                       bool fp64, std::vector<uint8_t> &fix) {
  const int *L = c->lsize;
  const int nb = c->lnb;
  const int64_t C = (int64_t)L[0] * L[1] * L[2];
  const int64_t n = c->n;
  fix.assign((size_t)n, 0);
  std::atomic<int> bad{0};
  parallel_for(n, [&](int64_t d) {
    const int b = (int)(d / C);
    const int64_t cellidx = d % C;
    int cell[3];
    cell[2] = (int)(cellidx % L[2]);
    cell[1] = (int)((cellidx / L[2]) % L[1]);
    cell[0] = (int)(cellidx / ((int64_t)L[2] * L[1]));
    const int64_t i = cellidx * nb + b;
    const auto &sb = sten[b];
    const std::vector<int> &o =
        fp64 ? seqs[wseq[((size_t)b * L[0] + cell[0]) * L[1] + cell[1]]] : seqs[b];
    thread_local std::vector<std::pair<int64_t, int32_t>> got;
    got.clear();
    bool ordered = true;
    int64_t last = -1;
    for (int k : o) {
      const int64_t j = stencil_col(c, sb[k], cell);
      if (j < 0) continue;
      if (j <= last) ordered = false;
      last = j;
      got.push_back({j, sb[k].t});
    }
    if (!ordered) std::sort(got.begin(), got.end());
    const int64_t len = x.ptr[i + 1] - x.ptr[i];
    bool ok = len == (int64_t)got.size();
    for (int64_t q = 0; ok && q < len; ++q)
      ok = got[q].first == x.col[x.ptr[i] + q] && got[q].second == x.typ[x.ptr[i] + q];
    if (!ok) bad.store(1, std::memory_order_relaxed);
    if (fp64 && !ordered) fix[d] = 1;
  });
  return bad.load() == 0;
}

// ---- STENCIL layout ------------------------------------------------------
//
// Lz a multiple of 64.  Builds the sequences, the tile geometry and, in
// FP64, the rows the row-list launch computes; returns false (and leaves the
// context without it) otherwise.
bool build_stencil(jh_ctx *c, const XHost &x, bool diag) {
  std::vector<std::vector<StEnt>> sten;
  int lo[kStMaxBasis][3], hi[kStMaxBasis][3];
  if (!crystal_stencil(c, x, sten, lo, hi)) return false;
  const int L[3] = {c->lsize[0], c->lsize[1], c->lsize[2]};
  const int nb = c->lnb;
  const int64_t C = (int64_t)L[0] * L[1] * L[2];
  const bool fp64 = c->prec == PR_FP64;
  const int elem = fp64 ? 24 : 16;
  // workgroup shape: with jh_set_block_size the shape of that many threads,
  // else the most waves whose tile leaves the LDS for two workgroups per CU
  // (one workgroup's tile load then overlaps the other's sums), else one
  static const int kShapes[][2] = {{4, 4}, {2, 4}, {4, 2}, {2, 2}, {1, 4}, {4, 1},
                                   {1, 2}, {2, 1}, {1, 1}};
  int wx = 0, wy = 0;
  for (int pass = 0; pass < 2 && !wx; ++pass)
  for (const auto &sh : kShapes) {
    if (L[0] % sh[0] || L[1] % sh[1]) continue;
    if (c->block && sh[0] * sh[1] * 64 != c->block) continue;
    size_t need = 0;
    for (int b = 0; b < nb; ++b)
      need = std::max(need, (size_t)nb * (sh[0] + hi[b][0] - lo[b][0]) *
                                (sh[1] + hi[b][1] - lo[b][1]) * (64 + hi[b][2] - lo[b][2]) * elem);
    const size_t cap = (pass == 0 && !c->block) ? kStMaxTileLds / 2 : kStMaxTileLds;
    if (need <= cap) { wx = sh[0]; wy = sh[1]; break; }
  }
  if (!wx) return false;
  StView v{};
  v.lx = L[0]; v.ly = L[1]; v.lz = L[2]; v.nb = nb;
  for (int a = 0; a < 3; ++a) v.per[a] = c->lper[a] ? 1 : 0;
  v.wx = wx; v.wy = wy;
  v.nbx = L[0] / wx; v.nby = L[1] / wy; v.nbz = L[2] / 64;
  for (int b = 0; b < nb; ++b) {
    v.ox[b] = lo[b][0]; v.oy[b] = lo[b][1]; v.oz[b] = lo[b][2];
    v.tx[b] = wx + hi[b][0] - lo[b][0];
    v.ty[b] = wy + hi[b][1] - lo[b][1];
    v.tz[b] = 64 + hi[b][2] - lo[b][2];
  }
  // sequences: FP32 form, the stencil order of each basis site (any order
  // serves); FP64, per line (b, x, y) the column order of its rows away from
  // the z ends
  std::vector<std::vector<int>> seqs;
  std::vector<int32_t> wseq;
  std::vector<int> seq_basis;
  if (!fp64) {
    for (int b = 0; b < nb; ++b) {
      std::vector<int> o(sten[b].size());
      for (size_t k = 0; k < o.size(); ++k) o[k] = (int)k;
      seqs.push_back(o);
      seq_basis.push_back(b);
    }
  } else if (!line_orders(c, sten, seqs, wseq, seq_basis)) {
    return false;
  }
  std::vector<uint8_t> fix;
  if (!check_rows(c, x, sten, seqs, wseq, fp64, fix)) return false;
  const int64_t n = c->n;
  // tables
  std::vector<int32_t> seqv, offv;
  std::vector<double> valv;
  const int V = diag ? 3 : 9;
  for (size_t s = 0; s < seqs.size(); ++s) {
    const int b = seq_basis[s];
    seqv.push_back((int32_t)offv.size());
    seqv.push_back((int32_t)seqs[s].size());
    for (int k : seqs[s]) {
      const StEnt &e = sten[b][k];
      const int64_t el = (((int64_t)e.bb * v.tx[b] + e.d[0] - v.ox[b]) * v.ty[b] + e.d[1] - v.oy[b]) *
                             v.tz[b] + e.d[2] - v.oz[b];
      // FP64: byte offset in the z plane (the (x, y) plane's is twice it);
      // FP32: byte offset of the 16-byte element
      offv.push_back((int32_t)(el * (fp64 ? 8 : 16)));
      for (int q = 0; q < V; ++q) valv.push_back(diag ? x.tab[9 * e.t + 4 * q] : x.tab[9 * e.t + q]);
    }
  }
  c->st_seq.alloc(seqv.size());
  HIPCHK(hipMemcpy(c->st_seq.p, seqv.data(), sizeof(int32_t) * seqv.size(), hipMemcpyHostToDevice));
  c->st_off.alloc(std::max<size_t>(offv.size(), 1));
  if (!offv.empty())
    HIPCHK(hipMemcpy(c->st_off.p, offv.data(), sizeof(int32_t) * offv.size(), hipMemcpyHostToDevice));
  c->st_val.alloc(std::max<size_t>(valv.size(), 1));
  if (!valv.empty())
    HIPCHK(hipMemcpy(c->st_val.p, valv.data(), sizeof(double) * valv.size(), hipMemcpyHostToDevice));
  c->st_wseq.release();
  c->st_rows.release();
  c->st_fixmask.release();
  c->st_nfix = 0;
  if (fp64) {
    c->st_wseq.alloc(wseq.size());
    HIPCHK(hipMemcpy(c->st_wseq.p, wseq.data(), sizeof(int32_t) * wseq.size(), hipMemcpyHostToDevice));
    std::vector<int32_t> rows;
    for (int64_t d = 0; d < n; ++d)
      if (fix[d]) rows.push_back((int32_t)d);
    if (!rows.empty()) {
      // per wave of the stencil launch: the lanes the row list owns
      const int wpg = wx * wy;
      const int64_t nwg = (int64_t)nb * v.nbx * v.nby * v.nbz;
      std::vector<uint64_t> mask((size_t)(nwg * wpg), 0);
      for (int32_t d : rows) {
        const int b = (int)(d / C);
        const int64_t cellidx = d % C;
        const int z = (int)(cellidx % L[2]);
        const int yy = (int)((cellidx / L[2]) % L[1]);
        const int xx = (int)(cellidx / ((int64_t)L[2] * L[1]));
        const int64_t lb = (((int64_t)b * v.nbx + xx / wx) * v.nby + yy / wy) * v.nbz + z / 64;
        const int wave = (xx % wx) * wy + (yy % wy);
        mask[(size_t)(lb * wpg + wave)] |= 1ull << (z % 64);
      }
      c->st_rows.alloc(rows.size());
      HIPCHK(hipMemcpy(c->st_rows.p, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice));
      c->st_fixmask.alloc(mask.size());
      HIPCHK(hipMemcpy(c->st_fixmask.p, mask.data(), sizeof(uint64_t) * mask.size(), hipMemcpyHostToDevice));
      c->st_nfix = (int64_t)rows.size();
    }
  }
  c->st = v;
  c->st_nseq = (int)seqs.size();
  return true;
}

// ---- LINE layout ---------------------------------------------------------

// FP32-form pair program of one sequence (entries o of stencil sb): every
// tensor J I + [D]x, and every line's entries matched with those of one
// line (itself or another) whose tensors are their transposes, J I - [D]x,
// at z offsets dz' = -dz + delta with |delta| <= 1.  A pair (e, e') adds
// J (s_e + s_e') + D x (s_e - s_e'), which is T_e s_e + T_e' s_e'.  One
// group per pair of lines (slot 0 the line of e, slot 1 of e'); its word 0
// is the bitmask of the pair codes (dz + 2) * 3 + delta + 1 present, with
// bit 31 marking a pair group; word 1 the index of its first pair's values
// (J, D) in pval, 4 doubles a pair.  False if the sequence does not pair.
static bool line_pair_program(const std::vector<StEnt> &sb, const std::vector<int> &o,
                              const XHost &x, std::vector<int32_t> &seqv,
                              std::vector<int32_t> &prog, std::vector<double> &pval,
                              int &ngroups) {
  struct PE {
    int32_t lw;
    int dz;
    double J, D[3];
    bool used;
  };
  std::vector<PE> es;
  std::vector<int32_t> lines;  // in sequence order of first appearance
  for (int k : o) {
    const StEnt &e = sb[k];
    const double *T = &x.tab[9 * (size_t)e.t];
    if (!(T[0] == T[4] && T[0] == T[8] && T[1] == -T[3] && T[2] == -T[6] && T[5] == -T[7]))
      return false;
    const int32_t lw = (int32_t)(e.bb | ((e.d[0] + 2048) << 4) | ((e.d[1] + 2048) << 16));
    es.push_back(PE{lw, e.d[2], T[0], {T[7], T[2], T[3]}, false});
    if (std::find(lines.begin(), lines.end(), lw) == lines.end()) lines.push_back(lw);
  }
  auto partners = [](const PE &a, const PE &b) {
    const int delta = a.dz + b.dz;
    return a.J == b.J && a.D[0] == -b.D[0] && a.D[1] == -b.D[1] && a.D[2] == -b.D[2] &&
           delta >= -1 && delta <= 1;
  };
  seqv.push_back((int32_t)(prog.size() / kLnGroupWords));
  int groups = 0;
  for (int32_t L : lines) {
    std::vector<int> uL;
    for (int k = 0; k < (int)es.size(); ++k)
      if (es[k].lw == L && !es[k].used) uL.push_back(k);
    if (uL.empty()) continue;
    bool found = false;
    std::vector<std::pair<int, int>> pairs;
    int32_t Mfound = 0;
    // the partner line: L itself first, then the others in order
    std::vector<int32_t> cand{L};
    for (int32_t M : lines)
      if (M != L) cand.push_back(M);
    for (int32_t M : cand) {
      std::vector<int> uM;
      for (int k = 0; k < (int)es.size(); ++k)
        if (es[k].lw == M && !es[k].used) uM.push_back(k);
      pairs.clear();
      std::vector<char> taken(es.size(), 0);
      bool ok = M == L ? uL.size() % 2 == 0 : uM.size() == uL.size();
      for (size_t u = 0; ok && u < uL.size(); ++u) {
        const int a = uL[u];
        if (taken[a]) continue;
        taken[a] = 1;
        int pick = -1;
        for (int bidx : uM)
          if (!taken[bidx] && partners(es[a], es[bidx])) { pick = bidx; break; }
        if (pick < 0) { ok = false; break; }
        taken[pick] = 1;
        pairs.push_back({a, pick});
      }
      if (ok) { found = true; Mfound = M; break; }
    }
    if (!found) return false;
    int32_t w[kLnGroupWords] = {0};
    uint32_t mask = 0;
    std::vector<std::pair<int, std::pair<int, int>>> coded;
    for (const auto &pr : pairs) {
      const int code = (es[pr.first].dz + 2) * 3 + (es[pr.first].dz + es[pr.second].dz + 1);
      if (mask & (1u << code)) return false;
      mask |= 1u << code;
      coded.push_back({code, pr});
      es[pr.first].used = es[pr.second].used = true;
    }
    std::sort(coded.begin(), coded.end());
    w[0] = (int32_t)(mask | 0x80000000u);
    w[1] = (int32_t)(pval.size() / 4);
    w[2] = 3;  // both slots
    for (const auto &cp : coded) {
      const PE &a = es[cp.second.first];
      pval.push_back(a.J);
      for (int m = 0; m < 3; ++m) pval.push_back(a.D[m]);
    }
    w[4] = L;
    w[5] = Mfound;
    prog.insert(prog.end(), w, w + kLnGroupWords);
    ++groups;
  }
  seqv.push_back(groups);
  ngroups += groups;
  return true;
}

//
// A fully periodic crystal with Lz = 64 or 128 and at most two basis sites
// whose stencils reach at most two cells along z.  Each sequence becomes a
// program of groups, every group loading exactly two lines (a repeat pads a
// group that needs one) and then summing the entries that read them, in the
// sequence's order.  FP64: the per-line column orders and the row list as
// in STENCIL; FP32 form: one sequence per basis site, entries grouped by
// column so that each line is loaded once.
bool build_lines(jh_ctx *c, const XHost &x, bool diag) {
  std::vector<std::vector<StEnt>> sten;
  int lo[kStMaxBasis][3], hi[kStMaxBasis][3];
  if (!crystal_stencil(c, x, sten, lo, hi)) return false;
  const int L[3] = {c->lsize[0], c->lsize[1], c->lsize[2]};
  const int nb = c->lnb;
  if (nb > kLnSlots || (L[2] != 64 && L[2] != 128)) return false;
  // the kernels address cells with 32-bit offsets (3 N doubles per array)
  if (c->n >= ((int64_t)1 << 31) / 3) return false;
  if (!c->lper[0] || !c->lper[1] || !c->lper[2]) return false;
  const int P = L[2] / 64;
  for (int b = 0; b < nb; ++b) {
    if (lo[b][2] < -2 || hi[b][2] > 2) return false;
    for (int a = 0; a < 2; ++a)
      if (lo[b][a] < -2000 || hi[b][a] > 2000) return false;
  }
  const bool fp64 = c->prec == PR_FP64;
  std::vector<std::vector<int>> seqs;
  std::vector<int32_t> wseq;
  std::vector<int> seq_basis;
  if (!fp64) {
    // FP32 form: column by column (dx, dy ascending), so each column's lines
    // are loaded once
    for (int b = 0; b < nb; ++b) {
      std::vector<int> o(sten[b].size());
      for (size_t k = 0; k < o.size(); ++k) o[k] = (int)k;
      const auto &sb = sten[b];
      std::stable_sort(o.begin(), o.end(), [&](int p, int q) {
        if (sb[p].d[0] != sb[q].d[0]) return sb[p].d[0] < sb[q].d[0];
        if (sb[p].d[1] != sb[q].d[1]) return sb[p].d[1] < sb[q].d[1];
        return sb[p].bb < sb[q].bb;
      });
      seqs.push_back(o);
      seq_basis.push_back(b);
    }
  } else if (!line_orders(c, sten, seqs, wseq, seq_basis)) {
    return false;
  }
  std::vector<uint8_t> fix;
  if (!check_rows(c, x, sten, seqs, wseq, fp64, fix)) return false;
  // FP32 form with isotropic (+ DMI) tensors whose entries pair up as
  // (T, T^T): pair programs (one group per pair of lines)
  std::vector<int32_t> pseqv, pprog;
  std::vector<double> ppval;
  int pgroups = 0;
  bool pairs_ok = !fp64 && kLnPairs;
  for (size_t s = 0; s < seqs.size() && pairs_ok; ++s)
    pairs_ok = line_pair_program(sten[seq_basis[s]], seqs[s], x, pseqv, pprog, ppval, pgroups);
  // programs: per sequence a list of groups of kLnGroupWords words (a
  // bitmask of the entry codes present, the index of its first entry's
  // values, two unused words, the two line words).  The code of an entry is
  // slot | (dz + 2) << 1, and a group's entries are summed in ascending code
  // order, so each code occurs at most once in a group (it names one
  // neighbour).  FP64: a group is one (dx, dy) column with the line of basis
  // site b' in slot b', whose column order (z' then b') is the code order;
  // FP32 form: columns are packed two lines to a group.
  const int V = diag ? 3 : 9;
  std::vector<int32_t> seqv, prog;
  std::vector<double> valv;
  int ngroups = 0;
  auto line_word = [](const StEnt &e) {
    return (int32_t)(e.bb | ((e.d[0] + 2048) << 4) | ((e.d[1] + 2048) << 16));
  };
  for (size_t s = 0; s < seqs.size(); ++s) {
    const int b = seq_basis[s];
    const auto &sb = sten[b];
    seqv.push_back((int32_t)(prog.size() / kLnGroupWords));
    int groups = 0;
    const auto &o = seqs[s];
    std::vector<int32_t> lines;           // the group's lines by slot (-1: free)
    std::vector<std::pair<int, const StEnt *>> ents;  // (code, entry)
    bool ok = true;
    auto close_group = [&]() {
      if (ents.empty()) return;
      int32_t fill = -1;
      for (int32_t lw : lines)
        if (lw != -1) fill = lw;
      int32_t w[kLnGroupWords] = {0};
      uint32_t mask = 0;
      // FP64 keeps the sequence order, which must be the code order
      for (size_t k = 1; k < ents.size(); ++k)
        if (ents[k].first <= ents[k - 1].first) {
          if (fp64) ok = false;
        }
      if (!fp64) std::sort(ents.begin(), ents.end(),
                           [](const auto &p, const auto &q) { return p.first < q.first; });
      w[1] = (int32_t)(valv.size() / V);
      for (const auto &pe : ents) {
        if (mask & (1u << pe.first)) ok = false;  // one neighbour per code
        mask |= 1u << pe.first;
        const StEnt &e = *pe.second;
        for (int m = 0; m < V; ++m) valv.push_back(diag ? x.tab[9 * e.t + 4 * m] : x.tab[9 * e.t + m]);
      }
      w[0] = (int32_t)mask;
      // the slots the entries read: a slot no entry reads is not loaded
      for (const auto &pe : ents) w[2] |= 1 << (pe.first & 1);
      for (int k = 0; k < kLnSlots; ++k) w[4 + k] = lines[k] != -1 ? lines[k] : fill;
      prog.insert(prog.end(), w, w + kLnGroupWords);
      ++groups;
      ents.clear();
      lines.assign(kLnSlots, -1);
    };
    lines.assign(kLnSlots, -1);
    size_t k = 0;
    while (k < o.size()) {
      // the run of consecutive entries in one (dx, dy) column
      size_t q = k;
      while (q < o.size() && sb[o[q]].d[0] == sb[o[k]].d[0] && sb[o[q]].d[1] == sb[o[k]].d[1]) ++q;
      // the slots its lines take: FP64 slot b'; FP32 the free slots in turn
      std::vector<int32_t> want;
      for (size_t u = k; u < q; ++u) {
        const int32_t lw = line_word(sb[o[u]]);
        if (std::find(want.begin(), want.end(), lw) == want.end()) want.push_back(lw);
      }
      auto place = [&](bool dry) {
        std::vector<int32_t> L = lines;
        for (int32_t lw : want) {
          if (std::find(L.begin(), L.end(), lw) != L.end()) continue;
          const int bb = lw & 15;
          int slot = -1;
          if (fp64) {
            slot = bb;
            if (L[slot] != -1) return false;
          } else {
            for (int u = 0; u < kLnSlots && slot < 0; ++u)
              if (L[u] == -1) slot = u;
            if (slot < 0) return false;
          }
          L[slot] = lw;
        }
        if (!dry) lines = L;
        return true;
      };
      // FP64: one column per group
      if (fp64 || !place(true)) close_group();
      if (!place(false)) return false;
      for (size_t u = k; u < q; ++u) {
        const StEnt &e = sb[o[u]];
        const int slot = (int)(std::find(lines.begin(), lines.end(), line_word(e)) - lines.begin());
        ents.push_back({slot | ((e.d[2] + 2) << 1), &e});
      }
      k = q;
    }
    close_group();
    if (!ok) return false;
    seqv.push_back(groups);
    ngroups += groups;
  }
  if (pairs_ok) {
    seqv = pseqv;
    prog = pprog;
    ngroups = pgroups;
  }
  // one spare entry of values: the kernels read the next entry's values
  // while summing the current one
  for (int m = 0; m < V; ++m) valv.push_back(0.0);
  if (valv.size() / V >= ((size_t)1 << 23) || prog.size() >= ((size_t)1 << 30)) return false;
  const int64_t n = c->n;
  c->ln_seq.alloc(seqv.size());
  HIPCHK(hipMemcpy(c->ln_seq.p, seqv.data(), sizeof(int32_t) * seqv.size(), hipMemcpyHostToDevice));
  c->ln_prog.alloc(std::max<size_t>(prog.size(), 1));
  if (!prog.empty())
    HIPCHK(hipMemcpy(c->ln_prog.p, prog.data(), sizeof(int32_t) * prog.size(), hipMemcpyHostToDevice));
  c->ln_val.alloc(std::max<size_t>(valv.size(), 1));
  if (!valv.empty())
    HIPCHK(hipMemcpy(c->ln_val.p, valv.data(), sizeof(double) * valv.size(), hipMemcpyHostToDevice));
  c->ln_pval.release();
  c->ln_pairs = pairs_ok;
  if (pairs_ok) {
    c->ln_pval.alloc(std::max<size_t>(ppval.size(), 1));
    if (!ppval.empty())
      HIPCHK(hipMemcpy(c->ln_pval.p, ppval.data(), sizeof(double) * ppval.size(), hipMemcpyHostToDevice));
  }
  c->ln_wseq.release();
  c->ln_fixmask.release();
  c->st_rows.release();
  c->st_nfix = 0;
  if (fp64) {
    c->ln_wseq.alloc(wseq.size());
    HIPCHK(hipMemcpy(c->ln_wseq.p, wseq.data(), sizeof(int32_t) * wseq.size(), hipMemcpyHostToDevice));
    // the row list owns whole lanes: a lane with a wrapped row gives the
    // list all its P cells (the line kernel stores a lane's cells together)
    std::vector<uint64_t> mask((size_t)(n / L[2]), 0);
    for (int64_t d = 0; d < n; ++d)
      if (fix[d]) mask[(size_t)(d / L[2])] |= 1ull << ((d % L[2]) / P);
    std::vector<int32_t> rows;
    for (int64_t d = 0; d < n; ++d)
      if ((mask[(size_t)(d / L[2])] >> ((d % L[2]) / P)) & 1ull) rows.push_back((int32_t)d);
    if (!rows.empty()) {
      c->st_rows.alloc(rows.size());
      HIPCHK(hipMemcpy(c->st_rows.p, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice));
      c->ln_fixmask.alloc(mask.size());
      HIPCHK(hipMemcpy(c->ln_fixmask.p, mask.data(), sizeof(uint64_t) * mask.size(), hipMemcpyHostToDevice));
      c->st_nfix = (int64_t)rows.size();
    }
  }
  LnView v{};
  v.lx = L[0]; v.ly = L[1]; v.lz = L[2]; v.nb = nb;
  v.per[0] = v.per[1] = 1;
  v.ty = L[1] % kLnBand == 0 ? kLnBand : L[1];
  c->ln = v;
  c->ln_nseq = (int)seqs.size();
  c->ln_ngroups = ngroups;
  return true;
}

void prepare_exchange(jh_ctx *c) {
  c->xv = XV_NONE;
  c->st_on = false;
  c->ln_on = false;
  c->ntab = 0;
  c->nent = 0;
  if (c->xh.empty()) return;
  const XHost &x = c->xh[0];
  const int64_t n = c->n;
  const int ntypes = (int)(x.tab.size() / 9);
  bool diag = true;
  for (int t = 0; t < ntypes && diag; ++t)
    for (int m = 0; m < 3; ++m)
      for (int q = 0; q < 3; ++q)
        if (m != q && x.tab[9 * t + 3 * m + q] != 0.0) diag = false;
  const int vpt = diag ? 3 : 9;
  const int ntab = ntypes + 1;  // + trailing zero tensor used by padding
  int colbits = 1;
  while (colbits < 32 && ((int64_t)1 << colbits) < n) ++colbits;
  int tbits = 0;
  while ((1 << tbits) < ntab) ++tbits;
  const size_t lds_need = (size_t)ntab * vpt * 8;
  bool typed_ok = (colbits + tbits <= 32) && lds_need <= kMaxTableLds;
  // row patterns: identical sequences of (column - row, tensor id)
  std::vector<uint32_t> pofs, pat;
  const bool by_tensor = c->prec == PR_FP32;
  const bool want_pattern = c->xfmt_req == JH_XFMT_AUTO || c->xfmt_req == JH_XFMT_PATTERN ||
                            c->xfmt_req == JH_XFMT_PATTERN_GLOBAL ||
                            c->xfmt_req == JH_XFMT_STENCIL || c->xfmt_req == JH_XFMT_LINE;
  bool pattern_ok = typed_ok && want_pattern &&
                    build_patterns(c, x, tbits, by_tensor, pofs, pat, c->npatterns);
  int fmt = c->xfmt_req;
  // AUTO keeps the gather kernels: at the bcc 128^3 x 2 five-shell workload the
  // STENCIL launch measured slower (its tile staging costs more VALU than the
  // texture-address time it saves); jh_autotune times it beside the others
  const bool want_stencil = fmt == JH_XFMT_STENCIL;
  // AUTO takes LINE for a lattice whose rows hold at least kLnAutoMinRow
  // neighbours: measured at bcc 128^3 x 2 with 58 neighbours it runs the
  // step 1.25x (FP64) and 1.15x (FP32 form) faster than the gathers, at
  // bcc 64^3 x 2 with 14 it runs 13 % and 5 % slower
  const bool want_line = fmt == JH_XFMT_LINE ||
                         (fmt == JH_XFMT_AUTO && c->have_lattice && c->nrep == 0 &&
                          c->integrator == JH_INTEGRATOR_HEUN &&
                          c->step_launch != JH_STEP_LAUNCH_PERSISTENT && n > 0 &&
                          x.ptr[n / 2 + 1] - x.ptr[n / 2] >= kLnAutoMinRow);
  if ((want_stencil || want_line) && !pattern_ok)
    JH_THROW(JH_ERR_UNSUPPORTED, "STENCIL and LINE exchange formats need the row patterns of a crystal");
  if (fmt == JH_XFMT_AUTO || want_stencil || want_line)
    fmt = pattern_ok ? JH_XFMT_PATTERN : typed_ok ? JH_XFMT_TYPED : JH_XFMT_BLOCK;
  if (fmt != JH_XFMT_BLOCK && !typed_ok)
    JH_THROW(JH_ERR_UNSUPPORTED, "TYPED exchange format does not fit (too many distinct tensors)");
  if ((fmt == JH_XFMT_PATTERN || fmt == JH_XFMT_PATTERN_GLOBAL) && !pattern_ok)
    JH_THROW(JH_ERR_UNSUPPORTED, "PATTERN exchange format does not fit (too many distinct rows)");
  c->xfmt = fmt;
  if (fmt == JH_XFMT_PATTERN || fmt == JH_XFMT_PATTERN_GLOBAL) {
    upload_table(c, x, ntypes, ntab, vpt, diag);
    c->pofs.alloc(pofs.size());
    c->pat.alloc(pat.size());
    HIPCHK(hipMemcpy(c->pofs.p, pofs.data(), sizeof(uint32_t) * pofs.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->pat.p, pat.data(), sizeof(uint32_t) * pat.size(), hipMemcpyHostToDevice));
    c->pat_words = (int)pat.size();
    c->pat_lds = (fmt == JH_XFMT_PATTERN &&
                  ((lds_need + 15) & ~(size_t)15) + pat.size() * 4 <= kMaxTableLds) ? 1 : 0;
    c->tbits = tbits;
    c->ntab = ntab;
    c->nent = x.ptr[n];
    c->colbits = 0;
    c->xv = c->pat_lds ? (diag ? XV_PDIAG : XV_PFULL) : (diag ? XV_GDIAG : XV_GFULL);
    c->nslices = (n + kSlice - 1) / kSlice;
    c->nuniform = 0;
    if (by_tensor && c->pat_lds) {
      std::vector<int32_t> sseq, slice_seq;
      std::vector<double> seqtab;
      std::vector<std::vector<uint32_t>> seq_ids;
      const int64_t nu = build_uniform(x, n, tbits, diag, pofs, pat, sseq, seqtab, seq_ids,
                                       slice_seq);
      // worth its own variant when at least half of the slices are uniform
      // byte offsets of the gather copy must fit 31 bits
      const bool fits = (double)n * kG32 * sizeof(float) < 2147483648.0;
      if (nu >= 0 && nu * 2 >= c->nslices && fits) {
        // each pattern's words as byte offsets (its length word stays)
        std::vector<int32_t> up(pat.size());
        std::vector<char> done(pat.size(), 0);
        for (int64_t d = 0; d < n; ++d) {
          const uint32_t o = pofs[d];
          if (done[o]) continue;
          done[o] = 1;
          const uint32_t len = pat[o];
          up[o] = (int32_t)len;
          for (uint32_t k = 1; k <= len; ++k)
            up[o + k] = (((int32_t)pat[o + k]) >> tbits) * (int32_t)(kG32 * sizeof(float));
        }
        c->upat.alloc(up.size());
        HIPCHK(hipMemcpy(c->upat.p, up.data(), sizeof(int32_t) * up.size(), hipMemcpyHostToDevice));
        c->nuniform = nu;
        c->sseq.alloc(sseq.size());
        HIPCHK(hipMemcpy(c->sseq.p, sseq.data(), sizeof(int32_t) * sseq.size(), hipMemcpyHostToDevice));
        c->seqtab.alloc(std::max<size_t>(seqtab.size(), 1));
        if (!seqtab.empty())
          HIPCHK(hipMemcpy(c->seqtab.p, seqtab.data(), sizeof(double) * seqtab.size(), hipMemcpyHostToDevice));
        c->xv = diag ? XV_UDIAG : XV_UFULL;
        // pair programs (isotropic + DMI tensors), when most slices have one;
        // with isotropic tensors alone (no DMI) the plain wave-uniform
        // variant is kept: its three products per neighbour leave the pair
        // sums little to save
        std::vector<int32_t> spair, pgrp;
        std::vector<double> pval;
        bool dmi = false;
        const int64_t npr = kPairPrograms ? build_pairs(x, seq_ids, slice_seq, spair, pgrp, pval, dmi) : 0;
        if (npr * 2 >= c->nslices && !pgrp.empty() && dmi) {
          c->spair.alloc(spair.size());
          HIPCHK(hipMemcpy(c->spair.p, spair.data(), sizeof(int32_t) * spair.size(), hipMemcpyHostToDevice));
          c->pgrp.alloc(pgrp.size());
          HIPCHK(hipMemcpy(c->pgrp.p, pgrp.data(), sizeof(int32_t) * pgrp.size(), hipMemcpyHostToDevice));
          c->pval.alloc(pval.size());
          HIPCHK(hipMemcpy(c->pval.p, pval.data(), sizeof(double) * pval.size(), hipMemcpyHostToDevice));
          c->npaired = npr;
          c->xv = dmi ? XV_QDMI : XV_QISO;
        }
      }
    }
    if (!xv_pair(c->xv)) {
      c->spair.release();
      c->pgrp.release();
      c->pval.release();
      c->npaired = 0;
    }
    if (!xv_uniform(c->xv)) {
      c->sseq.release();
      c->seqtab.release();
      c->upat.release();
    }
    c->slice_off.release();
    c->ent.release();
    c->val64.release();
    c->val32.release();
    c->st_on = want_stencil && build_stencil(c, x, diag);
    if (c->xfmt_req == JH_XFMT_STENCIL && !c->st_on)
      JH_THROW(JH_ERR_UNSUPPORTED,
               "STENCIL exchange format does not apply (needs jh_set_lattice, one crystal whose "
               "rows all translate one stencil per basis site, Lz a multiple of 64, Heun)");
    if (c->st_on) c->xfmt = JH_XFMT_STENCIL;
    c->ln_on = want_line && build_lines(c, x, diag);
    if (c->xfmt_req == JH_XFMT_LINE && !c->ln_on)
      JH_THROW(JH_ERR_UNSUPPORTED,
               "LINE exchange format does not apply (needs jh_set_lattice, one periodic crystal "
               "of at most two basis sites, one with a dipole term, whose rows all translate "
               "one stencil reaching at most two cells along z, Lz = 64 or 128, Heun)");
    if (c->ln_on) c->xfmt = JH_XFMT_LINE;
    return;
  }
  c->pofs.release();
  c->pat.release();
  c->sseq.release();
  c->seqtab.release();
  c->nuniform = 0;
  c->pat_words = 0;
  const bool typed = fmt == JH_XFMT_TYPED;
  if (!typed) colbits = 32;
  const uint32_t colmask = colbits >= 32 ? 0xFFFFFFFFu : ((1u << colbits) - 1u);

  const int64_t nsl = (n + kSlice - 1) / kSlice;
  std::vector<int64_t> off((size_t)nsl + 1, 0);
  for (int64_t s = 0; s < nsl; ++s) {
    int64_t w = 0;
    for (int64_t d = s * kSlice; d < std::min(n, (s + 1) * kSlice); ++d) {
      const int64_t i = ref_of(c, d);
      w = std::max(w, x.ptr[i + 1] - x.ptr[i]);
    }
    off[s + 1] = off[s] + w * kSlice;
  }
  const int64_t E = off[nsl];
  std::vector<uint32_t> ent((size_t)E);
  std::vector<double> val64;
  std::vector<float> val32;
  if (!typed) {
    if (c->prec == PR_FP64) val64.assign((size_t)E * vpt, 0.0);
    else val32.assign((size_t)E * vpt, 0.f);
  }
  const uint32_t zero_t = (uint32_t)ntypes;
  parallel_for(nsl, [&](int64_t s) {
    const int64_t w = (off[s + 1] - off[s]) / kSlice;
    for (int64_t k = 0; k < w; ++k) {
      for (int l = 0; l < kSlice; ++l) {
        const int64_t d = s * kSlice + l;
        const int64_t i = d < n ? ref_of(c, d) : 0;
        const int64_t e = off[s] + k * kSlice + l;
        const int64_t len = d < n ? x.ptr[i + 1] - x.ptr[i] : 0;
        uint32_t j, t;
        if (k < len) {
          j = (uint32_t)dev_of(c, x.col[x.ptr[i] + k]);
          t = (uint32_t)x.typ[x.ptr[i] + k];
        } else {
          j = (uint32_t)(d < n ? d : 0);
          t = zero_t;
        }
        ent[e] = typed ? (j | (t << colbits)) : j;
        if (!typed && t != zero_t) {
          const double *T = &x.tab[(size_t)t * 9];
          for (int q = 0; q < vpt; ++q) {
            const double v = diag ? T[4 * q] : T[q];
            const size_t vi = (size_t)(off[s] + k * kSlice) * vpt + (size_t)q * kSlice + l;
            if (c->prec == PR_FP64) val64[vi] = v; else val32[vi] = (float)v;
          }
        }
      }
    }
  });
  c->slice_off.alloc((size_t)nsl + 1);
  HIPCHK(hipMemcpy(c->slice_off.p, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice));
  c->ent.alloc((size_t)std::max<int64_t>(E, 1));
  if (E > 0) HIPCHK(hipMemcpy(c->ent.p, ent.data(), sizeof(uint32_t) * E, hipMemcpyHostToDevice));
  if (typed) {
    upload_table(c, x, ntypes, ntab, vpt, diag);
    c->val64.release();
    c->val32.release();
    c->xv = diag ? XV_TDIAG : XV_TFULL;
  } else {
    if (c->prec == PR_FP64) {
      c->val64.alloc(std::max<size_t>(val64.size(), 1));
      if (!val64.empty())
        HIPCHK(hipMemcpy(c->val64.p, val64.data(), sizeof(double) * val64.size(), hipMemcpyHostToDevice));
      c->val32.release();
    } else {
      c->val32.alloc(std::max<size_t>(val32.size(), 1));
      if (!val32.empty())
        HIPCHK(hipMemcpy(c->val32.p, val32.data(), sizeof(float) * val32.size(), hipMemcpyHostToDevice));
      c->val64.release();
    }
    c->xv = diag ? XV_BDIAG : XV_BFULL;
  }
  c->colbits = colbits;
  c->ntab = typed ? ntab : 0;
  c->nent = E;
}

void prepare_classes(jh_ctx *c) {
  const int64_t n = c->n;
  const TermHost *zee = nullptr, *ani = nullptr;
  for (auto &t : c->terms) {
    if (t.kind == TK_ZEEMAN) zee = &t;
    if (t.kind == TK_UNIAXIAL) ani = &t;
  }
  c->has_ac = (zee && !zee->ac.empty()) ? 1 : 0;
  c->ani_power = ani ? ani->power : 2;
  std::vector<double> row(CF_STRIDE);
  std::map<std::vector<uint64_t>, int> map;
  std::vector<double> table;
  std::vector<uint16_t> ids((size_t)n);
  for (int64_t d = 0; d < n; ++d) {
    const int64_t i = ref_of(c, d);
    std::fill(row.begin(), row.end(), 0.0);
    row[CF_MUS] = c->mus[i];
    row[CF_GYRO] = c->gyro[i];
    row[CF_ALPHA] = c->alpha[i];
    if (c->have_dt) {
      // solvers/cpu_llg_heun.cc:36-43 and cuda_thermostat_classical.cc:33-45
      double denominator = 1.0;
      if (c->gilbert) denominator = 1.0 + c->alpha[i] * c->alpha[i];
      row[CF_SIGMA] = std::sqrt((2.0 * kBoltzmannIU * c->alpha[i]) /
                                (c->mus[i] * c->gyro[i] * c->dt * denominator));
    }
    if (zee) {
      for (int q = 0; q < 3; ++q) row[CF_ZDC + q] = zee->dc[3 * i + q];
      if (!zee->ac.empty()) {
        for (int q = 0; q < 3; ++q) row[CF_ZAC + q] = zee->ac[3 * i + q];
        row[CF_ZOMEGA] = zee->omega[i];
      }
    }
    if (ani) {
      row[CF_K] = ani->K[i];
      for (int q = 0; q < 3; ++q) row[CF_AXIS + q] = ani->axis[3 * i + q];
    }
    // thermostat of the spin's system: sqrt(T), seed, first index
    double T = c->temperature;
    uint64_t seed = c->seed;
    int64_t roff = 0;
    if (c->nrep > 0) {
      const int r = (int)(std::upper_bound(c->rep_off.begin(), c->rep_off.end(), i) -
                          c->rep_off.begin()) - 1;
      T = c->rep_T[r];
      seed = c->rep_seed[r];
      roff = c->rep_off[r];
    }
    row[CF_SQRTT] = std::sqrt(T);
    std::memcpy(&row[CF_SEED], &seed, sizeof seed);
    std::memcpy(&row[CF_ROFF], &roff, sizeof roff);
    std::vector<uint64_t> key(CF_STRIDE);
    std::memcpy(key.data(), row.data(), sizeof(double) * CF_STRIDE);
    auto it = map.find(key);
    int id;
    if (it == map.end()) {
      id = (int)map.size();
      if (id > 65535) JH_THROW(JH_ERR_UNSUPPORTED, "more than 65536 distinct site classes");
      map.emplace(key, id);
      table.insert(table.end(), row.begin(), row.end());
    } else {
      id = it->second;
    }
    ids[d] = (uint16_t)id;
  }
  c->nclass = (int)map.size();
  c->ctab_host = table;
  c->ctab.alloc(table.size());
  HIPCHK(hipMemcpy(c->ctab.p, table.data(), sizeof(double) * table.size(), hipMemcpyHostToDevice));
  if (c->nclass > 1) {
    c->cls.alloc((size_t)n);
    HIPCHK(hipMemcpy(c->cls.p, ids.data(), sizeof(uint16_t) * n, hipMemcpyHostToDevice));
  } else {
    c->cls.release();
  }
}

void sync_f32(jh_ctx *c) {
  if (c->prec != PR_FP32) return;
  c->s32.alloc((size_t)(kG32 * c->n));
  c->sp32.alloc((size_t)(kG32 * c->n));
  HIPCHK(launch_to_f32(c->s.p, c->s32.p, c->n, c->stream));
}

int dipole_term(const jh_ctx *c) {
  for (auto &t : c->terms) if (t.kind == TK_DIPOLE) return t.dipole;
  return -1;
}

// Why the context cannot run a Heun step as one persistent launch (null: it
// can).  The dipole transforms run between the phases, and the STENCIL and
// LINE layouts have launches of their own.
const char *persistent_unfit(const jh_ctx *c) {
  if (c->integrator != JH_INTEGRATOR_HEUN) return "the Heun integrator";
  if (dipole_term(c) >= 0) return "no dipole term (its transforms run between the phases)";
  if (c->st_on || c->ln_on) return "a gathering exchange format (PATTERN, TYPED or BLOCK)";
  return nullptr;
}

// The grid barrier's words, zero before the first persistent launch (set up
// outside any captured graph: prepare runs before a capture)
void ensure_step_sync(jh_ctx *c) {
  if (c->step_sync.p) return;
  c->step_sync.alloc(4);
  HIPCHK(hipMemsetAsync(c->step_sync.p, 0, 4 * sizeof(unsigned), c->stream));
}

// A persistent step whose barrier timed out left its spins invalid: report
// it (once) when the host next reads the state or synchronizes.
void check_step_sync(jh_ctx *c) {
  if (!c->step_sync.p) return;
  unsigned w[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(w, c->step_sync.p, sizeof w, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (w[2]) {
    HIPCHK(hipMemsetAsync(c->step_sync.p + 2, 0, sizeof(unsigned), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    JH_THROW(JH_ERR_STATE,
             "a persistent step's grid barrier timed out (its workgroups were not all resident "
             "at once); the spins are invalid");
  }
}

void prepare(jh_ctx *c) {
  require_sites(c);
  if (!c->spins_set) JH_THROW(JH_ERR_STATE, "spins have not been set (jh_set_spins)");
  if (!c->dirty) return;
  if (c->xh.size() > 1)
    JH_THROW(JH_ERR_UNSUPPORTED, "more than one exchange term; merge them into one list");
  int nzee = 0, nani = 0;
  for (auto &t : c->terms) {
    nzee += t.kind == TK_ZEEMAN;
    nani += t.kind == TK_UNIAXIAL;
  }
  if (nzee > 1 || nani > 1)
    JH_THROW(JH_ERR_UNSUPPORTED, "at most one zeeman and one uniaxial term per context");
  if (c->terms.size() > 4) JH_THROW(JH_ERR_UNSUPPORTED, "at most 4 Hamiltonian terms");
  // device order: basis-major when rows repeat with the basis period (the
  // dipole term works on the reference order)
  const int nb = (!c->xh.empty() && dipole_term(c) < 0 && c->nrep == 0) ? choose_order(c->xh[0]) : 1;
  if (nb != c->order_nb) {
    std::vector<double> tmp_s((size_t)(3 * c->n)), tmp_w;
    download_vec(c, c->s.p, tmp_s.data(), true);
    const bool inj = c->noise_mode == NM_INJECTED && c->winj.p;
    if (inj) {
      tmp_w.resize((size_t)(3 * c->n));
      download_vec(c, c->winj.p, tmp_w.data(), false);
    }
    c->order_nb = nb;
    c->order_cells = c->n / nb;
    upload_vec(c, tmp_s.data(), c->s.p, true);
    if (inj) upload_vec(c, tmp_w.data(), c->winj.p, false);
  }
  c->drop_graph();
  if (c->nrep > 0)
    for (const XHost &x : c->xh)
      for (int r = 0; r < c->nrep; ++r)
        for (int64_t i = c->rep_off[r]; i < c->rep_off[r + 1]; ++i)
          for (int64_t q = x.ptr[i]; q < x.ptr[i + 1]; ++q)
            if (x.col[q] < c->rep_off[r] || x.col[q] >= c->rep_off[r + 1])
              JH_THROW(JH_ERR_INVALID, "an exchange pair couples two replicas");
  if (c->nrep > 0 && dipole_term(c) >= 0)
    JH_THROW(JH_ERR_UNSUPPORTED, "the dipole term is not available with replicas");
  if (c->noise_mode != NM_INJECTED) {
    bool hot = c->temperature > 0.0;
    if (c->nrep > 0) {
      hot = false;
      for (double T : c->rep_T) hot = hot || T > 0.0;
    }
    c->noise_mode = hot ? NM_PHILOX : NM_NONE;
  }
  prepare_exchange(c);
  prepare_classes(c);
  if (c->step_launch == JH_STEP_LAUNCH_PERSISTENT) {
    const char *why = persistent_unfit(c);
    if (why) JH_THROW(JH_ERR_UNSUPPORTED, std::string("one persistent launch per step needs ") + why);
    ensure_step_sync(c);
  }
  if (c->noise_mode == NM_PHILOX) c->wbuf.alloc((size_t)(3 * c->n));
  if (c->integrator == JH_INTEGRATOR_RK4) {
    c->sq.alloc((size_t)(3 * c->n));
    if (c->prec == PR_FP32) c->sq32.alloc((size_t)(kG32 * c->n));
  } else {
    c->sq.release();
    c->sq32.release();
  }
  sync_f32(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->dirty = false;
}

KParams make_params(jh_ctx *c) {
  KParams p;
  std::memset(&p, 0, sizeof p);
  p.n = c->n;
  p.s = c->s.p;
  p.sp = c->sp.p;
  // with no dipole term nothing but the corrector reads s*, and the FP32
  // form's corrector forms it from s and ds
  p.spfree = c->prec == PR_FP32 && dipole_term(c) < 0 && c->integrator == JH_INTEGRATOR_HEUN;
  p.ds = c->ds.p;
  p.snew = c->s.p;
  p.cls = c->nclass > 1 ? c->cls.p : nullptr;
  p.ctab = c->ctab.p;
  p.x.slice_off = c->slice_off.p;
  p.x.ent = c->ent.p;
  p.x.colbits = c->colbits;
  p.x.colmask = c->colbits >= 32 ? 0xFFFFFFFFu : ((1u << c->colbits) - 1u);
  p.x.ntab = c->ntab;
  p.x.tab64 = c->tab64.p;
  p.x.val64 = c->val64.p;
  p.x.val32 = c->val32.p;
  p.x.pofs = c->pofs.p;
  p.x.pat = c->pat.p;
  p.x.pat_words = c->pat_words;
  p.x.tbits = c->tbits;
  p.x.sseq = reinterpret_cast<const int2 *>(c->sseq.p);
  p.x.seqtab = c->seqtab.p;
  p.x.upat = c->upat.p;
  p.x.spair = reinterpret_cast<const int2 *>(c->spair.p);
  p.x.pgrp = reinterpret_cast<const int4 *>(c->pgrp.p);
  p.x.pval = c->pval.p;
  p.rowlist = nullptr;
  p.nrowlist = 0;
  p.nterms = (int)c->terms.size();
  for (int q = 0; q < p.nterms; ++q) p.kind[q] = c->terms[q].kind;
  p.hext = c->hext.p;
  p.has_ac = c->has_ac;
  p.ani_power = c->ani_power;
  p.noise = c->noise_mode;
  p.dstep = nullptr;
  if (c->nclass == 1) {
    // the class-0 row as prepare_classes stored it
    const double *row = c->ctab_host.data();
    std::memcpy(&p.seed, &row[CF_SEED], sizeof p.seed);
    std::memcpy(&p.roff, &row[CF_ROFF], sizeof p.roff);
    p.sqrtT = row[CF_SQRTT];
  }
  p.winj = c->winj.p;
  p.wbuf = c->wbuf.p;
  p.dt = c->dt;
  p.order_nb = c->order_nb;
  p.order_cells = c->order_cells;
  p.order_rcells = c->order_cells > 0 ? 1.0 / (double)c->order_cells : 0.0;
  // non-temporal step-kernel stores from kNtStoreMin spins: the 256-replica
  // workload (6.4 M spins) FP64 step 561 -> 529 µs, the C3 LINE step 3 %
  // faster; the C2 system (0.5 M spins, whose outputs the next launch still
  // finds in the caches) 3-7 % slower with them
  p.ntstore = c->n >= kNtStoreMin ? 1 : 0;
  // short rows (the two-shell bcc workload): alternating the basis sites'
  // workgroups keeps each XCD's gathers of the other site in its L2 (C2: HBM
  // reads 40 % lower in the predictor, 19 % in the corrector, same speed);
  // with long rows the lines run on LINE, and the gathering kernels measured
  // 4 % slower interleaved
  p.basis_ilv = 0;
  if (c->order_nb > 1 && !c->xh.empty()) {
    const XHost &x = c->xh[0];
    const int64_t m = c->n / 2;
    p.basis_ilv = (x.ptr[m + 1] - x.ptr[m] < kLnAutoMinRow) ? 1 : 0;
  }
  return p;
}


// One RK4 step's launches: stage k evaluates the fields at its input (s,
// then sp, sq, sp) and writes the next input (sp, sq, sp) or, at stage 3, s.
void enqueue_rk4(jh_ctx *c, KParams p) {
  const int dip = dipole_term(c);
  const bool f32 = c->prec == PR_FP32;
  double *in64[4] = {c->s.p, c->sp.p, c->sq.p, c->sp.p};
  float *in32[4] = {c->s32.p, c->sp32.p, c->sq32.p, c->sp32.p};
  double *out64[3] = {c->sp.p, c->sq.p, c->sp.p};
  float *out32[4] = {c->sp32.p, c->sq32.p, c->sp32.p, c->s32.p};
  for (int k = 0; k < 4; ++k) {
    if (dip >= 0) HIPCHK(c->dipoles[dip]->field(in64[k], c->hext.p, f32, c->stream));
    p.stage = k;
    p.sp = in64[k];
    p.sy = k < 3 ? out64[k] : nullptr;
    p.g32 = in32[k];
    p.w32 = out32[k];
    HIPCHK(launch_rk4(p, c->prec, c->xv, c->launch_block(), c->stream));
  }
}

// One Heun step's launches (dipole fields, predictor, corrector) for step
// counter / time in p (or on the device when p.dstep is set).
// One Heun phase: the LINE or STENCIL launch and, in FP64, the row-list
// launch of its wrapped rows; or one launch of the context's exchange
// variant.
void launch_phase(jh_ctx *c, const KParams &p, int phase) {
  if (c->ln_on) {
    LnView v = c->ln;
    v.seq = reinterpret_cast<const int2 *>(c->ln_seq.p);
    v.prog = c->ln_prog.p;
    v.val = c->ln_val.p;
    v.wseq = c->ln_wseq.p;
    v.fixmask = c->ln_fixmask.p;
    v.pval = c->ln_pairs ? c->ln_pval.p : nullptr;
    // the row list rides in the same launch (its first workgroups)
    const int64_t lines = (int64_t)v.nb * v.lx * v.ly;
    const int bs = lines % (kLnBlock / 64) == 0 ? kLnBlock : 64;
    v.rows = c->st_nfix > 0 ? c->st_rows.p : nullptr;
    v.nrows = c->st_nfix;
    v.rlblocks = (int)((c->st_nfix + bs - 1) / bs);
    HIPCHK(launch_heun_ln(p, v, c->prec, xv_diag(c->xv), phase, c->stream));
    return;
  } else if (c->st_on) {
    StView v = c->st;
    v.seq = reinterpret_cast<const int2 *>(c->st_seq.p);
    v.off = c->st_off.p;
    v.val = c->st_val.p;
    v.wseq = c->st_wseq.p;
    v.fixmask = c->st_fixmask.p;
    HIPCHK(launch_heun_st(p, v, c->prec, xv_diag(c->xv), phase, c->stream));
  } else {
    HIPCHK(launch_heun(p, c->prec, c->xv, phase, c->launch_block(), c->stream));
    return;
  }
  if (c->st_nfix > 0) {
    KParams q = p;
    q.rowlist = c->st_rows.p;
    q.nrowlist = c->st_nfix;
    HIPCHK(launch_heun(q, c->prec, c->xv, phase, c->launch_block(), c->stream));
  }
}

void enqueue_step(jh_ctx *c, KParams p) {
  if (c->integrator == JH_INTEGRATOR_RK4) {
    enqueue_rk4(c, p);
    return;
  }
  p.g32 = c->s32.p;
  p.w32 = c->sp32.p;
  if (c->step_launch == JH_STEP_LAUNCH_PERSISTENT) {
    HIPCHK(launch_heun_step(p, c->prec, c->xv, c->launch_block(), c->cus, c->step_sync.p,
                            kStepPollLimit, c->stream));
    return;
  }
  const int dip = dipole_term(c);
  if (dip >= 0) HIPCHK(c->dipoles[dip]->field(c->s.p, c->hext.p, c->prec == PR_FP32, c->stream));
  launch_phase(c, p, 0);
  // corrector at t + dt
  if (dip >= 0) HIPCHK(c->dipoles[dip]->field(c->sp.p, c->hext.p, c->prec == PR_FP32, c->stream));
  p.g32 = c->sp32.p;
  p.w32 = c->s32.p;
  launch_phase(c, p, 1);
}

// Capture graph_steps steps.  Their kernels read the step counter as
// *c->dstep + k (k, the step's place in the graph, is baked into its
// parameters), and the host sets *c->dstep once per replay with a one-thread
// kernel outside the graph (no memset nodes).
void build_graph(jh_ctx *c) {
  c->drop_graph();
  if (!c->dstep.p) c->dstep.alloc(1);
  KParams p = make_params(c);
  p.dstep = c->dstep.p;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int k = 0; k < c->graph_steps; ++k) {
      p.step = (uint64_t)k;
      enqueue_step(c, p);
    }
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIPCHK(hipStreamEndCapture(c->stream, &c->graph));
  HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
}

void do_steps(jh_ctx *c, int64_t nsteps) {
  int64_t done = 0;
  if (c->graph_steps > 0 && nsteps >= c->graph_steps) {
    if (!c->gexec) build_graph(c);
    for (; done + c->graph_steps <= nsteps; done += c->graph_steps) {
      HIPCHK(launch_set_step(c->dstep.p, (uint64_t)(c->iteration + done), c->stream));
      HIPCHK(hipGraphLaunch(c->gexec, c->stream));
    }
    c->iteration += done;
    c->time = (double)c->iteration * c->dt;
  }
  KParams p = make_params(c);
  for (; done < nsteps; ++done) {
    p.step = (uint64_t)c->iteration;
    p.t = c->time;
    enqueue_step(c, p);
    c->iteration += 1;
    c->time = (double)c->iteration * c->dt;
  }
}

template <class F>
int guarded(jh_ctx *c, F &&f) {
  if (!c) return JH_ERR_INVALID;
  try {
    f();
    return JH_OK;
  } catch (const JhError &e) {
    c->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    c->err = "host allocation failed";
    return JH_ERR_NOMEM;
  } catch (const std::exception &e) {
    c->err = e.what();
    return JH_ERR_INVALID;
  }
}

void field_into(jh_ctx *c, int term, double t, double *dst) {
  KParams p = make_params(c);
  p.t = t;
  p.g32 = c->s32.p;
  const int dip = dipole_term(c);
  if (dip >= 0 && (term < 0 || c->terms[term].kind == TK_DIPOLE))
    HIPCHK(c->dipoles[dip]->field(c->s.p, c->hext.p, c->prec == PR_FP32, c->stream));
  HIPCHK(launch_field(p, c->prec, c->xv, term, dst, c->stream));
}

}  // namespace

// ===========================================================================
// C ABI

extern "C" {

int jh_version(void) { return 10000; }

int jh_create(int device, jh_ctx **out) {
  if (!out) return JH_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    g_create_error = std::string("no HIP device available: ") + hipGetErrorString(e);
    return JH_ERR_HIP;
  }
  if (device < 0 || device >= count) {
    g_create_error = "device index out of range";
    return JH_ERR_INVALID;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return JH_ERR_HIP;
  }
  e = bind_scratch(device);
  if (e != hipSuccess) {
    g_create_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return JH_ERR_HIP;
  }
  jh_ctx *c = new (std::nothrow) jh_ctx();
  if (!c) return JH_ERR_NOMEM;
  c->device = device;
  if ((e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device)) !=
      hipSuccess) {
    g_create_error = std::string("compute unit count: ") + hipGetErrorString(e);
    delete c;
    return JH_ERR_HIP;
  }
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
    g_create_error = std::string("stream/event creation: ") + hipGetErrorString(e);
    delete c;
    return JH_ERR_HIP;
  }
  *out = c;
  return JH_OK;
}

int jh_destroy(jh_ctx *c) {
  if (!c) return JH_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->drop_graph();
  c->dipoles.clear();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return JH_OK;
}

const char *jh_last_error(const jh_ctx *c) {
  return c ? c->err.c_str() : g_create_error.c_str();
}

int jh_set_sites(jh_ctx *c, int64_t n, const double *mus, const double *gyro, const double *alpha) {
  return guarded(c, [&] {
    if (n <= 0 || !mus || !gyro || !alpha) JH_THROW(JH_ERR_INVALID, "jh_set_sites: bad arguments");
    if (n > (int64_t)1 << 31) JH_THROW(JH_ERR_INVALID, "jh_set_sites: at most 2^31 spins");
    HIPCHK(hipSetDevice(c->device));
    c->n = n;
    c->mus.assign(mus, mus + n);
    c->gyro.assign(gyro, gyro + n);
    c->alpha.assign(alpha, alpha + n);
    for (int64_t i = 0; i < n; ++i)
      if (!(c->mus[i] != 0.0)) JH_THROW(JH_ERR_INVALID, "jh_set_sites: mus must be non-zero");
    c->terms.clear();
    c->xh.clear();
    c->dipoles.clear();
    c->s.alloc((size_t)(3 * n));
    c->sp.alloc((size_t)(3 * n));
    c->ds.alloc((size_t)(3 * n));
    c->fout.alloc((size_t)(3 * n));
    const int64_t nb = (n + kBlock - 1) / kBlock;
    c->red_partial.alloc((size_t)(4 * nb));
    c->red_out.alloc(4);
    c->spins_set = false;
    c->have_lattice = false;
    c->order_nb = 1;
    c->order_cells = n;
    c->noise_mode = NM_NONE;
    c->nrep = 0;
    c->rep_off.clear();
    c->rep_T.clear();
    c->rep_seed.clear();
    c->iteration = 0;
    c->time = 0.0;
    c->dirty = true;
  });
}

int jh_set_lattice(jh_ctx *c, const int32_t size[3], const int32_t periodic[3], int nbasis) {
  return guarded(c, [&] {
    require_sites(c);
    if (!size || !periodic || nbasis < 1)
      JH_THROW(JH_ERR_INVALID, "jh_set_lattice: bad arguments");
    int64_t cells = 1;
    for (int a = 0; a < 3; ++a) {
      if (size[a] < 1) JH_THROW(JH_ERR_INVALID, "jh_set_lattice: sizes must be positive");
      cells *= size[a];
    }
    if (cells * nbasis != c->n)
      JH_THROW(JH_ERR_INVALID, "jh_set_lattice: size[0] size[1] size[2] nbasis must equal the number of spins");
    c->have_lattice = true;
    for (int a = 0; a < 3; ++a) {
      c->lsize[a] = size[a];
      c->lper[a] = periodic[a] ? 1 : 0;
    }
    c->lnb = nbasis;
    c->dirty = true;
  });
}

int jh_set_spins(jh_ctx *c, const double *s) {
  return guarded(c, [&] {
    require_sites(c);
    if (!s) JH_THROW(JH_ERR_INVALID, "jh_set_spins: null");
    upload_vec(c, s, c->s.p, true);
    c->spins_set = true;
    if (!c->dirty) sync_f32(c);
  });
}

int jh_get_spins(jh_ctx *c, double *s) {
  return guarded(c, [&] {
    require_sites(c);
    if (!s) JH_THROW(JH_ERR_INVALID, "jh_get_spins: null");
    check_step_sync(c);
    download_vec(c, c->s.p, s, true);
  });
}

int jh_set_time_step(jh_ctx *c, double dt) {
  return guarded(c, [&] {
    if (!(dt > 0.0)) JH_THROW(JH_ERR_INVALID, "time step must be positive");
    c->dt = dt;
    c->have_dt = true;
    c->time = (double)c->iteration * dt;
    c->dirty = true;
  });
}

int jh_add_exchange_csr(jh_ctx *c, int64_t nrows, int64_t nnz, const int64_t *rowptr,
                        const int32_t *col, const double *val, int *term_id) {
  return guarded(c, [&] {
    require_sites(c);
    if (!rowptr || (nnz > 0 && (!col || !val)) || nnz < 0)
      JH_THROW(JH_ERR_INVALID, "jh_add_exchange_csr: bad arguments");
    c->xh.push_back(xhost_from_csr(c->n, nrows, nnz, rowptr, col, val));
    TermHost t;
    t.kind = TK_EXCHANGE;
    t.xidx = (int)c->xh.size() - 1;
    c->terms.push_back(std::move(t));
    if (term_id) *term_id = (int)c->terms.size() - 1;
    c->dirty = true;
  });
}

int jh_add_exchange_list(jh_ctx *c, int64_t np, const int32_t *pi, const int32_t *pj,
                         const int32_t *vi, int64_t nv, const double *v9, int *term_id) {
  return guarded(c, [&] {
    require_sites(c);
    if (np < 0 || nv < 0 || (np > 0 && (!pi || !pj || !vi || !v9)))
      JH_THROW(JH_ERR_INVALID, "jh_add_exchange_list: bad arguments");
    c->xh.push_back(xhost_from_list(c->n, np, pi, pj, vi, nv, v9));
    TermHost t;
    t.kind = TK_EXCHANGE;
    t.xidx = (int)c->xh.size() - 1;
    c->terms.push_back(std::move(t));
    if (term_id) *term_id = (int)c->terms.size() - 1;
    c->dirty = true;
  });
}

int jh_add_zeeman(jh_ctx *c, const double *dc, const double *ac, const double *omega, int *term_id) {
  return guarded(c, [&] {
    require_sites(c);
    if (!dc || ((ac == nullptr) != (omega == nullptr)))
      JH_THROW(JH_ERR_INVALID, "jh_add_zeeman: dc required; ac and omega together");
    TermHost t;
    t.kind = TK_ZEEMAN;
    t.dc.assign(dc, dc + 3 * c->n);
    if (ac) {
      t.ac.assign(ac, ac + 3 * c->n);
      t.omega.assign(omega, omega + c->n);
    }
    c->terms.push_back(std::move(t));
    if (term_id) *term_id = (int)c->terms.size() - 1;
    c->dirty = true;
  });
}

int jh_add_uniaxial(jh_ctx *c, const double *K, const double *axis, int power, int *term_id) {
  return guarded(c, [&] {
    require_sites(c);
    if (!K || !axis) JH_THROW(JH_ERR_INVALID, "jh_add_uniaxial: null");
    if (power != 2 && power != 4 && power != 6)
      JH_THROW(JH_ERR_INVALID, "jh_add_uniaxial: power must be 2, 4 or 6");
    TermHost t;
    t.kind = TK_UNIAXIAL;
    t.K.assign(K, K + c->n);
    t.axis.assign(axis, axis + 3 * c->n);
    t.power = power;
    c->terms.push_back(std::move(t));
    if (term_id) *term_id = (int)c->terms.size() - 1;
    c->dirty = true;
  });
}

int jh_add_dipole_fft(jh_ctx *c, const int32_t size[3], const int32_t periodic[3], int nbasis,
                      const double cell[9], const double *basis_frac, double a_m, double r_cutoff,
                      int *term_id) {
  return guarded(c, [&] {
    require_sites(c);
    if (!size || !periodic || !cell || !basis_frac || nbasis <= 0)
      JH_THROW(JH_ERR_INVALID, "jh_add_dipole_fft: bad arguments");
    for (auto &t : c->terms)
      if (t.kind == TK_DIPOLE) JH_THROW(JH_ERR_UNSUPPORTED, "at most one dipole term");
    const int64_t cells = (int64_t)size[0] * size[1] * size[2];
    if (cells * nbasis != c->n) JH_THROW(JH_ERR_INVALID, "dipole lattice does not match N");
    std::string msg;
    std::unique_ptr<DipoleFFT> d(DipoleFFT::create(size, periodic, nbasis, cell, basis_frac, a_m,
                                                   r_cutoff, c->mus.data(), c->n, c->stream, &msg));
    if (!d) JH_THROW(JH_ERR_INVALID, "dipole-fft: " + msg);
    c->dipoles.push_back(std::move(d));
    c->hext.alloc((size_t)(3 * c->n));
    TermHost t;
    t.kind = TK_DIPOLE;
    t.dipole = (int)c->dipoles.size() - 1;
    c->terms.push_back(std::move(t));
    if (term_id) *term_id = (int)c->terms.size() - 1;
    c->dirty = true;
  });
}

int jh_num_terms(const jh_ctx *c) { return c ? (int)c->terms.size() : 0; }

int jh_term_kind(const jh_ctx *c, int term_id) {
  if (!c || term_id < 0 || term_id >= (int)c->terms.size()) return -1;
  return c->terms[term_id].kind;
}

int jh_set_thermostat(jh_ctx *c, double temperature, uint64_t seed, int gilbert) {
  return guarded(c, [&] {
    require_sites(c);
    if (!(temperature >= 0.0)) JH_THROW(JH_ERR_INVALID, "temperature must be >= 0");
    c->temperature = temperature;
    c->seed = seed;
    c->gilbert = gilbert ? 1 : 0;
    c->noise_mode = temperature > 0.0 ? NM_PHILOX : NM_NONE;
    c->dirty = true;
  });
}

int jh_set_noise(jh_ctx *c, const double *w) {
  return guarded(c, [&] {
    require_sites(c);
    if (!w) JH_THROW(JH_ERR_INVALID, "jh_set_noise: null");
    c->winj.alloc((size_t)(3 * c->n));
    upload_vec(c, w, c->winj.p, false);
    c->noise_mode = NM_INJECTED;
  });
}

int jh_get_noise(jh_ctx *c, int64_t step, double *w) {
  return guarded(c, [&] {
    prepare(c);
    if (!c->have_dt) JH_THROW(JH_ERR_STATE, "time step not set");
    KParams p = make_params(c);
    p.step = (uint64_t)step;
    HIPCHK(launch_noise(p, c->fout.p, c->stream));
    download_vec(c, c->fout.p, w, false);
  });
}

int jh_set_precision(jh_ctx *c, int mode) {
  return guarded(c, [&] {
    if (mode != JH_PREC_FP64 && mode != JH_PREC_FP32_FIELD)
      JH_THROW(JH_ERR_INVALID, "unknown precision mode");
    c->prec = mode == JH_PREC_FP64 ? PR_FP64 : PR_FP32;
    c->dirty = true;
  });
}

int jh_set_integrator(jh_ctx *c, int integrator) {
  return guarded(c, [&] {
    if (integrator != JH_INTEGRATOR_HEUN && integrator != JH_INTEGRATOR_RK4)
      JH_THROW(JH_ERR_INVALID, "unknown integrator");
    c->integrator = integrator;
    c->dirty = true;
    c->drop_graph();
  });
}

int jh_set_exchange_format(jh_ctx *c, int fmt) {
  return guarded(c, [&] {
    if (fmt != JH_XFMT_AUTO && fmt != JH_XFMT_TYPED && fmt != JH_XFMT_BLOCK && fmt != JH_XFMT_PATTERN &&
        fmt != JH_XFMT_PATTERN_GLOBAL && fmt != JH_XFMT_STENCIL && fmt != JH_XFMT_LINE)
      JH_THROW(JH_ERR_INVALID, "unknown exchange format");
    c->xfmt_req = fmt;
    c->dirty = true;
  });
}

int jh_step(jh_ctx *c, int64_t nsteps) {
  return guarded(c, [&] {
    if (nsteps < 0) JH_THROW(JH_ERR_INVALID, "nsteps must be >= 0");
    if (!c->have_dt) JH_THROW(JH_ERR_STATE, "time step not set");
    HIPCHK(hipSetDevice(c->device));
    prepare(c);
    do_steps(c, nsteps);
  });
}

int jh_get_iteration(const jh_ctx *c, int64_t *it, double *t) {
  if (!c) return JH_ERR_INVALID;
  if (it) *it = c->iteration;
  if (t) *t = c->time;
  return JH_OK;
}

int jh_set_iteration(jh_ctx *c, int64_t it) {
  return guarded(c, [&] {
    if (it < 0) JH_THROW(JH_ERR_INVALID, "iteration must be >= 0");
    c->iteration = it;
    c->time = (double)it * c->dt;
  });
}

int jh_compute_term_field(jh_ctx *c, int term, double t, double *out) {
  return guarded(c, [&] {
    prepare(c);
    if (term < 0 || term >= (int)c->terms.size()) JH_THROW(JH_ERR_INVALID, "term id out of range");
    field_into(c, term, t, c->fout.p);
    download_vec(c, c->fout.p, out, false);
  });
}

int jh_compute_fields(jh_ctx *c, double t, double *out) {
  return guarded(c, [&] {
    prepare(c);
    field_into(c, -1, t, c->fout.p);
    download_vec(c, c->fout.p, out, false);
  });
}

int jh_term_energy(jh_ctx *c, int term, double t, double *energy) {
  return guarded(c, [&] {
    prepare(c);
    if (term < 0 || term >= (int)c->terms.size()) JH_THROW(JH_ERR_INVALID, "term id out of range");
    KParams p = make_params(c);
    p.t = t;
    const int kind = c->terms[term].kind;
    if (kind == TK_UNIAXIAL) {
      HIPCHK(launch_reduce(p, 2, nullptr, c->red_partial.p, c->red_out.p, c->stream));
    } else {
      field_into(c, term, t, c->fout.p);
      HIPCHK(launch_reduce(p, 1, c->fout.p, c->red_partial.p, c->red_out.p, c->stream));
    }
    double r[4];
    HIPCHK(hipMemcpyAsync(r, c->red_out.p, sizeof r, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // exchange and dipole: -1/2 sum s.h (sparse_interaction.cc:86-100,
    // dipole_bruteforce.cc:55-58); zeeman: -sum s.h; uniaxial: -sum K (s.e)^p
    *energy = (kind == TK_EXCHANGE || kind == TK_DIPOLE) ? 0.5 * r[0] : r[0];
  });
}

int jh_reduce_magnetisation(jh_ctx *c, double out[4]) {
  return guarded(c, [&] {
    prepare(c);
    KParams p = make_params(c);
    HIPCHK(launch_reduce(p, 0, nullptr, c->red_partial.p, c->red_out.p, c->stream));
    HIPCHK(hipMemcpyAsync(out, c->red_out.p, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int jh_reduce_magnetisation_noalloc(jh_ctx *c, double out[4]) {
  return guarded(c, [&] {
    require_sites(c);
    if (!out) JH_THROW(JH_ERR_INVALID, "jh_reduce_magnetisation_noalloc: out is null");
    if (c->n > kNoAllocMaxSpins)
      JH_THROW(JH_ERR_INVALID, "jh_reduce_magnetisation_noalloc: " + std::to_string(c->n) +
                                   " spins exceed its scratch limit of " +
                                   std::to_string(kNoAllocMaxSpins) +
                                   "; use jh_reduce_magnetisation");
    if (!c->spins_set) JH_THROW(JH_ERR_STATE, "spins have not been set (jh_set_spins)");
    if (c->dirty)
      JH_THROW(JH_ERR_STATE,
               "jh_reduce_magnetisation_noalloc: the context has set-up changes that are not laid "
               "out on the device yet (that allocates); call jh_step, jh_compute_fields or "
               "jh_reduce_magnetisation once first");
    double *scr = scratch_on(c->device);
    if (!scr) JH_THROW(JH_ERR_STATE, "jh_reduce_magnetisation_noalloc: no scratch address for this device");
    HIPCHK(hipSetDevice(c->device));
    KParams p = make_params(c);
    const int64_t nb = (c->n + kBlock - 1) / kBlock;
    double *result = scr + 4 * nb;
    HIPCHK(launch_reduce(p, 0, nullptr, scr, result, c->stream));
    HIPCHK(hipMemcpyAsync(out, result, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int jh_time_steps(jh_ctx *c, int64_t nsteps, double *ms) {
  return guarded(c, [&] {
    if (!c->have_dt) JH_THROW(JH_ERR_STATE, "time step not set");
    prepare(c);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    do_steps(c, nsteps);
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = f;
  });
}

int jh_time_phase(jh_ctx *c, int phase, int reps, double *ms) {
  return guarded(c, [&] {
    if (!c->have_dt) JH_THROW(JH_ERR_STATE, "time step not set");
    if (reps <= 0 || (phase != 0 && phase != 1)) JH_THROW(JH_ERR_INVALID, "bad phase or reps");
    if (c->integrator != JH_INTEGRATOR_HEUN)
      JH_THROW(JH_ERR_UNSUPPORTED, "phase timing covers the Heun phases; time RK4 with jh_time_steps");
    prepare(c);
    // The corrector overwrites s; keep a copy and restore it afterwards so the
    // timing leaves the simulation state unchanged.
    DevBuf<double> save;
    DevBuf<float> save32;
    save.alloc((size_t)(3 * c->n));
    HIPCHK(hipMemcpyAsync(save.p, c->s.p, sizeof(double) * 3 * c->n, hipMemcpyDeviceToDevice, c->stream));
    if (c->prec == PR_FP32) {
      save32.alloc((size_t)(kG32 * c->n));
      HIPCHK(hipMemcpyAsync(save32.p, c->s32.p, sizeof(float) * kG32 * c->n, hipMemcpyDeviceToDevice, c->stream));
    }
    KParams p = make_params(c);
    p.step = (uint64_t)c->iteration;
    p.t = c->time;
    p.g32 = phase == 0 ? c->s32.p : c->sp32.p;
    p.w32 = phase == 0 ? c->sp32.p : c->s32.p;
    // one untimed launch of each phase so the corrector reads a valid s*
    {
      KParams q = p;
      q.g32 = c->s32.p;
      q.w32 = c->sp32.p;
      launch_phase(c, q, 0);
    }
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; ++r) launch_phase(c, p, phase);
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = (double)f / reps;
    HIPCHK(hipMemcpyAsync(c->s.p, save.p, sizeof(double) * 3 * c->n, hipMemcpyDeviceToDevice, c->stream));
    if (c->prec == PR_FP32)
      HIPCHK(hipMemcpyAsync(c->s32.p, save32.p, sizeof(float) * kG32 * c->n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int jh_phase_bytes(jh_ctx *c, int phase, double *fmt_bytes, double *csr_bytes) {
  return guarded(c, [&] {
    prepare(c);
    const double n = (double)c->n;
    const double sv = c->prec == PR_FP64 ? 8.0 : 4.0;
    const int vpb = (c->xv == XV_BDIAG || c->xv == XV_TDIAG) ? 3 : 9;
    double mat = 0.0, gather = 0.0;
    if (c->ln_on) {
      // LINE: the programs are per sequence, not per spin
      mat = 0.0;
      gather = c->prec == PR_FP64 ? 3.0 * n * 8 : kG32 * n * 4;
    } else if (xv_pattern(c->xv)) {
      mat = n * 4.0 + (c->pat_lds ? 0.0 : (double)c->pat_words * 4.0);
      gather = c->prec == PR_FP64 ? 3.0 * n * 8 : kG32 * n * 4;
    } else if (c->xv != XV_NONE) {
      mat = (double)c->nent * 4.0 + (double)(c->slice_off.n) * 8.0;
      if (c->xv == XV_BDIAG || c->xv == XV_BFULL) mat += (double)c->nent * vpb * sv;
      // every neighbour vector read once (ideal reuse of the gathers)
      gather = c->prec == PR_FP64 ? 3.0 * n * 8 : kG32 * n * 4;
    }
    // per-spin state traffic: the FP64 vector the field is evaluated at (its
    // gathers are counted above: in FP64 they are the same array, so only
    // the FP32 form adds its own FP64 read), plus the Heun state
    // (the FP32 form without a dipole term neither stores s* nor reads it back)
    const bool spfree = make_params(c).spfree != 0;
    double state = c->prec == PR_FP32 || c->xv == XV_NONE ? 3 * n * 8 : 0.0;
    if (phase == 1 && spfree) state = 0.0;
    if (phase == 0) state += (spfree ? 0.0 : 3 * n * 8 /*s* w*/) + 3 * n * 8 /*ds w*/;
    else state += 3 * n * 8 /*s_old r*/ + 3 * n * 8 /*ds r*/ + 3 * n * 8 /*s w*/;
    if (c->prec == PR_FP32) state += kG32 * n * 4;  // FP32 gather copy written
    if (c->nclass > 1) state += 2 * n;
    if (c->noise_mode == NM_INJECTED) state += 3 * n * 8;
    if (c->noise_mode == NM_PHILOX) state += 3 * n * 8;  // noise written, then read back
    if (dipole_term(c) >= 0) state += 3 * n * 8;
    *fmt_bytes = mat + gather + state;
    // reference CSR SpMV (FP64 values, int32 indices): nnz (8+4) + (3N+1) 4 + 3N 8 + 3N 8
    const double nnz = c->xh.empty() ? 0.0 : (double)c->xh[0].csr_nnz;
    *csr_bytes = nnz * 12.0 + (3 * n + 1) * 4.0 + 3 * n * 8.0 + 3 * n * 8.0;
  });
}

int jh_time_term(jh_ctx *c, int term, int reps, double *ms) {
  return guarded(c, [&] {
    prepare(c);
    if (term < 0 || term >= (int)c->terms.size()) JH_THROW(JH_ERR_INVALID, "term id out of range");
    if (reps <= 0) JH_THROW(JH_ERR_INVALID, "reps must be positive");
    KParams p = make_params(c);
    p.t = c->time;
    p.g32 = c->s32.p;
    const bool dip = c->terms[term].kind == TK_DIPOLE;
    auto once = [&] {
      if (dip)
        HIPCHK(c->dipoles[c->terms[term].dipole]->field(c->s.p, c->hext.p, c->prec == PR_FP32, c->stream));
      else
        HIPCHK(launch_field(p, c->prec, c->xv, term, c->fout.p, c->stream));
    };
    once();
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; ++r) once();
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = (double)f / reps;
  });
}

int jh_term_bytes(jh_ctx *c, int term, double *bytes) {
  return guarded(c, [&] {
    prepare(c);
    if (term < 0 || term >= (int)c->terms.size()) JH_THROW(JH_ERR_INVALID, "term id out of range");
    const double n = (double)c->n;
    const int kind = c->terms[term].kind;
    if (kind == TK_DIPOLE) {
      *bytes = c->dipoles[c->terms[term].dipole]->bytes(c->prec == PR_FP32);
    } else {
      // the field kernel: spins read once, field written; exchange adds its format
      *bytes = 24.0 * n + 24.0 * n;
      if (kind == TK_EXCHANGE) *bytes += xv_pattern(c->xv) ? 4.0 * n : (double)c->nent * 4.0;
    }
  });
}

int jh_set_replicas(jh_ctx *c, int nrep, const int64_t *offsets, const double *temperature,
                    const uint64_t *seeds) {
  return guarded(c, [&] {
    require_sites(c);
    if (nrep < 0) JH_THROW(JH_ERR_INVALID, "nrep must be >= 0");
    c->dirty = true;
    if (nrep == 0) {
      c->nrep = 0;
      c->rep_off.clear();
      c->rep_T.clear();
      c->rep_seed.clear();
      return;
    }
    if (!offsets || !temperature || !seeds) JH_THROW(JH_ERR_INVALID, "null replica arrays");
    if (offsets[0] != 0 || offsets[nrep] != c->n)
      JH_THROW(JH_ERR_INVALID, "replica offsets must start at 0 and end at N");
    for (int r = 0; r < nrep; ++r) {
      if (offsets[r + 1] <= offsets[r]) JH_THROW(JH_ERR_INVALID, "replica offsets must increase");
      if (!(temperature[r] >= 0.0)) JH_THROW(JH_ERR_INVALID, "temperature must be >= 0");
    }
    c->nrep = nrep;
    c->rep_off.assign(offsets, offsets + nrep + 1);
    c->rep_T.assign(temperature, temperature + nrep);
    c->rep_seed.assign(seeds, seeds + nrep);
    c->d_rep_off.alloc((size_t)nrep + 1);
    HIPCHK(hipMemcpy(c->d_rep_off.p, offsets, sizeof(int64_t) * (nrep + 1), hipMemcpyHostToDevice));
    c->rep_out.alloc((size_t)nrep * 4);
    if (c->noise_mode == NM_INJECTED) c->noise_mode = NM_NONE;
  });
}

int jh_reduce_magnetisation_replicas(jh_ctx *c, double *out) {
  return guarded(c, [&] {
    prepare(c);
    if (c->nrep == 0) JH_THROW(JH_ERR_STATE, "no replicas (jh_set_replicas)");
    KParams p = make_params(c);
    HIPCHK(launch_reduce_replicas(p, c->d_rep_off.p, c->nrep, c->rep_out.p, c->stream));
    HIPCHK(hipMemcpyAsync(out, c->rep_out.p, sizeof(double) * 4 * c->nrep, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int jh_set_graph_steps(jh_ctx *c, int steps) {
  return guarded(c, [&] {
    if (steps < 0) JH_THROW(JH_ERR_INVALID, "steps must be >= 0");
    c->graph_steps = steps;
    c->drop_graph();
  });
}

int jh_set_block_size(jh_ctx *c, int threads) {
  return guarded(c, [&] {
    if (threads != 0 && (threads < 64 || threads > kStMaxBlock || threads % 64 != 0))
      JH_THROW(JH_ERR_INVALID, "workgroup size must be 0 or a multiple of 64 in [64, 1024]");
    if (threads != c->block && (c->st_on || c->have_lattice)) c->dirty = true;  // tile shape
    c->block = threads;
    c->drop_graph();
  });
}

int jh_autotune(jh_ctx *c, int steps, char *report, int64_t len) {
  return guarded(c, [&] {
    if (!c->have_dt) JH_THROW(JH_ERR_STATE, "time step not set");
    if (steps <= 0) JH_THROW(JH_ERR_INVALID, "steps must be positive");
    prepare(c);
    // the state every candidate starts from, restored afterwards
    DevBuf<double> save;
    save.alloc((size_t)(3 * c->n));
    HIPCHK(hipMemcpyAsync(save.p, c->s.p, sizeof(double) * 3 * c->n, hipMemcpyDeviceToDevice, c->stream));
    const int64_t it0 = c->iteration;
    const double t0 = c->time;
    const int fmt0 = c->xfmt_req, bs0 = c->block, gs0 = c->graph_steps, sl0 = c->step_launch;
    c->graph_steps = 0;
    c->step_launch = JH_STEP_LAUNCH_TWO;
    std::string rep;
    double best = 1e300;
    int best_fmt = fmt0, best_bs = bs0, best_sl = sl0;
    for (int fmt : {JH_XFMT_PATTERN, JH_XFMT_TYPED, JH_XFMT_BLOCK, JH_XFMT_STENCIL, JH_XFMT_LINE}) {
      if (c->xh.empty() && fmt != JH_XFMT_PATTERN) continue;
      const bool stencil = fmt == JH_XFMT_STENCIL;
      const bool line = fmt == JH_XFMT_LINE;
      if ((stencil || line) && (!c->have_lattice || c->integrator != JH_INTEGRATOR_HEUN)) continue;
      c->xfmt_req = fmt;
      c->dirty = true;
      if (!stencil) {
        try {
          prepare(c);
        } catch (const JhError &) {
          continue;  // this format does not fit the system
        }
      }
      for (int bs : {128, 256, 512, 1024}) {
        if (bs == 1024 && !stencil) continue;
        if (line && bs != kLnBlock) continue;  // one shape: a wave per line
        c->block = bs;
        if (stencil) {
          // the block size is the tile shape: lay it out again
          c->dirty = true;
          try {
            prepare(c);
          } catch (const JhError &) {
            continue;  // no stencil of this many threads fits
          }
        }
        // the gathering formats also as one persistent launch per step
        const bool persist = persistent_unfit(c) == nullptr;
        for (int sl : {JH_STEP_LAUNCH_TWO, JH_STEP_LAUNCH_PERSISTENT}) {
          if (sl == JH_STEP_LAUNCH_PERSISTENT && !persist) continue;
          c->step_launch = sl;
          if (sl == JH_STEP_LAUNCH_PERSISTENT) ensure_step_sync(c);
          HIPCHK(hipMemcpyAsync(c->s.p, save.p, sizeof(double) * 3 * c->n, hipMemcpyDeviceToDevice, c->stream));
          sync_f32(c);
          c->iteration = it0;
          c->time = t0;
          do_steps(c, 2);
          HIPCHK(hipEventRecord(c->ev0, c->stream));
          do_steps(c, steps);
          HIPCHK(hipEventRecord(c->ev1, c->stream));
          HIPCHK(hipEventSynchronize(c->ev1));
          check_step_sync(c);
          float f = 0.f;
          HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
          const double ms = (double)f / steps;
          char line[128];
          std::snprintf(line, sizeof line, "format=%d block=%d launch=%s %.4f ms/step\n", fmt, bs,
                        sl == JH_STEP_LAUNCH_PERSISTENT ? "persistent" : "two", ms);
          rep += line;
          if (ms < best) {
            best = ms;
            best_fmt = fmt;
            best_bs = bs;
            best_sl = sl;
          }
        }
        c->step_launch = JH_STEP_LAUNCH_TWO;
      }
    }
    c->xfmt_req = best_fmt;
    c->block = best_bs;
    c->step_launch = best_sl;
    c->graph_steps = gs0;
    c->dirty = true;
    prepare(c);
    HIPCHK(hipMemcpyAsync(c->s.p, save.p, sizeof(double) * 3 * c->n, hipMemcpyDeviceToDevice, c->stream));
    sync_f32(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->iteration = it0;
    c->time = t0;
    char line[128];
    std::snprintf(line, sizeof line, "chosen format=%d block=%d launch=%s", best_fmt, best_bs,
                  best_sl == JH_STEP_LAUNCH_PERSISTENT ? "persistent" : "two");
    rep += line;
    if (report && len > 0) {
      std::strncpy(report, rep.c_str(), (size_t)len - 1);
      report[len - 1] = 0;
    }
  });
}

int jh_describe(jh_ctx *c, char *buf, int64_t buflen) {
  return guarded(c, [&] {
    prepare(c);
    static const char *xvn[] = {"none", "typed-diag", "typed-full", "block-diag", "block-full",
                                "pattern-diag", "pattern-full", "pattern-diag-global",
                                "pattern-full-global", "pattern-uniform-diag",
                                "pattern-uniform-full", "pattern-uniform-pairs-iso",
                                "pattern-uniform-pairs-dmi"};
    char tmp[512];
    std::snprintf(tmp, sizeof tmp,
                  "n=%lld prec=%s exchange=%s entries=%lld tensors=%d colbits=%d patterns=%d "
                  "pattern_words=%d lds=%d uniform_slices=%lld/%lld paired_slices=%lld "
                  "order=%d classes=%d "
                  "terms=%d noise=%d block=%d replicas=%d graph=%d integrator=%s",
                  (long long)c->n, c->prec == PR_FP64 ? "fp64" : "fp32-field", xvn[c->xv],
                  (long long)c->nent, c->ntab > 0 ? c->ntab - 1 : 0, c->colbits,
                  c->npatterns, c->pat_words, c->pat_lds, (long long)c->nuniform,
                  (long long)c->nslices, (long long)c->npaired, c->order_nb, c->nclass,
                  (int)c->terms.size(), c->noise_mode, c->launch_block(), c->nrep, c->graph_steps,
                  c->integrator == JH_INTEGRATOR_RK4 ? "rk4" : "heun");
    std::string out = tmp;
    if (c->st_on) {
      char st[160];
      std::snprintf(st, sizeof st, " stencil=%dx%dx64 tile_lds=%zu sequences=%d wrapped_rows=%lld",
                    c->st.wx, c->st.wy, st_tile_bytes(c->st, c->prec, 0), c->st_nseq,
                    (long long)c->st_nfix);
      out += st;
    } else {
      out += " stencil=off";
    }
    if (c->ln_on) {
      char ln[160];
      std::snprintf(ln, sizeof ln, " line=%dx64 sequences=%d groups=%d pairs=%d wrapped_rows=%lld",
                    c->ln.lz / 64, c->ln_nseq, c->ln_ngroups, c->ln_pairs ? 1 : 0,
                    (long long)c->st_nfix);
      out += ln;
    }
    const int dip = dipole_term(c);
    if (dip >= 0) out += " " + c->dipoles[dip]->describe(c->prec == PR_FP32);
    out += c->step_launch == JH_STEP_LAUNCH_PERSISTENT ? " launch=persistent" : " launch=two";
    if (buf && buflen > 0) {
      std::strncpy(buf, out.c_str(), (size_t)buflen - 1);
      buf[buflen - 1] = 0;
    }
  });
}

int jh_synchronize(jh_ctx *c) {
  return guarded(c, [&] {
    HIPCHK(hipStreamSynchronize(c->stream));
    check_step_sync(c);
  });
}

int jh_set_step_launch(jh_ctx *c, int mode) {
  return guarded(c, [&] {
    if (mode != JH_STEP_LAUNCH_TWO && mode != JH_STEP_LAUNCH_PERSISTENT)
      JH_THROW(JH_ERR_INVALID, "step launch mode must be JH_STEP_LAUNCH_TWO or _PERSISTENT");
    if (mode != c->step_launch) {
      c->step_launch = mode;
      c->dirty = true;
      c->drop_graph();
    }
  });
}

}  // extern "C"
