// jh_internal.h -- data structures shared by the engine's host code and its
// kernels.
//
// Device layout (HBM):
//  - FP64 spin arrays (s, predictor s*, RK4 stage inputs) are one
//    allocation of 3N doubles with (x, y) of spin i at [2i], [2i + 1] and z
//    at [2N + i] (jh_layout.h);
//  - the other vectors (ds, noise, fields) are SoA FP64: component c of
//    spin i at [c*N + i];
//  - the exchange operator is a sliced-ELL over spins: slices of 64
//    consecutive spins (one wavefront), each padded to the longest row in the
//    slice and stored column-major, so lane l of slot k reads entry
//    slice_off[s] + 64 k + l (coalesced).  TYPED entries pack the neighbour
//    column and the index of its 3x3 tensor in a table held in LDS; BLOCK
//    entries hold the column and the explicit tensor (FP64 or FP32);
//  - PATTERN entries: rows with the same sequence of (column - row, tensor)
//    pairs (every interior site of a crystal, and each kind of boundary site)
//    share one pattern; a spin stores only the word offset of its pattern in
//    a small table (LDS-resident when it fits), so a translation-invariant
//    lattice costs 4 bytes per spin and no dependent load precedes the
//    neighbour gathers;
//  - device spin order is basis-major when the rows repeat with a period nb
//    (one pattern per basis site of a crystal): device index d holds
//    reference spin ((d mod C) nb + d / C), C = N / nb, so the lanes of a
//    wave read contiguous neighbours; the reference (AoS) order is restored
//    at the C ABI;
//  - per-spin constants (mus, gyro, alpha, sigma, Zeeman and anisotropy
//    parameters, and the temperature, noise seed and first index of the
//    spin's replica) are de-duplicated into a small table of site classes
//    with a 16-bit class index per spin (no index at all with one class);
//  - replicas (independent systems batched in one context) are contiguous
//    index ranges; nothing else distinguishes them on the device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jh_layout.h"

namespace jh {

constexpr int kBlock = 256;          // threads per workgroup of the helper kernels
constexpr int kMaxBlock = 512;       // largest step-kernel workgroup (autotuned)
constexpr int kSlice = 64;           // spins per ELL slice (one wavefront)
constexpr uint32_t kPad = 0xFFFFFFFFu;
#ifndef JH_G32
#define JH_G32 3
#endif
constexpr int kG32 = JH_G32;         // floats per spin in the FP32 gather copy

enum TermKind { TK_EXCHANGE = 0, TK_ZEEMAN = 1, TK_UNIAXIAL = 2, TK_DIPOLE = 3 };
enum NoiseMode { NM_NONE = 0, NM_PHILOX = 1, NM_INJECTED = 2 };
enum Prec { PR_FP64 = 0, PR_FP32 = 1 };
// exchange kernel variants
// (PDIAG/PFULL keep the pattern table in LDS, GDIAG/GFULL in global memory;
// UDIAG/UFULL are PDIAG/PFULL plus wave-uniform tensor sequences, FP32 form)
// QISO/QDMI: UDIAG/UFULL plus pair programs for isotropic (+ DMI) tensors
enum XVariant {
  XV_NONE = 0, XV_TDIAG = 1, XV_TFULL = 2, XV_BDIAG = 3, XV_BFULL = 4,
  XV_PDIAG = 5, XV_PFULL = 6, XV_GDIAG = 7, XV_GFULL = 8, XV_UDIAG = 9, XV_UFULL = 10,
  XV_QISO = 11, XV_QDMI = 12
};
constexpr bool xv_pair(int v) { return v == XV_QISO || v == XV_QDMI; }

// STENCIL (crystal lattices, Heun step kernels): a workgroup owns wx x wy
// runs of 64 cells along z of one basis site and stages the spins its rows
// read (the runs plus the stencil's halo, every basis site) in LDS once;
// each neighbour is then one LDS read at a per-entry offset from the lane's
// own tile position, and the tensor of entry k of the wave's sequence is
// wave-uniform (scalar loads).  FP64 sums each row in the reference column
// order: the waves of one (basis, x, y) line share it unless a z neighbour
// wraps around a periodic boundary, and those few rows are left to a second
// launch over a row list with their own row patterns.
constexpr int kStMaxBasis = 4;
constexpr int kStMaxBlock = 1024;
struct StView {
  int lx, ly, lz, nb;      // cells per axis, basis sites
  int per[3];              // periodic axes
  int wx, wy;              // waves per workgroup along x and y
  int nbx, nby, nbz;       // workgroups along x, y and z (nbz = lz / 64)
  // per source basis site b: cell offset of tile element 0 from the
  // workgroup's first cell (the lowest dx, dy, dz of its stencil) and tile
  // extents; the tile holds every basis site: nb * tx * ty * tz elements
  // ((bb * tx + tx_) * ty + ty_) * tz + tz_
  int ox[kStMaxBasis], oy[kStMaxBasis], oz[kStMaxBasis];
  int tx[kStMaxBasis], ty[kStMaxBasis], tz[kStMaxBasis];
  const int2 *seq;         // per sequence: (first entry, length)
  const int32_t *off;      // per entry: byte offset from the lane's own element
  const double *val;       // per entry: 3 (diagonal) or 9 tensor values
  const int32_t *wseq;     // FP64: sequence of line (b, x, y) at (b * lx + x) * ly + y;
                           // null: sequence b (FP32 form: one order per basis site)
  const uint64_t *fixmask; // FP64: per wave, the lanes whose row the row-list
                           // launch computes (their z neighbours wrap)
};
// LINE (crystal lattices with periodic z, Lz = 64 P, P = 1 or 2, at most two
// basis sites; Heun step kernels): a wave holds one whole (basis site, x, y)
// line along z, cell z = P lane + h in lane `lane` (h < P).  A neighbour
// line (b', x + dx, y + dy) is loaded once, coalesced, into registers, and
// the neighbour at z + dz is read from the lane that holds it by a
// wave-wide lane rotation (DPP), so a wave issues one load per neighbour
// line instead of one gather per neighbour.  The wave's program, wave-
// uniform, is a list of groups: the lines of one (x + dx, y + dy) column
// (one per source basis site, at most two) and the entries that read them,
// in the order the sum takes.  FP64 sums in the reference column order,
// which lines share except in rows whose z neighbours wrap round the
// periodic boundary; those rows are left to the row-list launch, as in
// STENCIL.
constexpr int kLnSlots = 2;          // lines of one group (basis sites)
constexpr int kLnBlock = 256;        // threads per workgroup: 4 lines
constexpr int kLnGroupWords = 8;     // program words per group
struct LnView {
  int lx, ly, lz, nb;
  int per[2];              // periodic x and y (z is periodic)
  int ty;                  // rows of y per band of the wave order (divides ly)
  // per sequence: (first group, number of groups).  A group is 8 words: a
  // bitmask of the entry codes slot | (dz + 2) << 1 present (one neighbour
  // each), the index of its first entry's values, two unused words and its
  // two lines b' | (dx + 2048) << 4 | (dy + 2048) << 16; its entries are
  // summed in ascending code order, the k-th reading values first + k
  const int2 *seq;
  const int32_t *prog;
  const double *val;       // per value index: 3 (diagonal) or 9 tensor values,
                           // and one spare index at the end
  // FP32 form, pair groups (word 0 bit 31): bit (dz + 2) * 3 + delta + 1 of
  // word 0 marks the pair of slot 0's neighbour at dz and slot 1's at
  // -dz + delta; pval holds (J, D) per pair, 4 doubles, from index word 1
  const double *pval;
  const int32_t *wseq;     // FP64: sequence of line (b, x, y) at (b lx + x) ly + y;
                           // null: sequence b (FP32 form)
  const uint64_t *fixmask; // FP64: per line, bit l marks lane l's cells as
                           // the row list's
  // FP64 row list, computed by the launch's first `rlblocks` workgroups
  // (gathers over the row patterns, one row per thread) while the others
  // run the lines
  const int32_t *rows;
  int64_t nrows;
  int rlblocks;
};
constexpr bool xv_uniform(int v) { return v == XV_UDIAG || v == XV_UFULL || xv_pair(v); }
constexpr bool xv_pattern(int v) { return (v >= XV_PDIAG && v <= XV_GFULL) || xv_uniform(v); }
constexpr bool xv_pattern_lds(int v) {
  return v == XV_PDIAG || v == XV_PFULL || xv_uniform(v);
}
constexpr bool xv_typed(int v) { return v == XV_TDIAG || v == XV_TFULL || xv_pattern(v); }
constexpr bool xv_diag(int v) {
  return v == XV_TDIAG || v == XV_BDIAG || v == XV_PDIAG || v == XV_GDIAG || v == XV_UDIAG ||
         v == XV_QISO;
}

// site-class table: CF_STRIDE doubles per class
enum ClassField {
  CF_MUS = 0, CF_GYRO = 1, CF_ALPHA = 2, CF_SIGMA = 3,
  CF_ZDC = 4,   // 4..6
  CF_ZAC = 7,   // 7..9
  CF_ZOMEGA = 10, CF_K = 11,
  CF_AXIS = 12,  // 12..14
  CF_SQRTT = 15,  // sqrt(temperature) of the spin's system (replica)
  CF_SEED = 16,   // its noise seed (the 64 bits stored in the double)
  CF_ROFF = 17,   // index of its first spin (int64 bits): the noise counter
                  // is index - ROFF
  CF_STRIDE = 18
};

struct XView {
  const int64_t *slice_off;   // nslices + 1 entry offsets (multiples of 64)
  const uint32_t *ent;        // packed entries
  uint32_t colmask;
  int colbits;
  int ntab;                   // table rows incl. the trailing zero tensor
  const double *tab64;        // TYPED: ntab x (3|9) tensors (FP64 in both forms)
  const double *val64;        // BLOCK: per entry (3|9) values, slice-major
  const float *val32;
  // PATTERN: pofs[i] = word offset of spin i's pattern in pat; pat[o] = length,
  // pat[o+1..o+len] = (offset << tbits) | tensor id, offset = column - row
  const uint32_t *pofs;
  const uint32_t *pat;
  int pat_words;              // table size in 32-bit words
  int tbits;
  // UNIFORM (FP32 form): rows are ordered by tensor id, so in a slice whose
  // rows all carry the same tensor sequence the tensor of slot k is the same
  // for every lane.  sseq[slice] = (offset of that sequence's tensors in
  // seqtab, length), or (-1, 0) when the slice's rows differ.
  const int2 *sseq;
  const double *seqtab;       // per sequence: len x (3|9) tensor values (FP64)
  const int32_t *upat;        // pat with each word replaced by its byte offset
                              // (column - row) * bytes per spin of the FP32 copy
  // pair programs (QISO / QDMI): spair[slice] = (first group, groups) or
  // (-1, 0); group g: pgrp[g] = slots (k1A, k2A, k1B, k2B), k1B < 0 for one
  // pair; pval[8 g ..] = (J, DA, DB, 0)
  const int2 *spair;
  const int4 *pgrp;
  const double *pval;
};

struct KParams {
  int64_t n;
  const double *s;     // current spins (phase 0 source, phase 1 s_old)
  double *sp;          // predictor spins
  // FP32 form without a precomputed field term: the predictor does not store
  // s* in FP64 and the corrector recomputes it from s and ds (bit for bit)
  int spfree;
  double *ds;          // ds/dt accumulator
  double *snew;        // phase 1 output (== s)
  double *sy;          // RK4: the next stage's input written by stages 0-2
  int stage;           // RK4 stage 0..3
  const float *g32;    // FP32 gather source, kG32 floats per spin (s32 / sp32)
  float *w32;          // FP32 copy written by the phase (sp32 / s32)
  const uint16_t *cls; // per-spin class or null
  const double *ctab;  // class table
  XView x;
  int nterms;
  int kind[4];
  const double *hext;  // precomputed field term (dipole), SoA, meV
  int has_ac;
  int ani_power;
  int noise;
  uint64_t step;          // the step counter, or with dstep its offset
  // thermostat of the only site class when cls is null (the CF_SQRTT,
  // CF_SEED and CF_ROFF entries of class 0)
  uint64_t seed;
  double sqrtT;
  uint64_t roff;
  const uint64_t *dstep;  // graph replay: the replay's first step, on the device
  const double *winj;  // injected noise SoA
  double *wbuf;        // noise of the current step: written by the predictor,
                       // read by the corrector (SoA)
  double dt;
  double t;
  int order_nb;         // basis-major period (1: reference order)
  int basis_ilv;        // gathering step kernels: workgroups alternate basis sites
  int ntstore;          // step kernels store their outputs non-temporally
  int64_t order_cells;  // N / order_nb
  double order_rcells;  // 1 / order_cells, rounded
  // row-list launch (the FP64 stencil's wrapped rows): thread j computes
  // device row rowlist[j] (pattern format) instead of row j
  const int32_t *rowlist;
  int64_t nrowlist;
};

// reference spin index of device index d (the noise counter uses it, so the
// stream does not depend on the device layout).  In the basis-major order
// d = b cells + cell and the reference index is cell nb + b.  b = d / cells
// comes from the double quotient d (1 / cells), truncated and corrected once:
// for d < 2^53 that quotient is within 2^-21 of d / cells (< nb < 2^31), so
// the truncation is the floor or one off it.  On the GPU this is a few vector
// operations against several tens for a 64-bit integer division.
__host__ __device__ inline int64_t ref_index(const KParams &p, int64_t d) {
  if (p.order_nb == 1) return d;
  int b = (int)((double)d * p.order_rcells);
  int64_t r = d - (int64_t)b * p.order_cells;
  if (r < 0) {
    b -= 1;
    r += p.order_cells;
  } else if (r >= p.order_cells) {
    b += 1;
    r -= p.order_cells;
  }
  return r * p.order_nb + b;
}

}  // namespace jh
