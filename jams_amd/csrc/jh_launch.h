// jh_launch.h -- host-callable launchers of the step kernels (jh_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "jh_internal.h"

namespace jh {

// bs: threads per workgroup of the step kernels (64 .. 512, multiple of 64)
hipError_t launch_heun(const KParams &p, int prec, int xv, int phase, int bs, hipStream_t st);
// one whole Heun step (predictor, grid barrier, corrector) in one persistent
// launch of the gathering kernels; cus: compute units of the device; sync:
// the barrier's 3 words (zero before the first launch); limit: polls before
// a workgroup gives the barrier up and sets sync[2]
hipError_t launch_heun_step(const KParams &p, int prec, int xv, int bs, int cus, unsigned *sync,
                            unsigned limit, hipStream_t st);
// one Heun phase over the STENCIL layout v, full or diagonal tensors
hipError_t launch_heun_st(const KParams &p, const StView &v, int prec, bool diag, int phase,
                          hipStream_t st);
// LDS bytes of a STENCIL workgroup of source basis site b
size_t st_tile_bytes(const StView &v, int prec, int b);
// one Heun phase over the LINE layout v, full or diagonal tensors
hipError_t launch_heun_ln(const KParams &p, const LnView &v, int prec, bool diag, int phase,
                          hipStream_t st);
// one RK4 stage (p.stage)
hipError_t launch_rk4(const KParams &p, int prec, int xv, int bs, hipStream_t st);
hipError_t launch_reduce_replicas(const KParams &p, const int64_t *off, int nrep, double *out,
                                  hipStream_t st);
hipError_t launch_set_step(uint64_t *step, uint64_t value, hipStream_t st);
hipError_t launch_field(const KParams &p, int prec, int xv, int term, double *out,
                        hipStream_t st);
hipError_t launch_noise(const KParams &p, double *out, hipStream_t st);
hipError_t launch_to_f32(const double *a, float *b, int64_t m, hipStream_t st);
hipError_t launch_reduce(const KParams &p, int mode, const double *h, double *partial,
                         double *out, hipStream_t st);

}  // namespace jh
