// Launch wrappers of the device kernels (mq_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "mq_plan.hpp"

namespace mq {

hipError_t launch_copy2d(const Copy2D *tasks, int n, const double *src, double *dst, hipStream_t s);
hipError_t launch_wapply(const Strip *strips, int n, const Src *srcs, const double *T, double *C, hipStream_t s);
hipError_t launch_gemm(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const int *split_gids, int nsplit,
                       const double *A, const double *B, double *C, double *W, hipStream_t s);
hipError_t launch_multidot(const double *const *xs, const double *const *ys, int npairs, int64_t n, int nblocks,
                           double *partial, hipStream_t s);
hipError_t launch_lincomb(const double *const *vs, const double *coef, int nv, int64_t n, double beta, double *out,
                          hipStream_t s);

}  // namespace mq
