// Launch wrappers of the device kernels (mq_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "mq_plan.hpp"

namespace mq {

// Pipeline stage of a launch: each stage gets its own kernel instantiation so
// that profiles attribute time per stage.
enum Stage { ST_SIGMA = 0, ST_SIGMA_OUT = 1, ST_STEP = 2, ST_STEP_OUT = 3, ST_FUSE = 4, ST_MISC = 5 };

// sigma: the instantiation of the sigma's psi reshape (its own kernel name in profiles)
hipError_t launch_copy2d(const Copy2D *tasks, int n, const double *src, double *dst, hipStream_t s,
                         bool sigma = false);
hipError_t launch_copy2d(const Copy2D *tasks, int n, const double *src, float *dst, hipStream_t s,
                         bool sigma = false);
hipError_t launch_to_f32(const double *src, float *dst, int64_t n, hipStream_t s);
// one workgroup per pack of strips (StripList::packs, npacks = packs.size() - 1)
// nvec > 1: the same task list for nvec vectors, vector v at T + v * src_stride and
// C + v * dst_stride
hipError_t launch_wapply(const Strip *strips, const int *packs, int npacks, const Src *srcs, const double *T,
                         double *C, int stage, hipStream_t s, int nvec = 1, int64_t src_stride = 0,
                         int64_t dst_stride = 0);
hipError_t launch_wapply(const Strip *strips, const int *packs, int npacks, const Src *srcs, const float *T, float *C,
                         int stage, hipStream_t s, int nvec = 1, int64_t src_stride = 0, int64_t dst_stride = 0);
// grouped GEMM over a plan's tile list (FP64; FP32 operands with FP32 or FP64 output,
// multiplied in FP32 or, with bf16x3, as three bf16 products accumulated in FP32);
// rowmap: the row placements of groups with rowmap > 0 (may be null when none has)
hipError_t launch_gemm(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const int *split_gids, int nsplit,
                       const double *A, const double *B, double *C, double *W, const RowMap *rowmap, int stage,
                       hipStream_t s);
hipError_t launch_gemm(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const int *split_gids, int nsplit,
                       const float *A, const float *B, float *C, float *W, const RowMap *rowmap, int stage,
                       hipStream_t s, bool bf16x3 = false);
hipError_t launch_gemm(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const int *split_gids, int nsplit,
                       const float *A, const float *B, double *C, float *W, const RowMap *rowmap, int stage,
                       hipStream_t s, bool bf16x3 = false);
hipError_t launch_multidot(const double *const *xs, const double *const *ys, int npairs, int64_t n, int nblocks,
                           double *partial, hipStream_t s);
// out = (x - f u) / (theta - D), elementwise, leaving the elements where theta == D
// undivided (MultDiagonal::mult_diag); out may alias x
hipError_t launch_diag_scale(const double *x, const double *u, double f, const double *D, double theta, double *out,
                             int64_t n, hipStream_t s);
hipError_t launch_lincomb(const double *const *vs, const double *coef, int nv, int64_t n, double beta, double *out,
                          hipStream_t s);
// two-site split: gather psi into the SVD working matrices (X must be zeroed
// first) and the assembly of the two site tensors
hipError_t launch_split_gather(const SplitSlab *slabs, int n, const double *psi, double *X, hipStream_t s);
// Jacobi SVD of every block: init (V = I, flags), one launch per round of a sweep
// (one wave per column pair of every block), the end-of-sweep flag update, and
// the sorted singular values with their permutation
hipError_t launch_jacobi_init(const SvdBlk *blks, int nblk, double *V, int32_t *done, int32_t *rot, int32_t *sweeps,
                              hipStream_t s);
hipError_t launch_jacobi_round(const SvdBlk *blks, const int32_t *slot_blk, const int32_t *slot_k, int nslots, int t,
                               double *X, double *V, const int32_t *done, int32_t *rot, hipStream_t s);
hipError_t launch_jacobi_sweep_end(int nblk, int sweep, int32_t *done, int32_t *rot, int32_t *sweeps,
                                   hipStream_t s);
hipError_t launch_jacobi_finish(const SvdBlk *blks, int nblk, int max_c, const double *X, double *S, int32_t *perm,
                                const int32_t *done, int32_t *sweeps, hipStream_t s);
hipError_t launch_split_assemble(const SvdBlk *blks, int nblk, int max_keep, const int32_t *keep,
                                 const int64_t *a_off, const int64_t *b_off, int l2r, const double *X,
                                 const double *V, const double *S, const int32_t *perm, double *A, double *B,
                                 hipStream_t s);

}  // namespace mq
