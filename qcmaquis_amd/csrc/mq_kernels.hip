// HIP kernels of the hot path, gfx950 (CDNA4, wave64).
#include <hip/hip_runtime.h>

#include "mq_kernels.hpp"

namespace mq {

// ------------------------------------------------------------------ copy / scatter
// One workgroup per task; element e -> (i = e % rows, j = e / rows).
__global__ void __launch_bounds__(256) k_copy2d(const Copy2D *__restrict__ tasks, const double *__restrict__ src,
                                                double *__restrict__ dst) {
    const Copy2D t = tasks[blockIdx.x];
    const int n = t.rows * t.cols;
    if (t.trans) {
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            int i = e % t.rows, j = e / t.rows;
            dst[t.dst + i + (int64_t)j * t.ldd] = t.scale * src[t.src + j + (int64_t)i * t.lds];
        }
    } else {
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            int i = e % t.rows, j = e / t.rows;
            dst[t.dst + i + (int64_t)j * t.ldd] = t.scale * src[t.src + i + (int64_t)j * t.lds];
        }
    }
}

// ------------------------------------------------------------------ W application
// dst strip = sum_s coef_s * src_s, summed in task order (deterministic, no atomics).
__global__ void __launch_bounds__(256) k_wapply(const Strip *__restrict__ strips, const Src *__restrict__ srcs,
                                                const double *__restrict__ T, double *__restrict__ C) {
    const Strip st = strips[blockIdx.x];
    const int n = st.rows * st.cols;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        int i = e % st.rows, j = e / st.rows;
        double acc = 0.0;
        for (int s = st.s0; s < st.s1; ++s) {
            const Src sr = srcs[s];
            acc += sr.coef * T[sr.off + i + (int64_t)j * sr.ld];
        }
        C[st.dst + i + (int64_t)j * st.ldd] = acc;
    }
}

// ------------------------------------------------------------------ grouped GEMM (FP64, MFMA)
// Block tile 64x64, K step 16, 256 threads = 4 waves in a 2x2 grid, each
// wave owning a 32x32 sub-tile computed as 2x2 v_mfma_f64_16x16x4 products.
// The MFMA is fed with (B^T, A^T) so that its output lanes run along the rows
// of the column-major C, which makes the stores 128-byte segments.  Operand
// layout of v_mfma_f64_16x16x4: the first operand holds X[lane % 16][lane / 16],
// the second Y[lane / 16][lane % 16], and accumulator register r of a lane
// holds D[4 * r + lane / 16][lane % 16]; with X = B^T, Y = A^T that element is
// C(row lane % 16, column 4 * r + lane / 16).  Operands are staged through a
// double-buffered LDS tile with zero padding at the edges, so any M, N, K
// (including K = 0) is handled; the next K tile is fetched into registers
// while the current one is multiplied.
constexpr int TM = 64, TN = 64, TK = 16;
typedef double dbl4 __attribute__((ext_vector_type(4)));

struct GemmStage {
    double a[4], b[4];
};

__device__ __forceinline__ void gemm_fetch(const GemmGroup &g, const double *__restrict__ A,
                                           const double *__restrict__ B, int m0, int n0, int k0, int kend,
                                           GemmStage &st) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = threadIdx.x + 256 * r;
        int i, k;
        if (!g.transA) {
            i = e % TM;
            k = e / TM;
        } else {
            k = e % TK;
            i = e / TK;
        }
        const int gm = m0 + i, gk = k0 + k;
        double v = 0.0;
        if (gm < g.M && gk < kend) v = g.transA ? A[g.a + gk + (int64_t)gm * g.lda] : A[g.a + gm + (int64_t)gk * g.lda];
        st.a[r] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = threadIdx.x + 256 * r;
        int j, k;
        if (!g.transB) {
            k = e % TK;
            j = e / TK;
        } else {
            j = e % TN;
            k = e / TN;
        }
        const int gn = n0 + j, gk = k0 + k;
        double v = 0.0;
        if (gn < g.N && gk < kend) v = g.transB ? B[g.b + gn + (int64_t)gk * g.ldb] : B[g.b + gk + (int64_t)gn * g.ldb];
        st.b[r] = v;
    }
}

__device__ __forceinline__ void gemm_stash(const GemmGroup &g, const GemmStage &st, double (*As)[TM + 2],
                                           double (*Bs)[TN + 2]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = threadIdx.x + 256 * r;
        if (!g.transA)
            As[e / TM][e % TM] = st.a[r];
        else
            As[e % TK][e / TK] = st.a[r];
        if (!g.transB)
            Bs[e % TK][e / TK] = st.b[r];
        else
            Bs[e / TN][e % TN] = st.b[r];
    }
}

__global__ void __launch_bounds__(256) k_gemm_f64(const GemmGroup *__restrict__ groups,
                                                  const GemmTile *__restrict__ tiles, const double *__restrict__ A,
                                                  const double *__restrict__ B, double *__restrict__ C,
                                                  double *__restrict__ W) {
    const GemmTile t = tiles[blockIdx.x];
    const GemmGroup g = groups[t.g];
    const int m0 = t.mt * TM, n0 = t.nt * TN;
    const int kbeg = t.ks * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    __shared__ double As[2][TK][TM + 2];
    __shared__ double Bs[2][TK][TN + 2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    dbl4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
    GemmStage st;
    int buf = 0;
    if (kbeg < kend) {
        gemm_fetch(g, A, B, m0, n0, kbeg, kend, st);
        gemm_stash(g, st, As[0], Bs[0]);
    }
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        const bool more = k0 + TK < kend;
        if (more) gemm_fetch(g, A, B, m0, n0, k0 + TK, kend, st);
#pragma unroll
        for (int kk = 0; kk < TK; kk += 4) {
            double a0 = As[buf][kk + lk][wm + lr], a1 = As[buf][kk + lk][wm + 16 + lr];
            double b0 = Bs[buf][kk + lk][wn + lr], b1 = Bs[buf][kk + lk][wn + 16 + lr];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[1][1], 0, 0, 0);
        }
        if (more) gemm_stash(g, st, As[buf ^ 1], Bs[buf ^ 1]);
        __syncthreads();
        buf ^= 1;
    }
    double *out;
    int64_t ld;
    if (g.split == 1) {
        out = C + g.c;
        ld = g.ldc;
    } else {
        out = W + g.part + (int64_t)t.ks * g.M * g.N;
        ld = g.M;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int gm = m0 + wm + 16 * i + lr;
        if (gm >= g.M) continue;
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gn = n0 + wn + 16 * j + 4 * r + lk;
                if (gn < g.N) out[gm + gn * ld] = acc[i][j][r];
            }
    }
}

// C = sum over splits of the partial tiles, in split order (deterministic).
__global__ void __launch_bounds__(256) k_split_reduce(const GemmGroup *__restrict__ groups,
                                                      const int *__restrict__ gids, const double *__restrict__ W,
                                                      double *__restrict__ C) {
    const GemmGroup g = groups[gids[blockIdx.y]];
    const int64_t n = (int64_t)g.M * g.N;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int s = 0; s < g.split; ++s) acc += W[g.part + s * n + e];
        int i = (int)(e % g.M), j = (int)(e / g.M);
        C[g.c + i + (int64_t)j * g.ldc] = acc;
    }
}

// ------------------------------------------------------------------ vector kernels (Davidson)
// Partial dot products of K vector pairs; block b writes its partial for pair p.
__global__ void __launch_bounds__(256) k_multidot(const double *const *__restrict__ xs,
                                                  const double *const *__restrict__ ys, int64_t n,
                                                  double *__restrict__ partial) {
    const int p = blockIdx.y;
    const double *x = xs[p], *y = ys[p];
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        acc = fma(x[e], y[e], acc);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[p * gridDim.x + blockIdx.x] = red[0];
}

// out = sum_j c_j * v_j  (+ beta * out if beta != 0)
__global__ void __launch_bounds__(256) k_lincomb(const double *const *__restrict__ vs,
                                                 const double *__restrict__ coef, int nv, int64_t n, double beta,
                                                 double *__restrict__ out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        double acc = (beta != 0.0) ? beta * out[e] : 0.0;
        for (int j = 0; j < nv; ++j) acc = fma(coef[j], vs[j][e], acc);
        out[e] = acc;
    }
}

// ------------------------------------------------------------------ launchers
hipError_t launch_copy2d(const Copy2D *tasks, int n, const double *src, double *dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy2d, dim3(n), dim3(256), 0, s, tasks, src, dst);
    return hipGetLastError();
}
hipError_t launch_wapply(const Strip *strips, int n, const Src *srcs, const double *T, double *C, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_wapply, dim3(n), dim3(256), 0, s, strips, srcs, T, C);
    return hipGetLastError();
}
hipError_t launch_gemm(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const int *split_gids,
                       int nsplit, const double *A, const double *B, double *C, double *W, hipStream_t s) {
    if (ntiles > 0) hipLaunchKernelGGL(k_gemm_f64, dim3(ntiles), dim3(256), 0, s, groups, tiles, A, B, C, W);
    if (nsplit > 0) hipLaunchKernelGGL(k_split_reduce, dim3(64, nsplit), dim3(256), 0, s, groups, split_gids, W, C);
    return hipGetLastError();
}
hipError_t launch_multidot(const double *const *xs, const double *const *ys, int npairs, int64_t n, int nblocks,
                           double *partial, hipStream_t s) {
    hipLaunchKernelGGL(k_multidot, dim3(nblocks, npairs), dim3(256), 0, s, xs, ys, n, partial);
    return hipGetLastError();
}
hipError_t launch_lincomb(const double *const *vs, const double *coef, int nv, int64_t n, double beta, double *out,
                          hipStream_t s) {
    int nb = (int)std::min<int64_t>((n + 255) / 256, 4096);
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_lincomb, dim3(nb), dim3(256), 0, s, vs, coef, nv, n, beta, out);
    return hipGetLastError();
}

}  // namespace mq
