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

// ------------------------------------------------------------------ grouped GEMM (FP64)
// Tile 64x64, K step 16, 256 threads; each thread owns a 4x4 register block
// spread with stride 16 over the tile.  Operands staged through LDS with
// zero padding at the edges, so any M, N, K (including K = 0) is handled.
constexpr int TM = 64, TN = 64, TK = 16;

__global__ void __launch_bounds__(256) k_gemm_f64(const GemmGroup *__restrict__ groups,
                                                  const GemmTile *__restrict__ tiles, const double *__restrict__ A,
                                                  const double *__restrict__ B, double *__restrict__ C,
                                                  double *__restrict__ W) {
    const GemmTile t = tiles[blockIdx.x];
    const GemmGroup g = groups[t.g];
    const int m0 = t.mt * TM, n0 = t.nt * TN;
    const int kbeg = t.ks * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    __shared__ double As[TK][TM + 1];
    __shared__ double Bs[TK][TN + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        if (!g.transA) {
            for (int e = threadIdx.x; e < TM * TK; e += 256) {
                int i = e % TM, k = e / TM;
                int gm = m0 + i, gk = k0 + k;
                As[k][i] = (gm < g.M && gk < kend) ? A[g.a + gm + (int64_t)gk * g.lda] : 0.0;
            }
        } else {
            for (int e = threadIdx.x; e < TM * TK; e += 256) {
                int k = e % TK, i = e / TK;
                int gm = m0 + i, gk = k0 + k;
                As[k][i] = (gm < g.M && gk < kend) ? A[g.a + gk + (int64_t)gm * g.lda] : 0.0;
            }
        }
        if (!g.transB) {
            for (int e = threadIdx.x; e < TK * TN; e += 256) {
                int k = e % TK, j = e / TK;
                int gn = n0 + j, gk = k0 + k;
                Bs[k][j] = (gn < g.N && gk < kend) ? B[g.b + gk + (int64_t)gn * g.ldb] : 0.0;
            }
        } else {
            for (int e = threadIdx.x; e < TK * TN; e += 256) {
                int j = e % TN, k = e / TN;
                int gn = n0 + j, gk = k0 + k;
                Bs[k][j] = (gn < g.N && gk < kend) ? B[g.b + gn + (int64_t)gk * g.ldb] : 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TK; ++k) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[k][tx + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[k][ty + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    if (g.split == 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int gn = n0 + ty + 16 * j;
            if (gn >= g.N) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                int gm = m0 + tx + 16 * i;
                if (gm < g.M) C[g.c + gm + (int64_t)gn * g.ldc] = acc[i][j];
            }
        }
    } else {
        double *P = W + g.part + (int64_t)t.ks * g.M * g.N;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int gn = n0 + ty + 16 * j;
            if (gn >= g.N) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                int gm = m0 + tx + 16 * i;
                if (gm < g.M) P[gm + (int64_t)gn * g.M] = acc[i][j];
            }
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
