// HIP kernels of the hot path, gfx950 (CDNA4, wave64).
//
// Every kernel is instantiated for FP64 data (the plain form) and for FP32
// data (the mixed form, whose output sigma is still accumulated into FP64).
// The STAGE template parameter only names the instantiation, one per
// pipeline stage, so that profiles attribute time per stage.
#include <hip/hip_runtime.h>

#include "mq_kernels.hpp"

namespace mq {

// ------------------------------------------------------------------ copy / scatter
// One workgroup per task; element e -> (i = e % rows, j = e / rows).  The
// destination may be of a narrower type (FP64 -> FP32 conversion).
template <int STAGE, class TS, class TD>
__global__ void __launch_bounds__(256) k_copy2d(const Copy2D *__restrict__ tasks, const TS *__restrict__ src,
                                                TD *__restrict__ dst) {
    const Copy2D t = tasks[blockIdx.x];
    const int n = t.rows * t.cols;
    if (t.trans) {
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            int i = e % t.rows, j = e / t.rows;
            dst[t.dst + i + (int64_t)j * t.ldd] = (TD)(t.scale * src[t.src + j + (int64_t)i * t.lds]);
        }
    } else {
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            int i = e % t.rows, j = e / t.rows;
            dst[t.dst + i + (int64_t)j * t.ldd] = (TD)(t.scale * src[t.src + i + (int64_t)j * t.lds]);
        }
    }
}

// flat FP64 -> FP32 conversion of an arena
__global__ void __launch_bounds__(256) k_to_f32(const double *__restrict__ src, float *__restrict__ dst, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = (float)src[e];
}

// ------------------------------------------------------------------ W application
// dst strip = sum_s coef_s * src_s, summed in task order (deterministic, no
// atomics).  Each thread owns WU elements of a pass, so every source issues
// WU independent loads before they are accumulated; element e of the strip
// is (i, j) = (e mod rows, e / rows), the quotient taken from a float
// reciprocal and corrected (exact for strips below 2^24 elements).
constexpr int WU = 4;

// One workgroup per pack of consecutive strips (see StripList): a pack of one
// large strip runs the strip with all threads; a pack of several small strips
// runs their elements as one index range, each element finding its strip with a
// binary search over the pack's element offsets in LDS.
// A large strip gives each thread WS = 8 elements per pass, so that a wave has
// 2 KB (FP32) or 4 KB (FP64) of loads in flight per source: the W application is
// bound by the bytes in flight per wave more than by instructions (measured at
// the L=24 centre: FP32 1.06 -> 1.00 ms from 4 to 8 elements, FP64 1.77 -> 1.72 ms;
// 16 FP32 elements gained nothing more).
template <class T>
constexpr int strip_units() {
    return 8;
}
template <class T>
__device__ __forceinline__ void wapply_strip(const Strip &st, const Src *__restrict__ srcs, const T *__restrict__ src,
                                             T *__restrict__ C) {
    constexpr int WS = strip_units<T>();
    const int rows = st.rows, n = st.rows * st.cols;
    const float inv = 1.0f / (float)max(rows, 1);
    for (int base = 0; base < n; base += 256 * WS) {
        int ii[WS], jj[WS];
        T acc[WS];
#pragma unroll
        for (int u = 0; u < WS; ++u) {
            const int e = base + threadIdx.x + 256 * u;
            int j = (int)((float)e * inv);
            if (j * rows > e) --j;
            else if ((j + 1) * rows <= e) ++j;
            ii[u] = e - j * rows;
            jj[u] = e < n ? j : -1;
            acc[u] = 0;
        }
        for (int s = st.s0; s < st.s1; ++s) {
            const Src sr = srcs[s];
            const T c = (T)sr.coef;
            T v[WS];
#pragma unroll
            for (int u = 0; u < WS; ++u) v[u] = jj[u] >= 0 ? src[sr.off + ii[u] + (int64_t)jj[u] * sr.ld] : (T)0;
#pragma unroll
            for (int u = 0; u < WS; ++u) acc[u] += c * v[u];
        }
#pragma unroll
        for (int u = 0; u < WS; ++u)
            if (jj[u] >= 0) C[st.dst + ii[u] + (int64_t)jj[u] * st.ldd] = acc[u];
    }
}

static_assert(kPackElems <= 256 * WU, "a pack of small strips is run in one pass");

// nvec vectors share one task list: vector v reads src + v * sstride and writes
// C + v * dstride (the batched sigma), so the strip and source records are read
// once per workgroup for all of them.
// (MULTI false: one vector, compiled without the vector loop.)
template <int STAGE, class T, bool MULTI>
__global__ void __launch_bounds__(256) k_wapply(const Strip *__restrict__ strips, const int *__restrict__ packs,
                                                const Src *__restrict__ srcs, const T *__restrict__ src,
                                                T *__restrict__ C, int nvec_arg, int64_t sstride, int64_t dstride) {
    const int nvec = MULTI ? nvec_arg : 1;
    const int p0 = packs[blockIdx.x], ns = packs[blockIdx.x + 1] - p0;
    if (ns == 1) {
        const Strip st = strips[p0];
        for (int v = 0; v < nvec; ++v) wapply_strip<T>(st, srcs, src + v * sstride, C + v * dstride);
        return;
    }
    __shared__ Strip sst[kPackStrips];
    __shared__ int pre[kPackStrips + 1];
    if ((int)threadIdx.x < ns) sst[threadIdx.x] = strips[p0 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < ns; ++k) {
            pre[k] = acc;
            acc += sst[k].rows * sst[k].cols;
        }
        pre[ns] = acc;
    }
    __syncthreads();
    const int total = pre[ns];  // at most kPackElems = 256 * WU: one pass
    int s0[WU], s1[WU], ii[WU], jj[WU];
    int64_t dst[WU];
    int nsrc = 0;
#pragma unroll
    for (int u = 0; u < WU; ++u) {
        const int e = threadIdx.x + 256 * u;
        s0[u] = s1[u] = 0;
        ii[u] = jj[u] = 0;
        dst[u] = -1;
        if (e < total) {
            int lo = 0, hi = ns - 1;  // last k with pre[k] <= e
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pre[mid] <= e) lo = mid;
                else hi = mid - 1;
            }
            const Strip &st = sst[lo];
            const int el = e - pre[lo];
            jj[u] = el / st.rows;
            ii[u] = el - jj[u] * st.rows;
            s0[u] = st.s0;
            s1[u] = st.s1;
            dst[u] = st.dst + ii[u] + (int64_t)jj[u] * st.ldd;
            nsrc = max(nsrc, st.s1 - st.s0);
        }
    }
    // the elements' sources in lock step, so their loads are issued together
    for (int vec = 0; vec < nvec; ++vec) {
        const T *sv = src + vec * sstride;
        T *Cv = C + vec * dstride;
        T acc[WU];
#pragma unroll
        for (int u = 0; u < WU; ++u) acc[u] = 0;
        for (int q = 0; q < nsrc; ++q) {
            T v[WU], c[WU];
#pragma unroll
            for (int u = 0; u < WU; ++u) {
                v[u] = 0;
                c[u] = 0;
                if (s0[u] + q < s1[u]) {
                    const Src sr = srcs[s0[u] + q];
                    c[u] = (T)sr.coef;
                    v[u] = sv[sr.off + ii[u] + (int64_t)jj[u] * sr.ld];
                }
            }
#pragma unroll
            for (int u = 0; u < WU; ++u) acc[u] += c[u] * v[u];
        }
#pragma unroll
        for (int u = 0; u < WU; ++u)
            if (dst[u] >= 0) Cv[dst[u]] = acc[u];
    }
}

// ------------------------------------------------------------------ grouped GEMM (MFMA)
// Block tile TM x TN (TM 64 or 128, TN a multiple of 16 up to 128, chosen per
// group to limit padding), K step 16, 256 threads = 4 waves stacked along M,
// each wave owning a TM/4 x TN sub-tile computed with 16x16x4 MFMA products (v_mfma_f64_16x16x4 for FP64,
// v_mfma_f32_16x16x4 for FP32 operands).  The MFMA is fed with (B^T, A^T) so
// that its output lanes run along the rows of the column-major C, which makes
// the stores 128-byte segments.  Operand layout of both instructions: the
// first operand holds X[lane % 16][lane / 16], the second Y[lane / 16][lane % 16].
// Accumulator register r of a lane holds D[4 * r + lane / 16][lane % 16] for
// the FP64 instruction and D[4 * (lane / 16) + r][lane % 16] for the FP32 one;
// with X = B^T, Y = A^T, D's row index is C's column and its column index
// C's row.  Operands are staged through a double-buffered LDS tile with zero
// padding at the edges, so any M, N, K (including K = 0) is handled; the next
// K tile is fetched into registers while the current one is multiplied.
constexpr int TMAX = 128;
// K step of a tile: 16 for FP64 operands, KSTEP_F32 for FP32 ones (the LDS tile
// holds the same bytes either way).
constexpr int KSTEP_F32 = 16;
template <class T>
constexpr int kstep() {
    return sizeof(T) == 4 ? KSTEP_F32 : 16;
}
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <class T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef dbl4 acc;
    static __device__ __forceinline__ acc run(double x, double y, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int dcol(int r, int lq) { return 4 * r + lq; }
};
template <>
struct Mfma<float> {
    typedef flt4 acc;
    static __device__ __forceinline__ acc run(float x, float y, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int dcol(int r, int lq) { return 4 * lq + r; }
};

// BF16x3 form of an FP32 product: each operand element x is split into bf16
// parts hi = bf16(x) and lo = bf16(x - hi) (round to nearest even) when it is
// staged into LDS, and C += A_lo B_hi + A_hi B_lo + A_hi B_hi is accumulated in
// FP32 with v_mfma_f32_16x16x16_bf16 (operand lanes: X[lane % 16][4 * (lane / 16) + q]
// and Y[4 * (lane / 16) + q][lane % 16] for q = 0..3; accumulators as the FP32
// instruction).  The dropped A_lo B_lo term and the split leave a relative error
// of about 2^-16 per product.  LDS holds, per row of A (or column of B) and
// group of four k, the four hi parts followed by the four lo parts (16 bytes),
// rows padded to five such groups.
typedef short sh4 __attribute__((ext_vector_type(4)));
struct Bf16Quad {
    sh4 hi, lo;
};
constexpr int X3_QUADS = 5;  // 4 groups of k per K step of 16, plus one of padding
__device__ __forceinline__ unsigned short bf16_rne(float x) {
    const unsigned u = __float_as_uint(x);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ void bf16_split(float x, short &hi, short &lo) {
    const unsigned short h = bf16_rne(x);
    hi = (short)h;
    lo = (short)bf16_rne(x - __uint_as_float((unsigned)h << 16));
}

template <int TM, int TN, class T>
struct GemmStage {
    T a[TM * kstep<T>() / 256], b[TN * kstep<T>() / 256];
};

// Row and k of A-tile slot r of this thread, column-major A (transA false).
// QUADS: the slots of a thread form whole groups of four consecutive k of one
// row (group kq + (256 / TM) * (r / 4), kq = thread / TM), so the BF16x3 stash
// writes one 16-byte LDS quad per group; the lanes of one load still read
// consecutive rows of one column.  Otherwise k = (thread + 256 r) / TM.
template <int TM, bool QUADS>
__device__ __forceinline__ void a_slot(int r, int &i, int &k) {
    i = threadIdx.x % TM;
    if (QUADS)
        k = 4 * ((int)threadIdx.x / TM + (256 / TM) * (r >> 2)) + (r & 3);
    else
        k = ((int)threadIdx.x + 256 * r) / TM;
}

template <int TM, int TN, class T, bool QUADS = false>
__device__ __forceinline__ void gemm_fetch(const GemmGroup &g, const T *__restrict__ A, const T *__restrict__ B,
                                           int m0, int n0, int k0, int kend, GemmStage<TM, TN, T> &st) {
    constexpr int TK = kstep<T>();
#pragma unroll
    for (int r = 0; r < TM * TK / 256; ++r) {
        const int e = threadIdx.x + 256 * r;
        int i, k;
        if (!g.transA) {
            a_slot<TM, QUADS>(r, i, k);
        } else {
            k = e % TK;
            i = e / TK;
        }
        const int gm = m0 + i, gk = k0 + k;
        T v = 0;
        if (gm < g.M && gk < kend) v = g.transA ? A[g.a + gk + (int64_t)gm * g.lda] : A[g.a + gm + (int64_t)gk * g.lda];
        st.a[r] = v;
    }
#pragma unroll
    for (int r = 0; r < TN * TK / 256; ++r) {
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
        T v = 0;
        if (gn < g.N && gk < kend) v = g.transB ? B[g.b + gn + (int64_t)gk * g.ldb] : B[g.b + gk + (int64_t)gn * g.ldb];
        st.b[r] = v;
    }
}

template <int TM, int TN, class T>
__device__ __forceinline__ void gemm_stash(const GemmGroup &g, const GemmStage<TM, TN, T> &st, T (*As)[TM + 2],
                                           T (*Bs)[TN + 2]) {
    constexpr int TK = kstep<T>();
#pragma unroll
    for (int r = 0; r < TM * TK / 256; ++r) {
        const int e = threadIdx.x + 256 * r;
        if (!g.transA)
            As[e / TM][e % TM] = st.a[r];
        else
            As[e % TK][e / TK] = st.a[r];
    }
#pragma unroll
    for (int r = 0; r < TN * TK / 256; ++r) {
        const int e = threadIdx.x + 256 * r;
        if (!g.transB)
            Bs[e % TK][e / TK] = st.b[r];
        else
            Bs[e / TN][e % TN] = st.b[r];
    }
}

// Split the staged FP32 tile into the bf16 hi/lo groups of the BF16x3 form.
template <int TM, int TN>
__device__ __forceinline__ void gemm_stash_x3(const GemmGroup &g, const GemmStage<TM, TN, float> &st,
                                              Bf16Quad (*Ax)[X3_QUADS], Bf16Quad (*Bx)[X3_QUADS]) {
    constexpr int TK = kstep<float>();
    if (!g.transA) {
        // whole quads (the QUADS slots of gemm_fetch): one 16-byte write each
#pragma unroll
        for (int q = 0; q < TM * TK / 1024; ++q) {
            int i, k;
            a_slot<TM, true>(4 * q, i, k);
            Bf16Quad v;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                short h, l;
                bf16_split(st.a[4 * q + u], h, l);
                v.hi[u] = h;
                v.lo[u] = l;
            }
            Ax[i][k >> 2] = v;
        }
    } else {
#pragma unroll
        for (int r = 0; r < TM * TK / 256; ++r) {
            const int e = threadIdx.x + 256 * r;
            const int i = e / TK, k = e % TK;
            short h, l;
            bf16_split(st.a[r], h, l);
            Ax[i][k >> 2].hi[k & 3] = h;
            Ax[i][k >> 2].lo[k & 3] = l;
        }
    }
#pragma unroll
    for (int r = 0; r < TN * TK / 256; ++r) {
        const int e = threadIdx.x + 256 * r;
        const int j = !g.transB ? e / TK : e % TN, k = !g.transB ? e % TK : e / TN;
        short h, l;
        bf16_split(st.b[r], h, l);
        Bx[j][k >> 2].hi[k & 3] = h;
        Bx[j][k >> 2].lo[k & 3] = l;
    }
}

// Row-mapped output tiles (the T panels, each (b1, lb) row block stored
// contiguously) are written through LDS: a column block of the tile at a time is
// put into LDS column-major, then each wave stores 64 consecutive tile rows of one
// column per instruction -- runs of a row block that are contiguous in memory --
// instead of the accumulator layout's 16-row pieces of four columns.  At the L=24
// m=512 centre this took the BF16x3 T GEMM from 1.14 to 1.03 ms.
// FP32 outputs only (the FP64 kernel would lose its second wave per SIMD to the
// staging registers).  LDS leading dimension TM + 4 puts the four accumulator
// column groups of a store on distinct banks; 64 columns per LDS pass.
constexpr int epi_ld(int tm) { return tm + 4; }
constexpr int EPI_COLS = 64;
constexpr int EPI_LDS_BYTES = EPI_COLS * (TMAX + 4) * 4;
template <int TM, int TN, class T, class TO>
__device__ __forceinline__ void store_rowmap_staged(const GemmGroup &g, int m0, int n0, int wm, int lr, int lq,
                                                    const typename Mfma<T>::acc (&acc)[TM / 64][TN / 16],
                                                    TO *__restrict__ C, const RowMap *__restrict__ rowmap,
                                                    double *lds) {
    typedef Mfma<T> MF;
    constexpr int MI = TM / 64, NJ = TN / 16;
    static_assert(sizeof(TO) == 4, "staged stores are for FP32 outputs");
    constexpr int LD = epi_ld(TM), CH = EPI_COLS, CSTEP = 256 / TM;
    TO *st = reinterpret_cast<TO *>(lds);
    const int row = threadIdx.x % TM, cfirst = threadIdx.x / TM;
    const int gm = m0 + row;
    RowMap rm{0, 0, 0};
    if (gm < g.M) rm = rowmap[g.rowmap - 1 + gm];
    TO *out = C + g.c + rm.off;
#pragma unroll
    for (int c0 = 0; c0 < TN; c0 += CH) {
        __syncthreads();  // LDS free: operand tiles, or the previous column block, consumed
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            if (16 * j < c0 || 16 * j >= c0 + CH) continue;
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) st[(16 * j + MF::dcol(r, lq) - c0) * LD + wm + 16 * i + lr] = acc[i][j][r];
        }
        __syncthreads();
        const int ncol = TN - c0 < CH ? TN - c0 : CH;
#pragma unroll
        for (int k = 0; k < ncol / CSTEP; ++k) {
            const int c = cfirst + k * CSTEP;
            const int gn = n0 + c0 + c;
            if (gm < g.M && gn < g.N) out[(int64_t)gn * rm.ld] = st[c * LD + row];
        }
    }
}

// T: operand type; TO: output type; partial tiles of split-K groups are
// stored in T and summed by k_split_reduce.  X3 (T = float only): BF16x3 products.
// STAGED: row-mapped groups store through LDS (store_rowmap_staged).
template <int TM, int TN, class T, class TO, bool X3, bool STAGED>
__device__ __forceinline__ void gemm_tile(const GemmGroup &g, const GemmTile &t, const T *__restrict__ A,
                                          const T *__restrict__ B, TO *__restrict__ C, T *__restrict__ W,
                                          const RowMap *__restrict__ rowmap, double *lds) {
    static_assert(!X3 || sizeof(T) == 4, "BF16x3 products take FP32 operands");
    typedef Mfma<T> MF;
    constexpr int TK = kstep<T>();
    // the four waves split the tile along M; each holds TM/4 rows x TN columns
    constexpr int MI = TM / 64, NJ = TN / 16;  // 16x16 MFMA blocks per wave
    T(*As)[TK][TM + 2] = reinterpret_cast<T(*)[TK][TM + 2]>(lds);
    T(*Bs)[TK][TN + 2] = reinterpret_cast<T(*)[TK][TN + 2]>(reinterpret_cast<T *>(lds) + 2 * TK * (TM + 2));
    Bf16Quad(*Ax)[TM][X3_QUADS] = reinterpret_cast<Bf16Quad(*)[TM][X3_QUADS]>(lds);
    Bf16Quad(*Bx)[TN][X3_QUADS] = reinterpret_cast<Bf16Quad(*)[TN][X3_QUADS]>(
        reinterpret_cast<Bf16Quad *>(lds) + 2 * TM * X3_QUADS);
    const int m0 = t.mt * TM, n0 = t.nt * TN;
    const int kbeg = t.ks * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave * (TM / 4);
    const int lr = lane & 15, lq = lane >> 4;
    typename MF::acc acc[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = typename MF::acc{0, 0, 0, 0};
    GemmStage<TM, TN, T> st;
    int buf = 0;
    if (kbeg < kend) {
        gemm_fetch<TM, TN, T, X3>(g, A, B, m0, n0, kbeg, kend, st);
        if constexpr (X3)
            gemm_stash_x3<TM, TN>(g, st, Ax[0], Bx[0]);
        else
            gemm_stash<TM, TN, T>(g, st, As[0], Bs[0]);
    }
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        const bool more = k0 + TK < kend;
        if (more) gemm_fetch<TM, TN, T, X3>(g, A, B, m0, n0, k0 + TK, kend, st);
        if constexpr (X3) {
            // one 16x16x16 step covers the K tile: lane group lq holds k = 4 lq .. 4 lq + 3
            Bf16Quad av[MI];
#pragma unroll
            for (int i = 0; i < MI; ++i) av[i] = Ax[buf][wm + 16 * i + lr][lq];
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const Bf16Quad bj = Bx[buf][16 * j + lr][lq];
#pragma unroll
                for (int i = 0; i < MI; ++i) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bj.hi, av[i].lo, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bj.lo, av[i].hi, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bj.hi, av[i].hi, acc[i][j], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < TK; kk += 4) {
                T av[MI];
#pragma unroll
                for (int i = 0; i < MI; ++i) av[i] = As[buf][kk + lq][wm + 16 * i + lr];
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    T bj = Bs[buf][kk + lq][16 * j + lr];
#pragma unroll
                    for (int i = 0; i < MI; ++i) acc[i][j] = MF::run(bj, av[i], acc[i][j]);
                }
            }
        }
        if (more) {
            if constexpr (X3)
                gemm_stash_x3<TM, TN>(g, st, Ax[buf ^ 1], Bx[buf ^ 1]);
            else
                gemm_stash<TM, TN, T>(g, st, As[buf ^ 1], Bs[buf ^ 1]);
        }
        __syncthreads();
        buf ^= 1;
    }
    if constexpr (STAGED) {
        if (g.split == 1 && g.rowmap > 0) {  // uniform over the workgroup
            store_rowmap_staged<TM, TN, T, TO>(g, m0, n0, wm, lr, lq, acc, C, rowmap, lds);
            return;
        }
    }
    if (g.split == 1) {
#pragma unroll
        for (int i = 0; i < MI; ++i) {
            const int gm = m0 + wm + 16 * i + lr;
            if (gm >= g.M) continue;
            // row gm of C: column-major with ldc, or placed by the group's row map
            TO *out = C + g.c + gm;
            int64_t ld = g.ldc;
            if (g.rowmap > 0) {
                const RowMap rm = rowmap[g.rowmap - 1 + gm];
                out = C + g.c + rm.off;
                ld = rm.ld;
            }
#pragma unroll
            for (int j = 0; j < NJ; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int gn = n0 + 16 * j + MF::dcol(r, lq);
                    if (gn < g.N) out[gn * ld] = (TO)acc[i][j][r];
                }
        }
    } else {
        T *out = W + g.part + (int64_t)t.ks * g.M * g.N;
        const int64_t ld = g.M;
#pragma unroll
        for (int i = 0; i < MI; ++i) {
            const int gm = m0 + wm + 16 * i + lr;
            if (gm >= g.M) continue;
#pragma unroll
            for (int j = 0; j < NJ; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int gn = n0 + 16 * j + MF::dcol(r, lq);
                    if (gn < g.N) out[gm + gn * ld] = acc[i][j][r];
                }
        }
    }
}

// Column tile widths are multiples of 16 up to 128.
template <int TM, class T, class TO, bool X3, bool STAGED>
__device__ __forceinline__ void gemm_tile_n(const GemmGroup &g, const GemmTile &t, const T *__restrict__ A,
                                            const T *__restrict__ B, TO *__restrict__ C, T *__restrict__ W,
                                            const RowMap *__restrict__ rowmap, double *lds) {
    switch (g.tn) {
        case 16: gemm_tile<TM, 16, T, TO, X3, STAGED>(g, t, A, B, C, W, rowmap, lds); break;
        case 32: gemm_tile<TM, 32, T, TO, X3, STAGED>(g, t, A, B, C, W, rowmap, lds); break;
        case 48: gemm_tile<TM, 48, T, TO, X3, STAGED>(g, t, A, B, C, W, rowmap, lds); break;
        case 64: gemm_tile<TM, 64, T, TO, X3, STAGED>(g, t, A, B, C, W, rowmap, lds); break;
        case 80: gemm_tile<TM, 80, T, TO, X3, STAGED>(g, t, A, B, C, W, rowmap, lds); break;
        case 96: gemm_tile<TM, 96, T, TO, X3, STAGED>(g, t, A, B, C, W, rowmap, lds); break;
        case 112: gemm_tile<TM, 112, T, TO, X3, STAGED>(g, t, A, B, C, W, rowmap, lds); break;
        default: gemm_tile<TM, 128, T, TO, X3, STAGED>(g, t, A, B, C, W, rowmap, lds); break;
    }
}

// LDS bytes of a tile pair (two buffers): FP64/FP32 K-major tiles, or the
// bf16 hi/lo groups of the BF16x3 form.
template <class T, bool X3>
constexpr int gemm_lds_bytes() {
    return X3 ? 2 * 2 * TMAX * X3_QUADS * (int)sizeof(Bf16Quad) : 2 * kstep<T>() * (TMAX + 2) * 2 * (int)sizeof(T);
}
// The FP32 T GEMMs (sigma and left step) stage their row-mapped stores through LDS.
template <int STAGE, class T, class TO>
constexpr bool gemm_staged() {
    return (STAGE == ST_SIGMA || STAGE == ST_STEP) && sizeof(T) == 4 && sizeof(TO) == 4;
}
template <int STAGE, class T, class TO, bool X3>
constexpr int gemm_kernel_lds_bytes() {
    return gemm_staged<STAGE, T, TO>() && EPI_LDS_BYTES > gemm_lds_bytes<T, X3>() ? EPI_LDS_BYTES
                                                                                  : gemm_lds_bytes<T, X3>();
}

// One launch per stage: each tile runs the body of its group's tile shape.
// The BF16x3 body asks for 3 waves per SIMD: left to itself the compiler gives it
// 122 VGPRs + 64 AGPRs (2 waves); within 170 registers it needs no spills, and the
// third wave took the L=24 T GEMM from 1.33 to 1.16 ms (4 waves would spill).
template <int STAGE, class T, class TO, bool X3>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(X3 ? 3 : 1))) k_gemm(const GemmGroup *__restrict__ groups, const GemmTile *__restrict__ tiles,
                                              const T *__restrict__ A, const T *__restrict__ B, TO *__restrict__ C,
                                              T *__restrict__ W, const RowMap *__restrict__ rowmap) {
    constexpr bool staged = gemm_staged<STAGE, T, TO>();
    __shared__ double lds[(gemm_kernel_lds_bytes<STAGE, T, TO, X3>() + 7) / 8];
    const GemmTile t = tiles[blockIdx.x];
    if (t.g < 0) return;  // padding tile of the XCD-ordered list
    const GemmGroup g = groups[t.g];
    if (g.tm == 128)
        gemm_tile_n<128, T, TO, X3, staged>(g, t, A, B, C, W, rowmap, lds);
    else
        gemm_tile_n<64, T, TO, X3, staged>(g, t, A, B, C, W, rowmap, lds);
}

// C = sum over splits of the partial tiles, in split order (deterministic),
// accumulated in FP64.
template <int STAGE, class T, class TO>
__global__ void __launch_bounds__(256) k_split_reduce(const GemmGroup *__restrict__ groups,
                                                      const int *__restrict__ gids, const T *__restrict__ W,
                                                      TO *__restrict__ C) {
    const GemmGroup g = groups[gids[blockIdx.y]];
    const int64_t n = (int64_t)g.M * g.N;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int s = 0; s < g.split; ++s) acc += (double)W[g.part + s * n + e];
        int i = (int)(e % g.M), j = (int)(e / g.M);
        C[g.c + i + (int64_t)j * g.ldc] = (TO)acc;
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
template <class TD>
static hipError_t copy2d_launch(const Copy2D *tasks, int n, const double *src, TD *dst, hipStream_t s, bool sigma) {
    if (n == 0) return hipSuccess;
    if (sigma)
        hipLaunchKernelGGL((k_copy2d<ST_SIGMA, double, TD>), dim3(n), dim3(256), 0, s, tasks, src, dst);
    else
        hipLaunchKernelGGL((k_copy2d<ST_MISC, double, TD>), dim3(n), dim3(256), 0, s, tasks, src, dst);
    return hipGetLastError();
}
hipError_t launch_copy2d(const Copy2D *tasks, int n, const double *src, double *dst, hipStream_t s, bool sigma) {
    return copy2d_launch(tasks, n, src, dst, s, sigma);
}
hipError_t launch_copy2d(const Copy2D *tasks, int n, const double *src, float *dst, hipStream_t s, bool sigma) {
    return copy2d_launch(tasks, n, src, dst, s, sigma);
}
hipError_t launch_to_f32(const double *src, float *dst, int64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    int nb = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

template <class T, bool MULTI>
static void wapply_stage(const Strip *strips, const int *packs, int npacks, const Src *srcs, const T *src, T *C,
                         int stage, hipStream_t s, int nvec, int64_t ss, int64_t ds) {
    const dim3 g(npacks), b(256);
    switch (stage) {
        case ST_SIGMA:
            hipLaunchKernelGGL((k_wapply<ST_SIGMA, T, MULTI>), g, b, 0, s, strips, packs, srcs, src, C, nvec, ss, ds);
            break;
        case ST_FUSE:
            hipLaunchKernelGGL((k_wapply<ST_FUSE, T, MULTI>), g, b, 0, s, strips, packs, srcs, src, C, nvec, ss, ds);
            break;
        case ST_MISC:
            hipLaunchKernelGGL((k_wapply<ST_MISC, T, MULTI>), g, b, 0, s, strips, packs, srcs, src, C, nvec, ss, ds);
            break;
        default:
            hipLaunchKernelGGL((k_wapply<ST_STEP, T, MULTI>), g, b, 0, s, strips, packs, srcs, src, C, nvec, ss, ds);
            break;
    }
}
template <class T>
static hipError_t wapply_launch(const Strip *strips, const int *packs, int npacks, const Src *srcs, const T *src,
                                T *C, int stage, hipStream_t s, int nvec, int64_t ss, int64_t ds) {
    if (npacks == 0 || nvec < 1) return hipSuccess;
    if (nvec == 1)
        wapply_stage<T, false>(strips, packs, npacks, srcs, src, C, stage, s, 1, 0, 0);
    else
        wapply_stage<T, true>(strips, packs, npacks, srcs, src, C, stage, s, nvec, ss, ds);
    return hipGetLastError();
}
hipError_t launch_wapply(const Strip *strips, const int *packs, int npacks, const Src *srcs, const double *T,
                         double *C, int stage, hipStream_t s, int nvec, int64_t src_stride, int64_t dst_stride) {
    return wapply_launch<double>(strips, packs, npacks, srcs, T, C, stage, s, nvec, src_stride, dst_stride);
}
hipError_t launch_wapply(const Strip *strips, const int *packs, int npacks, const Src *srcs, const float *T, float *C,
                         int stage, hipStream_t s, int nvec, int64_t src_stride, int64_t dst_stride) {
    return wapply_launch<float>(strips, packs, npacks, srcs, T, C, stage, s, nvec, src_stride, dst_stride);
}

template <int STAGE, class T, class TO, bool X3>
static void gemm_stage(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const T *A, const T *B, TO *C,
                       T *W, const RowMap *rowmap, hipStream_t s) {
    if (ntiles > 0)
        hipLaunchKernelGGL((k_gemm<STAGE, T, TO, X3>), dim3(ntiles), dim3(256), 0, s, groups, tiles, A, B, C, W,
                           rowmap);
}

template <class T, class TO, bool X3>
static hipError_t gemm_launch(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const int *split_gids,
                              int nsplit, const T *A, const T *B, TO *C, T *W, const RowMap *rowmap, int stage,
                              hipStream_t s) {
    switch (stage) {
        case ST_SIGMA: gemm_stage<ST_SIGMA, T, TO, X3>(groups, tiles, ntiles, A, B, C, W, rowmap, s); break;
        case ST_SIGMA_OUT: gemm_stage<ST_SIGMA_OUT, T, TO, X3>(groups, tiles, ntiles, A, B, C, W, rowmap, s); break;
        case ST_STEP: gemm_stage<ST_STEP, T, TO, X3>(groups, tiles, ntiles, A, B, C, W, rowmap, s); break;
        case ST_STEP_OUT: gemm_stage<ST_STEP_OUT, T, TO, X3>(groups, tiles, ntiles, A, B, C, W, rowmap, s); break;
        default: gemm_stage<ST_MISC, T, TO, X3>(groups, tiles, ntiles, A, B, C, W, rowmap, s); break;
    }
    if (nsplit > 0) {
        // the sigma's reduction has its own instantiation (profiles attribute it per stage)
        if (stage == ST_SIGMA_OUT)
            hipLaunchKernelGGL((k_split_reduce<ST_SIGMA_OUT, T, TO>), dim3(64, nsplit), dim3(256), 0, s, groups,
                               split_gids, W, C);
        else
            hipLaunchKernelGGL((k_split_reduce<ST_MISC, T, TO>), dim3(64, nsplit), dim3(256), 0, s, groups,
                               split_gids, W, C);
    }
    return hipGetLastError();
}

hipError_t launch_gemm(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const int *split_gids, int nsplit,
                       const double *A, const double *B, double *C, double *W, const RowMap *rowmap, int stage,
                       hipStream_t s) {
    return gemm_launch<double, double, false>(groups, tiles, ntiles, split_gids, nsplit, A, B, C, W, rowmap, stage,
                                              s);
}
hipError_t launch_gemm(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const int *split_gids, int nsplit,
                       const float *A, const float *B, float *C, float *W, const RowMap *rowmap, int stage,
                       hipStream_t s, bool bf16x3) {
    if (bf16x3)
        return gemm_launch<float, float, true>(groups, tiles, ntiles, split_gids, nsplit, A, B, C, W, rowmap, stage,
                                               s);
    return gemm_launch<float, float, false>(groups, tiles, ntiles, split_gids, nsplit, A, B, C, W, rowmap, stage, s);
}
hipError_t launch_gemm(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const int *split_gids, int nsplit,
                       const float *A, const float *B, double *C, float *W, const RowMap *rowmap, int stage,
                       hipStream_t s, bool bf16x3) {
    if (bf16x3)
        return gemm_launch<float, double, true>(groups, tiles, ntiles, split_gids, nsplit, A, B, C, W, rowmap, stage,
                                                s);
    return gemm_launch<float, double, false>(groups, tiles, ntiles, split_gids, nsplit, A, B, C, W, rowmap, stage,
                                             s);
}

hipError_t launch_multidot(const double *const *xs, const double *const *ys, int npairs, int64_t n, int nblocks,
                           double *partial, hipStream_t s) {
    hipLaunchKernelGGL(k_multidot, dim3(nblocks, npairs), dim3(256), 0, s, xs, ys, n, partial);
    return hipGetLastError();
}
__global__ void __launch_bounds__(256) k_diag_scale(const double *x, const double *u, double f,
                                                    const double *__restrict__ D, double theta, double *out,
                                                    int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const double v = x[e] - f * u[e];
        const double d = theta - D[e];
        out[e] = d != 0.0 ? v / d : v;
    }
}

hipError_t launch_diag_scale(const double *x, const double *u, double f, const double *D, double theta, double *out,
                             int64_t n, hipStream_t s) {
    int nb = (int)std::min<int64_t>((n + 255) / 256, 4096);
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_diag_scale, dim3(nb), dim3(256), 0, s, x, u, f, D, theta, out, n);
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

// ------------------------------------------------------------------ two-site split
namespace mq {

// psi -> SVD working matrices, one workgroup per slab (SplitSlab)
__global__ void __launch_bounds__(256) k_split_gather(const SplitSlab *__restrict__ slabs,
                                                      const double *__restrict__ psi, double *__restrict__ X) {
    const SplitSlab t = slabs[blockIdx.x];
    const int n = t.rows * t.cols;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const int i = e % t.rows, j = e / t.rows;
        const int64_t so = i + (int64_t)j * t.src_cs;
        double v = t.c0 * psi[t.src0 + so];
        if (t.src1 >= 0) v += t.c1 * psi[t.src1 + so];
        X[t.dst + (int64_t)i * t.dst_rs + (int64_t)j * t.dst_cs] = v;
    }
}

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One-sided (Hestenes) Jacobi SVD of the working matrices X (R x C, C <= R,
// columns contiguous), all blocks together: sweeps of plane rotations that make
// the columns of each X pairwise orthogonal while its V (C x C, from the
// identity) collects them, so that X V_final = U diag(sigma) with sigma_j the
// final column norms.  A sweep visits every column pair of a block once, in
// round-robin order: round t pairs column Ce-1 (a padding column when C is odd)
// with column t, and column (t + k) mod (Ce-1) with (t - k) mod (Ce-1) for
// k = 1 .. Ce/2-1.  The pairs of a round are disjoint, so one launch runs round
// t of every block with one wave per pair, across the whole GPU; a pair is
// rotated when |x_p . x_q| > eps sqrt(R) |x_p| |x_q|, and a block is done after
// a sweep that rotates none of its pairs.
__device__ inline void jacobi_pair(int t, int k, int Ce, int &p, int &q) {
    if (k == 0) {
        p = t;
        q = Ce - 1;
    } else {
        p = (t + k) % (Ce - 1);
        q = (t - k + Ce - 1) % (Ce - 1);
    }
    if (p > q) {
        const int z = p;
        p = q;
        q = z;
    }
}

// V = I for every block; done / rotated flags cleared (blocks of one column are done)
__global__ void __launch_bounds__(256) k_jacobi_init(const SvdBlk *__restrict__ blks, double *__restrict__ V,
                                                     int32_t *__restrict__ done, int32_t *__restrict__ rot,
                                                     int32_t *__restrict__ sweeps) {
    const SvdBlk b = blks[blockIdx.x];
    double *v = V + b.v_off;
    const int C = b.C;
    for (int e = threadIdx.x; e < C * C; e += blockDim.x) v[e] = (e % (C + 1) == 0) ? 1.0 : 0.0;
    if (threadIdx.x == 0) {
        done[blockIdx.x] = C < 2;
        rot[blockIdx.x] = 0;
        sweeps[blockIdx.x] = 0;
    }
}

// round t of a sweep: wave w of the launch takes pair slot w (block slot_blk[w],
// pair index slot_k[w])
__global__ void __launch_bounds__(256) k_jacobi_round(const SvdBlk *__restrict__ blks,
                                                      const int32_t *__restrict__ slot_blk,
                                                      const int32_t *__restrict__ slot_k, int nslots, int t,
                                                      double *__restrict__ X, double *__restrict__ V,
                                                      const int32_t *__restrict__ done, int32_t *__restrict__ rot) {
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (w >= nslots) return;
    const int bi = slot_blk[w];
    if (done[bi]) return;
    const SvdBlk b = blks[bi];
    const int R = b.R, C = b.C, Ce = C + (C & 1);
    if (t >= Ce - 1) return;
    int p, q;
    jacobi_pair(t, slot_k[w], Ce, p, q);
    if (q >= C) return;
    double *xp = X + b.x_off + (int64_t)p * R, *xq = X + b.x_off + (int64_t)q * R;
    double a = 0.0, bb = 0.0, g = 0.0;
    for (int i = lane; i < R; i += 64) {
        const double u = xp[i], v = xq[i];
        a += u * u;
        bb += v * v;
        g += u * v;
    }
    a = wave_sum(a);
    bb = wave_sum(bb);
    g = wave_sum(g);
    const double tol = 2.220446049250313e-16 * sqrt((double)R);
    if (!(fabs(g) > tol * sqrt(a * bb))) return;
    const double zeta = (bb - a) / (2.0 * g);
    const double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    if (tt == 0.0) return;
    const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
    for (int i = lane; i < R; i += 64) {
        const double u = xp[i], v = xq[i];
        xp[i] = cs * u - sn * v;
        xq[i] = sn * u + cs * v;
    }
    double *vp = V + b.v_off + (int64_t)p * C, *vq = V + b.v_off + (int64_t)q * C;
    for (int i = lane; i < C; i += 64) {
        const double u = vp[i], v = vq[i];
        vp[i] = cs * u - sn * v;
        vq[i] = sn * u + cs * v;
    }
    if (lane == 0) rot[bi] = 1;
}

// end of sweep s: blocks that rotated nothing are done after s + 1 sweeps
__global__ void __launch_bounds__(256) k_jacobi_sweep_end(int nblk, int s, int32_t *__restrict__ done,
                                                          int32_t *__restrict__ rot, int32_t *__restrict__ sweeps) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    if (!done[b] && !rot[b]) {
        done[b] = 1;
        sweeps[b] = s + 1;
    }
    rot[b] = 0;
}

// singular values (final column norms) sorted descending, ties by column, with
// their column permutation; a block that did not converge reports sweeps = -1
__global__ void __launch_bounds__(256) k_jacobi_finish(const SvdBlk *__restrict__ blks, const double *__restrict__ X,
                                                       double *__restrict__ S, int32_t *__restrict__ perm,
                                                       const int32_t *__restrict__ done,
                                                       int32_t *__restrict__ sweeps) {
    extern __shared__ double sig[];
    const SvdBlk b = blks[blockIdx.x];
    const int R = b.R, C = b.C;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    for (int j = wave; j < C; j += nw) {
        const double *xj = X + b.x_off + (int64_t)j * R;
        double a = 0.0;
        for (int i = lane; i < R; i += 64) a += xj[i] * xj[i];
        a = wave_sum(a);
        if (lane == 0) sig[j] = sqrt(a);
    }
    __syncthreads();
    for (int j = tid; j < C; j += blockDim.x) {
        const double sj = sig[j];
        int r = 0;
        for (int k = 0; k < C; ++k) {
            const double sk = sig[k];
            r += (sk > sj) || (sk == sj && k < j);
        }
        S[b.s_off + r] = sj;
        perm[b.s_off + r] = j;
    }
    if (tid == 0 && !done[blockIdx.x]) sweeps[blockIdx.x] = -1;
}

// The two site tensors from the decomposition: for kept triplet k of block b
// (column j = perm[k] of the working matrices) the left vector u (X_j / sigma
// or V_j) goes to column k of A and the right vector w (V_j or X_j / sigma) to
// row k of B, with sigma on B (l2r) or on A (r2l).
__global__ void __launch_bounds__(256) k_split_assemble(const SvdBlk *__restrict__ blks,
                                                        const int32_t *__restrict__ keep,
                                                        const int64_t *__restrict__ a_off,
                                                        const int64_t *__restrict__ b_off, int l2r,
                                                        const double *__restrict__ X, const double *__restrict__ V,
                                                        const double *__restrict__ S, const int32_t *__restrict__ perm,
                                                        double *__restrict__ A, double *__restrict__ B) {
    const int bi = blockIdx.y;
    const SvdBlk b = blks[bi];
    const int K = keep[bi];
    for (int k = blockIdx.x; k < K; k += gridDim.x) {
        const int j = perm[b.s_off + k];
        const double s = S[b.s_off + k];
        const double inv = s > 0.0 ? 1.0 / s : 0.0;
        const double *xj = X + b.x_off + (int64_t)j * b.R;
        const double *vj = V + b.v_off + (int64_t)j * b.C;
        const double *lu = b.trans ? vj : xj;
        const double *rw = b.trans ? xj : vj;
        double ls = b.trans ? 1.0 : inv, rs = b.trans ? inv : 1.0;
        if (l2r)
            rs *= s;
        else
            ls *= s;
        double *a = A + a_off[bi] + (int64_t)k * b.rows;
        for (int i = threadIdx.x; i < b.rows; i += blockDim.x) a[i] = ls * lu[i];
        double *w = B + b_off[bi] + (int64_t)k * b.cols;
        for (int i = threadIdx.x; i < b.cols; i += blockDim.x) w[i] = rs * rw[i];
    }
}

hipError_t launch_split_gather(const SplitSlab *slabs, int n, const double *psi, double *X, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_split_gather, dim3(n), dim3(256), 0, s, slabs, psi, X);
    return hipGetLastError();
}

hipError_t launch_jacobi_init(const SvdBlk *blks, int nblk, double *V, int32_t *done, int32_t *rot, int32_t *sweeps,
                              hipStream_t s) {
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_jacobi_init, dim3(nblk), dim3(256), 0, s, blks, V, done, rot, sweeps);
    return hipGetLastError();
}

hipError_t launch_jacobi_round(const SvdBlk *blks, const int32_t *slot_blk, const int32_t *slot_k, int nslots, int t,
                               double *X, double *V, const int32_t *done, int32_t *rot, hipStream_t s) {
    if (nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_jacobi_round, dim3((nslots + 3) / 4), dim3(256), 0, s, blks, slot_blk, slot_k, nslots, t, X,
                       V, done, rot);
    return hipGetLastError();
}

hipError_t launch_jacobi_sweep_end(int nblk, int sweep, int32_t *done, int32_t *rot, int32_t *sweeps,
                                   hipStream_t s) {
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_jacobi_sweep_end, dim3((nblk + 255) / 256), dim3(256), 0, s, nblk, sweep, done, rot, sweeps);
    return hipGetLastError();
}

hipError_t launch_jacobi_finish(const SvdBlk *blks, int nblk, int max_c, const double *X, double *S, int32_t *perm,
                                const int32_t *done, int32_t *sweeps, hipStream_t s) {
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_jacobi_finish, dim3(nblk), dim3(256), sizeof(double) * std::max(1, max_c), s, blks, X, S,
                       perm, done, sweeps);
    return hipGetLastError();
}

hipError_t launch_split_assemble(const SvdBlk *blks, int nblk, int max_keep, const int32_t *keep,
                                 const int64_t *a_off, const int64_t *b_off, int l2r, const double *X,
                                 const double *V, const double *S, const int32_t *perm, double *A, double *B,
                                 hipStream_t s) {
    if (nblk <= 0 || max_keep <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_split_assemble, dim3(std::min(max_keep, 256), nblk), dim3(256), 0, s, blks, keep, a_off,
                       b_off, l2r, X, V, S, perm, A, B);
    return hipGetLastError();
}

}  // namespace mq
