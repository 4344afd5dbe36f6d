// Execution plans of the hot path: task lists for the device kernels,
// built once per site problem on the host.
#pragma once

#include "mq_host.hpp"

namespace mq {

// ---- device task records (plain old data, copied to the GPU as arrays) ----

// Buffers of task lists of at least kTaskCacheMin bytes come from a small
// process-wide cache of released buffers (task_alloc / task_free): the lists of
// consecutive plans have similar sizes (hundreds of MB at a chain's centre), and a
// reused buffer has its pages mapped already instead of faulting fresh ones in.
constexpr size_t kTaskCacheMin = size_t(1) << 20;
void *task_alloc(size_t bytes);
void task_free(void *p, size_t bytes) noexcept;

// Allocator whose resize() leaves the records uninitialised: the builder writes
// every record of a task list it sizes, often from several threads, and zero-
// filling lists of 10^6-10^7 records first would be one more pass over them on
// one thread.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    T *allocate(size_t n) {
        if (n * sizeof(T) >= kTaskCacheMin) return static_cast<T *>(task_alloc(n * sizeof(T)));
        return std::allocator<T>::allocate(n);
    }
    void deallocate(T *p, size_t n) noexcept {
        if (n * sizeof(T) >= kTaskCacheMin) task_free(p, n * sizeof(T));
        else std::allocator<T>::deallocate(p, n);
    }
    template <class U>
    void construct(U *p) noexcept {
        ::new ((void *)p) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
template <class T>
using TaskVec = std::vector<T, NoInitAlloc<T>>;

// dst[i + j*ldd] = scale * (trans ? src[j + i*lds] : src[i + j*lds]), i < rows, j < cols
struct Copy2D {
    int64_t dst, src;
    int rows, cols, ldd, lds;
    int trans;
    double scale;
};

// C(M x N) = op(A)(M x K) * op(B)(K x N), column-major with leading dims.
// transA: A stored K x M; transB: B stored N x K.  split > 1: K is cut into
// `split` chunks whose partial tiles go to the workspace at `part`.  tm, tn:
// tile height (64 or 128) and width (a multiple of 16 up to 128).  rowmap > 0:
// row r of C is stored at c + map[rowmap - 1 + r].off + col * map[rowmap - 1 + r].ld
// (row blocks of different heights each stored contiguously; such groups are
// never split); rowmap == 0: C is column-major with ldc.
struct GemmGroup {
    int64_t a, b, c, part;
    int M, N, K;
    int lda, ldb, ldc;
    int transA, transB;
    int split;
    int kchunk;
    int tm, tn;
    int64_t rowmap;
};
struct RowMap {
    int64_t off;  // relative to the group's c
    int ld;
    int pad;
};
// one output tile; g < 0 marks a padding tile that does nothing
struct GemmTile {
    int g;
    int mt, nt, ks;
};
// partial-sum reduction of one split group: C = sum_s part[s]
struct ReduceTask {
    int g;
};

// dst(rows x cols, ldd) = sum over sources of coef * src(rows x cols, lds)
struct Strip {
    int64_t dst;
    int rows, cols, ldd;
    int s0, s1;
};
struct Src {
    int64_t off;
    int ld;
    double coef;
};

// Tiles are ordered for the XCDs: workgroup i of a launch runs on XCD i % 8,
// and the tiles of one (group, K chunk, row block) -- which share their A
// panel -- are placed on one XCD, consecutively, so the panel is read from
// HBM once and then served by that XCD's L2.
struct GemmList {
    std::vector<GemmGroup> groups;
    std::vector<GemmTile> tiles;
    TaskVec<RowMap> rowmap;  // row placements of the groups with rowmap >= 0
    int64_t ws = 0;  // workspace elements for split partials
    double flops() const {
        double f = 0;
        for (auto &g : groups) f += 2.0 * g.M * g.N * g.K;
        return f;
    }
};
// Consecutive strips are run in packs of one workgroup each: a pack is either
// one strip of more than kPackElems elements, or up to kPackStrips strips of
// together at most kPackElems elements (the fragmented blocks of small sectors).
constexpr int kPackStrips = 64, kPackElems = 1024;
struct StripList {
    TaskVec<Strip> strips;
    TaskVec<Src> srcs;
    std::vector<int> packs;  // pack p runs strips [packs[p], packs[p + 1])
};
// (re)compute the packs of a strip list, keeping the strip order
void make_packs(StripList &sl);

enum PlanKind { PLAN_SIGMA = 0, PLAN_LSTEP = 1, PLAN_RSTEP = 2, PLAN_DIAG = 3 };

// Host-side plan; the API layer uploads the task arrays.
struct HostPlan {
    int kind = PLAN_SIGMA;
    int rbtm = 0;  // sigma plans: 1 for the rbtm schedule (build_sigma_plan_rbtm)
    int symm = MQ_SYMM_SU2U1;
    int64_t vec_len = 0;
    // arena sizes (elements)
    int64_t n_la = 0, n_rb = 0, n_prp = 0, n_t = 0, n_c = 0, n_tmp = 0, n_out = 0;
    // stage task lists
    std::vector<Copy2D> arrL;   // L boundary -> LA  (sigma, lstep)
    StripList combL;            // fused rows of LA: weighted sums of arranged rows (sigma, lstep)
    std::vector<Copy2D> arrR;   // R boundary -> RB  (sigma) / RC (rstep)
    std::vector<Copy2D> reshape;  // x -> PRP
    GemmList g1;                // T = LA * PRP    (sigma, lstep) ; t = X * RC (rstep)
    StripList wapply;           // T -> C          / t -> BUF
    GemmList g2;                // y = C * RB      / tmp = C^T X  / tmp = BUF * PRP^T
    std::vector<Copy2D> scatter;  // tmp -> output boundary (steps)
    // output boundary structure (steps)
    int out_aux = 0;
    std::vector<int32_t> out_bptr, out_lc, out_rc, out_rows, out_cols;
    std::vector<int64_t> out_off;
    // sigma plans: live[i] = 1 where H_eff couples coordinate i of the vector, 0
    // where row i and column i of H_eff are structurally zero (no operator term
    // reaches that (phys, left, right) slot of the reduced two-site tensor)
    std::vector<uint8_t> live;
    // algorithmic accounting (reference schedule)
    double ref_flops = 0, min_bytes = 0;
    // executed FLOPs of this implementation (for reporting)
    double exec_flops = 0;
};

// Which MPO columns are summed on the left boundary first ("fused"): mode 0 none,
// 1 the columns whose operator entries number at least `ratio` times their
// distinct positions and at least `min_entries`, 2 every non-empty column.
struct FuseRule {
    int mode = 1;
    int ratio = 4;
    int min_entries = 64;
};
// ref_count: also count the reference schedule's FLOPs (ref_flops; -1 when not
// counted).  The count walks the reference schedule once more and is only needed
// for measurement.
// Schedule of a sigma plan: SCHED_LBTM the lbtm form (with fused columns) at
// every site; SCHED_REFERENCE the reference's choice per site (site_hamil.hpp:35):
// rbtm (build_sigma_plan_rbtm, no fused rows) where (rows - one_rows) >=
// (cols - one_cols) for SU2U1, lbtm elsewhere.
enum SigmaSchedule { SCHED_LBTM = 0, SCHED_REFERENCE = 1 };
HostPlan build_sigma_plan(const mq_site_desc &d, int tile_m, int tile_n, int kchunk_target, FuseRule fuse = {},
                          bool ref_count = true, int schedule = SCHED_LBTM);
// The rbtm schedule (site_hamil.hpp:149-204) on every site: t = X * RC, the rbtm
// W application into BUF, S = LA^T * BUF, y = S reshaped to the left-paired layout.
HostPlan build_sigma_plan_rbtm(const mq_site_desc &d, int tile_m, int tile_n, int kchunk_target,
                               bool ref_count = true);
HostPlan build_lstep_plan(const mq_site_desc &d, int tile_m, int tile_n, int kchunk_target, FuseRule fuse = {},
                          bool ref_count = true);
HostPlan build_rstep_plan(const mq_site_desc &d, int tile_m, int tile_n, int kchunk_target);
// Diagonal of H_eff (SU2 only): arrL extracts the diagonals of the L[b1](lc,lc)
// blocks into LA, arrR those of R[b2](rc,rc) into RB, the W-application strips
// combine LA columns into T, and g2 writes the diagonal = T * RB per row slice.
HostPlan build_diag_plan(const mq_site_desc &d, int tile_m, int tile_n, int kchunk_target);

// One GEMM group split into tiles (and K chunks) as the plans do.  tile_m: 64,
// 128 or 0 (128 rows unless M fits 64); tile_n: a multiple of 16 up to 128, or 0
// (the padding-minimising column tile).  The plan builders take the same tile_m
// and tile_n for every group they add.
void tile_gemm(GemmList &gl, GemmGroup g, int tile_m, int tile_n, int kchunk_target);
// XCD-aware tile order (call once every group has been added).
void order_tiles(GemmList &gl);

// Task lists for nvec vectors in one launch per stage: vector j's copies are the
// single-vector tasks with their per-vector arena offsets advanced by j steps
// (a step of 0 marks an operand all vectors share).  The W application needs no
// copies: its kernel runs the plan's one strip list for every vector, with
// per-vector strides.
std::vector<Copy2D> batch_copies(const std::vector<Copy2D> &v, int nvec, int64_t src_step, int64_t dst_step);
GemmList batch_gemm(const GemmList &gl, int nvec, int64_t a_step, int64_t b_step, int64_t c_step);

// FLOPs of the reference schedule (site_hamil.hpp:28-204 for sigma, move_boundary.hpp for steps)
double reference_sigma_flops(const mq_site_desc &d);

// ---- two-site split ----
// One slab of the map from the two-site tensor psi (canonical left-paired,
// column-major blocks) to the working matrices of the block SVD:
// dst[i*dst_rs + j*dst_cs] = c0 * psi[src0 + i + j*src_cs] (+ c1 * psi[src1 + ...]
// when src1 >= 0), i < rows, j < cols.
struct SplitSlab {
    int64_t dst, src0, src1;
    double c0, c1;
    int32_t rows, cols;
    int32_t dst_rs, dst_cs, src_cs;
    int32_t pad;
};
// One both-paired block (rows x cols) and its SVD working arrays: the R x C
// working matrix X (C = min(rows, cols), columns contiguous) at x_off is the
// block itself (trans = 0, rows >= cols) or its transpose (trans = 1); V
// (C x C, columns contiguous) at v_off; singular values and their column
// permutation at s_off.
struct SvdBlk {
    int64_t x_off, v_off, s_off;
    int32_t R, C;
    int32_t rows, cols;
    int32_t trans;
    int32_t pad;
};
struct SplitMap {
    std::vector<SplitSlab> slabs;
    std::vector<SvdBlk> blk;                  // descending charge
    std::vector<int32_t> charge, rows, cols;  // per block, in blk order
    std::vector<int32_t> order;               // block indices, largest working matrix first (pair-slot order)
    int64_t psi_len = 0, x_total = 0, v_total = 0, s_total = 0;
    int max_c = 0;
};
// The map of TwoSiteTensor::make_both_paired (ts_reduction.h unreduce_left for
// SU2U1, then ts_reshape.h:88-145 reshape_left_to_both).
SplitMap build_split_map(const mq_split_desc &d);
// The same map applied on the host (test introspection): both-paired blocks,
// row-major, one after another in blk order.
void apply_split_map_host(const SplitMap &m, const double *psi, double *both);

}  // namespace mq
