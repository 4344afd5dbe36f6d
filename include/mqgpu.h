/*
 * mqgpu.h -- C ABI of the MI355X DMRG sweep hot path.
 *
 * The library replaces, for one site problem at a time, the three C++
 * template entry points of the reference's contraction engine and local
 * eigensolver:
 *
 *   contraction::Engine<M,M,S>::site_hamil2(ket, L, R, W)
 *       dmrg/framework/dmrg/mp_tensors/contractions/non-abelian/engine.hpp:198
 *       dmrg/framework/dmrg/mp_tensors/contractions/abelian/engine.hpp:197
 *       called from ietl::mult, dmrg/framework/dmrg/optimize/ietl_lanczos_solver.h:109-115
 *     -> mq_sigma
 *   contraction::Engine<M,M,S>::overlap_mpo_left_step(bra, ket, L, W)
 *       non-abelian/engine.hpp:105, common/move_boundary.hpp:130-187
 *     -> mq_lstep_plan_create + mq_boundary_step
 *   contraction::Engine<M,M,S>::overlap_mpo_right_step(bra, ket, R, W)
 *       non-abelian/engine.hpp:116, common/move_boundary.hpp:191-229
 *     -> mq_rstep_plan_create + mq_boundary_step
 *   solve_ietl_jcd(SiteProblem&, MPSTensor const&, BaseParameters&)
 *       dmrg/framework/dmrg/optimize/ietl_jacobi_davidson.h:18-77 with
 *       ietl::jacobi_davidson::calculate_eigenvalue, dmrg/framework/ietl/jacobi.h:364-451
 *     -> mq_jacobi_davidson
 *   solve_ietl_davidson(SiteProblem&, MPSTensor const&, BaseParameters&)
 *       dmrg/framework/dmrg/optimize/ietl_davidson.h:121-160 with
 *       ietl::davidson::calculate_eigenvalue, optimize/davidson.h:64-161, and
 *       MultDiagonal (ietl_davidson.h:78-118), whose diagonal is
 *       contraction::SU2::diagonal_hamiltonian, non-abelian/h_diag.hpp:120-152
 *     -> mq_diag_plan_create + mq_diagonal + mq_davidson
 *
 * All data crosses the boundary as flat integer metadata plus FP64 arrays.
 * Every function returns MQ_OK (0) or a negative error code; a message for
 * the last error of a context is available from mq_last_error().  No C++
 * exception crosses the ABI.  A context owns one HIP stream; calls on one
 * context must not overlap.
 *
 * Data conventions (shared by every entry point):
 *  - A charge is two int32: SU2U1 (N, 2S), TwoU1 (N_alpha, N_beta).
 *  - An mq_index lists (charge, size) in descending lexicographic charge
 *    order (dmrg/framework/dmrg/block_matrix/indexing_stable.hpp:82-130).
 *  - Dense blocks are column-major.
 *  - An MPS site tensor (ket psi, sigma, Davidson vectors) is stored
 *    LEFT-PAIRED and CANONICAL: one block per charge c of the right index
 *    that also occurs in phys*left, in descending charge order, with
 *    rows = sum over (s, l) with s+l=c of size(s)*size(l) (the fused
 *    row offsets follow ProductBasis(phys, left), indexing_stable.hpp:305-343)
 *    and cols = size(c); blocks are contiguous in that order.
 *  - A boundary is a vector over the MPO bond index b of block matrices;
 *    block (ket charge, bra charge) holds rows = ket size, cols = bra size
 *    (common/move_boundary.hpp:165, non-abelian/gemm.hpp:193-201).  The
 *    boundary descriptor gives, per bond b, a range of blocks with explicit
 *    element offsets into one flat FP64 arena.  Bonds whose hermitian
 *    partner is stored instead (MPO herm_info left_skip/right_skip) carry no
 *    blocks (common/boundary_times_mps.hpp:196-207).
 */
#ifndef MQGPU_H
#define MQGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MQ_OK 0
#define MQ_ERR_ARG (-1)
#define MQ_ERR_HIP (-2)
#define MQ_ERR_PLAN (-3)
#define MQ_ERR_ALLOC (-4)
#define MQ_ERR_STATE (-5)

enum mq_symm { MQ_SYMM_TWOU1 = 0, MQ_SYMM_SU2U1 = 1 };
enum mq_prec { MQ_PREC_FP64 = 0, MQ_PREC_FP32 = 1, MQ_PREC_BF16X3 = 2 };

typedef struct mq_index {
    int32_t n;
    const int32_t *charges; /* [2n] */
    const int32_t *sizes;   /* [n]  */
} mq_index;

/* Block list: block i = (lc[2i..], rc[2i..]) of rows[i] x cols[i] at data + offset[i]. */
typedef struct mq_blocks {
    int32_t n;
    const int32_t *lc;      /* [2n] */
    const int32_t *rc;      /* [2n] */
    const int32_t *rows;    /* [n]  */
    const int32_t *cols;    /* [n]  */
    const int64_t *offset;  /* [n]  element offset into the arena */
} mq_blocks;

/* Boundary structure: blocks of bond b are blocks[bptr[b] .. bptr[b+1]). */
typedef struct mq_boundary_desc {
    int32_t aux_dim;
    const int32_t *bptr;    /* [aux_dim+1] */
    mq_blocks blocks;
    int64_t total;          /* arena length in elements */
} mq_boundary_desc;

/*
 * MPO tensor (dmrg/framework/dmrg/mp_tensors/mpotensor.h:24-107):
 * column-compressed (b1,b2) -> list of (operator, scale) terms, operator
 * table with SU(2) sparse entries (block_matrix/sparse_operator.h:105-168),
 * bond spins and hermitian partner info (mpotensor_detail.h:118-168).
 */
typedef struct mq_mpo_desc {
    int32_t row_dim, col_dim;
    const int32_t *col_ptr;     /* [col_dim+1] */
    const int32_t *row_idx;     /* [nnz], ascending b1 within a column */
    const int32_t *term_ptr;    /* [nnz+1] */
    const int32_t *term_op;     /* [nterms] */
    const double  *term_scale;  /* [nterms] */
    int32_t n_ops;
    const int32_t *op_spin;     /* [n_ops] 2S of the reduced operator */
    const int32_t *op_bptr;     /* [n_ops+1] operator blocks */
    const int32_t *opb_in;      /* [2*nb] physical input (ket) charge */
    const int32_t *opb_out;     /* [2*nb] physical output (bra) charge */
    const int32_t *opb_eptr;    /* [nb+1] sparse entries of each block */
    const int32_t *ent_row;     /* [ne] row inside the physical in-block */
    const int32_t *ent_col;     /* [ne] col inside the physical out-block */
    const int32_t *ent_rspin;   /* [ne] */
    const int32_t *ent_cspin;   /* [ne] */
    const double  *ent_coeff;   /* [ne] */
    const int32_t *left_spin;   /* [row_dim] 2S of MPO bond b1 */
    const int32_t *right_spin;  /* [col_dim] */
    const int32_t *left_herm;   /* [row_dim] hermitian partner of b1 */
    const int32_t *right_herm;  /* [col_dim] */
    const int32_t *left_phase;  /* [row_dim] */
    const int32_t *right_phase; /* [col_dim] */
} mq_mpo_desc;

/* One site problem: ket tensor indices, MPO tensor, boundary structures. */
typedef struct mq_site_desc {
    int32_t symm;               /* enum mq_symm */
    mq_index phys, left, right; /* ket (= bra) tensor indices */
    mq_mpo_desc mpo;
    mq_boundary_desc L;         /* aux_dim = mpo.row_dim (unused by right steps) */
    mq_boundary_desc R;         /* aux_dim = mpo.col_dim (unused by left steps)  */
} mq_site_desc;

typedef struct mq_ctx mq_ctx;
typedef struct mq_plan mq_plan;

/* -- context ---------------------------------------------------------- */
int mq_ctx_create(int device, mq_ctx **out);
int mq_ctx_destroy(mq_ctx *ctx);
const char *mq_last_error(const mq_ctx *ctx);
/* Plan-construction options of a context (no reference counterpart), applied
 * to plans created afterwards.  MQ_OPT_FUSE: how MPO columns with many terms
 * are evaluated -- 0 term by term, 1 (default) columns whose operator entries
 * repeat at few positions are summed on the left boundary first, 2 every
 * column is summed that way (testing aid).  Under mode 1 a column is summed
 * that way when its entries number at least MQ_OPT_FUSE_RATIO (1..255,
 * default 4) times its distinct positions and at least MQ_OPT_FUSE_MIN
 * (1..4095, default 64).  MQ_OPT_KCHUNK: target K chunk of split-K GEMM
 * groups (256..65536, a multiple of 16, default 2048).  Results agree to
 * round-off.  MQ_OPT_GRAPH (0 default / 1): the Davidson drivers run their
 * sigma vectors through mq_sigma_graph's captured graph (same results).
 * MQ_OPT_TILE_M: row tile of every GEMM group, 64 or 128, or 0 (default:
 * 128 rows unless the group's M fits 64).  MQ_OPT_REF_FLOPS (1 default / 0):
 * count the reference schedule's FLOPs for mq_plan_flops; with 0 sigma and
 * left-step plans skip that count (it walks the reference schedule once more)
 * and mq_plan_flops reports -1.  MQ_OPT_SCHEDULE: how sigma plans contract,
 * MQ_SCHED_LBTM (0, default) the lbtm form with fused columns at every site,
 * MQ_SCHED_REFERENCE (1) the reference's choice per site
 * (non-abelian/site_hamil.hpp:35): the rbtm form of site_hamil_rbtm
 * (site_hamil.hpp:149-204) where (rows - one_rows) >= (cols - one_cols) for
 * SU2U1, lbtm elsewhere. */
/* MQ_OPT_TILE_N: column tile of every GEMM group, a multiple of 16 from 16 to
 * 128, or 0 (default: per group, the width that wastes the fewest padded
 * columns).  Tile shapes change the launch, not the results. */
enum mq_option { MQ_OPT_FUSE = 0, MQ_OPT_FUSE_RATIO = 1, MQ_OPT_FUSE_MIN = 2, MQ_OPT_KCHUNK = 3,
                 MQ_OPT_GRAPH = 4, MQ_OPT_TILE_M = 5, MQ_OPT_REF_FLOPS = 6, MQ_OPT_SCHEDULE = 7,
                 MQ_OPT_TILE_N = 8 };
enum mq_schedule { MQ_SCHED_LBTM = 0, MQ_SCHED_REFERENCE = 1 };
int mq_ctx_set_option(mq_ctx *ctx, int option, int value);
/* HIP stream (hipStream_t) the context launches on. */
void *mq_ctx_stream(mq_ctx *ctx);
int mq_ctx_synchronize(mq_ctx *ctx);

/* -- sigma vector ------------------------------------------------------- */
/* Plan for site_hamil2 on one site problem; keeps device task lists. */
int mq_sigma_plan_create(mq_ctx *ctx, const mq_site_desc *desc, mq_plan **out);
int mq_plan_destroy(mq_plan *plan);
/* Elements of one MPS vector (canonical left-paired layout). */
int64_t mq_plan_vec_len(const mq_plan *plan);
/* Contraction form of a sigma plan: 0 lbtm, 1 rbtm (see MQ_OPT_SCHEDULE). */
int mq_plan_schedule(const mq_plan *plan);
/* Algorithmic FLOPs of the reference schedule (2mnk per block GEMM + 2 len
 * per axpy, lbtm/rbtm chosen as site_hamil.hpp:35) and minimum HBM bytes. */
int mq_plan_flops(const mq_plan *plan, double *flops, double *bytes);
/* Device scratch bytes the plan holds between launches. */
int64_t mq_plan_workspace_bytes(const mq_plan *plan);
/* FLOPs this implementation executes per sigma (differs from the
 * algorithmic count: no lazy products, no hermitian-skip savings). */
double mq_plan_exec_flops(const mq_plan *plan);
/* y = H_eff x.  L, R: device boundary arenas; x, y: device vectors.
 * The boundaries are rearranged into GEMM panels at the first call with a
 * given (L, R) pointer pair and reused while the pair stays the same. */
int mq_sigma(mq_plan *plan, const double *L, const double *R,
             const double *x, double *y, int prec);
/* Y(:, j) = H_eff X(:, j), j < nvec (1..64): nvec device vectors stored one
 * after another (vector j at X + j*vec_len), with one launch per stage for all
 * of them; the arranged boundary panels are read once per tile for all
 * vectors.  Each Y(:, j) equals mq_sigma of X(:, j) (no reference
 * counterpart: the several Davidson roots of a block solver in one pass).
 * The plan keeps per-vector work arenas for the last nvec used. */
int mq_sigma_batch(mq_plan *plan, const double *L, const double *R,
                   const double *X, double *Y, int nvec, int prec);
/* y = H_eff x like mq_sigma, through a HIP graph of the sigma stages captured
 * on the plan's own staging vectors (x is copied in, y copied out); the graph
 * is captured on the first call and again when L, R or prec change, and
 * replayed otherwise (no reference counterpart: for callers that replay the
 * same problem many times). */
int mq_sigma_graph(mq_plan *plan, const double *L, const double *R,
                   const double *x, double *y, int prec);
/* Drop the rearranged boundaries (call after changing L or R in place). */
int mq_plan_rebind(mq_plan *plan);
/* Rearrange the boundaries L, R into the plan's GEMM panels now and make them
 * the current binding (asynchronous on the plan's stream).  The implicit
 * binding of mq_sigma and the solvers is keyed on the (L, R) pointer pair
 * only, so a caller that frees a boundary and allocates a new one (which may
 * land at the same address) calls mq_plan_bind (or mq_plan_rebind) before the
 * next sigma.  Stands in for the boundary arguments of site_hamil2
 * (non-abelian/engine.hpp:198), which the reference reads afresh every call. */
int mq_plan_bind(mq_plan *plan, const double *L, const double *R);
/* mask (host, vec_len bytes, written): 1 for the coordinates of the
 * left-paired vector that H_eff couples, 0 for those whose row and column of
 * H_eff are structurally zero (no MPO term reaches that slot of the reduced
 * two-site tensor: e.g. the spin-1 pair of two singly occupied sites next to a
 * spin-0 bond state, which ts_reduction.h:103-105 never fills and
 * apply_op.hpp:75 filters).  Start vectors of an eigensolver that are not
 * built from an MPS must vanish there, or the zero eigenvalues of those
 * coordinates become roots.  Sigma plans only. */
int mq_plan_live_mask(const mq_plan *plan, uint8_t *mask);
/* Instrumented mq_sigma (no reference counterpart): HIP events around each
 * stage on the plan's stream.  ms[0] boundary arrangement, ms[1] reshape,
 * ms[2] T GEMM, ms[3] W application, ms[4] output GEMM.  work[0..2] are the
 * stage FLOPs (T GEMM, W application axpys, output GEMM), work[3..5] the
 * minimum HBM bytes of the same three stages. */
int mq_sigma_timed(mq_plan *plan, const double *L, const double *R,
                   const double *x, double *y, int prec, float ms[5], double work[6]);

/* -- boundary steps ----------------------------------------------------- */
/* Plan for overlap_mpo_left_step (desc.L = input) or right step (desc.R). */
int mq_lstep_plan_create(mq_ctx *ctx, const mq_site_desc *desc, mq_plan **out);
int mq_rstep_plan_create(mq_ctx *ctx, const mq_site_desc *desc, mq_plan **out);
/* Output boundary structure: aux_dim, per-bond block ranges and offsets.
 * Arrays stay owned by the plan. */
int mq_step_out_desc(const mq_plan *plan, mq_boundary_desc *out);
/* out = step(in, A).  A: device MPS site tensor (canonical left-paired). */
int mq_boundary_step(mq_plan *plan, const double *in, const double *A, double *out);

/* -- local eigensolver -------------------------------------------------- */
/* Jacobi-Davidson with the reference semantics: modified Gram-Schmidt with
 * one refinement, Rayleigh matrix, lowest Ritz pair, residual test
 * |r| <= tol*|theta| or |r| < tol, correction t = -r + (r.u)/(u.u) u,
 * at most max_iter sigma vectors (1..1000), no restart.  x0 (device, read) is the
 * initial guess; u (device, written) the Ritz vector; *theta the eigenvalue.
 * n_ortho vectors (device, each vec_len) are projected out as
 * SingleSiteVS::project does (ietl_lanczos_solver.h:90-95). */
/* With prec != MQ_PREC_FP64 the iterations run in that form and the returned
 * eigenvalue is the FP64 Rayleigh quotient of the Ritz vector.  With tol > 0
 * those iterations stop at tol or at the form's rounding floor, whichever is
 * larger (|r| <= 4e-6 |theta| for MQ_PREC_FP32, 2e-5 |theta| for
 * MQ_PREC_BF16X3), or when the largest open residual has not halved for 8
 * iterations; when the FP64 residual of the pair then misses tol, the
 * solve continues in FP64 from the Ritz vector (up to max_iter more
 * iterations, counted in *iterations).  The same holds for mq_davidson and
 * mq_block_davidson. */
int mq_jacobi_davidson(mq_plan *plan, const double *L, const double *R,
                       const double *x0, double *u, double *theta,
                       int max_iter, double tol, int prec,
                       const double *const *ortho, int n_ortho,
                       int *iterations);

/* Plan for the diagonal of H_eff (SU2U1 only; other symmetries fail with
 * MQ_ERR_PLAN, as MultDiagonal throws for them, ietl_davidson.h:60-76).
 * desc is the sigma plan's descriptor. */
int mq_diag_plan_create(mq_ctx *ctx, const mq_site_desc *desc, mq_plan **out);
/* diag (device, vec_len, written) = diagonal of H_eff in the left-paired
 * layout, as contraction::SU2::diagonal_hamiltonian computes it
 * (non-abelian/h_diag.hpp:20-152): from the diagonals of the stored
 * L[b1](lc,lc) and R[b2](rc,rc) blocks and the diagonal entries of the
 * phys-diagonal operator blocks.  Boundary slices that are not stored
 * (hermitian partners) contribute nothing.  Asynchronous on the plan's
 * stream. */
int mq_diagonal(mq_plan *plan, const double *L, const double *R, double *diag);
/* Davidson with the diagonal preconditioner (SU2U1 sigma plans): Gram-Schmidt
 * with one refinement, lowest Ritz pair of the Rayleigh matrix, residual
 * test |r| <= tol*|theta| or |r| < tol, correction
 * t = (r - (V.r/(V.V)) V) / (theta - D) with V = u / (theta - D)
 * (elements with theta == D are left undivided), restart to the two lowest
 * Ritz vectors when the basis reaches 20, at most max_iter sigma vectors.
 * diag: device, from mq_diagonal.  x0, u, ortho, prec, iterations as for
 * mq_jacobi_davidson; the caller orthogonalises ortho beforehand
 * (ietl_davidson.h:126-129). */
int mq_davidson(mq_plan *plan, const double *L, const double *R,
                const double *diag, const double *x0, double *u, double *theta,
                int max_iter, double tol, int prec,
                const double *const *ortho, int n_ortho, int *iterations);

/* The nroots (1..16) lowest eigenpairs of H_eff together (no counterpart in
 * the reference, which finds excited states one at a time under
 * orthogonality constraints): a block Davidson whose sigma vectors for a
 * whole block go through one batched launch (mq_sigma_batch).  Corrections
 * are t = -r + (r.u)/(u.u) u (as mq_jacobi_davidson) or, when diag (from
 * mq_diagonal) is given, the diagonal preconditioner of mq_davidson; the
 * basis holds up to 60 vectors and restarts from the lowest 2*nroots Ritz
 * vectors.  X0: nroots device start vectors one after another (linearly
 * independent); U: the nroots Ritz vectors, same layout; theta: nroots
 * eigenvalues, ascending.  Convergence per root as mq_jacobi_davidson;
 * max_iter counts block iterations.  With prec != MQ_PREC_FP64 every
 * eigenvalue is the FP64 Rayleigh quotient of its Ritz vector.  ortho as for
 * mq_jacobi_davidson. */
int mq_block_davidson(mq_plan *plan, const double *L, const double *R,
                      const double *diag, const double *X0, double *U,
                      double *theta, int nroots, int max_iter, double tol,
                      int prec, const double *const *ortho, int n_ortho,
                      int *iterations);

/* The last eigensolver call on this plan (no reference counterpart: the
 * record of a precision form's two phases): info[0] iterations in the
 * requested form, info[1] FP64 continuation iterations that followed them,
 * info[2] the form's last Ritz value (lowest root of a block solve), info[3]
 * the FP64 Rayleigh quotient of that Ritz vector (the defect-corrected value
 * before any continuation).  For MQ_PREC_FP64 calls info[1] = 0 and
 * info[2] = info[3] = the returned eigenvalue. */
int mq_plan_solve_info(const mq_plan *plan, double info[4]);

/* Free the work arrays a plan keeps between calls (batched task lists and
 * arenas, a sigma plan's FP64 intermediates, Davidson vectors, the captured
 * sigma graph); they are rebuilt by the next call that needs them.  No
 * reference counterpart: for callers that hold many plans at once and need the
 * device memory back.  A sigma plan allocates its FP64 intermediates with its
 * first FP64 sigma, and a batched sigma whose per-vector arenas would not fit
 * in free device memory runs its vectors one after another (same results). */
int mq_plan_release(mq_plan *plan);

/* -- two-site split on the device ---------------------------------------- */
/* TwoSiteTensor::split_mps_l2r / split_mps_r2l (dmrg/framework/dmrg/
 * mp_tensors/twositetensor.hpp:141-181) with the block SVD of svd_truncate
 * (block_matrix/block_matrix_algorithms.h:211-333).  The two-site tensor psi
 * is a device vector in the canonical left-paired layout of
 * (phys1*phys2, left, right) -- the layout of the sigma vectors, for
 * MQ_SYMM_SU2U1 in the two-site reduced coupling (ts_reduction.h) -- and is
 * split along the bond between the two sites:
 *   1. mq_split_create builds (host) the map from psi to the blocks of its
 *      both-paired matrix [(phys1, left), (-phys2, right)] (ts_reduction.h
 *      unreduce_left, then ts_reshape.h:88-145), one block per charge K;
 *   2. mq_split_decompose runs the map and a one-sided Jacobi SVD of every
 *      block on the device (FP64) and returns the singular values;
 *   3. the caller picks the number of singular triplets kept per block -- the
 *      reference's global rule: the largest Mmax overall, at least
 *      cutoff * the largest -- and mq_split_assemble writes the two site
 *      tensors to device arrays. */
typedef struct mq_split mq_split;
typedef struct mq_split_desc {
    int32_t symm;            /* enum mq_symm */
    mq_index phys1, phys2;   /* physical indices of the first and second site */
    mq_index left, right;    /* left index of the first site, right index of the second */
} mq_split_desc;
int mq_split_create(mq_ctx *ctx, const mq_split_desc *desc, mq_split **out);
int mq_split_destroy(mq_split *sp);
/* The blocks, in descending charge order: block b has charge charge[2b..2b+1],
 * rows[b] = size of (phys1, left) at that charge, cols[b] = size of
 * (-phys2, right), and min(rows[b], cols[b]) singular values; *vec_len is the
 * length of psi.  Arrays stay owned by the handle. */
int mq_split_blocks(const mq_split *sp, int32_t *nblocks, const int32_t **charge, const int32_t **rows,
                    const int32_t **cols, int64_t *vec_len);
/* Block SVD of psi (device, read).  s (host, written): block after block, the
 * min(rows, cols) singular values of each in descending order.  Fails with
 * MQ_ERR_STATE when a block's Jacobi sweeps do not converge.  Synchronous. */
int mq_split_decompose(mq_split *sp, const double *psi, double *s);
/* Jacobi sweeps each block took in the last decomposition (diagnostic). */
int mq_split_sweeps(const mq_split *sp, int32_t *sweeps);
/* Site tensors from the last decomposition, keeping keep[b] (0..min(rows,
 * cols)) leading singular triplets of block b.  l2r != 0 (split_mps_l2r):
 * A = U, B = S V^T; l2r == 0 (split_mps_r2l): A = U S, B = V^T.  A (device,
 * written): block b column-major rows[b] x keep[b] at a_off[b] -- the
 * canonical left-paired layout of (phys1, left, kept bond) when a_off follows
 * it; B (device, written): block b row-major keep[b] x cols[b] at b_off[b]
 * (the right-paired block of (phys2, kept bond, right)).  Synchronous. */
int mq_split_assemble(mq_split *sp, const int32_t *keep, const int64_t *a_off, const int64_t *b_off, int l2r,
                      double *A, double *B);

/* -- diagnostics (no reference counterpart) ------------------------------ */
/* C(M x N, ldc) = op(A) op(B) with the library's grouped FP64 GEMM kernel on
 * device arrays (column-major; transA: A stored K x M, transB: B stored
 * N x K), K cut into chunks of about kchunk (>= 16) when the output is small.
 * Synchronous.  Used by the kernel tests. */
int mq_gemm_f64(mq_ctx *ctx, int transA, int transB, int M, int N, int K,
                const double *A, int lda, const double *B, int ldb,
                double *C, int ldc, int kchunk);
/* mq_gemm_f64 without any allocation (no reference counterpart: for callers
 * that cannot allocate device memory, e.g. code under HIP graph capture or
 * tight serving loops).  The task lists and split-K partial tiles go to a
 * fixed per-device scratch array of 32 MiB held by the library until the
 * process ends; a product that needs more fails with MQ_ERR_ARG and a message
 * naming the size.  Calls that overlap on one GPU share that array and race.
 * Same arguments, kernels and results (bit for bit) as mq_gemm_f64. */
int mq_gemm_f64_noalloc(mq_ctx *ctx, int transA, int transB, int M, int N, int K,
                        const double *A, int lda, const double *B, int ldb,
                        double *C, int ldc, int kchunk);
/* The same for the FP32-operand kernel of the mixed forms, with the tile shape
 * tm x tn (tm 64 or 128, tn a multiple of 16 up to 128): FP32 products, or the
 * BF16x3 products of MQ_PREC_BF16X3 when bf16x3 != 0. */
int mq_gemm_f32(mq_ctx *ctx, int transA, int transB, int M, int N, int K,
                const float *A, int lda, const float *B, int ldb,
                float *C, int ldc, int kchunk, int tm, int tn, int bf16x3);

/* -- host-side plan introspection (no GPU needed) ------------------------ */
/* Builds the same task lists the device plans execute, on the host only, so
 * that schedules, FLOP counts and workspace sizes can be inspected (and the
 * task lists checked) on machines without a GPU.  kind: 0 sigma, 1 left
 * step, 2 right step, 3 diagonal, plus MQ_HOST_FUSE(v), MQ_HOST_FUSE_RATIO(r)
 * and MQ_HOST_FUSE_MIN(m) to build with those MQ_OPT_FUSE* values (0 keeps a
 * default), and MQ_HOST_SCHED_REFERENCE for the reference's per-site schedule.  Arrays stay owned by the handle. */
#define MQ_HOST_FUSE(v) (((v) + 1) << 8)
#define MQ_HOST_FUSE_RATIO(r) ((r) << 10) /* 1..255 */
#define MQ_HOST_FUSE_MIN(m) ((m) << 18)   /* 1..4095 */
#define MQ_HOST_SCHED_REFERENCE (1 << 30) /* sigma plans with MQ_SCHED_REFERENCE */
typedef struct mq_host_plan mq_host_plan;
enum mq_plan_part {
    MQ_PART_ARR_L = 0, MQ_PART_ARR_R = 1, MQ_PART_RESHAPE = 2, MQ_PART_G1_GROUPS = 3,
    MQ_PART_G1_TILES = 4, MQ_PART_STRIPS = 5, MQ_PART_SRCS = 6, MQ_PART_G2_GROUPS = 7,
    MQ_PART_G2_TILES = 8, MQ_PART_SCATTER = 9, MQ_PART_COMB_STRIPS = 10, MQ_PART_COMB_SRCS = 11,
    MQ_PART_G1_ROWMAP = 12, MQ_PART_G2_ROWMAP = 13,
    MQ_PART_LIVE = 14 /* sigma plans: the mq_plan_live_mask bytes */
};
int mq_host_plan_create(const mq_site_desc *desc, int kind, mq_host_plan **out);
int mq_host_plan_part(const mq_host_plan *hp, int part, const void **ptr, int64_t *count, int64_t *elem_bytes);
/* arena sizes in elements: [vec, LA, RB, PRP, T, C, TMP, OUT, WS1, WS2] */
int mq_host_plan_sizes(const mq_host_plan *hp, int64_t sizes[10]);
int mq_host_plan_flops(const mq_host_plan *hp, double *ref_flops, double *exec_flops, double *min_bytes);
int mq_host_plan_out_desc(const mq_host_plan *hp, mq_boundary_desc *out);
/* 0 lbtm, 1 rbtm (sigma plans; see MQ_OPT_SCHEDULE). */
int mq_host_plan_schedule(const mq_host_plan *hp);
int mq_host_plan_destroy(mq_host_plan *hp);
/* Host-side check of mq_split's gather map (no GPU): the both-paired blocks
 * of psi (host), block after block in mq_split_blocks order, each row-major
 * rows x cols, computed on the host from the same task list the device runs.
 * *nblocks and *both_len (the sum of rows[b] * cols[b]) are set; with
 * both == NULL nothing else is written. */
int mq_host_split_both(const mq_split_desc *desc, const double *psi, int32_t *nblocks, int64_t *both_len,
                       double *both);
const char *mq_host_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MQGPU_H */
