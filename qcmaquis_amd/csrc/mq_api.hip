// C ABI of the MI355X hot path (declared in include/mqgpu.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "mq_kernels.hpp"

using namespace mq;

// Scratch of the allocation-free entry point (mq_gemm_f64_noalloc): one fixed
// array per GPU, held until the process ends.  Host code never uses the
// array's name as an address: set-up (mq_ctx_create, once per device, with the
// device current) asks for its device address with hipGetSymbolAddress and
// keeps it per device; kernels and copies get that address.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device -> address of `scratch` on that device

// device address of the scratch array on `device` (which must be current)
void *scratch_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) return it->second;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return ptr;
}
}  // namespace

struct mq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    void *scratch = nullptr;  // this device's `scratch` array (device address)
    std::string err;
    FuseRule fuse;  // MQ_OPT_FUSE, MQ_OPT_FUSE_RATIO, MQ_OPT_FUSE_MIN
    int kchunk = 2048;  // MQ_OPT_KCHUNK
    int tile_m = 0;     // MQ_OPT_TILE_M (0: automatic)
    int tile_n = 0;     // MQ_OPT_TILE_N (0: automatic)
    int ref_flops = 1;  // MQ_OPT_REF_FLOPS
    int graph = 0;      // MQ_OPT_GRAPH
    int schedule = SCHED_LBTM;  // MQ_OPT_SCHEDULE
};

namespace {

template <class T>
struct DevArray {
    T *p = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray &) = delete;  // owns its device memory
    DevArray &operator=(const DevArray &) = delete;
    void alloc(size_t count) {
        free_();
        n = count;
        if (count) {
            if (hipMalloc((void **)&p, sizeof(T) * count) != hipSuccess) {
                p = nullptr;
                throw std::runtime_error("hipMalloc failed");
            }
        }
    }
    template <class A>
    void upload(const std::vector<T, A> &v, hipStream_t s) {
        alloc(v.size());
        if (!v.empty() && hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s) != hipSuccess)
            throw std::runtime_error("hipMemcpyAsync failed");
    }
    void free_() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevArray() { free_(); }
};

struct DevGemm {
    DevArray<GemmGroup> groups;
    DevArray<GemmTile> tiles;
    DevArray<int> split;
    DevArray<RowMap> rowmap;
    int ntiles = 0, nsplit = 0;
    void upload(const GemmList &g, hipStream_t s) {
        for (auto &x : g.groups) {
            if ((x.tm != 64 && x.tm != 128) || x.tn < 16 || x.tn > 128 || x.tn % 16 != 0)
                throw std::runtime_error("GEMM group with an unsupported tile shape");
            if (x.rowmap < 0 || (x.rowmap > 0 && (x.split != 1 || x.rowmap - 1 + x.M > (int64_t)g.rowmap.size())))
                throw std::runtime_error("GEMM group with an invalid row map");
        }
        rowmap.upload(g.rowmap, s);
        for (auto &t : g.tiles)
            if (t.g >= (int)g.groups.size()) throw std::runtime_error("GEMM tile outside the group list");
        groups.upload(g.groups, s);
        tiles.upload(g.tiles, s);
        std::vector<int> sp;
        for (size_t i = 0; i < g.groups.size(); ++i)
            if (g.groups[i].split > 1) sp.push_back((int)i);
        split.upload(sp, s);
        ntiles = (int)g.tiles.size();
        nsplit = (int)sp.size();
    }
};

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// packs must cover the strips in order, each with 1..kPackStrips strips
void check_packs(const StripList &sl) {
    const auto &p = sl.packs;
    if (sl.strips.empty() && p.size() <= 1) return;
    if (p.size() < 2 || p.front() != 0 || p.back() != (int)sl.strips.size())
        throw std::runtime_error("strip packs do not cover the strip list");
    for (size_t i = 0; i + 1 < p.size(); ++i)
        if (p[i + 1] - p[i] < 1 || p[i + 1] - p[i] > kPackStrips) throw std::runtime_error("strip pack out of range");
}
int npacks(const StripList &sl) { return sl.packs.empty() ? 0 : (int)sl.packs.size() - 1; }

}  // namespace

struct mq_plan {
    mq_ctx *ctx = nullptr;
    HostPlan hp;
    DevArray<Copy2D> arrL, arrR, reshape, scatter;
    DevArray<Strip> strips, cstrips;
    DevArray<Src> srcs, csrcs;
    DevArray<int> packs, cpacks;  // strip packs of wapply / combL
    DevGemm g1, g2;
    DevArray<double> la, rb, prp, t, c, tmp, ws;
    double w_src_el = 0, w_dst_el = 0;  // W-application source / destination elements
    const double *bound_L = nullptr, *bound_R = nullptr;
    // mixed form: FP32 copies of the panels and FP32 intermediates (allocated on first use)
    DevArray<float> la32, rb32, prp32, t32, c32, ws32;
    bool conv32 = false;  // la32 / rb32 hold the current binding
    // Davidson workspace
    DevArray<double> jd_vecs, jd_partial, jd_coef;
    DevArray<const double *> jd_ptrs;
    int jd_cap = 0;
    // pinned host staging of the Davidson coefficients / pointers / partial sums
    double *h_coef = nullptr, *h_partial = nullptr;
    const double **h_ptrs = nullptr;
    // several vectors per launch (mq_sigma_batch): replicated task lists per batch
    // size, and per-vector copies of the psi, T and C arenas for the largest size used
    struct BatchLists {  // the W application reuses the plan's own strips for every vector
        int64_t ncopies = 0, nscatter = 0, ws = 0;
        DevArray<Copy2D> reshape, scatter;  // scatter: rbtm plans' right-to-left reshape
        DevGemm g1, g2;
    };
    std::map<int, std::unique_ptr<BatchLists>> blists;
    struct BatchArenas {  // vectors the FP64 (cap) and FP32 (cap32) arenas hold
        int cap = 0, cap32 = 0, tcap = 0;
        int64_t ws = 0;
        DevArray<double> prp, t, c, ws64, tmp;  // tmp: rbtm plans' right-paired sigma
        DevArray<float> prp32, t32, c32, ws32;
    } barena;
    // captured sigma stages on fixed staging vectors (mq_sigma_graph)
    hipGraphExec_t gexec = nullptr;
    int gprec = -1;
    const double *gL = nullptr, *gR = nullptr;
    DevArray<double> gx, gy;
    // last eigensolver call (mq_plan_solve_info): iterations in the requested form,
    // FP64 continuation iterations, the form's last Ritz value (lowest root) and the
    // FP64 Rayleigh quotient of its Ritz vector
    double solve_info[4] = {0, 0, 0, 0};
    ~mq_plan() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (h_coef) (void)hipHostFree(h_coef);
        if (h_partial) (void)hipHostFree(h_partial);
        if (h_ptrs) (void)hipHostFree((void *)h_ptrs);
    }
};

static int fail(mq_ctx *ctx, const std::exception &e, int code) {
    if (ctx) ctx->err = e.what();
    return code;
}

extern "C" {

int mq_ctx_create(int device, mq_ctx **out) {
    if (!out) return MQ_ERR_ARG;
    try {
        std::unique_ptr<mq_ctx> c(new mq_ctx());
        c->device = device;
        check(hipSetDevice(device), "hipSetDevice");
        check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        c->scratch = scratch_setup(device);
        *out = c.release();
        return MQ_OK;
    } catch (std::exception &e) {
        fprintf(stderr, "mq_ctx_create: %s\n", e.what());
        return MQ_ERR_HIP;
    }
}

int mq_ctx_destroy(mq_ctx *ctx) {
    if (!ctx) return MQ_ERR_ARG;
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MQ_OK;
}

const char *mq_last_error(const mq_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mq_ctx_set_option(mq_ctx *ctx, int option, int value) {
    if (!ctx) return MQ_ERR_ARG;
    if (option == MQ_OPT_FUSE && value >= 0 && value <= 2) {
        ctx->fuse.mode = value;
        return MQ_OK;
    }
    if (option == MQ_OPT_FUSE_RATIO && value >= 1 && value <= 255) {
        ctx->fuse.ratio = value;
        return MQ_OK;
    }
    if (option == MQ_OPT_FUSE_MIN && value >= 1 && value <= 4095) {
        ctx->fuse.min_entries = value;
        return MQ_OK;
    }
    if (option == MQ_OPT_KCHUNK && value >= 256 && value <= 65536 && value % 16 == 0) {
        ctx->kchunk = value;
        return MQ_OK;
    }
    if (option == MQ_OPT_REF_FLOPS && (value == 0 || value == 1)) {
        ctx->ref_flops = value;
        return MQ_OK;
    }
    if (option == MQ_OPT_TILE_M && (value == 0 || value == 64 || value == 128)) {
        ctx->tile_m = value;
        return MQ_OK;
    }
    if (option == MQ_OPT_TILE_N && (value == 0 || (value >= 16 && value <= 128 && value % 16 == 0))) {
        ctx->tile_n = value;
        return MQ_OK;
    }
    if (option == MQ_OPT_GRAPH && (value == 0 || value == 1)) {
        ctx->graph = value;
        return MQ_OK;
    }
    if (option == MQ_OPT_SCHEDULE && (value == SCHED_LBTM || value == SCHED_REFERENCE)) {
        ctx->schedule = value;
        return MQ_OK;
    }
    ctx->err = "mq_ctx_set_option: unknown option or value";
    return MQ_ERR_ARG;
}

void *mq_ctx_stream(mq_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int mq_ctx_synchronize(mq_ctx *ctx) {
    if (!ctx) return MQ_ERR_ARG;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ctx->err = hipGetErrorString(e);
        return MQ_ERR_HIP;
    }
    return MQ_OK;
}

static int make_plan(mq_ctx *ctx, const mq_site_desc *desc, mq_plan **out, int kind) {
    if (!ctx || !desc || !out) return MQ_ERR_ARG;
    try {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        std::unique_ptr<mq_plan> p(new mq_plan());
        p->ctx = ctx;
        const int tm = ctx->tile_m, tn = ctx->tile_n, kc = ctx->kchunk;
        if (kind == PLAN_SIGMA)
            p->hp = build_sigma_plan(*desc, tm, tn, kc, ctx->fuse, ctx->ref_flops != 0, ctx->schedule);
        else if (kind == PLAN_LSTEP) p->hp = build_lstep_plan(*desc, tm, tn, kc, ctx->fuse, ctx->ref_flops != 0);
        else if (kind == PLAN_RSTEP) p->hp = build_rstep_plan(*desc, tm, tn, kc);
        else p->hp = build_diag_plan(*desc, tm, tn, kc);
        hipStream_t s = ctx->stream;
        HostPlan &h = p->hp;
        p->arrL.upload(h.arrL, s);
        p->arrR.upload(h.arrR, s);
        p->reshape.upload(h.reshape, s);
        p->scatter.upload(h.scatter, s);
        p->strips.upload(h.wapply.strips, s);
        p->srcs.upload(h.wapply.srcs, s);
        p->cstrips.upload(h.combL.strips, s);
        p->csrcs.upload(h.combL.srcs, s);
        check_packs(h.wapply);
        check_packs(h.combL);
        p->packs.upload(h.wapply.packs, s);
        p->cpacks.upload(h.combL.packs, s);
        p->g1.upload(h.g1, s);
        p->g2.upload(h.g2, s);
        p->la.alloc(h.n_la);
        p->rb.alloc(h.n_rb);
        p->prp.alloc(h.n_prp);
        // sigma plans allocate the FP64 T and C arenas with their first FP64 sigma
        // (the mixed forms use FP32 ones), so a plan run only in a mixed form never
        // holds them
        if (kind != PLAN_SIGMA) {
            p->t.alloc(h.n_t);
            p->c.alloc(h.n_c);
        }
        p->tmp.alloc(h.n_tmp);
        p->ws.alloc(std::max(h.g1.ws, h.g2.ws));
        for (auto &st : h.wapply.strips) {
            p->w_src_el += (double)st.rows * st.cols * (st.s1 - st.s0);
            p->w_dst_el += (double)st.rows * st.cols;
        }
        // diagonal plans: T columns that no strip writes stay zero for the plan's life
        if (kind == PLAN_DIAG && h.n_t) check(hipMemsetAsync(p->t.p, 0, sizeof(double) * h.n_t, s), "clear T");
        check(hipStreamSynchronize(s), "plan upload");
        *out = p.release();
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(ctx, e, MQ_ERR_PLAN);
    }
}

int mq_sigma_plan_create(mq_ctx *ctx, const mq_site_desc *desc, mq_plan **out) {
    return make_plan(ctx, desc, out, PLAN_SIGMA);
}
int mq_lstep_plan_create(mq_ctx *ctx, const mq_site_desc *desc, mq_plan **out) {
    return make_plan(ctx, desc, out, PLAN_LSTEP);
}
int mq_rstep_plan_create(mq_ctx *ctx, const mq_site_desc *desc, mq_plan **out) {
    return make_plan(ctx, desc, out, PLAN_RSTEP);
}
int mq_diag_plan_create(mq_ctx *ctx, const mq_site_desc *desc, mq_plan **out) {
    return make_plan(ctx, desc, out, PLAN_DIAG);
}

int mq_plan_destroy(mq_plan *plan) {
    if (!plan) return MQ_ERR_ARG;
    (void)hipStreamSynchronize(plan->ctx->stream);
    delete plan;
    return MQ_OK;
}

int64_t mq_plan_vec_len(const mq_plan *plan) { return plan ? plan->hp.vec_len : -1; }

int mq_plan_flops(const mq_plan *plan, double *flops, double *bytes) {
    if (!plan) return MQ_ERR_ARG;
    if (flops) *flops = plan->hp.ref_flops;
    if (bytes) *bytes = plan->hp.min_bytes;
    return MQ_OK;
}

int64_t mq_plan_workspace_bytes(const mq_plan *plan) {
    if (!plan) return -1;
    const HostPlan &h = plan->hp;
    return 8 * (h.n_la + h.n_rb + h.n_prp + h.n_t + h.n_c + h.n_tmp + std::max(h.g1.ws, h.g2.ws));
}

// Executed (not algorithmic) FLOPs of this implementation.
double mq_plan_exec_flops(const mq_plan *plan) { return plan ? plan->hp.exec_flops : -1; }

// L boundary -> LA panels: copies of the stored blocks, then the fused rows.
static void arrange_left(mq_plan *p, const double *L, hipStream_t s) {
    HostPlan &h = p->hp;
    check(launch_copy2d(p->arrL.p, (int)h.arrL.size(), L, p->la.p, s), "arrange L");
    check(launch_wapply(p->cstrips.p, p->cpacks.p, npacks(h.combL), p->csrcs.p, p->la.p, p->la.p, ST_FUSE, s), "fuse L rows");
}

// (Re)bind the boundaries: arrange them into the FP64 panels when the pointer
// pair changes.
static void bind(mq_plan *p, const double *L, const double *R, hipStream_t s) {
    if (L == p->bound_L && R == p->bound_R) return;
    arrange_left(p, L, s);
    check(launch_copy2d(p->arrR.p, (int)p->hp.arrR.size(), R, p->rb.p, s), "arrange R");
    p->bound_L = L;
    p->bound_R = R;
    p->conv32 = false;
}

// Frees the batched arenas (rebuilt by the next batched sigma that needs them).
static void free_batch_arenas(mq_plan *p) {
    check(hipStreamSynchronize(p->ctx->stream), "release sync");
    auto &A = p->barena;
    A.prp.free_();
    A.t.free_();
    A.c.free_();
    A.ws64.free_();
    A.prp32.free_();
    A.t32.free_();
    A.c32.free_();
    A.ws32.free_();
    A.tmp.free_();
    A.cap = A.cap32 = A.tcap = 0;
    A.ws = 0;
}

// Allocates a plan's single-vector arenas with f; when device memory is short,
// frees the batched arenas (a block solve switching form holds the other form's)
// and tries once more.
static void alloc_or_free_batches(mq_plan *p, const std::function<void()> &f) {
    try {
        f();
    } catch (std::runtime_error &) {
        (void)hipGetLastError();  // the failed allocation's error must not reach the next launch check
        free_batch_arenas(p);
        f();
    }
}

// FP64 T and C arenas of a sigma plan (allocated with its first FP64 sigma)
static void ensure64(mq_plan *p) {
    HostPlan &h = p->hp;
    alloc_or_free_batches(p, [&] {
        if (!p->t.p && h.n_t) p->t.alloc(h.n_t);
        if (!p->c.p && h.n_c) p->c.alloc(h.n_c);
    });
}

// FP32 panels of the current binding (mixed form)
static void bind32(mq_plan *p, hipStream_t s) {
    HostPlan &h = p->hp;
    if (!p->t32.p && h.n_t) {
        alloc_or_free_batches(p, [&] {
            p->la32.alloc(h.n_la);
            p->rb32.alloc(h.n_rb);
            p->prp32.alloc(h.n_prp);
            p->t32.alloc(h.n_t);
            p->c32.alloc(h.n_c);
            p->ws32.alloc(std::max(h.g1.ws, h.g2.ws));
        });
    }
    if (p->conv32) return;
    check(launch_to_f32(p->la.p, p->la32.p, h.n_la, s), "LA to FP32");
    check(launch_to_f32(p->rb.p, p->rb32.p, h.n_rb, s), "RB to FP32");
    p->conv32 = true;
}

// Stages of one sigma after binding; ev (optional) records events between them.
static void sigma_stages(mq_plan *p, const double *x, double *y, int prec, hipEvent_t *ev) {
    hipStream_t s = p->ctx->stream;
    HostPlan &h = p->hp;
    const bool x3 = prec == MQ_PREC_BF16X3;
    if (prec == MQ_PREC_FP64 && ((!p->t.p && h.n_t) || (!p->c.p && h.n_c)))
        throw std::runtime_error("sigma: FP64 arenas not allocated");
    auto mark = [&](int i) {
        if (ev) check(hipEventRecord(ev[i], s), "record");
    };
    if (h.rbtm) {
        // rbtm schedule: t = X RC, W application into BUF, S = LA^T BUF, S -> y
        if (prec == MQ_PREC_FP64) {
            check(launch_copy2d(p->reshape.p, (int)h.reshape.size(), x, p->prp.p, s, true), "copy psi");
            mark(2);
            check(launch_gemm(p->g1.groups.p, p->g1.tiles.p, p->g1.ntiles, p->g1.split.p, p->g1.nsplit, p->prp.p,
                              p->rb.p, p->t.p, p->ws.p, p->g1.rowmap.p, ST_SIGMA, s),
                  "gemm t");
            mark(3);
            check(launch_wapply(p->strips.p, p->packs.p, npacks(h.wapply), p->srcs.p, p->t.p, p->c.p, ST_SIGMA, s),
                  "W apply");
            mark(4);
            check(launch_gemm(p->g2.groups.p, p->g2.tiles.p, p->g2.ntiles, p->g2.split.p, p->g2.nsplit, p->la.p,
                              p->c.p, p->tmp.p, p->ws.p, p->g2.rowmap.p, ST_SIGMA_OUT, s),
                  "gemm L");
        } else {
            check(launch_copy2d(p->reshape.p, (int)h.reshape.size(), x, p->prp32.p, s, true), "copy psi");
            mark(2);
            check(launch_gemm(p->g1.groups.p, p->g1.tiles.p, p->g1.ntiles, p->g1.split.p, p->g1.nsplit, p->prp32.p,
                              p->rb32.p, p->t32.p, p->ws32.p, p->g1.rowmap.p, ST_SIGMA, s, x3),
                  "gemm t");
            mark(3);
            check(launch_wapply(p->strips.p, p->packs.p, npacks(h.wapply), p->srcs.p, p->t32.p, p->c32.p, ST_SIGMA,
                                s),
                  "W apply");
            mark(4);
            check(launch_gemm(p->g2.groups.p, p->g2.tiles.p, p->g2.ntiles, p->g2.split.p, p->g2.nsplit, p->la32.p,
                              p->c32.p, p->tmp.p, p->ws32.p, p->g2.rowmap.p, ST_SIGMA_OUT, s, x3),
                  "gemm L");
        }
        check(launch_copy2d(p->scatter.p, (int)h.scatter.size(), p->tmp.p, y, s, true), "reshape sigma");
        mark(5);
        return;
    }
    if (prec == MQ_PREC_FP64) {
        check(launch_copy2d(p->reshape.p, (int)h.reshape.size(), x, p->prp.p, s, true), "reshape");
        mark(2);
        check(launch_gemm(p->g1.groups.p, p->g1.tiles.p, p->g1.ntiles, p->g1.split.p, p->g1.nsplit, p->la.p,
                          p->prp.p, p->t.p, p->ws.p, p->g1.rowmap.p, ST_SIGMA, s),
              "gemm T");
        mark(3);
        check(launch_wapply(p->strips.p, p->packs.p, npacks(h.wapply), p->srcs.p, p->t.p, p->c.p, ST_SIGMA, s),
              "W apply");
        mark(4);
        check(launch_gemm(p->g2.groups.p, p->g2.tiles.p, p->g2.ntiles, p->g2.split.p, p->g2.nsplit, p->c.p,
                          p->rb.p, y, p->ws.p, p->g2.rowmap.p, ST_SIGMA_OUT, s),
              "gemm R");
    } else {
        check(launch_copy2d(p->reshape.p, (int)h.reshape.size(), x, p->prp32.p, s, true), "reshape");
        mark(2);
        check(launch_gemm(p->g1.groups.p, p->g1.tiles.p, p->g1.ntiles, p->g1.split.p, p->g1.nsplit, p->la32.p,
                          p->prp32.p, p->t32.p, p->ws32.p, p->g1.rowmap.p, ST_SIGMA, s, x3),
              "gemm T");
        mark(3);
        check(launch_wapply(p->strips.p, p->packs.p, npacks(h.wapply), p->srcs.p, p->t32.p, p->c32.p, ST_SIGMA, s),
              "W apply");
        mark(4);
        check(launch_gemm(p->g2.groups.p, p->g2.tiles.p, p->g2.ntiles, p->g2.split.p, p->g2.nsplit, p->c32.p,
                          p->rb32.p, y, p->ws32.p, p->g2.rowmap.p, ST_SIGMA_OUT, s, x3),
              "gemm R");
    }
    mark(5);
}

static void sigma_launch(mq_plan *p, const double *L, const double *R, const double *x, double *y, int prec) {
    hipStream_t s = p->ctx->stream;
    bind(p, L, R, s);
    if (prec != MQ_PREC_FP64) bind32(p, s);
    else ensure64(p);
    sigma_stages(p, x, y, prec, nullptr);
}

// Task lists of a batch of nvec vectors (built once per size) and arenas for at
// least nvec vectors.
static mq_plan::BatchLists &batch_setup(mq_plan *p, int nvec, bool f32) {
    HostPlan &h = p->hp;
    hipStream_t s = p->ctx->stream;
    auto &slot = p->blists[nvec];
    if (!slot) {
        std::unique_ptr<mq_plan::BatchLists> B(new mq_plan::BatchLists());
        B->reshape.upload(batch_copies(h.reshape, nvec, h.vec_len, h.n_prp), s);
        B->ncopies = (int64_t)h.reshape.size() * nvec;
        // per-vector operands advance by their arena size; shared ones (panels) by 0
        GemmList g1 = h.rbtm ? batch_gemm(h.g1, nvec, h.n_prp, 0, h.n_t) : batch_gemm(h.g1, nvec, 0, h.n_prp, h.n_t);
        GemmList g2 = h.rbtm ? batch_gemm(h.g2, nvec, 0, h.n_c, h.n_tmp) : batch_gemm(h.g2, nvec, h.n_c, 0, h.vec_len);
        if (h.rbtm) {
            B->scatter.upload(batch_copies(h.scatter, nvec, h.n_tmp, h.vec_len), s);
            B->nscatter = (int64_t)h.scatter.size() * nvec;
        }
        B->g1.upload(g1, s);
        B->g2.upload(g2, s);
        B->ws = std::max(g1.ws, g2.ws);
        check(hipStreamSynchronize(s), "batch upload");
        slot = std::move(B);
    }
    auto &A = p->barena;
    if (A.ws < slot->ws) {
        A.ws = slot->ws;
        A.cap = A.cap32 = 0;  // the split-K workspaces are reallocated with the arenas
    }
    if (A.tcap < nvec) {  // FP64 output staging of rbtm plans, used by every form
        check(hipStreamSynchronize(s), "batch arenas");
        A.tmp.alloc((size_t)h.n_tmp * nvec);
        A.tcap = nvec;
    }
    if (!f32 && A.cap < nvec) {
        check(hipStreamSynchronize(s), "batch arenas");
        A.prp.alloc((size_t)h.n_prp * nvec);
        A.t.alloc((size_t)h.n_t * nvec);
        A.c.alloc((size_t)h.n_c * nvec);
        A.ws64.alloc((size_t)A.ws);
        A.cap = nvec;
    }
    if (f32 && A.cap32 < nvec) {
        check(hipStreamSynchronize(s), "batch arenas");
        A.prp32.alloc((size_t)h.n_prp * nvec);
        A.t32.alloc((size_t)h.n_t * nvec);
        A.c32.alloc((size_t)h.n_c * nvec);
        A.ws32.alloc((size_t)A.ws);
        A.cap32 = nvec;
    }
    return *slot;
}

// Device bytes the batch arenas of nvec vectors in a form would add to what the
// plan holds now.
static double batch_arena_bytes(mq_plan *p, int nvec, bool f32) {
    HostPlan &h = p->hp;
    auto &A = p->barena;
    const int have = f32 ? A.cap32 : A.cap;
    if (have >= nvec) return 0.0;
    return (double)nvec * (h.n_prp + h.n_t + h.n_c) * (f32 ? 4.0 : 8.0) + (double)nvec * h.n_tmp * 8.0;
}

// Y(:, j) = H X(:, j) for nvec vectors stored one after another: one launch per
// stage for all of them; the LA and RB panels are shared.
static void sigma_batch_launch(mq_plan *p, const double *L, const double *R, const double *X, double *Y, int nvec,
                               int prec) {
    hipStream_t s = p->ctx->stream;
    bind(p, L, R, s);
    if (prec != MQ_PREC_FP64) bind32(p, s);
    else ensure64(p);
    // when the batch arenas would not fit in free device memory (with 10 % to
    // spare), the vectors run one after another through the plan's own arenas:
    // the same task lists on the same data, so the results are identical
    const double need = batch_arena_bytes(p, nvec, prec != MQ_PREC_FP64);
    if (need > 0) {
        size_t free_b = 0, total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "memory info");
        if (need > 0.9 * (double)free_b) {
            const int64_t n = p->hp.vec_len;
            for (int j = 0; j < nvec; ++j) sigma_stages(p, X + j * n, Y + j * n, prec, nullptr);
            return;
        }
    }
    mq_plan::BatchLists &Bl = batch_setup(p, nvec, prec != MQ_PREC_FP64);
    auto &A = p->barena;
    struct {
        DevArray<Copy2D> &reshape;
        DevGemm &g1, &g2;
        int64_t ncopies;
        DevArray<double> &prp, &t, &c, &ws;
        DevArray<float> &prp32, &t32, &c32, &ws32;
    } B{Bl.reshape, Bl.g1, Bl.g2, Bl.ncopies, A.prp, A.t, A.c, A.ws64, A.prp32, A.t32, A.c32, A.ws32};
    HostPlan &h = p->hp;
    if (h.rbtm) {
        const bool x3 = prec == MQ_PREC_BF16X3;
        if (prec == MQ_PREC_FP64) {
            check(launch_copy2d(B.reshape.p, (int)B.ncopies, X, B.prp.p, s, true), "copy psi");
            check(launch_gemm(B.g1.groups.p, B.g1.tiles.p, B.g1.ntiles, B.g1.split.p, B.g1.nsplit, B.prp.p, p->rb.p,
                              B.t.p, B.ws.p, B.g1.rowmap.p, ST_SIGMA, s),
                  "gemm t");
            check(launch_wapply(p->strips.p, p->packs.p, npacks(h.wapply), p->srcs.p, B.t.p, B.c.p, ST_SIGMA, s, nvec,
                                h.n_t, h.n_c),
                  "W apply");
            check(launch_gemm(B.g2.groups.p, B.g2.tiles.p, B.g2.ntiles, B.g2.split.p, B.g2.nsplit, p->la.p, B.c.p,
                              A.tmp.p, B.ws.p, B.g2.rowmap.p, ST_SIGMA_OUT, s),
                  "gemm L");
        } else {
            check(launch_copy2d(B.reshape.p, (int)B.ncopies, X, B.prp32.p, s, true), "copy psi");
            check(launch_gemm(B.g1.groups.p, B.g1.tiles.p, B.g1.ntiles, B.g1.split.p, B.g1.nsplit, B.prp32.p,
                              p->rb32.p, B.t32.p, B.ws32.p, B.g1.rowmap.p, ST_SIGMA, s, x3),
                  "gemm t");
            check(launch_wapply(p->strips.p, p->packs.p, npacks(h.wapply), p->srcs.p, B.t32.p, B.c32.p, ST_SIGMA, s,
                                nvec, h.n_t, h.n_c),
                  "W apply");
            check(launch_gemm(B.g2.groups.p, B.g2.tiles.p, B.g2.ntiles, B.g2.split.p, B.g2.nsplit, p->la32.p,
                              B.c32.p, A.tmp.p, B.ws32.p, B.g2.rowmap.p, ST_SIGMA_OUT, s, x3),
                  "gemm L");
        }
        check(launch_copy2d(Bl.scatter.p, (int)Bl.nscatter, A.tmp.p, Y, s, true), "reshape sigma");
        return;
    }
    if (prec == MQ_PREC_FP64) {
        check(launch_copy2d(B.reshape.p, (int)B.ncopies, X, B.prp.p, s, true), "reshape");
        check(launch_gemm(B.g1.groups.p, B.g1.tiles.p, B.g1.ntiles, B.g1.split.p, B.g1.nsplit, p->la.p, B.prp.p,
                          B.t.p, B.ws.p, B.g1.rowmap.p, ST_SIGMA, s),
              "gemm T");
        check(launch_wapply(p->strips.p, p->packs.p, npacks(h.wapply), p->srcs.p, B.t.p, B.c.p, ST_SIGMA, s, nvec,
                            h.n_t, h.n_c),
              "W apply");
        check(launch_gemm(B.g2.groups.p, B.g2.tiles.p, B.g2.ntiles, B.g2.split.p, B.g2.nsplit, B.c.p, p->rb.p, Y,
                          B.ws.p, B.g2.rowmap.p, ST_SIGMA_OUT, s),
              "gemm R");
    } else {
        check(launch_copy2d(B.reshape.p, (int)B.ncopies, X, B.prp32.p, s, true), "reshape");
        check(launch_gemm(B.g1.groups.p, B.g1.tiles.p, B.g1.ntiles, B.g1.split.p, B.g1.nsplit, p->la32.p,
                          B.prp32.p, B.t32.p, B.ws32.p, B.g1.rowmap.p, ST_SIGMA, s, prec == MQ_PREC_BF16X3),
              "gemm T");
        check(launch_wapply(p->strips.p, p->packs.p, npacks(h.wapply), p->srcs.p, B.t32.p, B.c32.p, ST_SIGMA, s,
                            nvec, h.n_t, h.n_c),
              "W apply");
        check(launch_gemm(B.g2.groups.p, B.g2.tiles.p, B.g2.ntiles, B.g2.split.p, B.g2.nsplit, B.c32.p,
                          p->rb32.p, Y, B.ws32.p, B.g2.rowmap.p, ST_SIGMA_OUT, s, prec == MQ_PREC_BF16X3),
              "gemm R");
    }
}

// The sigma stages captured once into a HIP graph on the plan's staging vectors
// (gx -> gy) and replayed: the graph holds kernel launches only (no memset or
// host work), and is captured again when the boundaries or the precision change.
static void sigma_graph_launch(mq_plan *p, const double *L, const double *R, const double *x, double *y,
                               int prec) {
    hipStream_t s = p->ctx->stream;
    bind(p, L, R, s);
    if (prec != MQ_PREC_FP64) bind32(p, s);
    else ensure64(p);  // before the capture: no allocation inside it
    const int64_t n = p->hp.vec_len;
    if (!p->gx.p && n) {
        p->gx.alloc(n);
        p->gy.alloc(n);
    }
    if (!p->gexec || p->gL != L || p->gR != R || p->gprec != prec) {
        if (p->gexec) check(hipGraphExecDestroy(p->gexec), "graph destroy");
        p->gexec = nullptr;
        hipGraph_t graph = nullptr;
        check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            sigma_stages(p, p->gx.p, p->gy.p, prec, nullptr);
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        check(hipStreamEndCapture(s, &graph), "end capture");
        hipError_t e = hipGraphInstantiate(&p->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check(e, "graph instantiate");
        p->gL = L;
        p->gR = R;
        p->gprec = prec;
    }
    check(hipMemcpyAsync(p->gx.p, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s), "graph input");
    check(hipGraphLaunch(p->gexec, s), "graph launch");
    check(hipMemcpyAsync(y, p->gy.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s), "graph output");
}

// sigma of the Davidson drivers: eager launches, or the captured graph (MQ_OPT_GRAPH)
static void solver_sigma(mq_plan *p, const double *L, const double *R, const double *x, double *y, int prec) {
    if (p->ctx->graph) sigma_graph_launch(p, L, R, x, y, prec);
    else sigma_launch(p, L, R, x, y, prec);
}

static const char *prec_error(int prec) {
    if (prec == MQ_PREC_FP64 || prec == MQ_PREC_FP32 || prec == MQ_PREC_BF16X3) return nullptr;
    return "unknown precision";
}

int mq_sigma(mq_plan *plan, const double *L, const double *R, const double *x, double *y, int prec) {
    if (!plan || !x || !y) return MQ_ERR_ARG;
    if (plan->hp.kind != PLAN_SIGMA) {
        plan->ctx->err = "mq_sigma: plan is not a sigma plan";
        return MQ_ERR_ARG;
    }
    if (const char *e = prec_error(prec)) {
        plan->ctx->err = std::string("mq_sigma: ") + e;
        return MQ_ERR_ARG;
    }
    try {
        sigma_launch(plan, L, R, x, y, prec);
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

int mq_sigma_batch(mq_plan *plan, const double *L, const double *R, const double *X, double *Y, int nvec,
                   int prec) {
    if (!plan || !X || !Y || nvec < 1 || nvec > 64) return MQ_ERR_ARG;
    if (plan->hp.kind != PLAN_SIGMA) {
        plan->ctx->err = "mq_sigma_batch: plan is not a sigma plan";
        return MQ_ERR_ARG;
    }
    if (const char *e = prec_error(prec)) {
        plan->ctx->err = std::string("mq_sigma_batch: ") + e;
        return MQ_ERR_ARG;
    }
    try {
        sigma_batch_launch(plan, L, R, X, Y, nvec, prec);
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

int mq_sigma_graph(mq_plan *plan, const double *L, const double *R, const double *x, double *y, int prec) {
    if (!plan || !x || !y) return MQ_ERR_ARG;
    if (plan->hp.kind != PLAN_SIGMA) {
        plan->ctx->err = "mq_sigma_graph: plan is not a sigma plan";
        return MQ_ERR_ARG;
    }
    if (const char *e = prec_error(prec)) {
        plan->ctx->err = std::string("mq_sigma_graph: ") + e;
        return MQ_ERR_ARG;
    }
    try {
        sigma_graph_launch(plan, L, R, x, y, prec);
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

// Stage-timed sigma: HIP events on the plan's stream around each stage.
// ms[0] arrange L+R (0 if already bound), ms[1] reshape, ms[2] T GEMM,
// ms[3] W application, ms[4] output GEMM.  work[] returns the algorithmic
// FLOPs of stages 2..4 and the bytes each stage must move at minimum.
int mq_sigma_timed(mq_plan *plan, const double *L, const double *R, const double *x, double *y, int prec,
                   float ms[5], double work[6]) {
    if (!plan || !x || !y || !ms || plan->hp.kind != PLAN_SIGMA) return MQ_ERR_ARG;
    if (const char *e = prec_error(prec)) {
        plan->ctx->err = std::string("mq_sigma_timed: ") + e;
        return MQ_ERR_ARG;
    }
    try {
        hipStream_t s = plan->ctx->stream;
        HostPlan &h = plan->hp;
        struct Events {  // destroyed on every exit path
            hipEvent_t e[6] = {};
            ~Events() {
                for (auto &x : e)
                    if (x) (void)hipEventDestroy(x);
            }
        } evs;
        hipEvent_t *ev = evs.e;
        for (int i = 0; i < 6; ++i) check(hipEventCreate(&ev[i]), "event");
        check(hipEventRecord(ev[0], s), "record");
        bind(plan, L, R, s);
        if (prec != MQ_PREC_FP64) bind32(plan, s);
        else ensure64(plan);
        check(hipEventRecord(ev[1], s), "record");
        sigma_stages(plan, x, y, prec, ev);
        check(hipEventSynchronize(ev[5]), "sync");
        for (int i = 0; i < 5; ++i) check(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]), "elapsed");
        if (work) {
            const double es = prec == MQ_PREC_FP64 ? 8.0 : 4.0;  // bytes per stored element
            const double srcel = plan->w_src_el, dstel = plan->w_dst_el;
            work[0] = h.g1.flops();
            work[1] = 2.0 * srcel;
            work[2] = h.g2.flops();
            // bytes each stage must move: T GEMM reads LA + PRP and writes T; the W
            // application reads every source slice and writes C; the output GEMM
            // reads C + RB and writes y (FP64)
            work[4] = es * (srcel + dstel);
            if (h.rbtm) {  // t GEMM reads psi + RB; L GEMM reads LA + BUF and writes S (FP64), then S -> y
                work[3] = es * (h.n_prp + h.n_rb + h.n_t);
                work[5] = es * (h.n_c + h.n_la) + 8.0 * (2.0 * h.n_tmp + h.vec_len);
            } else {
                work[3] = es * (h.n_la + h.n_prp + h.n_t);
                work[5] = es * (h.n_c + h.n_rb) + 8.0 * h.vec_len;
            }
        }
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

// C = op(A) op(B) on device arrays with the grouped GEMM kernel (one group).
int mq_gemm_f64(mq_ctx *ctx, int transA, int transB, int M, int N, int K, const double *A, int lda,
                const double *B, int ldb, double *C, int ldc, int kchunk) {
    if (!ctx || M < 0 || N < 0 || K < 0 || kchunk < 16) return MQ_ERR_ARG;
    try {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        GemmGroup g{};
        g.M = M;
        g.N = N;
        g.K = K;
        g.lda = lda;
        g.ldb = ldb;
        g.ldc = ldc;
        g.transA = transA ? 1 : 0;
        g.transB = transB ? 1 : 0;
        GemmList gl;
        tile_gemm(gl, g, 0, 0, kchunk);
        order_tiles(gl);
        DevGemm dg;
        dg.upload(gl, ctx->stream);
        DevArray<double> ws;
        ws.alloc((size_t)gl.ws);
        check(launch_gemm(dg.groups.p, dg.tiles.p, dg.ntiles, dg.split.p, dg.nsplit, A, B, C, ws.p, dg.rowmap.p, ST_MISC,
                          ctx->stream),
              "gemm");
        check(hipStreamSynchronize(ctx->stream), "gemm sync");
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(ctx, e, MQ_ERR_HIP);
    }
}

// mq_gemm_f64 without allocation: the task lists and the split-K partial tiles
// go to this device's scratch array (kScratchBytes), laid out as
//   [group | split group ids | tiles | partials], each part 256-byte aligned.
// The partials need no clearing (the split tiles write every element of
// theirs).  Calls that overlap on one GPU share the array and race.
int mq_gemm_f64_noalloc(mq_ctx *ctx, int transA, int transB, int M, int N, int K, const double *A, int lda,
                        const double *B, int ldb, double *C, int ldc, int kchunk) {
    if (!ctx || M < 0 || N < 0 || K < 0 || kchunk < 16) return MQ_ERR_ARG;
    try {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (!ctx->scratch) throw std::runtime_error("context without scratch");
        GemmGroup g{};
        g.M = M;
        g.N = N;
        g.K = K;
        g.lda = lda;
        g.ldb = ldb;
        g.ldc = ldc;
        g.transA = transA ? 1 : 0;
        g.transB = transB ? 1 : 0;
        GemmList gl;
        tile_gemm(gl, g, 0, 0, kchunk);
        order_tiles(gl);
        std::vector<int> sp;
        for (size_t i = 0; i < gl.groups.size(); ++i)
            if (gl.groups[i].split > 1) sp.push_back((int)i);
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t o_grp = 0;
        const size_t o_sp = o_grp + up(sizeof(GemmGroup) * gl.groups.size());
        const size_t o_til = o_sp + up(sizeof(int) * sp.size());
        const size_t o_ws = o_til + up(sizeof(GemmTile) * gl.tiles.size());
        const size_t need = o_ws + sizeof(double) * (size_t)gl.ws;
        if (need > kScratchBytes) {
            ctx->err = "mq_gemm_f64_noalloc: this product needs " + std::to_string(need) +
                       " bytes of scratch (task lists and split-K partials); the maximum is " +
                       std::to_string(kScratchBytes);
            return MQ_ERR_ARG;
        }
        unsigned char *base = static_cast<unsigned char *>(ctx->scratch);
        hipStream_t s = ctx->stream;
        if (!gl.groups.empty())
            check(hipMemcpyAsync(base + o_grp, gl.groups.data(), sizeof(GemmGroup) * gl.groups.size(),
                                 hipMemcpyHostToDevice, s),
                  "group upload");
        if (!sp.empty())
            check(hipMemcpyAsync(base + o_sp, sp.data(), sizeof(int) * sp.size(), hipMemcpyHostToDevice, s),
                  "split upload");
        if (!gl.tiles.empty())
            check(hipMemcpyAsync(base + o_til, gl.tiles.data(), sizeof(GemmTile) * gl.tiles.size(),
                                 hipMemcpyHostToDevice, s),
                  "tile upload");
        check(launch_gemm(reinterpret_cast<const GemmGroup *>(base + o_grp),
                          reinterpret_cast<const GemmTile *>(base + o_til), (int)gl.tiles.size(),
                          reinterpret_cast<const int *>(base + o_sp), (int)sp.size(), A, B, C,
                          reinterpret_cast<double *>(base + o_ws), nullptr, ST_MISC, s),
              "gemm");
        // the host task lists must outlive their copies
        check(hipStreamSynchronize(s), "gemm sync");
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(ctx, e, MQ_ERR_HIP);
    }
}

// The FP32-operand GEMM of the mixed forms on one group with a given tile
// shape: FP32 products, or BF16x3 products (bf16x3 != 0).
int mq_gemm_f32(mq_ctx *ctx, int transA, int transB, int M, int N, int K, const float *A, int lda, const float *B,
                int ldb, float *C, int ldc, int kchunk, int tm, int tn, int bf16x3) {
    if (!ctx || M < 0 || N < 0 || K < 0 || kchunk < 16 || (tm != 64 && tm != 128) || tn < 16 || tn > 128 ||
        tn % 16)
        return MQ_ERR_ARG;
    try {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        GemmGroup g{};
        g.M = M;
        g.N = N;
        g.K = K;
        g.lda = lda;
        g.ldb = ldb;
        g.ldc = ldc;
        g.transA = transA ? 1 : 0;
        g.transB = transB ? 1 : 0;
        GemmList gl;
        tile_gemm(gl, g, tm, tn, kchunk);
        order_tiles(gl);
        DevGemm dg;
        dg.upload(gl, ctx->stream);
        DevArray<float> ws;
        ws.alloc((size_t)gl.ws);
        check(launch_gemm(dg.groups.p, dg.tiles.p, dg.ntiles, dg.split.p, dg.nsplit, A, B, C, ws.p, dg.rowmap.p, ST_MISC,
                          ctx->stream, bf16x3 != 0),
              "gemm");
        check(hipStreamSynchronize(ctx->stream), "gemm sync");
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(ctx, e, MQ_ERR_HIP);
    }
}

// Forget the arranged boundaries (call after modifying L or R in place).
int mq_plan_rebind(mq_plan *plan) {
    if (!plan) return MQ_ERR_ARG;
    plan->bound_L = plan->bound_R = nullptr;
    plan->gL = plan->gR = nullptr;
    return MQ_OK;
}

// Arrange L and R into the plan's panels now, whatever was bound before (a new
// boundary may sit at the address of a freed one), and make them the binding
// that later calls with the same pointers reuse.
int mq_plan_bind(mq_plan *plan, const double *L, const double *R) {
    if (!plan || !L || !R) return MQ_ERR_ARG;
    if (plan->hp.kind != PLAN_SIGMA) {
        plan->ctx->err = "mq_plan_bind: plan is not a sigma plan";
        return MQ_ERR_ARG;
    }
    try {
        check(hipSetDevice(plan->ctx->device), "hipSetDevice");
        plan->bound_L = plan->bound_R = nullptr;
        plan->gL = plan->gR = nullptr;
        bind(plan, L, R, plan->ctx->stream);
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

int mq_plan_schedule(const mq_plan *plan) { return plan ? plan->hp.rbtm : MQ_ERR_ARG; }

int mq_plan_solve_info(const mq_plan *plan, double info[4]) {
    if (!plan || !info) return MQ_ERR_ARG;
    for (int i = 0; i < 4; ++i) info[i] = plan->solve_info[i];
    return MQ_OK;
}

int mq_plan_live_mask(const mq_plan *plan, uint8_t *mask) {
    if (!plan || !mask || plan->hp.kind != PLAN_SIGMA) return MQ_ERR_ARG;
    std::copy(plan->hp.live.begin(), plan->hp.live.end(), mask);
    return MQ_OK;
}

// Free the work arrays a plan keeps between calls: the batched task lists and
// arenas of mq_sigma_batch / mq_block_davidson, the Davidson vectors and the
// captured graph with its staging vectors.  They are rebuilt on the next call
// that needs them.  The FP64 and FP32 panels stay.
int mq_plan_release(mq_plan *plan) {
    if (!plan) return MQ_ERR_ARG;
    try {
        check(hipSetDevice(plan->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(plan->ctx->stream), "release sync");
        plan->blists.clear();
        free_batch_arenas(plan);
        if (plan->hp.kind == PLAN_SIGMA) {  // re-allocated by the next FP64 sigma
            plan->t.free_();
            plan->c.free_();
        }
        plan->jd_vecs.free_();
        plan->jd_cap = 0;
        if (plan->gexec) (void)hipGraphExecDestroy(plan->gexec);
        plan->gexec = nullptr;
        plan->gprec = -1;
        plan->gL = plan->gR = nullptr;
        plan->gx.free_();
        plan->gy.free_();
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

int mq_step_out_desc(const mq_plan *plan, mq_boundary_desc *out) {
    if (!plan || !out || plan->hp.kind == PLAN_SIGMA) return MQ_ERR_ARG;
    const HostPlan &h = plan->hp;
    out->aux_dim = h.out_aux;
    out->bptr = h.out_bptr.data();
    out->blocks.n = (int32_t)h.out_rows.size();
    out->blocks.lc = h.out_lc.data();
    out->blocks.rc = h.out_rc.data();
    out->blocks.rows = h.out_rows.data();
    out->blocks.cols = h.out_cols.data();
    out->blocks.offset = h.out_off.data();
    out->total = h.n_out;
    return MQ_OK;
}

int mq_boundary_step(mq_plan *plan, const double *in, const double *A, double *out) {
    if (!plan || !in || !A || (!out && plan->hp.n_out > 0)) return MQ_ERR_ARG;
    HostPlan &h = plan->hp;
    if (h.kind == PLAN_SIGMA) return MQ_ERR_ARG;
    hipStream_t s = plan->ctx->stream;
    try {
        check(launch_copy2d(plan->reshape.p, (int)h.reshape.size(), A, plan->prp.p, s), "reshape");
        if (h.kind == PLAN_LSTEP) {
            arrange_left(plan, in, s);
            check(launch_gemm(plan->g1.groups.p, plan->g1.tiles.p, plan->g1.ntiles, plan->g1.split.p,
                              plan->g1.nsplit, plan->la.p, plan->prp.p, plan->t.p, plan->ws.p, plan->g1.rowmap.p, ST_STEP, s),
                  "gemm T");
            check(launch_wapply(plan->strips.p, plan->packs.p, npacks(h.wapply), plan->srcs.p, plan->t.p, plan->c.p,
                                ST_STEP, s),
                  "W apply");
            check(launch_gemm(plan->g2.groups.p, plan->g2.tiles.p, plan->g2.ntiles, plan->g2.split.p,
                              plan->g2.nsplit, plan->c.p, A, plan->tmp.p, plan->ws.p, plan->g2.rowmap.p, ST_STEP_OUT, s),
                  "gemm out");
        } else {
            check(launch_copy2d(plan->arrR.p, (int)h.arrR.size(), in, plan->rb.p, s), "arrange R");
            check(launch_gemm(plan->g1.groups.p, plan->g1.tiles.p, plan->g1.ntiles, plan->g1.split.p,
                              plan->g1.nsplit, A, plan->rb.p, plan->t.p, plan->ws.p, plan->g1.rowmap.p, ST_STEP, s),
                  "gemm t");
            check(launch_wapply(plan->strips.p, plan->packs.p, npacks(h.wapply), plan->srcs.p, plan->t.p, plan->c.p,
                                ST_STEP, s),
                  "W apply");
            check(launch_gemm(plan->g2.groups.p, plan->g2.tiles.p, plan->g2.ntiles, plan->g2.split.p,
                              plan->g2.nsplit, plan->c.p, plan->prp.p, plan->tmp.p, plan->ws.p, plan->g2.rowmap.p, ST_STEP_OUT, s),
                  "gemm out");
        }
        check(launch_copy2d(plan->scatter.p, (int)h.scatter.size(), plan->tmp.p, out, s), "scatter");
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

int mq_diagonal(mq_plan *plan, const double *L, const double *R, double *diag) {
    if (!plan || !L || !R || !diag) return MQ_ERR_ARG;
    HostPlan &h = plan->hp;
    if (h.kind != PLAN_DIAG) {
        plan->ctx->err = "mq_diagonal: plan is not a diagonal plan";
        return MQ_ERR_ARG;
    }
    hipStream_t s = plan->ctx->stream;
    try {
        check(launch_copy2d(plan->arrL.p, (int)h.arrL.size(), L, plan->la.p, s), "L diagonals");
        check(launch_copy2d(plan->arrR.p, (int)h.arrR.size(), R, plan->rb.p, s), "R diagonals");
        check(launch_wapply(plan->strips.p, plan->packs.p, npacks(h.wapply), plan->srcs.p, plan->la.p, plan->t.p,
                            ST_MISC, s),
              "diagonal W apply");
        check(launch_gemm(plan->g2.groups.p, plan->g2.tiles.p, plan->g2.ntiles, plan->g2.split.p, plan->g2.nsplit,
                          plan->t.p, plan->rb.p, diag, plan->ws.p, plan->g2.rowmap.p, ST_MISC, s),
              "diagonal gemm");
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

}  // extern "C"

// ------------------------------------------------------------------ Jacobi-Davidson
namespace {

// Davidson vector operations on the plan's stream.  Pointer lists and
// coefficients are staged in pinned host slots that stay untouched until the
// next synchronisation, so linear combinations are enqueued without waiting;
// only dot products synchronise (one launch for a whole batch of pairs).
constexpr int kJdSlots = 4096, kJdPartialPairs = 128, kJdBlocks = 256;

struct JD {
    mq_plan *p;
    hipStream_t s;
    int64_t n;
    int nblk;
    int ptr_used = 0, coef_used = 0;

    double *vec(int i) { return p->jd_vecs.p + (int64_t)i * n; }

    void init() {
        if (!p->h_coef) {
            check(hipHostMalloc((void **)&p->h_coef, sizeof(double) * kJdSlots), "pinned coef");
            check(hipHostMalloc((void **)&p->h_partial, sizeof(double) * kJdPartialPairs * kJdBlocks),
                  "pinned partials");
            check(hipHostMalloc((void **)&p->h_ptrs, sizeof(double *) * kJdSlots), "pinned ptrs");
            p->jd_partial.alloc((size_t)kJdPartialPairs * kJdBlocks);
            p->jd_coef.alloc(kJdSlots);
            p->jd_ptrs.alloc(kJdSlots);
        }
    }
    const double **stage_ptrs(const std::vector<const double *> &v, int *slot) {
        if (ptr_used + (int)v.size() > kJdSlots) sync();
        *slot = ptr_used;
        for (size_t i = 0; i < v.size(); ++i) p->h_ptrs[ptr_used + i] = v[i];
        check(hipMemcpyAsync(p->jd_ptrs.p + ptr_used, p->h_ptrs + ptr_used, sizeof(double *) * v.size(),
                             hipMemcpyHostToDevice, s),
              "stage ptrs");
        ptr_used += (int)v.size();
        return p->jd_ptrs.p + *slot;
    }
    void sync() {
        check(hipStreamSynchronize(s), "davidson sync");
        ptr_used = coef_used = 0;
    }
    // dots of pairs (x_k, y_k), summed on the host in a fixed order (one
    // synchronisation per kJdPartialPairs pairs)
    std::vector<double> dots(const std::vector<const double *> &xs, const std::vector<const double *> &ys) {
        const int np_all = (int)xs.size();
        std::vector<double> r(np_all, 0.0);
        for (int k0 = 0; k0 < np_all; k0 += kJdPartialPairs) {
            const int np = std::min(kJdPartialPairs, np_all - k0);
            std::vector<const double *> ptrs(xs.begin() + k0, xs.begin() + k0 + np);
            ptrs.insert(ptrs.end(), ys.begin() + k0, ys.begin() + k0 + np);
            int slot;
            const double *const *d = stage_ptrs(ptrs, &slot);
            check(launch_multidot(d, d + np, np, n, nblk, p->jd_partial.p, s), "multidot");
            check(hipMemcpyAsync(p->h_partial, p->jd_partial.p, sizeof(double) * np * nblk, hipMemcpyDeviceToHost,
                                 s),
                  "dot partials");
            sync();
            for (int k = 0; k < np; ++k) {
                long double acc = 0;
                for (int b = 0; b < nblk; ++b) acc += p->h_partial[(size_t)k * nblk + b];
                r[k0 + k] = (double)acc;
            }
        }
        return r;
    }
    double dot(const double *x, const double *y) { return dots({x}, {y})[0]; }
    // out = beta*out + sum c_j v_j, enqueued without waiting
    void lincomb(const std::vector<const double *> &vs, const std::vector<double> &c, double beta, double *out) {
        int nv = (int)vs.size();
        if (coef_used + nv > kJdSlots || ptr_used + nv > kJdSlots) sync();
        int slot;
        const double *const *d = stage_ptrs(vs, &slot);
        int cs = coef_used;
        for (int j = 0; j < nv; ++j) p->h_coef[cs + j] = c[j];
        check(hipMemcpyAsync(p->jd_coef.p + cs, p->h_coef + cs, sizeof(double) * nv, hipMemcpyHostToDevice, s),
              "stage coef");
        coef_used += nv;
        check(launch_lincomb(d, p->jd_coef.p + cs, nv, n, beta, out, s), "lincomb");
    }
};

// cyclic Jacobi eigen-decomposition of a small symmetric matrix (column-major)
void sym_eig(int n, std::vector<double> A, std::vector<double> &w, std::vector<double> &V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[(size_t)q * n + p] * A[(size_t)q * n + p];
        if (off < 1e-30) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[(size_t)q * n + p];
                if (std::fabs(apq) < 1e-300) continue;
                double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                double th = (aqq - app) / (2 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[(size_t)p * n + k], akq = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * akp - sn * akq;
                    A[(size_t)q * n + k] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[(size_t)k * n + p], aqk = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * apk - sn * aqk;
                    A[(size_t)k * n + q] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[(size_t)p * n + k], vkq = V[(size_t)q * n + k];
                    V[(size_t)p * n + k] = c * vkp - sn * vkq;
                    V[(size_t)q * n + k] = sn * vkp + c * vkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = A[(size_t)i * n + i];
}

}  // namespace

// Convergence threshold of the iterations in form prec: with tol > 0 the mixed
// forms stop at their rounding floor (the residual relative to |theta| that
// their sigma vectors cannot resolve: FP32 about 4e-6, BF16x3 about 2e-5) when
// that is above tol, and the FP64 continuation finishes the solve; tol <= 0
// (run max_iter iterations) is kept as given.
static double phase_tol(double tol, int prec) {
    if (tol <= 0 || prec == MQ_PREC_FP64) return tol;
    return std::max(tol, prec == MQ_PREC_BF16X3 ? 2e-5 : 4e-6);
}

// The mixed iterations also end when they stagnate above that floor (a root
// whose rounding floor is a little higher): the largest open residual, relative
// to |theta|, has not halved for kStallIters iterations.
struct Stall {
    static constexpr int kStallIters = 8;
    bool on;
    double ref = -1.0;
    int since = 0;
    explicit Stall(bool active) : on(active) {}
    bool update(double rel) {
        if (!on) return false;
        if (ref < 0 || rel < 0.5 * ref) {
            ref = rel;
            since = 0;
            return false;
        }
        return ++since >= kStallIters;
    }
};

extern "C" int mq_jacobi_davidson(mq_plan *plan, const double *L, const double *R, const double *x0, double *u_out,
                                  double *theta_out, int max_iter, double tol, int prec, const double *const *ortho,
                                  int n_ortho, int *iterations) {
    if (!plan || !x0 || !u_out || !theta_out) return MQ_ERR_ARG;
    if (max_iter < 1 || max_iter > 1000 || n_ortho < 0 || n_ortho > 60) {
        plan->ctx->err = "mq_jacobi_davidson: max_iter must be in 1..1000 and n_ortho in 0..60";
        return MQ_ERR_ARG;
    }
    if (plan->hp.kind != PLAN_SIGMA) {
        plan->ctx->err = "mq_jacobi_davidson: needs a sigma plan";
        return MQ_ERR_ARG;
    }
    if (const char *e = prec_error(prec)) {
        plan->ctx->err = std::string("mq_jacobi_davidson: ") + e;
        return MQ_ERR_ARG;
    }
    try {
        const double ptol = phase_tol(tol, prec);
        Stall stall(prec != MQ_PREC_FP64 && tol > 0);
        JD jd;
        jd.p = plan;
        jd.s = plan->ctx->stream;
        jd.n = plan->hp.vec_len;
        jd.nblk = (int)std::max<int64_t>(1, std::min<int64_t>(kJdBlocks, (jd.n + 255) / 256));
        jd.init();
        const int64_t n = jd.n;
        int need = 2 * max_iter + 2;
        if (plan->jd_cap < need) {
            plan->jd_vecs.alloc((size_t)need * n);
            plan->jd_cap = need;
        }
        auto V = [&](int i) { return jd.vec(i); };
        auto VA = [&](int i) { return jd.vec(max_iter + i); };
        double *u = jd.vec(2 * max_iter), *uA = jd.vec(2 * max_iter + 1);
        std::vector<double> oo(n_ortho);
        for (int k = 0; k < n_ortho; ++k) oo[k] = jd.dot(ortho[k], ortho[k]);
        auto project = [&](double *t) {  // SingleSiteVS::project (ietl_lanczos_solver.h:90-95)
            if (n_ortho == 0) return;
            std::vector<const double *> xs, ys;
            for (int k = 0; k < n_ortho; ++k) {
                xs.push_back(ortho[k]);
                ys.push_back(t);
            }
            std::vector<double> d = jd.dots(xs, ys), c(n_ortho);
            for (int k = 0; k < n_ortho; ++k) c[k] = -d[k] / oo[k];
            jd.lincomb(std::vector<const double *>(ortho, ortho + n_ortho), c, 1.0, t);
        };
        check(hipMemcpyAsync(V(0), x0, sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "x0");
        project(V(0));
        std::vector<double> M((size_t)max_iter * max_iter, 0.0);
        int it = 0;
        double theta = 0;
        while (true) {
            // orthogonalise t against V (jacobi.h:384-393: Gram-Schmidt, repeated when the
            // norm drops below a quarter of its value), then normalise
            double *t = V(it);
            std::vector<const double *> xs{t}, ys{t};
            for (int i = 0; i < it; ++i) {
                xs.push_back(V(i));
                ys.push_back(t);
            }
            std::vector<double> d = jd.dots(xs, ys);
            const double tau = std::sqrt(d[0]);
            double nt2 = d[0];
            for (int pass = 0; pass < 2 && it > 0; ++pass) {
                std::vector<const double *> vv;
                std::vector<double> c;
                for (int i = 0; i < it; ++i) {
                    vv.push_back(V(i));
                    c.push_back(-d[1 + i]);
                }
                jd.lincomb(vv, c, 1.0, t);
                d = jd.dots(xs, ys);
                nt2 = d[0];
                if (std::sqrt(nt2) >= 0.25 * tau) break;
            }
            if (n_ortho > 0) {
                project(t);
                nt2 = jd.dot(t, t);
            }
            const double nt = std::sqrt(nt2);
            jd.lincomb({t}, {1.0 / nt - 1.0}, 1.0, t);
            solver_sigma(plan, L, R, t, VA(it), prec);
            // Rayleigh column (jacobi.h:401-402)
            xs.clear();
            ys.clear();
            for (int i = 0; i <= it; ++i) {
                xs.push_back(V(i));
                ys.push_back(VA(it));
            }
            std::vector<double> col = jd.dots(xs, ys);
            for (int i = 0; i <= it; ++i) M[(size_t)it * max_iter + i] = col[i];
            int dim = it + 1;
            std::vector<double> A((size_t)dim * dim);
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j <= i; ++j) {
                    double v = M[(size_t)i * max_iter + j];
                    A[(size_t)i * dim + j] = v;
                    A[(size_t)j * dim + i] = v;
                }
            std::vector<double> w, Z;
            sym_eig(dim, A, w, Z);
            int lo = 0;
            for (int i = 1; i < dim; ++i)
                if (w[i] < w[lo]) lo = i;
            theta = w[lo];
            std::vector<double> sv(Z.begin() + (size_t)lo * dim, Z.begin() + (size_t)(lo + 1) * dim);
            int big = 0;
            for (int i = 1; i < dim; ++i)
                if (std::fabs(sv[i]) > std::fabs(sv[big])) big = i;
            if (sv[big] < 0)
                for (auto &v : sv) v = -v;
            std::vector<const double *> vv, va;
            for (int j = 0; j < dim; ++j) {
                vv.push_back(V(j));
                va.push_back(VA(j));
            }
            jd.lincomb(vv, sv, 0.0, u);
            jd.lincomb(va, sv, 0.0, uA);
            project(uA);
            jd.lincomb({u}, {-theta}, 1.0, uA);  // r = uA - theta u
            ++it;
            // |r|, r.u and u.u in one batch: convergence test and correction t = -r + (r.u)/(u.u) u
            std::vector<double> rr = jd.dots({uA, uA, u}, {uA, u, u});
            double rn = std::sqrt(rr[0]);
            if (rn <= ptol * std::fabs(theta) || rn < ptol || it >= max_iter ||
                stall.update(rn / std::max(std::fabs(theta), 1e-300)))
                break;
            jd.lincomb({uA, u}, {-1.0, rr[1] / rr[2]}, 0.0, V(it));
        }
        const double theta_form = theta;
        const int its_form = it;
        bool refine = false;
        if (prec != MQ_PREC_FP64) {
            // FP64 defect correction: the eigenvalue is the FP64 Rayleigh quotient of the
            // final Ritz vector (its error is second order in the vector's error) ...
            sigma_launch(plan, L, R, u, uA, MQ_PREC_FP64);
            std::vector<double> d = jd.dots({u, u}, {uA, u});
            theta = d[0] / d[1];
            // ... and when its FP64 residual misses the tolerance (the mixed iterations
            // stop at their rounding floor, e.g. next to a nearly degenerate state), the
            // solve continues in FP64 from that vector
            if (tol > 0) {
                jd.lincomb({u}, {-theta}, 1.0, uA);
                const double rn = std::sqrt(jd.dot(uA, uA));
                refine = !(rn <= tol * std::fabs(theta) || rn < tol);
            }
        }
        const double theta_rq = theta;
        check(hipMemcpyAsync(u_out, u, sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "u out");
        jd.sync();
        if (refine) {
            int its2 = 0;
            const int rc = mq_jacobi_davidson(plan, L, R, u_out, u_out, &theta, max_iter, tol, MQ_PREC_FP64, ortho,
                                              n_ortho, &its2);
            if (rc != MQ_OK) return rc;
            it += its2;
        }
        plan->solve_info[0] = its_form;
        plan->solve_info[1] = it - its_form;
        plan->solve_info[2] = theta_form;
        plan->solve_info[3] = theta_rq;
        *theta_out = theta;
        if (iterations) *iterations = it;
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

// ------------------------------------------------------------------ Davidson (diagonal preconditioner)
// optimize/davidson.h:64-161 with MultDiagonal::precondition (ietl_davidson.h:96-104).
// The reference rotates its whole basis into Ritz vectors every iteration
// (davidson.h:107-124); here the basis V, VA stays as built and the Ritz
// vectors are formed from the eigenvectors of its Rayleigh matrix (the same
// subspace and Ritz pairs), except at the restart, where the two lowest Ritz
// vectors and their images become the new basis (davidson.h:152-156).
extern "C" int mq_davidson(mq_plan *plan, const double *L, const double *R, const double *diag, const double *x0,
                           double *u_out, double *theta_out, int max_iter, double tol, int prec,
                           const double *const *ortho, int n_ortho, int *iterations) {
    constexpr int kMaxDim = 20;  // basis size that triggers the restart
    if (!plan || !diag || !x0 || !u_out || !theta_out || max_iter < 1 || max_iter > 100000 || n_ortho < 0 ||
        n_ortho > 60)
        return MQ_ERR_ARG;
    if (plan->hp.kind != PLAN_SIGMA) {
        plan->ctx->err = "mq_davidson: needs a sigma plan";
        return MQ_ERR_ARG;
    }
    if (plan->hp.symm != MQ_SYMM_SU2U1) {
        plan->ctx->err = "mq_davidson: Davidson only implemented for spin-adapted Hamiltonians";
        return MQ_ERR_ARG;
    }
    if (const char *e = prec_error(prec)) {
        plan->ctx->err = std::string("mq_davidson: ") + e;
        return MQ_ERR_ARG;
    }
    try {
        const double ptol = phase_tol(tol, prec);
        Stall stall(prec != MQ_PREC_FP64 && tol > 0);
        JD jd;
        jd.p = plan;
        jd.s = plan->ctx->stream;
        jd.n = plan->hp.vec_len;
        jd.nblk = (int)std::max<int64_t>(1, std::min<int64_t>(kJdBlocks, (jd.n + 255) / 256));
        jd.init();
        const int64_t n = jd.n;
        const int need = 2 * kMaxDim + 5;
        if (plan->jd_cap < need) {
            plan->jd_vecs.alloc((size_t)need * n);
            plan->jd_cap = need;
        }
        auto V = [&](int i) { return jd.vec(i); };
        auto VA = [&](int i) { return jd.vec(kMaxDim + i); };
        double *u = jd.vec(2 * kMaxDim), *r = jd.vec(2 * kMaxDim + 1), *vc = jd.vec(2 * kMaxDim + 2);
        double *w0 = jd.vec(2 * kMaxDim + 3), *w1 = jd.vec(2 * kMaxDim + 4);
        std::vector<double> oo(n_ortho);
        for (int k = 0; k < n_ortho; ++k) oo[k] = jd.dot(ortho[k], ortho[k]);
        auto project = [&](double *t) {  // SingleSiteVS::project (ietl_lanczos_solver.h:90-95)
            if (n_ortho == 0) return;
            std::vector<const double *> xs, ys;
            for (int k = 0; k < n_ortho; ++k) {
                xs.push_back(ortho[k]);
                ys.push_back(t);
            }
            std::vector<double> d = jd.dots(xs, ys), c(n_ortho);
            for (int k = 0; k < n_ortho; ++k) c[k] = -d[k] / oo[k];
            jd.lincomb(std::vector<const double *>(ortho, ortho + n_ortho), c, 1.0, t);
        };
        // Rayleigh matrix of the current basis, M[j * kMaxDim + i] = V_i . VA_j (i <= j)
        std::vector<double> M((size_t)kMaxDim * kMaxDim, 0.0);
        // ascending eigenpairs of M(0:dim), each eigenvector with its largest component positive
        auto ritz = [&](int dim, std::vector<double> &w, std::vector<double> &E) {
            std::vector<double> A((size_t)dim * dim), Z, wr;
            for (int j = 0; j < dim; ++j)
                for (int i = 0; i <= j; ++i) {
                    const double v = M[(size_t)j * kMaxDim + i];
                    A[(size_t)j * dim + i] = v;
                    A[(size_t)i * dim + j] = v;
                }
            sym_eig(dim, A, wr, Z);
            std::vector<int> order(dim);
            for (int i = 0; i < dim; ++i) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return wr[x] < wr[y]; });
            w.resize(dim);
            E.assign((size_t)dim * dim, 0.0);
            for (int j = 0; j < dim; ++j) {
                const double *z = Z.data() + (size_t)order[j] * dim;
                int big = 0;
                for (int i = 1; i < dim; ++i)
                    if (std::fabs(z[i]) > std::fabs(z[big])) big = i;
                const double sg = z[big] < 0 ? -1.0 : 1.0;
                for (int i = 0; i < dim; ++i) E[(size_t)j * dim + i] = sg * z[i];
                w[j] = wr[order[j]];
            }
        };
        check(hipMemcpyAsync(V(0), x0, sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "x0");
        project(V(0));
        int dim = 0, it = 0;
        double theta = 0;
        while (true) {
            // Gram-Schmidt with one refinement when the norm drops below kappa = 0.25 of
            // its value (davidson.h:93-99), projection, normalisation
            double *t = V(dim);
            std::vector<const double *> xs{t}, ys{t};
            for (int i = 0; i < dim; ++i) {
                xs.push_back(V(i));
                ys.push_back(t);
            }
            std::vector<double> d = jd.dots(xs, ys);
            const double tau = std::sqrt(d[0]);
            double nt2 = d[0];
            for (int pass = 0; pass < 2 && dim > 0; ++pass) {
                std::vector<const double *> vv;
                std::vector<double> c;
                for (int i = 0; i < dim; ++i) {
                    vv.push_back(V(i));
                    c.push_back(-d[1 + i]);
                }
                jd.lincomb(vv, c, 1.0, t);
                d = jd.dots(xs, ys);
                nt2 = d[0];
                if (std::sqrt(nt2) >= 0.25 * tau) break;
            }
            if (n_ortho > 0) {
                project(t);
                nt2 = jd.dot(t, t);
            }
            jd.lincomb({t}, {1.0 / std::sqrt(nt2) - 1.0}, 1.0, t);
            solver_sigma(plan, L, R, t, VA(dim), prec);
            ++dim;
            // new Rayleigh column (davidson.h:104-106)
            xs.clear();
            ys.clear();
            for (int i = 0; i < dim; ++i) {
                xs.push_back(V(i));
                ys.push_back(VA(dim - 1));
            }
            std::vector<double> col = jd.dots(xs, ys);
            for (int i = 0; i < dim; ++i) M[(size_t)(dim - 1) * kMaxDim + i] = col[i];
            std::vector<double> w, E;
            ritz(dim, w, E);
            theta = w[0];
            std::vector<const double *> vv, va;
            for (int j = 0; j < dim; ++j) {
                vv.push_back(V(j));
                va.push_back(VA(j));
            }
            std::vector<double> y0(E.begin(), E.begin() + dim);
            jd.lincomb(vv, y0, 0.0, u);
            jd.lincomb(va, y0, 0.0, r);
            jd.lincomb({u}, {-theta}, 1.0, r);  // r = VA y0 - theta V y0
            // Vcpy = u / (theta - D) for the preconditioner, formed before the test
            check(launch_diag_scale(u, u, 0.0, diag, theta, vc, n, jd.s), "precondition");
            ++it;
            std::vector<double> rr = jd.dots({r, vc, u}, {r, r, vc});
            const double rn = std::sqrt(rr[0]);
            if (rn <= ptol * std::fabs(theta) || rn < ptol || it >= max_iter ||
                stall.update(rn / std::max(std::fabs(theta), 1e-300)))
                break;
            if (dim >= kMaxDim) {
                // restart with the two lowest Ritz vectors and their images
                std::vector<double> y1(E.begin() + dim, E.begin() + 2 * dim);
                jd.lincomb(vv, y1, 0.0, w0);
                jd.lincomb(va, y0, 0.0, w1);
                check(hipMemcpyAsync(V(0), u, sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "restart");
                check(hipMemcpyAsync(V(1), w0, sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "restart");
                jd.lincomb(va, y1, 0.0, w0);
                check(hipMemcpyAsync(VA(0), w1, sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "restart");
                check(hipMemcpyAsync(VA(1), w0, sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "restart");
                dim = 2;
                std::vector<double> m = jd.dots({V(0), V(0), V(1)}, {VA(0), VA(1), VA(1)});
                M[0] = m[0];
                M[(size_t)kMaxDim] = m[1];
                M[(size_t)kMaxDim + 1] = m[2];
            }
            // t = (r - (Vcpy.r)/(u.Vcpy) u) / (theta - D), normalised by the next pass
            check(launch_diag_scale(r, u, rr[1] / rr[2], diag, theta, V(dim), n, jd.s), "precondition");
        }
        const double theta_form = theta;
        const int its_form = it;
        bool refine = false;
        if (prec != MQ_PREC_FP64) {
            // FP64 defect correction and, when the FP64 residual misses the tolerance,
            // FP64 iterations from the Ritz vector, as in the Jacobi-Davidson driver
            sigma_launch(plan, L, R, u, r, MQ_PREC_FP64);
            std::vector<double> d = jd.dots({u, u}, {r, u});
            theta = d[0] / d[1];
            if (tol > 0) {
                jd.lincomb({u}, {-theta}, 1.0, r);
                const double rn = std::sqrt(jd.dot(r, r));
                refine = !(rn <= tol * std::fabs(theta) || rn < tol);
            }
        }
        const double theta_rq = theta;
        check(hipMemcpyAsync(u_out, u, sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "u out");
        jd.sync();
        if (refine) {
            int its2 = 0;
            const int rc = mq_davidson(plan, L, R, diag, u_out, u_out, &theta, max_iter, tol, MQ_PREC_FP64, ortho,
                                       n_ortho, &its2);
            if (rc != MQ_OK) return rc;
            it += its2;
        }
        plan->solve_info[0] = its_form;
        plan->solve_info[1] = it - its_form;
        plan->solve_info[2] = theta_form;
        plan->solve_info[3] = theta_rq;
        *theta_out = theta;
        if (iterations) *iterations = it;
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

// ------------------------------------------------------------------ block Davidson
// The nroots lowest eigenpairs together: a Jacobi-Davidson iteration on a block
// (no counterpart in the reference, which solves excited states one at a time
// with orthogonality constraints).  Each iteration orthonormalises the new
// block against the basis (Gram-Schmidt with one refinement, as jacobi.h:384-393;
// vectors that vanish are dropped), applies H to the whole block with one
// batched sigma, extends the Rayleigh matrix, takes the lowest nroots Ritz pairs,
// and adds the correction t = -r + (r.u)/(u.u) u of every root that has not
// converged (|r| <= tol*|theta| or |r| < tol).  When the basis would exceed its
// capacity it restarts from the current Ritz vectors.
extern "C" int mq_block_davidson(mq_plan *plan, const double *L, const double *R, const double *diag,
                                 const double *X0, double *U, double *theta_out, int nroots, int max_iter, double tol,
                                 int prec, const double *const *ortho, int n_ortho, int *iterations) {
    if (!plan || !X0 || !U || !theta_out || nroots < 1 || nroots > 16 || max_iter < 1 || max_iter > 100000 ||
        n_ortho < 0 || n_ortho > 60)
        return MQ_ERR_ARG;
    if (plan->hp.kind != PLAN_SIGMA) {
        plan->ctx->err = "mq_block_davidson: needs a sigma plan";
        return MQ_ERR_ARG;
    }
    if (const char *e = prec_error(prec)) {
        plan->ctx->err = std::string("mq_block_davidson: ") + e;
        return MQ_ERR_ARG;
    }
    const int k = nroots;
    const int cap = 60;  // basis capacity
    try {
        const double ptol = phase_tol(tol, prec);
        Stall stall(prec != MQ_PREC_FP64 && tol > 0);
        JD jd;
        jd.p = plan;
        jd.s = plan->ctx->stream;
        jd.n = plan->hp.vec_len;
        jd.nblk = (int)std::max<int64_t>(1, std::min<int64_t>(kJdBlocks, (jd.n + 255) / 256));
        jd.init();
        const int64_t n = jd.n;
        const int need = 2 * cap + 7 * k;
        if (plan->jd_cap < need) {
            plan->jd_vecs.alloc((size_t)need * n);
            plan->jd_cap = need;
        }
        auto V = [&](int i) { return jd.vec(i); };
        auto VA = [&](int i) { return jd.vec(cap + i); };
        auto Ur = [&](int i) { return jd.vec(2 * cap + i); };          // Ritz vectors
        auto UA = [&](int i) { return jd.vec(2 * cap + k + i); };      // their images
        auto Rr = [&](int i) { return jd.vec(2 * cap + 2 * k + i); };  // residuals
        auto SV = [&](int i) { return jd.vec(2 * cap + 3 * k + i); };  // restart vectors / preconditioner scratch
        auto SA = [&](int i) { return jd.vec(2 * cap + 5 * k + i); };  // restart images
        std::vector<double> oo(n_ortho);
        for (int q = 0; q < n_ortho; ++q) oo[q] = jd.dot(ortho[q], ortho[q]);
        auto project = [&](double *t) {
            if (n_ortho == 0) return;
            std::vector<const double *> xs, ys;
            for (int q = 0; q < n_ortho; ++q) {
                xs.push_back(ortho[q]);
                ys.push_back(t);
            }
            std::vector<double> d = jd.dots(xs, ys), c(n_ortho);
            for (int q = 0; q < n_ortho; ++q) c[q] = -d[q] / oo[q];
            jd.lincomb(std::vector<const double *>(ortho, ortho + n_ortho), c, 1.0, t);
        };
        std::vector<double> M((size_t)cap * cap, 0.0);  // M[j * cap + i] = V_i . VA_j, i <= j
        // orthonormalise the block V(dim .. dim+nb) in place against V(0 .. dim) and
        // itself, compacting the kept vectors; returns how many were kept
        auto orthonormalise = [&](int dim, int nb) {
            int kept = 0;
            for (int b = 0; b < nb; ++b) {
                double *t = V(dim + b);
                const int m = dim + kept;
                std::vector<const double *> xs{t}, ys{t};
                for (int i = 0; i < m; ++i) {
                    xs.push_back(V(i));
                    ys.push_back(t);
                }
                std::vector<double> d = jd.dots(xs, ys);
                const double tau = std::sqrt(d[0]);
                double nt2 = d[0];
                for (int pass = 0; pass < 2 && m > 0; ++pass) {
                    std::vector<const double *> vv;
                    std::vector<double> c;
                    for (int i = 0; i < m; ++i) {
                        vv.push_back(V(i));
                        c.push_back(-d[1 + i]);
                    }
                    jd.lincomb(vv, c, 1.0, t);
                    d = jd.dots(xs, ys);
                    nt2 = d[0];
                    if (std::sqrt(nt2) >= 0.25 * tau) break;
                }
                if (n_ortho > 0) {
                    project(t);
                    nt2 = jd.dot(t, t);
                }
                if (!(std::sqrt(nt2) > 1e-10 * std::max(tau, 1e-300))) continue;  // vanished: dropped
                double *dst = V(dim + kept);
                if (dst != t) jd.lincomb({t}, {1.0 / std::sqrt(nt2)}, 0.0, dst);
                else jd.lincomb({t}, {1.0 / std::sqrt(nt2) - 1.0}, 1.0, t);
                ++kept;
            }
            return kept;
        };
        for (int j = 0; j < k; ++j)
            check(hipMemcpyAsync(V(j), X0 + (int64_t)j * n, sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "x0");
        int dim = 0, nb = k, it = 0;
        std::vector<double> theta(k, 0.0);
        while (true) {
            const int kept = orthonormalise(dim, nb);
            if (dim + kept < k) throw std::runtime_error("block Davidson: the start vectors are linearly dependent");
            if (kept > 0) sigma_batch_launch(plan, L, R, V(dim), VA(dim), kept, prec);
            for (int j = dim; j < dim + kept; ++j) {
                std::vector<const double *> xs, ys;
                for (int i = 0; i <= j; ++i) {
                    xs.push_back(V(i));
                    ys.push_back(VA(j));
                }
                std::vector<double> col = jd.dots(xs, ys);
                for (int i = 0; i <= j; ++i) M[(size_t)j * cap + i] = col[i];
            }
            dim += kept;
            // Ritz pairs of the basis
            std::vector<double> A((size_t)dim * dim), w, Z;
            for (int j = 0; j < dim; ++j)
                for (int i = 0; i <= j; ++i) {
                    A[(size_t)j * dim + i] = M[(size_t)j * cap + i];
                    A[(size_t)i * dim + j] = M[(size_t)j * cap + i];
                }
            sym_eig(dim, A, w, Z);
            std::vector<int> order(dim);
            for (int i = 0; i < dim; ++i) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return w[x] < w[y]; });
            std::vector<const double *> vv, va;
            for (int j = 0; j < dim; ++j) {
                vv.push_back(V(j));
                va.push_back(VA(j));
            }
            std::vector<const double *> xs, ys;
            for (int r = 0; r < k; ++r) {
                std::vector<double> y(Z.begin() + (size_t)order[r] * dim, Z.begin() + (size_t)(order[r] + 1) * dim);
                int big = 0;
                for (int i = 1; i < dim; ++i)
                    if (std::fabs(y[i]) > std::fabs(y[big])) big = i;
                if (y[big] < 0)
                    for (auto &v : y) v = -v;
                theta[r] = w[order[r]];
                jd.lincomb(vv, y, 0.0, Ur(r));
                jd.lincomb(va, y, 0.0, UA(r));
                jd.lincomb({UA(r), Ur(r)}, {1.0, -theta[r]}, 0.0, Rr(r));
                xs.insert(xs.end(), {Rr(r), Rr(r), Ur(r)});
                ys.insert(ys.end(), {Rr(r), Ur(r), Ur(r)});
            }
            ++it;
            std::vector<double> rr = jd.dots(xs, ys);
            std::vector<int> open;
            double worst = 0.0;
            for (int r = 0; r < k; ++r) {
                const double rn = std::sqrt(rr[3 * r]);
                if (!(rn <= ptol * std::fabs(theta[r]) || rn < ptol)) {
                    open.push_back(r);
                    worst = std::max(worst, rn / std::max(std::fabs(theta[r]), 1e-300));
                }
            }
            if (open.empty() || it >= max_iter || stall.update(worst)) break;
            if (dim + (int)open.size() > cap) {
                // thick restart: the lowest min(dim, 2 nroots) Ritz vectors and their images
                const int keep = std::min(dim, 2 * k);
                for (int r = 0; r < keep; ++r) {
                    std::vector<double> y(Z.begin() + (size_t)order[r] * dim,
                                          Z.begin() + (size_t)(order[r] + 1) * dim);
                    jd.lincomb(vv, y, 0.0, SV(r));
                    jd.lincomb(va, y, 0.0, SA(r));
                }
                for (int r = 0; r < keep; ++r) {
                    check(hipMemcpyAsync(V(r), SV(r), sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "restart");
                    check(hipMemcpyAsync(VA(r), SA(r), sizeof(double) * n, hipMemcpyDeviceToDevice, jd.s), "restart");
                }
                dim = keep;
                for (int j = 0; j < keep; ++j) {
                    std::vector<const double *> a, b;
                    for (int i = 0; i <= j; ++i) {
                        a.push_back(V(i));
                        b.push_back(VA(j));
                    }
                    std::vector<double> col = jd.dots(a, b);
                    for (int i = 0; i <= j; ++i) M[(size_t)j * cap + i] = col[i];
                }
            }
            nb = (int)open.size();
            if (diag) {
                // diagonal preconditioner (MultDiagonal::precondition, ietl_davidson.h:96-104)
                std::vector<const double *> px, py;
                for (int c = 0; c < nb; ++c) {
                    const int r = open[c];
                    check(launch_diag_scale(Ur(r), Ur(r), 0.0, diag, theta[r], SV(c), n, jd.s), "precondition");
                    px.insert(px.end(), {SV(c), Ur(r)});
                    py.insert(py.end(), {Rr(r), SV(c)});
                }
                std::vector<double> ab = jd.dots(px, py);
                for (int c = 0; c < nb; ++c) {
                    const int r = open[c];
                    check(launch_diag_scale(Rr(r), Ur(r), ab[2 * c] / ab[2 * c + 1], diag, theta[r], V(dim + c), n,
                                            jd.s),
                          "precondition");
                }
            } else {
                for (int c = 0; c < nb; ++c) {
                    const int r = open[c];
                    jd.lincomb({Rr(r), Ur(r)}, {-1.0, rr[3 * r + 1] / rr[3 * r + 2]}, 0.0, V(dim + c));
                }
            }
        }
        const double theta_form = theta[0];
        const int its_form = it;
        bool refine = false;
        if (prec != MQ_PREC_FP64) {
            // FP64 defect correction per root: Rayleigh quotients from one batched FP64
            // sigma; when a root's FP64 residual misses the tolerance, the block continues
            // in FP64 from the Ritz vectors
            sigma_batch_launch(plan, L, R, Ur(0), UA(0), k, MQ_PREC_FP64);
            std::vector<const double *> xs, ys;
            for (int r = 0; r < k; ++r) {
                xs.insert(xs.end(), {Ur(r), Ur(r)});
                ys.insert(ys.end(), {UA(r), Ur(r)});
            }
            std::vector<double> d = jd.dots(xs, ys);
            for (int r = 0; r < k; ++r) theta[r] = d[2 * r] / d[2 * r + 1];
            if (tol > 0) {
                xs.clear();
                for (int r = 0; r < k; ++r) {
                    jd.lincomb({UA(r), Ur(r)}, {1.0, -theta[r]}, 0.0, Rr(r));
                    xs.push_back(Rr(r));
                }
                std::vector<double> rr = jd.dots(xs, xs);
                for (int r = 0; r < k; ++r) {
                    const double rn = std::sqrt(rr[r]);
                    if (!(rn <= tol * std::fabs(theta[r]) || rn < tol)) refine = true;
                }
            }
        }
        const double theta_rq = theta[0];
        check(hipMemcpyAsync(U, Ur(0), sizeof(double) * n * k, hipMemcpyDeviceToDevice, jd.s), "U out");
        jd.sync();
        if (refine) {
            int its2 = 0;
            const int rc = mq_block_davidson(plan, L, R, diag, U, U, theta.data(), k, max_iter, tol, MQ_PREC_FP64,
                                             ortho, n_ortho, &its2);
            if (rc != MQ_OK) return rc;
            it += its2;
        }
        plan->solve_info[0] = its_form;
        plan->solve_info[1] = it - its_form;
        plan->solve_info[2] = theta_form;
        plan->solve_info[3] = theta_rq;
        for (int r = 0; r < k; ++r) theta_out[r] = theta[r];
        if (iterations) *iterations = it;
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(plan->ctx, e, MQ_ERR_HIP);
    }
}

// ------------------------------------------------------------------ two-site split
// Sweeps of the Jacobi SVD before a block is reported as not converged (the
// sweeps converge quadratically; blocks of a few hundred columns take 6-12).
static constexpr int kSvdMaxSweeps = 60;
// singular values of one block kept in LDS while they are sorted
static constexpr int kSvdMaxCols = 8192;

struct mq_split {
    mq_ctx *ctx = nullptr;
    SplitMap map;
    DevArray<SplitSlab> slabs;
    DevArray<SvdBlk> blk;
    DevArray<int32_t> perm, sweeps, keep;
    DevArray<int32_t> slot_blk, slot_k, done, rot;  // Jacobi rounds: one wave per column pair slot
    int nslots = 0, rounds = 0;                     // rounds per sweep: the largest block's Ce - 1
    DevArray<int64_t> aoff, boff;
    DevArray<double> X, V, S;
    std::vector<int32_t> last_sweeps;  // Jacobi sweeps per block of the last decomposition
    bool decomposed = false;
};

extern "C" {

int mq_split_create(mq_ctx *ctx, const mq_split_desc *desc, mq_split **out) {
    if (!ctx || !desc || !out) return MQ_ERR_ARG;
    try {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        std::unique_ptr<mq_split> sp(new mq_split());
        sp->ctx = ctx;
        sp->map = build_split_map(*desc);
        const SplitMap &m = sp->map;
        if (m.max_c > kSvdMaxCols) {
            ctx->err = "mq_split_create: a block has more than 8192 singular values";
            return MQ_ERR_ARG;
        }
        hipStream_t s = ctx->stream;
        sp->slabs.upload(m.slabs, s);
        sp->blk.upload(m.blk, s);
        // pair slots of every block, the largest blocks first
        std::vector<int32_t> sb, sk;
        for (int32_t b : m.order) {
            const int C = m.blk[b].C, Ce = C + (C & 1);
            if (C < 2) continue;
            for (int k = 0; k < Ce / 2; ++k) {
                sb.push_back(b);
                sk.push_back(k);
            }
            sp->rounds = std::max(sp->rounds, Ce - 1);
        }
        sp->nslots = (int)sb.size();
        sp->slot_blk.upload(sb, s);
        sp->slot_k.upload(sk, s);
        sp->done.alloc(m.blk.size());
        sp->rot.alloc(m.blk.size());
        sp->X.alloc((size_t)m.x_total);
        sp->V.alloc((size_t)m.v_total);
        sp->S.alloc((size_t)m.s_total);
        sp->perm.alloc((size_t)m.s_total);
        sp->sweeps.alloc(m.blk.size());
        check(hipStreamSynchronize(s), "split upload");
        *out = sp.release();
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(ctx, e, MQ_ERR_PLAN);
    }
}

int mq_split_destroy(mq_split *sp) {
    if (!sp) return MQ_ERR_ARG;
    (void)hipStreamSynchronize(sp->ctx->stream);
    delete sp;
    return MQ_OK;
}

int mq_split_blocks(const mq_split *sp, int32_t *nblocks, const int32_t **charge, const int32_t **rows,
                    const int32_t **cols, int64_t *vec_len) {
    if (!sp) return MQ_ERR_ARG;
    if (nblocks) *nblocks = (int32_t)sp->map.blk.size();
    if (charge) *charge = sp->map.charge.data();
    if (rows) *rows = sp->map.rows.data();
    if (cols) *cols = sp->map.cols.data();
    if (vec_len) *vec_len = sp->map.psi_len;
    return MQ_OK;
}

int mq_split_decompose(mq_split *sp, const double *psi, double *s_out) {
    if (!sp || !psi || !s_out) return MQ_ERR_ARG;
    const SplitMap &m = sp->map;
    hipStream_t s = sp->ctx->stream;
    try {
        if (m.x_total) check(hipMemsetAsync(sp->X.p, 0, sizeof(double) * m.x_total, s), "split clear");
        check(launch_split_gather(sp->slabs.p, (int)m.slabs.size(), psi, sp->X.p, s), "split gather");
        const int nb = (int)m.blk.size();
        check(launch_jacobi_init(sp->blk.p, nb, sp->V.p, sp->done.p, sp->rot.p, sp->sweeps.p, s), "jacobi init");
        std::vector<int32_t> dn(nb);
        for (int sw = 0; sw < kSvdMaxSweeps && sp->nslots > 0; ++sw) {
            for (int t = 0; t < sp->rounds; ++t)
                check(launch_jacobi_round(sp->blk.p, sp->slot_blk.p, sp->slot_k.p, sp->nslots, t, sp->X.p, sp->V.p,
                                          sp->done.p, sp->rot.p, s),
                      "jacobi round");
            check(launch_jacobi_sweep_end(nb, sw, sp->done.p, sp->rot.p, sp->sweeps.p, s), "jacobi sweep end");
            check(hipMemcpyAsync(dn.data(), sp->done.p, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s),
                  "jacobi done");
            check(hipStreamSynchronize(s), "jacobi sweep");
            if (std::all_of(dn.begin(), dn.end(), [](int32_t d) { return d != 0; })) break;
        }
        check(launch_jacobi_finish(sp->blk.p, nb, m.max_c, sp->X.p, sp->S.p, sp->perm.p, sp->done.p, sp->sweeps.p,
                                   s),
              "jacobi finish");
        std::vector<int32_t> sw(m.blk.size());
        if (!sw.empty())
            check(hipMemcpyAsync(sw.data(), sp->sweeps.p, sizeof(int32_t) * sw.size(), hipMemcpyDeviceToHost, s),
                  "split sweeps");
        if (m.s_total)
            check(hipMemcpyAsync(s_out, sp->S.p, sizeof(double) * m.s_total, hipMemcpyDeviceToHost, s), "split S");
        check(hipStreamSynchronize(s), "split decompose");
        sp->last_sweeps = sw;
        for (size_t b = 0; b < sw.size(); ++b)
            if (sw[b] < 0) {
                sp->ctx->err = "mq_split_decompose: the Jacobi sweeps of block " + std::to_string(b) +
                               " did not converge";
                sp->decomposed = false;
                return MQ_ERR_STATE;
            }
        sp->decomposed = true;
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(sp->ctx, e, MQ_ERR_HIP);
    }
}

int mq_split_sweeps(const mq_split *sp, int32_t *sweeps) {
    if (!sp || !sweeps) return MQ_ERR_ARG;
    if (!sp->decomposed) return MQ_ERR_STATE;
    std::copy(sp->last_sweeps.begin(), sp->last_sweeps.end(), sweeps);
    return MQ_OK;
}

int mq_split_assemble(mq_split *sp, const int32_t *keep, const int64_t *a_off, const int64_t *b_off, int l2r,
                      double *A, double *B) {
    if (!sp || !keep || !a_off || !b_off) return MQ_ERR_ARG;
    if (!sp->decomposed) {
        sp->ctx->err = "mq_split_assemble: no decomposition to assemble from";
        return MQ_ERR_STATE;
    }
    const SplitMap &m = sp->map;
    const size_t nb = m.blk.size();
    int max_keep = 0;
    for (size_t b = 0; b < nb; ++b) {
        if (keep[b] < 0 || keep[b] > m.blk[b].C) {
            sp->ctx->err = "mq_split_assemble: keep[" + std::to_string(b) + "] outside 0..min(rows, cols)";
            return MQ_ERR_ARG;
        }
        if (keep[b] > 0 && (a_off[b] < 0 || b_off[b] < 0)) {
            sp->ctx->err = "mq_split_assemble: negative offset";
            return MQ_ERR_ARG;
        }
        max_keep = std::max(max_keep, (int)keep[b]);
    }
    if (max_keep > 0 && (!A || !B)) return MQ_ERR_ARG;
    hipStream_t s = sp->ctx->stream;
    try {
        sp->keep.upload(std::vector<int32_t>(keep, keep + nb), s);
        sp->aoff.upload(std::vector<int64_t>(a_off, a_off + nb), s);
        sp->boff.upload(std::vector<int64_t>(b_off, b_off + nb), s);
        check(launch_split_assemble(sp->blk.p, (int)nb, max_keep, sp->keep.p, sp->aoff.p, sp->boff.p, l2r ? 1 : 0,
                                    sp->X.p, sp->V.p, sp->S.p, sp->perm.p, A, B, s),
              "split assemble");
        check(hipStreamSynchronize(s), "split assemble");
        return MQ_OK;
    } catch (std::exception &e) {
        return fail(sp->ctx, e, MQ_ERR_HIP);
    }
}

}  // extern "C"

// ------------------------------------------------------------------ host-side introspection
struct mq_host_plan {
    HostPlan hp;
};

static thread_local std::string g_host_err;

extern "C" {

const char *mq_host_last_error(void) { return g_host_err.c_str(); }

int mq_host_plan_create(const mq_site_desc *desc, int kind, mq_host_plan **out) {
    if (!desc || !out) return MQ_ERR_ARG;
    try {
        std::unique_ptr<mq_host_plan> h(new mq_host_plan());
        const int tm = 0, tn = 0, kc = 2048;
        const int fm = (kind >> 8) & 3, fr = (kind >> 10) & 255, fmin = (kind >> 18) & 4095, k = kind & 255;
        FuseRule fuse;
        if (fm) fuse.mode = fm - 1;
        if (fr) fuse.ratio = fr;
        if (fmin) fuse.min_entries = fmin;
        const int sched = (kind >> 30) & 1;
        if (k == PLAN_SIGMA) h->hp = build_sigma_plan(*desc, tm, tn, kc, fuse, true, sched);
        else if (k == PLAN_LSTEP) h->hp = build_lstep_plan(*desc, tm, tn, kc, fuse);
        else if (k == PLAN_RSTEP) h->hp = build_rstep_plan(*desc, tm, tn, kc);
        else if (k == PLAN_DIAG) h->hp = build_diag_plan(*desc, tm, tn, kc);
        else return MQ_ERR_ARG;
        *out = h.release();
        return MQ_OK;
    } catch (std::exception &e) {
        g_host_err = e.what();
        return MQ_ERR_PLAN;
    }
}

int mq_host_plan_part(const mq_host_plan *h, int part, const void **ptr, int64_t *count, int64_t *elem) {
    if (!h || !ptr || !count || !elem) return MQ_ERR_ARG;
    const HostPlan &p = h->hp;
    switch (part) {
        case MQ_PART_ARR_L: *ptr = p.arrL.data(); *count = p.arrL.size(); *elem = sizeof(Copy2D); break;
        case MQ_PART_ARR_R: *ptr = p.arrR.data(); *count = p.arrR.size(); *elem = sizeof(Copy2D); break;
        case MQ_PART_RESHAPE: *ptr = p.reshape.data(); *count = p.reshape.size(); *elem = sizeof(Copy2D); break;
        case MQ_PART_G1_GROUPS: *ptr = p.g1.groups.data(); *count = p.g1.groups.size(); *elem = sizeof(GemmGroup); break;
        case MQ_PART_G1_TILES: *ptr = p.g1.tiles.data(); *count = p.g1.tiles.size(); *elem = sizeof(GemmTile); break;
        case MQ_PART_STRIPS: *ptr = p.wapply.strips.data(); *count = p.wapply.strips.size(); *elem = sizeof(Strip); break;
        case MQ_PART_SRCS: *ptr = p.wapply.srcs.data(); *count = p.wapply.srcs.size(); *elem = sizeof(Src); break;
        case MQ_PART_G2_GROUPS: *ptr = p.g2.groups.data(); *count = p.g2.groups.size(); *elem = sizeof(GemmGroup); break;
        case MQ_PART_G2_TILES: *ptr = p.g2.tiles.data(); *count = p.g2.tiles.size(); *elem = sizeof(GemmTile); break;
        case MQ_PART_SCATTER: *ptr = p.scatter.data(); *count = p.scatter.size(); *elem = sizeof(Copy2D); break;
        case MQ_PART_COMB_STRIPS: *ptr = p.combL.strips.data(); *count = p.combL.strips.size(); *elem = sizeof(Strip); break;
        case MQ_PART_COMB_SRCS: *ptr = p.combL.srcs.data(); *count = p.combL.srcs.size(); *elem = sizeof(Src); break;
        case MQ_PART_G1_ROWMAP: *ptr = p.g1.rowmap.data(); *count = p.g1.rowmap.size(); *elem = sizeof(RowMap); break;
        case MQ_PART_G2_ROWMAP: *ptr = p.g2.rowmap.data(); *count = p.g2.rowmap.size(); *elem = sizeof(RowMap); break;
        case MQ_PART_LIVE: *ptr = p.live.data(); *count = p.live.size(); *elem = 1; break;
        default: return MQ_ERR_ARG;
    }
    return MQ_OK;
}

int mq_host_plan_sizes(const mq_host_plan *h, int64_t s[10]) {
    if (!h || !s) return MQ_ERR_ARG;
    const HostPlan &p = h->hp;
    int64_t v[10] = {p.vec_len, p.n_la, p.n_rb, p.n_prp, p.n_t, p.n_c, p.n_tmp, p.n_out, p.g1.ws, p.g2.ws};
    for (int i = 0; i < 10; ++i) s[i] = v[i];
    return MQ_OK;
}

int mq_host_plan_flops(const mq_host_plan *h, double *ref, double *exec, double *bytes) {
    if (!h) return MQ_ERR_ARG;
    if (ref) *ref = h->hp.ref_flops;
    if (exec) *exec = h->hp.exec_flops;
    if (bytes) *bytes = h->hp.min_bytes;
    return MQ_OK;
}

int mq_host_plan_out_desc(const mq_host_plan *h, mq_boundary_desc *out) {
    if (!h || !out || h->hp.kind == PLAN_SIGMA) return MQ_ERR_ARG;
    const HostPlan &p = h->hp;
    out->aux_dim = p.out_aux;
    out->bptr = p.out_bptr.data();
    out->blocks.n = (int32_t)p.out_rows.size();
    out->blocks.lc = p.out_lc.data();
    out->blocks.rc = p.out_rc.data();
    out->blocks.rows = p.out_rows.data();
    out->blocks.cols = p.out_cols.data();
    out->blocks.offset = p.out_off.data();
    out->total = p.n_out;
    return MQ_OK;
}

int mq_host_plan_schedule(const mq_host_plan *h) { return h ? h->hp.rbtm : MQ_ERR_ARG; }

int mq_host_plan_destroy(mq_host_plan *h) {
    delete h;
    return MQ_OK;
}

int mq_host_split_both(const mq_split_desc *desc, const double *psi, int32_t *nblocks, int64_t *both_len,
                       double *both) {
    if (!desc || !nblocks || !both_len) return MQ_ERR_ARG;
    try {
        const SplitMap m = build_split_map(*desc);
        *nblocks = (int32_t)m.blk.size();
        int64_t n = 0;
        for (auto &b : m.blk) n += (int64_t)b.rows * b.cols;
        *both_len = n;
        if (both) {
            if (!psi) return MQ_ERR_ARG;
            apply_split_map_host(m, psi, both);
        }
        return MQ_OK;
    } catch (std::exception &e) {
        g_host_err = e.what();
        return MQ_ERR_PLAN;
    }
}

}  // extern "C"
