// Plan construction: turns one site problem into task lists for the HIP
// kernels.  The sigma vector follows the lbtm form of
// contractions/non-abelian/site_hamil.hpp:57-147 (abelian/site_hamil.hpp for
// TwoU1) reorganised for the GPU:
//   PRP  = psi right-paired                         (reshapes.h:174-230)
//   T_lk = LA_lk * PRP_lk  one GEMM per ket-left charge, where LA_lk stacks
//          s*L[b1]^T (or the stored hermitian partner, boundary_times_mps.hpp:188-209)
//   C_al = W applied to T   one deterministic gather per output strip
//          (apply_op.hpp:23-101, micro_kernels.hpp:20-45)
//   y_al = C_al * RB_al     one GEMM per output charge, RB_al stacking
//          R[b2](ar,al) (or phase*R[conj]^T, site_hamil.hpp:129-134)
// The boundary steps (move_boundary.hpp:130-229) reuse the same stages.
#include "mq_plan.hpp"

#include <array>
#include <atomic>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>

namespace mq {

// ---------------- task-list buffer cache
// Released large buffers are kept (at most kTaskCacheKeep of them, kTaskCacheBytes in
// all) and handed out again for requests they fit without wasting more than half;
// every buffer is allocated with a capacity rounded up to 2 MiB, recorded by address.
namespace {
constexpr size_t kTaskCacheKeep = 16, kTaskCacheBytes = size_t(2) << 30, kTaskRound = size_t(2) << 20;
struct TaskCache {
    std::mutex mu;
    std::vector<std::pair<void *, size_t>> free_;  // (buffer, capacity), oldest first
    std::unordered_map<void *, size_t> cap;        // capacity of every buffer handed out
    size_t held = 0;
};
TaskCache &task_cache() {
    static TaskCache *c = new TaskCache();  // never destroyed: buffers may be released at exit
    return *c;
}
}  // namespace

void *task_alloc(size_t bytes) {
    TaskCache &tc = task_cache();
    {
        std::lock_guard<std::mutex> lk(tc.mu);
        size_t best = tc.free_.size();
        for (size_t i = 0; i < tc.free_.size(); ++i) {
            const size_t c = tc.free_[i].second;
            if (c >= bytes && c <= 2 * bytes && (best == tc.free_.size() || c < tc.free_[best].second)) best = i;
        }
        if (best < tc.free_.size()) {
            void *p = tc.free_[best].first;
            tc.held -= tc.free_[best].second;
            tc.cap[p] = tc.free_[best].second;
            tc.free_.erase(tc.free_.begin() + best);
            return p;
        }
    }
    const size_t c = (bytes + kTaskRound - 1) / kTaskRound * kTaskRound;
    void *p = std::malloc(c);
    if (!p) throw std::bad_alloc();
    std::lock_guard<std::mutex> lk(tc.mu);
    tc.cap[p] = c;
    return p;
}

void task_free(void *p, size_t) noexcept {
    if (!p) return;
    TaskCache &tc = task_cache();
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> lk(tc.mu);
        auto it = tc.cap.find(p);
        const size_t c = it == tc.cap.end() ? 0 : it->second;
        if (it != tc.cap.end()) tc.cap.erase(it);
        if (c == 0 || c > kTaskCacheBytes) {
            drop.push_back(p);
        } else {
            tc.free_.push_back({p, c});
            tc.held += c;
            while (tc.free_.size() > kTaskCacheKeep || tc.held > kTaskCacheBytes) {
                drop.push_back(tc.free_.front().first);
                tc.held -= tc.free_.front().second;
                tc.free_.erase(tc.free_.begin());
            }
        }
    }
    for (void *q : drop) std::free(q);
}


namespace {

inline uint64_t qkey(Q c) { return ((uint64_t)(uint32_t)c.a << 32) | (uint32_t)c.b; }

struct QMap {
    std::unordered_map<uint64_t, int> m;
    void build(const Idx &ix) {
        m.clear();
        m.reserve(ix.n() * 2 + 1);
        for (int i = 0; i < ix.n(); ++i) m[qkey(ix.q[i])] = i;
    }
    int get(Q c) const {
        auto it = m.find(qkey(c));
        return it == m.end() ? -1 : it->second;
    }
};

// Per-thread cache of SU(2) couplings for init = 1 (set_coupling is linear in init).
const std::array<double, 4> &coupling(int j, int ts, int jp, int a, int k, int ap, int i, int tsp, int ip) {
    thread_local std::unordered_map<uint64_t, std::array<double, 4>> cache;
    uint64_t key = 0;
    for (int v : {j, ts, jp, a, k, ap, i, tsp, ip}) {
        if (v < 0 || v > 127) throw std::runtime_error("coupling argument out of range");
        key = (key << 7) | (uint64_t)v;
    }
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::array<double, 4> c;
    set_coupling(j, ts, jp, a, k, ap, i, tsp, ip, 1.0, c.data());
    return cache.emplace(key, c).first->second;
}

struct Site {
    int symm;
    bool su2;
    Idx phys, left, right;
    Mpo W;
    VLayout X;
    QMap pi, li, ri;
    int np, nl, nr;
    std::vector<int> pbl;     // [p*nl + l] offset of (p, l) in fused charge p+l, -1 if none
    std::vector<int> pbl_sz;  // [fused index into X] rows
    std::vector<int> pbr;     // [p*nr + r] offset of (p, r) in right-fused charge r-p
    std::vector<int> pbr_sz_l;  // [l] right-paired columns of ket-left charge l (0 if none)
    std::vector<int> x_of_r;  // [r] index into X.blk or -1
    std::vector<int> lp_r;    // [l*np + p] right id of left charge l + phys p, -1 if none
    std::vector<int> rmp_l;   // [r*np + p] left id of right charge r - phys p, -1 if none
    // operators with physical charges resolved to phys ids
    struct OpB {
        int pin, pout;
        const OpBlk *b;
    };
    std::vector<std::vector<OpB>> opb;
    Site(const mq_site_desc &d)
        : symm(d.symm), su2(d.symm == MQ_SYMM_SU2U1), phys(to_idx(d.phys)), left(to_idx(d.left)),
          right(to_idx(d.right)), W(to_mpo(d.mpo)) {
        if (d.symm != MQ_SYMM_SU2U1 && d.symm != MQ_SYMM_TWOU1) throw std::runtime_error("unknown symmetry");
        X = vec_layout(phys, left, right);
        pi.build(phys);
        li.build(left);
        ri.build(right);
        np = phys.n();
        nl = left.n();
        nr = right.n();
        PB pl(phys, left, false), pr(phys, right, true);
        pbl.assign((size_t)np * nl, -1);
        pbr.assign((size_t)np * nr, -1);
        for (int p = 0; p < np; ++p) {
            for (int l = 0; l < nl; ++l) pbl[(size_t)p * nl + l] = pl.at(phys.q[p], left.q[l]);
            for (int r = 0; r < nr; ++r) pbr[(size_t)p * nr + r] = pr.at(phys.q[p], right.q[r]);
        }
        pbr_sz_l.assign(nl, 0);
        for (int l = 0; l < nl; ++l) pbr_sz_l[l] = pr.sz(left.q[l]);
        x_of_r.assign(nr, -1);
        for (size_t i = 0; i < X.blk.size(); ++i) x_of_r[ri.get(X.blk[i].c)] = (int)i;
        lp_r.assign((size_t)nl * np, -1);
        for (int l = 0; l < nl; ++l)
            for (int p = 0; p < np; ++p) lp_r[(size_t)l * np + p] = ri.get(left.q[l] + phys.q[p]);
        rmp_l.assign((size_t)nr * np, -1);
        for (int r = 0; r < nr; ++r)
            for (int p = 0; p < np; ++p) rmp_l[(size_t)r * np + p] = li.get(right.q[r] - phys.q[p]);
        opb.resize(W.ops.size());
        for (size_t o = 0; o < W.ops.size(); ++o)
            for (auto &b : W.ops[o].blocks) {
                int pin = pi.get(b.in), pout = pi.get(b.out);
                if (pin >= 0 && pout >= 0) opb[o].push_back({pin, pout, &b});
            }
    }
    int lsize(int l) const { return left.s[l]; }
    int rsize(int r) const { return right.s[r]; }
};

int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Host threads of the plan builder: MQ_HOST_THREADS (1..64) when set, else the
// hardware's, at most 16.  Plans do not depend on it.
int host_threads() {
    static const int nt = [] {
        const char *e = std::getenv("MQ_HOST_THREADS");
        if (e && std::atoi(e) >= 1) return std::min(64, std::atoi(e));
        return (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    }();
    return nt;
}

// f(i) for i < n on the host threads, at least `grain` indices per thread, taken
// `chunk` at a time; f must only write data owned by index i.  Exceptions are
// rethrown on the calling thread.
template <class F>
void parallel_for(int n, F &&f, int grain = 64, int chunk = 16) {
    int nt = std::min(host_threads(), std::max(1, n / grain));
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int> next(0);
    std::exception_ptr err;
    std::mutex mu;
    auto work = [&]() {
        try {
            for (;;) {
                int i0 = next.fetch_add(chunk);
                if (i0 >= n) break;
                for (int i = i0; i < std::min(n, i0 + chunk); ++i) f(i);
            }
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu);
            if (!err) err = std::current_exception();
            next.store(n);
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < nt; ++k) th.emplace_back(work);
    work();
    for (auto &x : th) x.join();
    if (err) std::rethrow_exception(err);
}

// Column tile width: a multiple of 16 up to 128 minimising the padded columns
// plus one 16-column block of fixed cost per tile (ties go to the wider tile).
int column_tile(int N) {
    if (N <= 128) return std::max(16, ceil_div(N, 16) * 16);
    int best = 128, cost = ceil_div(N, 128) * (128 + 16);
    for (int tn = 112; tn >= 64; tn -= 16) {
        const int c = ceil_div(N, tn) * (tn + 16);
        if (c < cost) best = tn, cost = c;
    }
    return best;
}

void add_gemm(GemmList &gl, GemmGroup g, int tm, int tn, int kchunk_target) {
    if (g.rowmap < 0 || (g.rowmap > 0 && g.rowmap - 1 + g.M > (int64_t)gl.rowmap.size()))
        throw std::runtime_error("GEMM row map out of range");
    if ((tm != 0 && tm != 64 && tm != 128) || (tn != 0 && (tn < 16 || tn > 128 || tn % 16)))
        throw std::runtime_error("GEMM tile shape out of range");
    g.split = 1;
    g.kchunk = std::max(g.K, 1);
    // 0: automatic (128 rows unless M fits 64; the column tile of column_tile)
    g.tm = tm ? tm : (g.M > 64 ? 128 : 64);
    g.tn = tn ? tn : column_tile(g.N);
    if (g.M == 0 || g.N == 0) return;
    int mt = ceil_div(g.M, g.tm), nt = ceil_div(g.N, g.tn);
    // split K when the output is small and K long: chunks of about kchunk_target
    if (g.K > 2 * kchunk_target && mt * nt < 64 && g.rowmap == 0) {
        int s = std::min(ceil_div(g.K, kchunk_target), 64);
        int kc = ceil_div(ceil_div(g.K, s), 16) * 16;
        s = ceil_div(g.K, kc);
        if (s > 1) {
            g.split = s;
            g.kchunk = kc;
            g.part = gl.ws;
            gl.ws += (int64_t)s * g.M * g.N;
        }
    }
    int gi = (int)gl.groups.size();
    gl.groups.push_back(g);
    // consecutive tiles share an operand panel: the A row block (runs along n), or the
    // B column block when the group has a single column tile (runs along m)
    if (nt == 1) {
        for (int ks = 0; ks < g.split; ++ks)
            for (int m = 0; m < mt; ++m) gl.tiles.push_back({gi, m, 0, ks});
    } else {
        for (int ks = 0; ks < g.split; ++ks)
            for (int m = 0; m < mt; ++m)
                for (int n = 0; n < nt; ++n) gl.tiles.push_back({gi, m, n, ks});
    }
}

constexpr int kXcds = 8;

std::vector<GemmTile> xcd_order(const std::vector<GemmTile> &tiles, const std::vector<GemmGroup> &groups) {
    // chains: consecutive tiles that read one operand block -- the same A row block
    // (A offset, row tile, K range), or, in groups with a single column tile, the
    // same B column block (B offset, K range; cut into chains of at most 16 tiles so
    // the chains still balance).  Tiles of different groups that read the same
    // block (the per-vector copies of a batched plan) join one chain.
    auto kbeg = [&](const GemmTile &t) { return (int64_t)t.ks * groups[t.g].kchunk; };
    auto same_a = [&](const GemmTile &x, const GemmTile &y) {
        const GemmGroup &gx = groups[x.g], &gy = groups[y.g];
        return gx.a == gy.a && gx.lda == gy.lda && gx.transA == gy.transA && gx.tm == gy.tm && x.mt == y.mt &&
               kbeg(x) == kbeg(y);
    };
    auto same_b = [&](const GemmTile &x, const GemmTile &y) {
        const GemmGroup &gx = groups[x.g], &gy = groups[y.g];
        return gx.b == gy.b && gx.ldb == gy.ldb && gx.transB == gy.transB && gx.tn == gy.tn && x.nt == y.nt &&
               kbeg(x) == kbeg(y);
    };
    std::vector<std::pair<size_t, size_t>> chains;
    for (size_t i = 0; i < tiles.size();) {
        size_t j = i + 1;
        const bool along_m = groups[tiles[i].g].N <= groups[tiles[i].g].tn;
        while (j < tiles.size() &&
               (along_m ? (same_b(tiles[j], tiles[i]) && j - i < 16) : same_a(tiles[j], tiles[i])))
            ++j;
        chains.push_back({i, j});
        i = j;
    }
    std::vector<std::vector<GemmTile>> q(kXcds);
    for (auto &c : chains) {
        size_t best = 0;
        for (size_t x = 1; x < q.size(); ++x)
            if (q[x].size() < q[best].size()) best = x;
        q[best].insert(q[best].end(), tiles.begin() + c.first, tiles.begin() + c.second);
    }
    size_t len = 0;
    for (auto &v : q) len = std::max(len, v.size());
    std::vector<GemmTile> out;
    out.reserve(len * kXcds);
    for (size_t i = 0; i < len; ++i)
        for (int x = 0; x < kXcds; ++x) out.push_back(i < q[x].size() ? q[x][i] : GemmTile{-1, 0, 0, 0});
    return out;
}

Copy2D mk_copy(int64_t dst, int64_t src, int rows, int cols, int ldd, int lds, bool trans, double scale) {
    Copy2D c;
    c.dst = dst;
    c.src = src;
    c.rows = rows;
    c.cols = cols;
    c.ldd = ldd;
    c.lds = lds;
    c.trans = trans ? 1 : 0;
    c.scale = scale;
    return c;
}

// ---------------- right-paired psi --------------------------------------
struct PRP {
    std::vector<int64_t> off;  // [l] (-1 if the ket has no block with left charge l)
    int64_t total = 0;
};

PRP prp_layout(const Site &s, std::vector<Copy2D> &tasks) {
    PRP p;
    p.off.assign(s.nl, -1);
    for (int l = 0; l < s.nl; ++l) {
        int n = s.pbr_sz_l[l];
        if (n == 0) continue;
        p.off[l] = p.total;
        p.total += (int64_t)s.lsize(l) * n;
    }
    for (auto &xb : s.X.blk) {
        int r = s.ri.get(xb.c);
        for (int pp = 0; pp < s.np; ++pp) {
            int l = s.li.get(xb.c - s.phys.q[pp]);
            if (l < 0) continue;
            int ld = s.lsize(l);
            for (int ss = 0; ss < s.phys.s[pp]; ++ss)
                tasks.push_back(mk_copy(p.off[l] + (int64_t)(s.pbr[(size_t)pp * s.nr + r] + ss * xb.cols) * ld,
                                        xb.off + s.pbl[(size_t)pp * s.nl + l] + ss * ld, ld, xb.cols, ld, xb.rows,
                                        false, 1.0));
        }
    }
    return p;
}

// ---------------- column program: how each MPO column reaches T ------------
// Column b2 of the W application sums, over its terms (b1, op, scale) and the
// sparse entries e of op, coef * T[b1](:, slice of the input state e.row)
// into the output state e.col (apply_op.hpp:23-101).  The coupling factor of
// an entry depends on b1 only through a = spin(b1) (set_coupling arguments,
// apply_op.hpp:90), so all entries of one column that share a position
//     P = (input block, output block, e.row, e.col, row spin, col spin, 2S(op), a)
// can be summed on the left boundary first:
//     sum_b1 w_b1 T[b1](slice) = (sum_b1 w_b1 L[b1])^T psi(slice),  w = e.coeff * scale.
// Quantum-chemistry MPOs have a few columns (complementary operators) with
// hundreds of terms but only a few hundred distinct positions; those columns
// are "fused": each position becomes a fused group whose boundary rows are
// combined once per boundary binding, and whose T slice is one small GEMM
// per sigma.  The other columns keep their terms ("direct").
struct CTerm {
    int row;  // MPO row b1
    int op;
    double scale;
    int spin;  // 2S of the left MPO bond (a)
};
struct FBlk {  // one (lb, lk) block of a fused group
    int lb, lk, ar, al;
    int64_t lf;  // LA-arena offset of its first combined row (column 0)
    int64_t tf;  // T-arena offset of its T block (stored contiguously, leading dimension |lb|)
    int ld;      // rows of the fused LA panel (its leading dimension)
};
struct FGroup {
    int b2, pin, pout, row, col, rspin, cspin, k, a;
    std::vector<std::pair<int, double>> mem;  // (b1, weight)
    std::vector<FBlk> blk;
};
struct ColProg {
    std::vector<std::vector<CTerm>> col;  // [b2] direct terms
    std::vector<std::vector<int>> fcol;   // [b2] fused group ids
    std::vector<FGroup> fg;
    std::vector<char> direct, member;     // [b1] row has direct terms / feeds a fused group
};

// fuse.mode: 0 never, 1 when a column's entries are at least fuse.ratio times its
// distinct positions and number at least fuse.min_entries, 2 every non-empty
// column (used by tests to exercise fused groups on small systems)
ColProg column_program(const Site &s, bool skip_rskip_cols, FuseRule fuse) {
    ColProg cp;
    cp.col.resize(s.W.cols);
    cp.fcol.resize(s.W.cols);
    cp.direct.assign(s.W.rows, 0);
    cp.member.assign(s.W.rows, 0);
    struct E {
        uint64_t key;
        int b1;
        double w;
    };
    // per column, in parallel: its direct terms (cp.col[b2]) or its fused groups
    std::vector<std::vector<FGroup>> groups(s.W.cols);
    std::vector<char> fused_col(s.W.cols, 0);
    parallel_for(
        s.W.cols,
        [&](int b2) {
            if (skip_rskip_cols && s.W.rskip(b2)) return;
            thread_local std::vector<E> ents;
            ents.clear();
            for (int b1 : s.W.col_b1[b2])
                for (auto &term : s.W.terms(b1, b2)) {
                    const Op &op = s.W.ops[term.op];
                    for (auto &ob : s.opb[term.op])
                        for (auto &e : ob.b->e) {
                            if (ob.pin > 255 || ob.pout > 255 || e.row > 255 || e.col > 255 || e.rspin > 3 ||
                                e.cspin > 3 || op.spin > 127 || s.W.lspin[b1] > 127)
                                throw std::runtime_error("operator entry out of the fused-key range");
                            uint64_t key = ((uint64_t)ob.pin << 56) | ((uint64_t)ob.pout << 48) |
                                           ((uint64_t)e.row << 40) | ((uint64_t)e.col << 32) |
                                           ((uint64_t)e.rspin << 30) | ((uint64_t)e.cspin << 28) |
                                           ((uint64_t)op.spin << 21) | ((uint64_t)s.W.lspin[b1] << 14);
                            ents.push_back({key, b1, e.coeff * term.scale});
                        }
                }
            std::stable_sort(ents.begin(), ents.end(), [](const E &x, const E &y) { return x.key < y.key; });
            size_t ngroups = 0;
            for (size_t i = 0; i < ents.size(); ++i)
                if (i == 0 || ents[i].key != ents[i - 1].key) ++ngroups;
            bool fused = (fuse.mode == 2 && !ents.empty()) ||
                         (fuse.mode == 1 && !ents.empty() && ents.size() >= (size_t)fuse.min_entries &&
                          ents.size() >= (size_t)fuse.ratio * ngroups);
            if (!fused) {
                for (int b1 : s.W.col_b1[b2])
                    for (auto &term : s.W.terms(b1, b2)) cp.col[b2].push_back({b1, term.op, term.scale, s.W.lspin[b1]});
                return;
            }
            fused_col[b2] = 1;
            size_t i = 0;
            while (i < ents.size()) {
                size_t j = i;
                while (j < ents.size() && ents[j].key == ents[i].key) ++j;
                uint64_t k = ents[i].key;
                FGroup g;
                g.b2 = b2;
                g.pin = (int)(k >> 56);
                g.pout = (int)((k >> 48) & 255);
                g.row = (int)((k >> 40) & 255);
                g.col = (int)((k >> 32) & 255);
                g.rspin = (int)((k >> 30) & 3);
                g.cspin = (int)((k >> 28) & 3);
                g.k = (int)((k >> 21) & 127);
                g.a = (int)((k >> 14) & 127);
                for (size_t q = i; q < j; ++q) g.mem.push_back({ents[q].b1, ents[q].w});
                groups[b2].push_back(std::move(g));
                i = j;
            }
        },
        8, 1);
    // in column order: the rows with direct terms, the group ids
    for (int b2 = 0; b2 < s.W.cols; ++b2) {
        if (skip_rskip_cols && s.W.rskip(b2)) continue;
        if (!fused_col[b2]) {
            for (int b1 : s.W.col_b1[b2]) cp.direct[b1] = 1;
            continue;
        }
        for (auto &g : groups[b2]) {
            for (auto &m : g.mem) cp.member[m.first] = 1;
            cp.fcol[b2].push_back((int)cp.fg.size());
            cp.fg.push_back(std::move(g));
        }
    }
    return cp;
}

// ---------------- T = L^T psi --------------------------------------------
struct TB {  // one T block (b1, lb, lk)
    int lb, lk;
    int64_t off;  // offset of its first row inside the T arena (column 0)
    int M;        // leading dimension (rows of T_lk)
};
struct TLay {
    std::vector<int> M, Mtot;      // [lk] rows of T_lk / rows of the arranged panel LA_lk
    std::vector<int64_t> la_off, t_off;
    std::vector<std::vector<TB>> tb;  // [b1]
    int64_t la_total = 0, t_total = 0;
    double gemm_flops = 0;
};

// BoundaryMPSProduct A-operand blocks (boundary_times_mps.hpp:188-209).
// With cp == nullptr every MPO row gets T rows (the reference's schedule, used
// for its FLOP count).  Otherwise:
//   LA arena = [LA_lk panels: direct rows, then staged rows | fused panels]
//   T arena  = [T_lk panels (direct rows)                   | fused T panels]
// A fused panel, keyed (lk, input block, input state), stacks the (group, lb)
// rows of every fused group at that input; its T GEMM reads only the psi
// columns of that input state (one right charge ar).  The combination strips
// (comb) build the fused panels from the arranged rows.
TLay t_layout(const Site &s, const BLayout &L, const PRP &prp, ColProg *cp, std::vector<Copy2D> *arr,
              StripList *comb, GemmList *g1, int tm, int tn, int kc) {
    struct A {
        int lb, lk;
        const BBlk *blk;
        bool trans;
        double scale;
        int row;
    };
    const int nb1 = s.W.rows;
    TLay t;
    t.M.assign(s.nl, 0);
    t.Mtot.assign(s.nl, 0);
    t.tb.resize(nb1);
    std::vector<std::vector<A>> blk(nb1);
    std::vector<char> real(nb1, 0);
    for (int b1 = 0; b1 < nb1; ++b1) {
        real[b1] = cp ? (cp->direct[b1] || cp->member[b1]) : 1;
        if (!real[b1]) continue;
        bool skip = s.W.lskip(b1);
        for (auto &x : L.b[skip ? s.W.lherm[b1] : b1]) {
            int p = s.li.get(x.lc), q = s.li.get(x.rc);
            if (p < 0 || q < 0 || prp.off[p] < 0 || prp.off[q] < 0) continue;
            int lb = skip ? p : q, lk = skip ? q : p;
            double sc = (skip && s.su2) ? conj_correction(x.lc, x.rc, s.W.lspin[b1]) * s.W.lphase[b1] : 1.0;
            if (x.rows * x.cols != s.lsize(lb) * s.lsize(lk)) throw std::runtime_error("boundary block size mismatch");
            blk[b1].push_back({lb, lk, &x, !skip, sc, -1});
        }
    }
    for (int b1 = 0; b1 < nb1; ++b1)
        if (!cp || cp->direct[b1])
            for (auto &a : blk[b1]) {
                a.row = t.M[a.lk];
                t.M[a.lk] += s.lsize(a.lb);
            }
    t.Mtot = t.M;
    for (int b1 = 0; b1 < nb1; ++b1)
        if (cp && !cp->direct[b1] && cp->member[b1])
            for (auto &a : blk[b1]) {
                a.row = t.Mtot[a.lk];
                t.Mtot[a.lk] += s.lsize(a.lb);
            }
    t.la_off.assign(s.nl, -1);
    t.t_off.assign(s.nl, -1);
    for (int lk = 0; lk < s.nl; ++lk) {
        if (t.Mtot[lk] == 0) continue;
        t.la_off[lk] = t.la_total;
        t.la_total += (int64_t)t.Mtot[lk] * s.lsize(lk);
        if (t.M[lk] == 0) continue;
        t.t_off[lk] = t.t_total;
        t.t_total += (int64_t)t.M[lk] * s.pbr_sz_l[lk];
        t.gemm_flops += 2.0 * t.M[lk] * s.pbr_sz_l[lk] * s.lsize(lk);
    }
    // T_lk stores each (b1, lb) row block contiguously (|lb| x N, leading dimension
    // |lb|), so a W-application source slice is one contiguous range
    for (int b1 = 0; b1 < nb1; ++b1) {
        if (!real[b1]) continue;
        bool in_t = !cp || cp->direct[b1];
        for (auto &a : blk[b1]) {
            if (in_t)
                t.tb[b1].push_back(
                    {a.lb, a.lk, t.t_off[a.lk] + (int64_t)a.row * s.pbr_sz_l[a.lk], s.lsize(a.lb)});
            if (arr) {
                int lbs = s.lsize(a.lb), lks = s.lsize(a.lk);
                arr->push_back(mk_copy(t.la_off[a.lk] + a.row, a.blk->off, lbs, lks, t.Mtot[a.lk],
                                       a.trans ? lks : lbs, a.trans, a.scale));
            }
        }
    }
    if (g1) {
        // row placements of the T_lk panels: rows [row, row + |lb|) of a direct block
        std::vector<std::vector<std::pair<int, int>>> rows_of(s.nl);  // [lk] (row, |lb|)
        for (int b1 = 0; b1 < nb1; ++b1)
            if (real[b1] && (!cp || cp->direct[b1]))
                for (auto &a : blk[b1]) rows_of[a.lk].push_back({a.row, s.lsize(a.lb)});
        std::vector<int64_t> map0(s.nl + 1, (int64_t)g1->rowmap.size());
        for (int lk = 0; lk < s.nl; ++lk) map0[lk + 1] = map0[lk] + t.M[lk];
        g1->rowmap.resize((size_t)map0[s.nl]);
        parallel_for(
            s.nl,
            [&](int lk) {
                RowMap *rm = g1->rowmap.data() + map0[lk];
                std::fill_n(rm, t.M[lk], RowMap{-1, 0, 0});
                for (auto &rb : rows_of[lk])
                    for (int r = 0; r < rb.second; ++r)
                        rm[rb.first + r] = {(int64_t)rb.first * s.pbr_sz_l[lk] + r, rb.second, 0};
                for (int r = 0; r < t.M[lk]; ++r)
                    if (rm[r].off < 0) throw std::runtime_error("T panel row without a block");
            },
            1, 1);
        for (int lk = 0; lk < s.nl; ++lk) {
            if (t.M[lk] == 0) continue;
            GemmGroup g{};
            g.a = t.la_off[lk];
            g.b = prp.off[lk];
            g.c = t.t_off[lk];
            g.M = t.M[lk];
            g.N = s.pbr_sz_l[lk];
            g.K = s.lsize(lk);
            g.lda = t.Mtot[lk];
            g.ldb = g.K;
            g.ldc = g.M;
            g.rowmap = map0[lk] + 1;
            add_gemm(*g1, g, tm, tn, kc);
        }
    }
    if (!cp || cp->fg.empty()) return t;

    // ---- fused groups: blocks, panels, combination strips, T GEMMs
    // Per-group work runs on a thread pool; every order-dependent step (panel and
    // row assignment, source offsets) is sequential in group order, so the plan
    // does not depend on the thread count.
    struct Panel {
        int lk, pin, row, ar, rows = 0;
        int64_t lf = 0, tf = 0;
        std::vector<std::pair<int, int>> blocks;  // (first row, rows) in row order
    };
    struct Cand {
        int lk, lb, ar, al;
    };
    const int nl = s.nl;
    const int ng = (int)cp->fg.size();
    // 1. (parallel) union of the members' (lb, lk) blocks, in (lk, lb) order, kept when
    //    the charges and spins let the block contribute
    std::vector<std::vector<Cand>> cand(ng);
    parallel_for(ng, [&](int gi) {
        thread_local std::vector<char> mark;
        mark.assign((size_t)nl * nl, 0);
        const FGroup &g = cp->fg[gi];
        const int ap = s.W.rspin[g.b2];
        for (auto &m : g.mem)
            for (auto &a : blk[m.first]) mark[(size_t)a.lk * nl + a.lb] = 1;
        for (int lk = 0; lk < nl; ++lk)
            for (int lb = 0; lb < nl; ++lb) {
                if (!mark[(size_t)lk * nl + lb]) continue;
                int ar = s.lp_r[(size_t)lk * s.np + g.pin];
                if (ar < 0 || s.pbr_sz_l[lk] == 0) continue;
                int al = s.lp_r[(size_t)lb * s.np + g.pout];
                if (al < 0 || s.x_of_r[al] < 0) continue;  // out_left_i
                if (s.su2 && !tri(spin2(s.right.q[ar]), ap, spin2(s.right.q[al]))) continue;
                cand[gi].push_back({lk, lb, ar, al});
            }
    });
    // 2. (sequential) panels keyed by (lk, input block, input state < 256), rows in group order
    const int nrow_states = 256;
    std::vector<int> panel_id((size_t)nl * s.np * nrow_states, -1);
    std::vector<Panel> panels;
    std::vector<std::vector<std::pair<int, int>>> fb_panel(ng);  // per group block: (panel, row)
    for (int gi = 0; gi < ng; ++gi) {
        FGroup &g = cp->fg[gi];
        g.blk.reserve(cand[gi].size());
        fb_panel[gi].reserve(cand[gi].size());
        for (auto &c : cand[gi]) {
            int &pid = panel_id[((size_t)c.lk * s.np + g.pin) * nrow_states + g.row];
            if (pid < 0) {
                pid = (int)panels.size();
                Panel pn;
                pn.lk = c.lk;
                pn.pin = g.pin;
                pn.row = g.row;
                pn.ar = c.ar;
                panels.push_back(pn);
            }
            g.blk.push_back({c.lb, c.lk, c.ar, c.al, 0, 0, 0});
            fb_panel[gi].push_back({pid, panels[pid].rows});
            panels[pid].blocks.push_back({panels[pid].rows, s.lsize(c.lb)});
            panels[pid].rows += s.lsize(c.lb);
        }
    }
    for (auto &pn : panels) {
        pn.lf = t.la_total;
        t.la_total += (int64_t)pn.rows * s.lsize(pn.lk);
        pn.tf = t.t_total;
        t.t_total += (int64_t)pn.rows * s.rsize(pn.ar);
    }
    // 3. (parallel) block offsets and, per block, the number of member sources
    std::vector<std::vector<int>> cnt(ng);
    parallel_for(ng, [&](int gi) {
        FGroup &g = cp->fg[gi];
        const size_t nb = g.blk.size();
        for (size_t k = 0; k < nb; ++k) {
            FBlk &fb = g.blk[k];
            const Panel &pn = panels[fb_panel[gi][k].first];
            int row = fb_panel[gi][k].second;
            fb.lf = pn.lf + row;
            fb.tf = pn.tf + (int64_t)row * s.rsize(pn.ar);
            fb.ld = pn.rows;
        }
        if (!comb) return;
        thread_local std::vector<int> slot;
        slot.assign((size_t)nl * nl, -1);
        for (size_t k = 0; k < nb; ++k) slot[(size_t)g.blk[k].lk * nl + g.blk[k].lb] = (int)k;
        cnt[gi].assign(nb, 0);
        for (auto &m : g.mem)
            for (auto &a : blk[m.first]) {
                int k = slot[(size_t)a.lk * nl + a.lb];
                if (k >= 0) ++cnt[gi][k];
            }
    });
    std::vector<uint64_t> comb_key;  // (lk, lb): strips reading one LA_lk row range run together
    if (comb) {
        // 4. strips and source ranges in group order: offsets by prefix sums, then
        //    (parallel) each group's strips
        std::vector<int> gbase(ng + 1, (int)comb->srcs.size());
        std::vector<size_t> sbase(ng + 1, comb->strips.size());
        for (int gi = 0; gi < ng; ++gi) {
            int64_t nsrc = 0;
            for (int c : cnt[gi]) nsrc += c;
            if (gbase[gi] + nsrc > INT32_MAX) throw std::runtime_error("fused-row combination: too many sources");
            gbase[gi + 1] = gbase[gi] + (int)nsrc;
            sbase[gi + 1] = sbase[gi] + cp->fg[gi].blk.size();
        }
        comb->strips.resize(sbase[ng]);
        comb_key.resize(sbase[ng] - sbase[0]);
        parallel_for(ng, [&](int gi) {
            const FGroup &g = cp->fg[gi];
            int base = gbase[gi];
            for (size_t k = 0; k < g.blk.size(); ++k) {
                const FBlk &fb = g.blk[k];
                Strip &st = comb->strips[sbase[gi] + k];
                st.dst = fb.lf;
                st.rows = s.lsize(fb.lb);
                st.cols = s.lsize(fb.lk);
                st.ldd = fb.ld;
                st.s0 = base;
                st.s1 = base + cnt[gi][k];
                base = st.s1;
                comb_key[sbase[gi] - sbase[0] + k] = ((uint64_t)fb.lb << 32) | (uint32_t)fb.lk;
            }
        });
        comb->srcs.resize(gbase[ng]);
        // 5. (parallel) sources of each block, in member order, into the group's range
        parallel_for(ng, [&](int gi) {
            const FGroup &g = cp->fg[gi];
            const size_t nb = g.blk.size();
            thread_local std::vector<int> slot, pos;
            slot.assign((size_t)nl * nl, -1);
            pos.assign(nb, 0);
            int base = gbase[gi];
            for (size_t k = 0; k < nb; ++k) {
                slot[(size_t)g.blk[k].lk * nl + g.blk[k].lb] = (int)k;
                pos[k] = base;
                base += cnt[gi][k];
            }
            for (auto &m : g.mem)
                for (auto &a : blk[m.first]) {
                    int k = slot[(size_t)a.lk * nl + a.lb];
                    if (k >= 0) comb->srcs[pos[k]++] = {t.la_off[a.lk] + a.row, t.Mtot[a.lk], m.second};
                }
        });
    }
    if (comb) {
        if (comb_key.size() != comb->strips.size()) throw std::runtime_error("combination strip list out of step");
        // stable counting sort by (lk, lb)
        std::vector<int> start((size_t)nl * nl + 1, 0);
        for (uint64_t k : comb_key) ++start[(size_t)(uint32_t)k * nl + (k >> 32) + 1];
        for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
        TaskVec<Strip> sorted(comb->strips.size());
        for (size_t i = 0; i < comb_key.size(); ++i)
            sorted[start[(size_t)(uint32_t)comb_key[i] * nl + (comb_key[i] >> 32)]++] = comb->strips[i];
        comb->strips.swap(sorted);
    }
    // row maps of the fused panels: each panel's rows at a known offset, filled in parallel
    std::vector<int64_t> map0(panels.size() + 1, g1 ? (int64_t)g1->rowmap.size() : 0);
    for (size_t k = 0; k < panels.size(); ++k) map0[k + 1] = map0[k] + panels[k].rows;
    if (g1) {
        g1->rowmap.resize((size_t)map0.back());
        parallel_for(
            (int)panels.size(),
            [&](int k) {
                const Panel &pn = panels[k];
                const int rsz = s.rsize(pn.ar);
                RowMap *rm = g1->rowmap.data() + map0[k];
                for (auto &rb : pn.blocks)
                    for (int r = 0; r < rb.second; ++r) *rm++ = {(int64_t)rb.first * rsz + r, rb.second, 0};
            },
            1, 1);
    }
    for (size_t k = 0; k < panels.size(); ++k) {
        const Panel &pn = panels[k];
        int rsz = s.rsize(pn.ar), ksz = s.lsize(pn.lk);
        t.gemm_flops += 2.0 * pn.rows * rsz * ksz;
        if (!g1) continue;
        GemmGroup g{};
        g.a = pn.lf;
        g.b = prp.off[pn.lk] + (int64_t)(s.pbr[(size_t)pn.pin * s.nr + pn.ar] + pn.row * rsz) * ksz;
        g.c = pn.tf;
        g.M = pn.rows;
        g.N = rsz;
        g.K = ksz;
        g.lda = pn.rows;
        g.ldb = ksz;
        g.ldc = pn.rows;
        g.rowmap = map0[k] + 1;
        add_gemm(*g1, g, tm, tn, kc);
    }
    return t;
}

// raw W-application contribution inside one MPO column
struct Raw {
    int al, ar;  // right-index ids of the output block (al = bra side, ar = ket side)
    int row0;    // row offset of the strip inside the al block
    int64_t off;
    int ld;
    double coef;
};

// Visits every (term, T block, operator block) of column b2 that passes the
// charge and spin-triangle checks of the lbtm kernel (apply_op.hpp:70-77;
// abelian apply_op.hpp:24-139): f(term, tb, ob, ar, al).
template <class F>
void for_column_blocks(const Site &s, const TLay &t, const std::vector<CTerm> &terms, int b2, F &&f) {
    int ap = s.W.rspin[b2];
    for (auto &term : terms) {
        const auto &tbs = t.tb[term.row];
        for (auto &tb : tbs)
            for (auto &ob : s.opb[term.op]) {
                int ar = s.lp_r[(size_t)tb.lk * s.np + ob.pin];
                if (ar < 0) continue;
                int al = s.lp_r[(size_t)tb.lb * s.np + ob.pout];
                if (al < 0 || s.x_of_r[al] < 0) continue;  // out_left_i
                if (s.su2 && !tri(spin2(s.right.q[ar]), ap, spin2(s.right.q[al]))) continue;
                f(term, tb, ob, ar, al);
            }
    }
}

// lbtm W application for column b2 (apply_op.hpp:23-101; abelian apply_op.hpp:24-139):
// raw sources of the direct terms and of the fused groups, sorted by strip.
void lbtm_column(const Site &s, const TLay &t, const ColProg *cp, const std::vector<CTerm> &terms, int b2,
                 std::vector<Raw> &raw) {
    raw.clear();
    int ap = s.W.rspin[b2];
    static const std::array<double, 4> one{1.0, 1.0, 1.0, 1.0};
    for_column_blocks(s, t, terms, b2, [&](const CTerm &term, const TB &tb, const Site::OpB &ob, int ar, int al) {
        const Op &op = s.W.ops[term.op];
        int i = spin2(s.left.q[tb.lb]), j = spin2(s.left.q[tb.lk]);
        int jp = spin2(s.right.q[ar]), ip = spin2(s.right.q[al]);
        int lsz = s.lsize(tb.lb), rsz = s.rsize(ar);
        const std::array<double, 4> &cpl =
            s.su2 ? coupling(j, std::abs(j - jp), jp, term.spin, op.spin, ap, i, std::abs(i - ip), ip) : one;
        int iro = s.pbr[(size_t)ob.pin * s.nr + ar];
        int olo = s.pbl[(size_t)ob.pout * s.nl + tb.lb];
        for (auto &e : ob.b->e) {
            double coef = e.coeff * term.scale * (s.su2 ? cpl[case_of(e.rspin, e.cspin)] : 1.0);
            raw.push_back({al, ar, olo + e.col * lsz, tb.off + (int64_t)(iro + e.row * rsz) * tb.M, tb.M, coef});
        }
    });
    if (cp)
        for (int gi : cp->fcol[b2]) {
            const FGroup &g = cp->fg[gi];
            for (auto &fb : g.blk) {
                int i = spin2(s.left.q[fb.lb]), j = spin2(s.left.q[fb.lk]);
                int jp = spin2(s.right.q[fb.ar]), ip = spin2(s.right.q[fb.al]);
                const std::array<double, 4> &cpl =
                    s.su2 ? coupling(j, std::abs(j - jp), jp, g.a, g.k, ap, i, std::abs(i - ip), ip) : one;
                double coef = s.su2 ? cpl[case_of(g.rspin, g.cspin)] : 1.0;
                int row0 = s.pbl[(size_t)g.pout * s.nl + fb.lb] + g.col * s.lsize(fb.lb);
                raw.push_back({fb.al, fb.ar, row0, fb.tf, s.lsize(fb.lb), coef});
            }
        }
    std::sort(raw.begin(), raw.end(), [](const Raw &x, const Raw &y) {
        if (x.al != y.al) return x.al < y.al;
        if (x.ar != y.ar) return x.ar < y.ar;
        if (x.row0 != y.row0) return x.row0 < y.row0;
        return x.off < y.off;
    });
    // merge identical T slices of one strip (several operators of one (row, b2) pair)
    size_t w = 0;
    for (size_t r = 0; r < raw.size(); ++r) {
        if (w > 0 && raw[w - 1].al == raw[r].al && raw[w - 1].ar == raw[r].ar && raw[w - 1].row0 == raw[r].row0 &&
            raw[w - 1].off == raw[r].off)
            raw[w - 1].coef += raw[r].coef;
        else
            raw[w++] = raw[r];
    }
    raw.resize(w);
}

// Column terms of the reference (every MPO term on its own row).
std::vector<CTerm> reference_terms(const Site &s, int b2) {
    std::vector<CTerm> v;
    for (int b1 : s.W.col_b1[b2])
        for (auto &term : s.W.terms(b1, b2)) v.push_back({b1, term.op, term.scale, s.W.lspin[b1]});
    return v;
}

// axpy FLOPs of the reference's lbtm kernel for column b2 (2 len per sparse entry)
double reference_axpy_flops(const Site &s, const TLay &tall, int b2) {
    double f = 0;
    for_column_blocks(s, tall, reference_terms(s, b2), b2,
                      [&](const CTerm &, const TB &tb, const Site::OpB &ob, int ar, int) {
                          f += 2.0 * s.lsize(tb.lb) * s.rsize(ar) * ob.b->e.size();
                      });
    return f;
}

// R operand of gemm_trim for (b2, ar, al): R[b2](ar,al) or phase*R[conj](al,ar)^T
bool r_operand(const Site &s, const BLayout &R, int b2, Q ar, Q al, const BBlk **blk, bool *trans, double *scale) {
    if (s.W.rskip(b2)) {
        int cj = s.W.rherm[b2];
        const BBlk *x = R.find(cj, al, ar);
        if (!x) return false;
        *blk = x;
        *trans = true;
        *scale = s.su2 ? conj_correction(ar, al, s.W.rspin[b2]) * s.W.rphase[cj] : 1.0;
        return true;
    }
    const BBlk *x = R.find(b2, ar, al);
    if (!x) return false;
    *blk = x;
    *trans = false;
    *scale = 1.0;
    return true;
}

struct Seg {
    int b2, al, ar, koff;
    const BBlk *rblk;
    bool rtrans;
    double rscale;
};

// Row strips (row0, rows) of each output block al, in row order: the (phys, lb,
// phys state) partition of the block's rows, computed once per block.
struct RowStrips {
    std::vector<std::vector<std::pair<int, int>>> of;
    std::vector<char> done;
    explicit RowStrips(const Site &s) : of(s.nr), done(s.nr, 0) {}
    const std::vector<std::pair<int, int>> &get(const Site &s, int al) {
        if (!done[al]) {
            auto &rs = of[al];
            for (int p = 0; p < s.np; ++p) {
                int lb = s.rmp_l[(size_t)al * s.np + p];
                if (lb < 0) continue;
                int lsz = s.lsize(lb);
                for (int ss = 0; ss < s.phys.s[p]; ++ss) rs.push_back({s.pbl[(size_t)p * s.nl + lb] + ss * lsz, lsz});
            }
            std::sort(rs.begin(), rs.end());
            done[al] = 1;
        }
        return of[al];
    }
};

// Walks the row strips of segment sg's output block alongside the segment's sources
// [r0, r1), which are sorted by row0: strip(row0, rows, q0, q1) receives the sources
// [q0, q1) whose row0 is the strip's.  Returns the sources placed.
template <class F>
int64_t walk_segment(const std::vector<std::pair<int, int>> &rs, const Raw *r0, const Raw *r1, F &&strip) {
    const Raw *q = r0;
    int64_t placed = 0;
    for (auto &r : rs) {
        while (q < r1 && q->row0 < r.first) ++q;
        const Raw *q0 = q;
        while (q < r1 && q->row0 == r.first) ++q;
        strip(r.first, r.second, q0, q);
        placed += q - q0;
    }
    if (q != r1) throw std::runtime_error("W-application source outside its segment");
    return placed;
}

// The W-application strips of every column (lbtm_column), in column order, which is
// the plan's strip and source order.  A batch of columns at a time:
//  1. (parallel) each column's sources are gathered and sorted, cut into segments
//     (al, ar), which accept(sg) keeps or drops (and may complete, e.g. with the R
//     operand), and counted; x = extra(b2) is computed beside;
//  2. (in column order) each kept segment gets its K offset inside its al block
//     (koff) and its first strip and source; on_seg(sg) sees them in order;
//  3. (parallel) the strips and sources are written at their places.
// Strip dst offsets are relative to the al block start (patch_strips adds the C
// arena offset of al once every K_al is known).  skip_rskip: the columns stored as
// their hermitian partner are left out.  Returns the sum of extra over the columns.
template <class Extra, class Accept, class OnSeg>
double column_strips(const Site &s, const TLay &t, const ColProg &cp, bool skip_rskip, Extra &&extra, Accept &&accept,
                   OnSeg &&on_seg, std::vector<int> &K, std::vector<Seg> &segs, StripList &sl,
                   std::vector<int> &strip_al) {
    RowStrips rows(s);
    for (int al = 0; al < s.nr; ++al)
        if (s.x_of_r[al] >= 0) rows.get(s, al);  // filled here: read-only in the parallel phases
    struct CSeg {
        Seg sg;
        size_t i, j;          // its sources in the column's raw list
        int64_t strip0, src0;  // its first strip and source in the plan's lists
    };
    const int batch = 512;
    const int nbuf = std::min(batch, s.W.cols);
    std::vector<std::vector<Raw>> raws(nbuf);
    std::vector<std::vector<CSeg>> cs(nbuf);
    std::vector<double> ex(nbuf, 0.0);
    double extra_sum = 0.0;
    for (int c0 = 0; c0 < s.W.cols; c0 += batch) {
        const int nb = std::min(batch, s.W.cols - c0);
        parallel_for(
            nb,
            [&](int k) {
                const int b2 = c0 + k;
                cs[k].clear();
                if (skip_rskip && s.W.rskip(b2)) return;
                std::vector<Raw> &raw = raws[k];
                lbtm_column(s, t, &cp, cp.col[b2], b2, raw);
                ex[k] = extra(b2);
                size_t i = 0;
                while (i < raw.size()) {
                    size_t j = i;
                    while (j < raw.size() && raw[j].al == raw[i].al && raw[j].ar == raw[i].ar) ++j;
                    CSeg c{Seg{b2, raw[i].al, raw[i].ar, 0, nullptr, false, 1.0}, i, j, 0, 0};
                    if (accept(c.sg)) cs[k].push_back(c);
                    i = j;
                }
            },
            8, 1);
        int64_t strip_n = (int64_t)sl.strips.size(), src_n = (int64_t)sl.srcs.size();
        for (int k = 0; k < nb; ++k) {
            const int b2 = c0 + k;
            if (skip_rskip && s.W.rskip(b2)) continue;
            for (auto &c : cs[k]) {
                const std::vector<std::pair<int, int>> &rs = rows.of[c.sg.al];
                c.sg.koff = K[c.sg.al];
                K[c.sg.al] += s.rsize(c.sg.ar);
                c.strip0 = strip_n;
                c.src0 = src_n;
                strip_n += (int64_t)rs.size();
                src_n += walk_segment(rs, raws[k].data() + c.i, raws[k].data() + c.j,
                                      [](int, int, const Raw *, const Raw *) {});
                segs.push_back(c.sg);
                on_seg(c.sg);
            }
            extra_sum += ex[k];
        }
        if (src_n > INT32_MAX) throw std::runtime_error("W application: too many sources");
        sl.strips.resize((size_t)strip_n);
        sl.srcs.resize((size_t)src_n);
        strip_al.resize((size_t)strip_n);
        parallel_for(
            nb,
            [&](int k) {
                for (const CSeg &c : cs[k]) {
                    const int rows_al = s.X.blk[s.x_of_r[c.sg.al]].rows;
                    const int rsz = s.rsize(c.sg.ar);
                    int64_t st_i = c.strip0, src_i = c.src0;
                    walk_segment(rows.of[c.sg.al], raws[k].data() + c.i, raws[k].data() + c.j,
                                 [&](int row0, int nrows, const Raw *q0, const Raw *q1) {
                                     Strip &st = sl.strips[st_i];
                                     st.dst = (int64_t)c.sg.koff * rows_al + row0;
                                     st.rows = nrows;
                                     st.cols = rsz;
                                     st.ldd = rows_al;
                                     st.s0 = (int)src_i;
                                     for (const Raw *q = q0; q < q1; ++q) sl.srcs[src_i++] = {q->off, q->ld, q->coef};
                                     st.s1 = (int)src_i;
                                     strip_al[st_i++] = c.sg.al;
                                 });
                }
            },
            8, 1);
    }
    return extra_sum;
}

void patch_strips(StripList &sl, const std::vector<int> &strip_al, const std::vector<int64_t> &c_off) {
    for (size_t i = 0; i < sl.strips.size(); ++i) sl.strips[i].dst += c_off[strip_al[i]];
}

}  // namespace

void tile_gemm(GemmList &gl, GemmGroup g, int tm, int tn, int kc) { add_gemm(gl, g, tm, tn, kc); }

void order_tiles(GemmList &gl) {
    std::vector<GemmTile> t;
    for (auto &x : gl.tiles)
        if (x.g >= 0) t.push_back(x);
    gl.tiles = xcd_order(t, gl.groups);
}

void make_packs(StripList &sl) {
    sl.packs.clear();
    const int n = (int)sl.strips.size();
    int i = 0;
    while (i < n) {
        sl.packs.push_back(i);
        int64_t el = (int64_t)sl.strips[i].rows * sl.strips[i].cols;
        int j = i + 1;
        if (el <= kPackElems)
            while (j < n && j - i < kPackStrips) {
                const int64_t e = (int64_t)sl.strips[j].rows * sl.strips[j].cols;
                if (el + e > kPackElems) break;
                el += e;
                ++j;
            }
        i = j;
    }
    sl.packs.push_back(n);
}

// ---------------- several vectors per launch ------------------------------
std::vector<Copy2D> batch_copies(const std::vector<Copy2D> &v, int nvec, int64_t src_step, int64_t dst_step) {
    std::vector<Copy2D> out;
    out.reserve(v.size() * nvec);
    for (int j = 0; j < nvec; ++j)
        for (Copy2D c : v) {
            c.src += j * src_step;
            c.dst += j * dst_step;
            out.push_back(c);
        }
    return out;
}

GemmList batch_gemm(const GemmList &gl, int nvec, int64_t a_step, int64_t b_step, int64_t c_step) {
    GemmList out;
    const int G = (int)gl.groups.size();
    for (int j = 0; j < nvec; ++j)
        for (GemmGroup g : gl.groups) {
            g.a += j * a_step;
            g.b += j * b_step;
            g.c += j * c_step;
            g.part += j * gl.ws;
            out.groups.push_back(g);
        }
    out.ws = gl.ws * nvec;
    out.rowmap = gl.rowmap;  // placements are relative to each group's c
    // the single-vector list is XCD-interleaved (tile i on XCD i % 8): recover each
    // XCD's sequence, put the per-vector copies of a tile next to each other, order again
    std::vector<GemmTile> seq;
    for (int x = 0; x < kXcds; ++x)
        for (size_t i = x; i < gl.tiles.size(); i += kXcds)
            if (gl.tiles[i].g >= 0)
                for (int j = 0; j < nvec; ++j) {
                    GemmTile t = gl.tiles[i];
                    t.g += j * G;
                    seq.push_back(t);
                }
    out.tiles = seq.empty() ? seq : xcd_order(seq, out.groups);
    return out;
}

// ============================================================ sigma
HostPlan build_sigma_plan(const mq_site_desc &d, int tm, int tn, int kc, FuseRule fuse, bool ref_count,
                          int schedule) {
    if (schedule == SCHED_REFERENCE && d.symm == MQ_SYMM_SU2U1) {
        Mpo w = to_mpo(d.mpo);
        if ((w.rows - w.one_rows) >= (w.cols - w.one_cols)) return build_sigma_plan_rbtm(d, tm, tn, kc, ref_count);
    }
    Site s(d);
    BLayout L = to_layout(d.L), R = to_layout(d.R);
    if (L.aux != s.W.rows || R.aux != s.W.cols) throw std::runtime_error("boundary aux_dim does not match the MPO");
    HostPlan P;
    P.kind = PLAN_SIGMA;
    P.symm = s.symm;
    P.vec_len = s.X.total;
    PRP prp = prp_layout(s, P.reshape);
    P.n_prp = prp.total;
    ColProg cp = column_program(s, false, fuse);
    TLay t = t_layout(s, L, prp, &cp, &P.arrL, &P.combL, &P.g1, tm, tn, kc);
    P.n_la = t.la_total;
    P.n_t = t.t_total;
    std::vector<int> K(s.nr, 0);
    std::vector<Seg> segs;
    std::vector<int> strip_al;
    // reference schedule (site_hamil.hpp:35): when it is lbtm its FLOPs are counted in this pass
    const bool ref_lbtm =
        ref_count && (s.symm != MQ_SYMM_SU2U1 || (s.W.rows - s.W.one_rows) < (s.W.cols - s.W.one_cols));
    TLay tall;
    if (ref_lbtm) tall = t_layout(s, L, prp, nullptr, nullptr, nullptr, nullptr, tm, tn, kc);
    double ref_axpy = 0, ref_trim = 0;
    ref_axpy = column_strips(
        s, t, cp, false, [&](int b2) { return ref_lbtm ? reference_axpy_flops(s, tall, b2) : 0.0; },
        [&](Seg &sg) {
            return r_operand(s, R, sg.b2, s.right.q[sg.ar], s.right.q[sg.al], &sg.rblk, &sg.rtrans, &sg.rscale);
        },
        [&](const Seg &sg) { ref_trim += 2.0 * s.X.blk[s.x_of_r[sg.al]].rows * s.rsize(sg.ar) * s.rsize(sg.al); }, K,
        segs, P.wapply, strip_al);
    std::vector<int64_t> c_off(s.nr, 0), rb_off(s.nr, 0);
    int64_t ct = 0, rbt = 0;
    for (auto &xb : s.X.blk) {
        int r = s.ri.get(xb.c);
        c_off[r] = ct;
        ct += (int64_t)xb.rows * K[r];
        rb_off[r] = rbt;
        rbt += (int64_t)K[r] * xb.cols;
    }
    // the coordinates H_eff couples: every row that a strip with a nonzero source
    // writes, in every column of its output block (strip dst is still relative to
    // its al block here: koff * rows + row0)
    P.live.assign((size_t)s.X.total, 0);
    {
        std::vector<std::vector<uint8_t>> row_live(s.nr);  // per al: rows any strip reaches
        for (size_t i = 0; i < P.wapply.strips.size(); ++i) {
            const Strip &st = P.wapply.strips[i];
            bool any = false;
            for (int k = st.s0; k < st.s1 && !any; ++k) any = P.wapply.srcs[k].coef != 0.0;
            if (!any) continue;
            const VBlk &xb = s.X.blk[s.x_of_r[strip_al[i]]];
            auto &rl = row_live[strip_al[i]];
            if (rl.empty()) rl.assign(xb.rows, 0);
            std::fill_n(rl.begin() + st.dst % xb.rows, st.rows, (uint8_t)1);
        }
        for (auto &xb : s.X.blk) {
            const auto &rl = row_live[s.ri.get(xb.c)];
            if (rl.empty()) continue;
            for (int c = 0; c < xb.cols; ++c)
                std::copy(rl.begin(), rl.end(), P.live.begin() + xb.off + (int64_t)c * xb.rows);
        }
    }
    patch_strips(P.wapply, strip_al, c_off);
    P.n_c = ct;
    P.n_rb = rbt;
    for (auto &sg : segs) {
        int rar = s.rsize(sg.ar), ral = s.rsize(sg.al);
        if (sg.rblk->rows * sg.rblk->cols != rar * ral) throw std::runtime_error("right boundary block size mismatch");
        P.arrR.push_back(mk_copy(rb_off[sg.al] + sg.koff, sg.rblk->off, rar, ral, K[sg.al], sg.rtrans ? ral : rar,
                                 sg.rtrans, sg.rscale));
    }
    for (auto &xb : s.X.blk) {
        int r = s.ri.get(xb.c);
        GemmGroup g{};
        g.a = c_off[r];
        g.b = rb_off[r];
        g.c = xb.off;
        g.M = xb.rows;
        g.N = xb.cols;
        g.K = K[r];
        g.lda = xb.rows;
        g.ldb = std::max(K[r], 1);
        g.ldc = xb.rows;
        add_gemm(P.g2, g, tm, tn, kc);
    }
    P.exec_flops = P.g1.flops() + P.g2.flops();
    for (auto &st : P.wapply.strips) P.exec_flops += 2.0 * st.rows * st.cols * (st.s1 - st.s0);
    if (!ref_count) {
        P.ref_flops = -1;
    } else if (ref_lbtm) {
        P.ref_flops = tall.gemm_flops + ref_axpy + ref_trim;
    } else {
        P.ref_flops = reference_sigma_flops(d);
    }
    P.min_bytes = 8.0 * (L.total + R.total + 2.0 * s.X.total);
    order_tiles(P.g1);
    order_tiles(P.g2);
    make_packs(P.wapply);
    make_packs(P.combL);
    return P;
}

// ============================================================ left step
HostPlan build_lstep_plan(const mq_site_desc &d, int tm, int tn, int kc, FuseRule fuse, bool ref_count) {
    Site s(d);
    BLayout L = to_layout(d.L);
    if (L.aux != s.W.rows) throw std::runtime_error("left boundary aux_dim does not match the MPO");
    HostPlan P;
    P.kind = PLAN_LSTEP;
    P.symm = s.symm;
    P.vec_len = s.X.total;
    PRP prp = prp_layout(s, P.reshape);
    P.n_prp = prp.total;
    ColProg cp = column_program(s, true, fuse);
    TLay t = t_layout(s, L, prp, &cp, &P.arrL, &P.combL, &P.g1, tm, tn, kc);
    TLay tall;
    if (ref_count) tall = t_layout(s, L, prp, nullptr, nullptr, nullptr, nullptr, tm, tn, kc);
    P.n_la = t.la_total;
    P.n_t = t.t_total;
    double axf = 0;
    std::vector<int> K(s.nr, 0);
    std::vector<Seg> segs;
    std::vector<int> strip_al;
    axf = column_strips(
        s, t, cp, true, [&](int b2) { return ref_count ? reference_axpy_flops(s, tall, b2) : 0.0; },
        [&](Seg &sg) {
            // spin filter of SU2::gemm(transpose(contr), bra_conj, ret[b2], spin(b2)) (move_boundary.hpp:165)
            return !s.su2 || tri(spin2(s.right.q[sg.ar]), s.W.rspin[sg.b2], spin2(s.right.q[sg.al]));
        },
        [](const Seg &) {}, K, segs, P.wapply, strip_al);
    std::vector<int64_t> c_off(s.nr, 0), tmp_off(s.nr, 0);
    int64_t ct = 0, tt = 0;
    for (auto &xb : s.X.blk) {
        int r = s.ri.get(xb.c);
        c_off[r] = ct;
        ct += (int64_t)xb.rows * K[r];
        tmp_off[r] = tt;
        tt += (int64_t)K[r] * xb.cols;
    }
    patch_strips(P.wapply, strip_al, c_off);
    P.n_c = ct;
    P.n_tmp = tt;
    for (auto &xb : s.X.blk) {
        int r = s.ri.get(xb.c);
        if (K[r] == 0) continue;
        GemmGroup g{};
        g.a = c_off[r];  // C_al (rows x K) used transposed
        g.b = xb.off;
        g.c = tmp_off[r];
        g.M = K[r];
        g.N = xb.cols;
        g.K = xb.rows;
        g.lda = xb.rows;
        g.ldb = xb.rows;
        g.ldc = K[r];
        g.transA = 1;
        add_gemm(P.g2, g, tm, tn, kc);
    }
    // output boundary: per b2, blocks (ar, al) in descending order
    P.out_aux = s.W.cols;
    std::vector<std::vector<const Seg *>> per_b2(s.W.cols);
    for (auto &sg : segs) per_b2[sg.b2].push_back(&sg);
    P.out_bptr.push_back(0);
    int64_t off = 0;
    double gf = 0;
    for (int b2 = 0; b2 < s.W.cols; ++b2) {
        auto v = per_b2[b2];
        std::sort(v.begin(), v.end(), [&](const Seg *x, const Seg *y) {
            Q xa = s.right.q[x->ar], ya = s.right.q[y->ar];
            if (xa != ya) return xa > ya;
            return s.right.q[x->al] > s.right.q[y->al];
        });
        for (auto *sg : v) {
            Q ar = s.right.q[sg->ar], al = s.right.q[sg->al];
            int rar = s.rsize(sg->ar), ral = s.rsize(sg->al);
            P.out_lc.push_back(ar.a);
            P.out_lc.push_back(ar.b);
            P.out_rc.push_back(al.a);
            P.out_rc.push_back(al.b);
            P.out_rows.push_back(rar);
            P.out_cols.push_back(ral);
            P.out_off.push_back(off);
            P.scatter.push_back(mk_copy(off, tmp_off[sg->al] + sg->koff, rar, ral, rar, K[sg->al], false, 1.0));
            off += (int64_t)rar * ral;
            gf += 2.0 * rar * s.X.blk[s.x_of_r[sg->al]].rows * ral;
        }
        P.out_bptr.push_back((int32_t)P.out_rows.size());
    }
    P.n_out = off;
    P.exec_flops = P.g1.flops() + P.g2.flops();
    for (auto &st : P.wapply.strips) P.exec_flops += 2.0 * st.rows * st.cols * (st.s1 - st.s0);
    for (auto &st : P.combL.strips) P.exec_flops += 2.0 * st.rows * st.cols * (st.s1 - st.s0);
    P.ref_flops = ref_count ? tall.gemm_flops + axf + gf : -1;
    P.min_bytes = 8.0 * (L.total + off + s.X.total);
    order_tiles(P.g1);
    order_tiles(P.g2);
    make_packs(P.wapply);
    make_packs(P.combL);
    return P;
}

// ============================================================ rbtm core
// Shared by the right step and the rbtm sigma: the MPSBoundaryProduct GEMMs
// t[b2] = X * Rop[b2] (boundary_times_mps.hpp:226-293) and the rbtm W
// application (apply_op.hpp:114-209) into BUF_or, which stacks the rows (b1, ol)
// over the right-paired columns of the output left charge or.
namespace {
struct RTB {  // block (c, c') of t[b2] = X_c Rop[b2](c, c')
    int c, cp;      // right-index ids
    int64_t off;    // offset of column 0 inside the t arena
    int ld;         // rows of X_c
};
struct RowBlk {
    int b1, ol, orr, row;
};
struct RbtmCore {
    std::vector<int> Mr;           // [l] rows of BUF_l
    std::vector<RowBlk> rowblks;   // in BUF row order per l
    std::vector<int64_t> buf_off;  // [l] offset of BUF_l in the C arena (-1: none)
    double axf = 0;                // axpy FLOPs of the reference's tasks
};

// g1 reads the X blocks at their canonical offsets in the A arena (psi itself for
// the right step, its copy for the sigma).  sigma: the sigma's rbtm (every MPO
// row, hermitian-partner rows included; a (b1, ol, or) group only when L[b1] has
// the block (ket ol, bra or), as rbtm_axpy_gemm skips the others); otherwise the
// right step (rows whose output boundary slice is stored).
RbtmCore rbtm_core(const Site &s, const BLayout &R, const BLayout *Lsig, HostPlan &P, int tm, int tn, int kc) {
    RbtmCore out;
    // MPSBoundaryProduct: t[b2] = X * Rop[b2] right-trimmed to X's right charges;
    // RC_c stacks Rop[b2](c, c') horizontally over (b2, c').
    struct Rop {
        int b2, c, cp, col;
        const BBlk *blk;
        bool trans;
        double scale;
    };
    std::vector<Rop> rops;
    std::vector<int> Nc(s.nr, 0);
    for (int b2 = 0; b2 < s.W.cols; ++b2) {
        if (s.W.col_nnz[b2] == 0) continue;
        bool skip = s.W.rskip(b2);
        int cj = skip ? s.W.rherm[b2] : b2;
        for (auto &x : R.b[cj]) {
            int c = s.ri.get(skip ? x.rc : x.lc), cp = s.ri.get(skip ? x.lc : x.rc);
            if (c < 0 || cp < 0 || s.x_of_r[c] < 0 || s.x_of_r[cp] < 0) continue;
            double sc = (skip && s.su2) ? conj_correction(s.right.q[c], s.right.q[cp], s.W.rspin[b2]) * s.W.rphase[cj]
                                        : 1.0;
            rops.push_back({b2, c, cp, Nc[c], &x, skip, sc});
            Nc[c] += s.rsize(cp);
        }
    }
    std::vector<int64_t> rc_off(s.nr, -1), t_off(s.nr, -1);
    int64_t rct = 0, tt = 0;
    for (auto &xb : s.X.blk) {
        int c = s.ri.get(xb.c);
        if (Nc[c] == 0) continue;
        rc_off[c] = rct;
        rct += (int64_t)xb.cols * Nc[c];
        t_off[c] = tt;
        tt += (int64_t)xb.rows * Nc[c];
    }
    P.n_rb = rct;
    P.n_t = tt;
    std::vector<std::vector<RTB>> tb(s.W.cols);
    for (auto &r : rops) {
        int rc_ = s.rsize(r.c), rcp = s.rsize(r.cp);
        if (r.blk->rows * r.blk->cols != rc_ * rcp) throw std::runtime_error("right boundary block size mismatch");
        P.arrR.push_back(
            mk_copy(rc_off[r.c] + (int64_t)r.col * rc_, r.blk->off, rc_, rcp, rc_, r.trans ? rcp : rc_, r.trans, r.scale));
        int ld = s.X.blk[s.x_of_r[r.c]].rows;
        tb[r.b2].push_back({r.c, r.cp, t_off[r.c] + (int64_t)r.col * ld, ld});
    }
    for (auto &xb : s.X.blk) {
        int c = s.ri.get(xb.c);
        if (Nc[c] == 0) continue;
        GemmGroup g{};
        g.a = xb.off;
        g.b = rc_off[c];
        g.c = t_off[c];
        g.M = xb.rows;
        g.N = Nc[c];
        g.K = xb.cols;
        g.lda = xb.rows;
        g.ldb = xb.cols;
        g.ldc = xb.rows;
        add_gemm(P.g1, g, tm, tn, kc);
    }
    // rbtm tasks (apply_op.hpp:114-209): buf[b1](ol, or) += coef * t rows.  Per b1: raw list
    // keyed by (or, ol, col0), BUF_or stacks rows over (b1, ol).
    struct RR {
        int orr, ol, col0;
        int64_t off;
        int ld;
        double coef;
    };
    out.Mr.assign(s.nl, 0);
    // per MPO row, in parallel a batch of rows at a time (bounded memory): the
    // row's sources, sorted and merged; then strips in row order
    const int batch = 512;
    std::vector<std::vector<RR>> raws(std::min(batch, std::max(s.W.rows, 1)));
    std::vector<double> axfs(raws.size(), 0.0);
    std::vector<int> order;
    for (int r0 = 0; r0 < s.W.rows; r0 += batch) {
        const int nb = std::min(batch, s.W.rows - r0);
        // the rows with the most MPO entries first, so that none is left to run alone
        order.resize(nb);
        for (int k = 0; k < nb; ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return s.W.row_b2[r0 + x].size() > s.W.row_b2[r0 + y].size();
        });
        parallel_for(
            nb,
            [&](int kk) {
                const int k = order[kk];
                const int b1 = r0 + k;
                std::vector<RR> &raw = raws[k];
                double &axf = axfs[k];
                raw.clear();
                axf = 0.0;
                const bool lskip = s.W.lskip(b1);
                if (!Lsig && lskip) return;
                // (bra, ket) left ids of the L blocks the sigma's charge_gemm can use
                std::unordered_map<uint64_t, char> lblk;
                if (Lsig)
                    for (auto &x : Lsig->b[lskip ? s.W.lherm[b1] : b1]) {
                        int p = s.li.get(x.lc), q = s.li.get(x.rc);
                        if (p < 0 || q < 0) continue;
                        int bra = lskip ? p : q, ket = lskip ? q : p;
                        lblk[((uint64_t)bra << 32) | (uint32_t)ket] = 1;
                    }
                int a = s.W.lspin[b1];
                for (int b2 : s.W.row_b2[b1]) {
                    int ap = s.W.rspin[b2];
                    for (auto &term : s.W.terms(b1, b2)) {
                        const Op &op = s.W.ops[term.op];
                        for (auto &t : tb[b2]) {
                            int jp = spin2(s.right.q[t.c]), ip = spin2(s.right.q[t.cp]);
                            for (auto &ob : s.opb[term.op]) {
                                const OpBlk &wb = *ob.b;
                                int pin = ob.pin, pout = ob.pout;
                                int ol = s.rmp_l[(size_t)t.c * s.np + pin];
                                if (ol < 0) continue;
                                int orr = s.rmp_l[(size_t)t.cp * s.np + pout];
                                if (orr < 0 || s.pbr_sz_l[orr] == 0) continue;  // out_right_i
                                int j = spin2(s.left.q[ol]), i = spin2(s.left.q[orr]);
                                if (s.su2 && !tri(j, a, i)) continue;
                                if (Lsig && !lblk.count(((uint64_t)orr << 32) | (uint32_t)ol)) continue;
                                int lsz = s.lsize(ol), rsz = s.rsize(t.cp);
                                static const std::array<double, 4> one{1.0, 1.0, 1.0, 1.0};
                                const std::array<double, 4> &cpl =
                                    s.su2 ? coupling(j, std::abs(j - jp), jp, a, op.spin, ap, i, std::abs(i - ip), ip)
                                          : one;
                                int ilo = s.pbl[(size_t)pin * s.nl + ol];
                                int oro = s.pbr[(size_t)pout * s.nr + t.cp];
                                for (auto &e : wb.e) {
                                    const double c = s.su2 ? cpl[case_of(e.rspin, e.cspin)] : 1.0;
                                    raw.push_back({orr, ol, oro + e.col * rsz, t.off + ilo + e.row * lsz, t.ld,
                                                   e.coeff * term.scale * c});
                                    axf += 2.0 * lsz * rsz;
                                }
                            }
                        }
                    }
                }
                std::sort(raw.begin(), raw.end(), [](const RR &x, const RR &y) {
                    if (x.orr != y.orr) return x.orr < y.orr;
                    if (x.ol != y.ol) return x.ol < y.ol;
                    if (x.col0 != y.col0) return x.col0 < y.col0;
                    return x.off < y.off;
                });
                size_t w = 0;
                for (size_t r = 0; r < raw.size(); ++r) {
                    if (w > 0 && raw[w - 1].orr == raw[r].orr && raw[w - 1].ol == raw[r].ol &&
                        raw[w - 1].col0 == raw[r].col0 && raw[w - 1].off == raw[r].off)
                        raw[w - 1].coef += raw[r].coef;
                    else
                        raw[w++] = raw[r];
                }
                raw.resize(w);
            },
            4, 1);
        for (int k = 0; k < nb; ++k) {
            const int b1 = r0 + k;
            if (!Lsig && s.W.lskip(b1)) continue;
            const std::vector<RR> &raw = raws[k];
            const int a = s.W.lspin[b1];
            out.axf += axfs[k];
            size_t i = 0;
            while (i < raw.size()) {
                size_t j = i;
                while (j < raw.size() && raw[j].orr == raw[i].orr && raw[j].ol == raw[i].ol) ++j;
                int orr = raw[i].orr, ol = raw[i].ol;
                // spin filter of SU2::gemm(ret[b1], transpose(bra_conj), tmp, spin(b1)) (move_boundary.hpp:224);
                // the sigma's tasks passed the same triangle test (apply_op.hpp:163)
                if (!s.su2 || tri(spin2(s.left.q[ol]), a, spin2(s.left.q[orr]))) {
                    RowBlk rb{b1, ol, orr, out.Mr[orr]};
                    out.Mr[orr] += s.lsize(ol);
                    out.rowblks.push_back(rb);
                    int lsz = s.lsize(ol);
                    Q orq = s.left.q[orr];
                    for (int p = 0; p < s.np; ++p) {
                        int r = s.ri.get(orq + s.phys.q[p]);
                        if (r < 0) continue;
                        int rsz = s.rsize(r);
                        for (int ss = 0; ss < s.phys.s[p]; ++ss) {
                            int col0 = s.pbr[(size_t)p * s.nr + r] + ss * rsz;
                            Strip st;
                            st.dst = 0;  // set below, once every BUF height is known
                            st.rows = lsz;
                            st.cols = rsz;
                            st.ldd = 0;
                            st.s0 = (int)P.wapply.srcs.size();
                            for (size_t q = i; q < j; ++q)
                                if (raw[q].col0 == col0) P.wapply.srcs.push_back({raw[q].off, raw[q].ld, raw[q].coef});
                            st.s1 = (int)P.wapply.srcs.size();
                            P.wapply.strips.push_back(st);
                        }
                    }
                }
                i = j;
            }
        }
    }
    out.buf_off.assign(s.nl, -1);
    int64_t bt = 0;
    for (int l = 0; l < s.nl; ++l) {
        if (out.Mr[l] == 0) continue;
        out.buf_off[l] = bt;
        bt += (int64_t)out.Mr[l] * s.pbr_sz_l[l];
    }
    P.n_c = bt;
    // strip destinations: row block rb.row of BUF_or, column col0 (leading dimension Mr[or]),
    // in the order the strips were emitted
    size_t k = 0;
    for (auto &rb : out.rowblks) {
        Q orq = s.left.q[rb.orr];
        for (int p = 0; p < s.np; ++p) {
            int r = s.ri.get(orq + s.phys.q[p]);
            if (r < 0) continue;
            int rsz = s.rsize(r);
            for (int ss = 0; ss < s.phys.s[p]; ++ss) {
                int col0 = s.pbr[(size_t)p * s.nr + r] + ss * rsz;
                Strip &st = P.wapply.strips[k++];
                st.dst = out.buf_off[rb.orr] + rb.row + (int64_t)col0 * out.Mr[rb.orr];
                st.ldd = out.Mr[rb.orr];
            }
        }
    }
    if (k != P.wapply.strips.size()) throw std::runtime_error("rbtm strip bookkeeping");
    return out;
}
}  // namespace

// ============================================================ right step
HostPlan build_rstep_plan(const mq_site_desc &d, int tm, int tn, int kc) {
    Site s(d);
    BLayout R = to_layout(d.R);
    if (R.aux != s.W.cols) throw std::runtime_error("right boundary aux_dim does not match the MPO");
    HostPlan P;
    P.kind = PLAN_RSTEP;
    P.symm = s.symm;
    P.vec_len = s.X.total;
    PRP prp = prp_layout(s, P.reshape);
    P.n_prp = prp.total;
    RbtmCore core = rbtm_core(s, R, nullptr, P, tm, tn, kc);
    const std::vector<int> &Mr = core.Mr;
    const std::vector<RowBlk> &rowblks = core.rowblks;
    const std::vector<int64_t> &buf_off = core.buf_off;
    const double axf = core.axf;
    std::vector<int64_t> tmp_off(s.nl, -1);
    int64_t mt = 0;
    for (int l = 0; l < s.nl; ++l) {
        if (Mr[l] == 0) continue;
        tmp_off[l] = mt;
        mt += (int64_t)Mr[l] * s.lsize(l);
    }
    P.n_tmp = mt;
    for (int l = 0; l < s.nl; ++l) {
        if (Mr[l] == 0) continue;
        GemmGroup g{};
        g.a = buf_off[l];
        g.b = prp.off[l];  // PRP block (bl x Nr) used transposed
        g.c = tmp_off[l];
        g.M = Mr[l];
        g.N = s.lsize(l);
        g.K = s.pbr_sz_l[l];
        g.lda = Mr[l];
        g.ldb = s.lsize(l);
        g.ldc = Mr[l];
        g.transB = 1;
        add_gemm(P.g2, g, tm, tn, kc);
    }
    P.out_aux = s.W.rows;
    std::vector<std::vector<const RowBlk *>> per_b1(s.W.rows);
    for (auto &rb : rowblks) per_b1[rb.b1].push_back(&rb);
    P.out_bptr.push_back(0);
    int64_t off = 0;
    double gf = 0;
    for (int b1 = 0; b1 < s.W.rows; ++b1) {
        auto v = per_b1[b1];
        std::sort(v.begin(), v.end(), [&](const RowBlk *x, const RowBlk *y) {
            Q xo = s.left.q[x->ol], yo = s.left.q[y->ol];
            if (xo != yo) return xo > yo;
            return s.left.q[x->orr] > s.left.q[y->orr];
        });
        for (auto *rb : v) {
            Q ol = s.left.q[rb->ol], orq = s.left.q[rb->orr];
            int r1 = s.lsize(rb->ol), c1 = s.lsize(rb->orr);
            P.out_lc.push_back(ol.a);
            P.out_lc.push_back(ol.b);
            P.out_rc.push_back(orq.a);
            P.out_rc.push_back(orq.b);
            P.out_rows.push_back(r1);
            P.out_cols.push_back(c1);
            P.out_off.push_back(off);
            P.scatter.push_back(mk_copy(off, tmp_off[rb->orr] + rb->row, r1, c1, r1, Mr[rb->orr], false, 1.0));
            off += (int64_t)r1 * c1;
            gf += 2.0 * r1 * s.pbr_sz_l[rb->orr] * c1;
        }
        P.out_bptr.push_back((int32_t)P.out_rows.size());
    }
    P.n_out = off;
    P.exec_flops = P.g1.flops() + P.g2.flops();
    for (auto &st : P.wapply.strips) P.exec_flops += 2.0 * st.rows * st.cols * (st.s1 - st.s0);
    P.ref_flops = P.g1.flops() + axf + gf;
    P.min_bytes = 8.0 * (R.total + off + s.X.total);
    order_tiles(P.g1);
    order_tiles(P.g2);
    make_packs(P.wapply);
    make_packs(P.combL);
    return P;
}


// ============================================================ sigma, rbtm schedule
// site_hamil_rbtm (non-abelian/site_hamil.hpp:149-204; abelian/site_hamil.hpp for
// TwoU1), the reference's schedule where (rows - one_rows) >= (cols - one_cols):
//   PRP  = psi copied as is (the FP32 forms read it in their operand type)
//   t    = X_c * RC_c              one GEMM per ket right charge (the right step's g1)
//   BUF  = rbtm W application      rows (b1, ol) per output left charge or
//   S_or = LA_or^T * BUF_or        LA_or stacks L[b1](ol, or) over the same rows
//                                  (or the conjugate of the stored partner)
//   y    = S reshaped from right-paired to left-paired (reshape_right_to_left_new)
HostPlan build_sigma_plan_rbtm(const mq_site_desc &d, int tm, int tn, int kc, bool ref_count) {
    Site s(d);
    BLayout L = to_layout(d.L), R = to_layout(d.R);
    if (L.aux != s.W.rows || R.aux != s.W.cols) throw std::runtime_error("boundary aux_dim does not match the MPO");
    HostPlan P;
    P.kind = PLAN_SIGMA;
    P.rbtm = 1;
    P.symm = s.symm;
    P.vec_len = s.X.total;
    for (auto &xb : s.X.blk) P.reshape.push_back(mk_copy(xb.off, xb.off, xb.rows, xb.cols, xb.rows, xb.rows, false, 1.0));
    P.n_prp = s.X.total;
    RbtmCore core = rbtm_core(s, R, &L, P, tm, tn, kc);
    // LA_or (K_or x |or|, leading dimension K_or = Mr[or]): row block rb.row holds
    // L[b1](ol, or) (ket rows, bra columns), or for a hermitian-partner row the
    // stored block (or, ol) of L[conj(b1)] transposed, scaled by the conjugate
    // correction and phase (rbtm_axpy_gemm with conjugate(left[left_conj(b1)]))
    std::vector<int64_t> la_off(s.nl, -1), s_off(s.nl, -1);
    int64_t lat = 0, st = 0;
    for (int l = 0; l < s.nl; ++l) {
        if (s.pbr_sz_l[l] == 0) continue;
        s_off[l] = st;
        st += (int64_t)s.lsize(l) * s.pbr_sz_l[l];
        if (core.Mr[l] == 0) continue;
        la_off[l] = lat;
        lat += (int64_t)core.Mr[l] * s.lsize(l);
    }
    P.n_la = lat;
    P.n_tmp = st;
    for (auto &rb : core.rowblks) {
        const bool skip = s.W.lskip(rb.b1);
        const int b = skip ? s.W.lherm[rb.b1] : rb.b1;
        const int ols = s.lsize(rb.ol), ors = s.lsize(rb.orr);
        const BBlk *x = skip ? L.find(b, s.left.q[rb.orr], s.left.q[rb.ol]) : L.find(b, s.left.q[rb.ol], s.left.q[rb.orr]);
        if (!x) throw std::runtime_error("rbtm row block without its boundary block");
        if (x->rows * x->cols != ols * ors) throw std::runtime_error("boundary block size mismatch");
        const double sc = (skip && s.su2) ? conj_correction(x->lc, x->rc, s.W.lspin[rb.b1]) * s.W.lphase[rb.b1] : 1.0;
        P.arrL.push_back(mk_copy(la_off[rb.orr] + rb.row, x->off, ols, ors, core.Mr[rb.orr], skip ? ors : ols, skip, sc));
    }
    // S_or = LA_or^T BUF_or; output charges no row reaches get K = 0 (zeros)
    for (int l = 0; l < s.nl; ++l) {
        if (s.pbr_sz_l[l] == 0) continue;
        GemmGroup g{};
        g.a = std::max<int64_t>(la_off[l], 0);
        g.b = std::max<int64_t>(core.buf_off[l], 0);
        g.c = s_off[l];
        g.M = s.lsize(l);
        g.N = s.pbr_sz_l[l];
        g.K = core.Mr[l];
        g.lda = std::max(core.Mr[l], 1);
        g.ldb = std::max(core.Mr[l], 1);
        g.ldc = s.lsize(l);
        g.transA = 1;
        add_gemm(P.g2, g, tm, tn, kc);
    }
    // right-paired S -> left-paired y: the inverse of the psi reshape (prp_layout)
    for (auto &xb : s.X.blk) {
        int r = s.ri.get(xb.c);
        for (int pp = 0; pp < s.np; ++pp) {
            int l = s.li.get(xb.c - s.phys.q[pp]);
            if (l < 0 || s_off[l] < 0) continue;
            int ld = s.lsize(l);
            for (int ss = 0; ss < s.phys.s[pp]; ++ss)
                P.scatter.push_back(mk_copy(xb.off + s.pbl[(size_t)pp * s.nl + l] + ss * ld,
                                            s_off[l] + (int64_t)(s.pbr[(size_t)pp * s.nr + r] + ss * xb.cols) * ld, ld,
                                            xb.cols, xb.rows, ld, false, 1.0));
        }
    }
    P.exec_flops = P.g1.flops() + P.g2.flops();
    for (auto &sp : P.wapply.strips) P.exec_flops += 2.0 * sp.rows * sp.cols * (sp.s1 - sp.s0);
    P.ref_flops = ref_count ? reference_sigma_flops(d) : -1;
    P.min_bytes = 8.0 * (L.total + R.total + 2.0 * s.X.total);
    // coordinates H_eff couples: the rows of y any nonzero strip source reaches go
    // through S_or = LA^T BUF, so they are the (or, phys, r) slots whose BUF rows
    // carry a source
    P.live.assign((size_t)s.X.total, 0);
    {
        std::vector<std::vector<char>> col_live(s.nl);
        for (int l = 0; l < s.nl; ++l) col_live[l].assign(s.pbr_sz_l[l], 0);
        size_t k = 0;
        for (auto &rb : core.rowblks) {
            Q orq = s.left.q[rb.orr];
            for (int p = 0; p < s.np; ++p) {
                int r = s.ri.get(orq + s.phys.q[p]);
                if (r < 0) continue;
                int rsz = s.rsize(r);
                for (int ss = 0; ss < s.phys.s[p]; ++ss) {
                    const Strip &sp = P.wapply.strips[k++];
                    bool any = false;
                    for (int q = sp.s0; q < sp.s1 && !any; ++q) any = P.wapply.srcs[q].coef != 0.0;
                    if (!any) continue;
                    int col0 = s.pbr[(size_t)p * s.nr + r] + ss * rsz;
                    for (int c = 0; c < rsz; ++c) col_live[rb.orr][col0 + c] = 1;
                }
            }
        }
        for (auto &xb : s.X.blk) {
            int r = s.ri.get(xb.c);
            for (int pp = 0; pp < s.np; ++pp) {
                int l = s.li.get(xb.c - s.phys.q[pp]);
                if (l < 0 || s_off[l] < 0) continue;
                int ld = s.lsize(l);
                for (int ss = 0; ss < s.phys.s[pp]; ++ss)
                    for (int c = 0; c < xb.cols; ++c)
                        if (col_live[l][s.pbr[(size_t)pp * s.nr + r] + ss * xb.cols + c])
                            std::fill_n(P.live.begin() + xb.off + s.pbl[(size_t)pp * s.nl + l] + ss * ld +
                                            (int64_t)c * xb.rows,
                                        ld, (uint8_t)1);
            }
        }
    }
    order_tiles(P.g1);
    order_tiles(P.g2);
    make_packs(P.wapply);
    return P;
}

// ============================================================ reference FLOPs of sigma
double reference_sigma_flops(const mq_site_desc &d) {
    Site s(d);
    BLayout L = to_layout(d.L), R = to_layout(d.R);
    std::vector<Copy2D> dummy;
    PRP prp = prp_layout(s, dummy);
    bool lbtm = s.symm != MQ_SYMM_SU2U1 || (s.W.rows - s.W.one_rows) < (s.W.cols - s.W.one_cols);
    double f = 0;
    if (lbtm) {
        TLay t = t_layout(s, L, prp, nullptr, nullptr, nullptr, nullptr, 64, 64, 1 << 30);
        f += t.gemm_flops;
        std::vector<Raw> raw;
        for (int b2 = 0; b2 < s.W.cols; ++b2) {
            std::vector<CTerm> terms = reference_terms(s, b2);
            lbtm_column(s, t, nullptr, terms, b2, raw);
            f += reference_axpy_flops(s, t, b2);
            size_t i = 0;
            while (i < raw.size()) {
                size_t j = i;
                while (j < raw.size() && raw[j].al == raw[i].al && raw[j].ar == raw[i].ar) ++j;
                const BBlk *blk;
                bool tr;
                double sc;
                if (r_operand(s, R, b2, s.right.q[raw[i].ar], s.right.q[raw[i].al], &blk, &tr, &sc))
                    f += 2.0 * s.X.blk[s.x_of_r[raw[i].al]].rows * s.rsize(raw[i].ar) * s.rsize(raw[i].al);
                i = j;
            }
        }
        return f;
    }
    // rbtm (site_hamil.hpp:149-204): MPSBoundaryProduct GEMMs (lazy single-use columns
    // counted when used), task axpys of groups with an L block, charge_gemm products.
    std::vector<std::vector<std::pair<int, int>>> tblocks(s.W.cols);  // (c, c') right ids
    auto tgemm = [&](int c, int cp) { return 2.0 * s.X.blk[s.x_of_r[c]].rows * s.rsize(cp) * s.rsize(c); };
    for (int b2 = 0; b2 < s.W.cols; ++b2) {
        bool skip = s.W.rskip(b2);
        for (auto &x : R.b[skip ? s.W.rherm[b2] : b2]) {
            int c = s.ri.get(skip ? x.rc : x.lc), cp = s.ri.get(skip ? x.lc : x.rc);
            if (c < 0 || cp < 0 || s.x_of_r[c] < 0 || s.x_of_r[cp] < 0) continue;
            tblocks[b2].push_back({c, cp});
        }
        if (s.W.col_nnz[b2] != 1)
            for (auto &p : tblocks[b2]) f += tgemm(p.first, p.second);
    }
    std::vector<char> lazy_used(s.W.cols, 0);
    for (int b1 = 0; b1 < s.W.rows; ++b1) {
        int a = s.W.lspin[b1];
        std::unordered_map<uint64_t, char> lblk;  // (bra, ket) left ids present in left_b1
        bool skip = s.W.lskip(b1);
        for (auto &x : L.b[skip ? s.W.lherm[b1] : b1]) {
            int p = s.li.get(x.lc), q = s.li.get(x.rc);
            if (p < 0 || q < 0) continue;
            int bra = skip ? p : q, ket = skip ? q : p;
            lblk[((uint64_t)bra << 32) | (uint32_t)ket] = 1;
        }
        std::unordered_map<uint64_t, std::vector<std::pair<double, int>>> groups;  // (or, ol)
        for (int b2 : s.W.row_b2[b1]) {
            for (auto &term : s.W.terms(b1, b2)) {
                for (auto &pr : tblocks[b2]) {
                    for (auto &ob : s.opb[term.op]) {
                        int ol = s.rmp_l[(size_t)pr.first * s.np + ob.pin];
                        if (ol < 0) continue;
                        int orr = s.rmp_l[(size_t)pr.second * s.np + ob.pout];
                        if (s.su2 && orr >= 0 && !tri(spin2(s.left.q[ol]), a, spin2(s.left.q[orr]))) continue;
                        if (orr < 0 || s.pbr_sz_l[orr] == 0) continue;
                        double fl = 2.0 * s.lsize(ol) * s.rsize(pr.second) * ob.b->e.size();
                        groups[((uint64_t)orr << 32) | (uint32_t)ol].push_back({fl, b2});
                    }
                }
            }
        }
        for (auto &g : groups) {
            int orr = (int)(g.first >> 32), ol = (int)(g.first & 0xffffffffu);
            if (!lblk.count(((uint64_t)orr << 32) | (uint32_t)ol)) continue;
            for (auto &x : g.second) {
                f += x.first;
                if (s.W.col_nnz[x.second] == 1) lazy_used[x.second] = 1;
            }
            f += 2.0 * s.lsize(orr) * s.pbr_sz_l[orr] * s.lsize(ol);
        }
    }
    for (int b2 = 0; b2 < s.W.cols; ++b2)
        if (lazy_used[b2])
            for (auto &p : tblocks[b2]) f += tgemm(p.first, p.second);
    return f;
}

// ============================================================ diagonal of H_eff
// non-abelian/h_diag.hpp:20-152 (the Davidson preconditioner).  For a row slice
// (right charge rc, phys p, left charge lc = rc - p, phys state ss) of the
// left-paired vector and column c:
//   diag = sum_b2 R[b2](rc,rc)[c,c] * sum_b1 alpha(b1,b2) * L[b1](lc,lc)[l,l]
// with alpha summed over the operator terms of (b1,b2), their phys-diagonal
// blocks and the entries with row == col == ss:
//   alpha += coeff * scale * coupling(j, |j-jp|, jp, a, k, ap, j, |j-jp|, jp)[rspin == 2 ? 3 : 0]
// (j = spin(lc), jp = spin(rc)).  Only stored boundary blocks contribute (hermitian
// partners are not stored).  Evaluation: LA = diagonals of the L blocks, RB_rc =
// (K_rc x cols) diagonals of the R blocks over the b2 of a non-empty MPO column,
// T_slice (l x K_rc) from W-application strips over LA, diag_slice = T_slice * RB_rc.
HostPlan build_diag_plan(const mq_site_desc &d, int tm, int tn, int kc) {
    Site s(d);
    if (!s.su2) throw std::runtime_error("Davidson only implemented for spin-adapted Hamiltonians");
    BLayout L = to_layout(d.L), R = to_layout(d.R);
    if (L.aux != s.W.rows || R.aux != s.W.cols) throw std::runtime_error("boundary aux_dim does not match the MPO");
    HostPlan P;
    P.kind = PLAN_DIAG;
    P.symm = s.symm;
    P.vec_len = s.X.total;
    // diagonals of L[b1](lc,lc)
    std::vector<int64_t> dl((size_t)s.W.rows * s.nl, -1);
    int64_t ndl = 0;
    for (int b1 = 0; b1 < s.W.rows; ++b1)
        for (const BBlk &bk : L.b[b1]) {
            if (!(bk.lc == bk.rc)) continue;
            const int l = s.li.get(bk.lc);
            if (l < 0) continue;
            if (bk.rows != s.lsize(l) || bk.cols != bk.rows)
                throw std::runtime_error("left boundary block does not match the left index");
            dl[(size_t)b1 * s.nl + l] = ndl;
            P.arrL.push_back({ndl, bk.off, 1, bk.rows, 1, bk.rows + 1, 0, 1.0});
            ndl += bk.rows;
        }
    // diagonals of R[b2](rc,rc) for the b2 of non-empty MPO columns, per right charge
    std::vector<std::vector<int>> kb2(s.nr);
    std::vector<int64_t> dr_off(s.nr, 0);
    int64_t ndr = 0;
    for (int r = 0; r < s.nr; ++r) {
        if (s.x_of_r[r] < 0) continue;
        const Q rc = s.right.q[r];
        const int cs = s.rsize(r);
        std::vector<const BBlk *> blks;
        for (int b2 = 0; b2 < s.W.cols; ++b2) {
            if (s.W.col_b1[b2].empty()) continue;
            const BBlk *bk = R.find(b2, rc, rc);
            if (!bk) continue;
            if (bk->rows != cs || bk->cols != cs)
                throw std::runtime_error("right boundary block does not match the right index");
            kb2[r].push_back(b2);
            blks.push_back(bk);
        }
        const int K = (int)kb2[r].size();
        dr_off[r] = ndr;
        for (int k = 0; k < K; ++k) P.arrR.push_back({ndr + k, blks[k]->off, 1, cs, K, blks[k]->rows + 1, 0, 1.0});
        ndr += (int64_t)K * cs;
    }
    P.n_la = ndl;
    P.n_rb = ndr;
    // row slices (r, p, ss) of every vector block; T offsets in that order
    struct Slice {
        int p, l, ss;
        int64_t t;
    };
    std::vector<std::vector<Slice>> slices(s.nr);
    std::vector<std::vector<int>> slice_of(s.nr);  // [p * maxss + ss] -> slice index
    int maxss = 1;
    for (int p = 0; p < s.np; ++p) maxss = std::max(maxss, s.phys.s[p]);
    int64_t nt = 0;
    for (int r = 0; r < s.nr; ++r) {
        if (s.x_of_r[r] < 0) continue;
        slice_of[r].assign((size_t)s.np * maxss, -1);
        const int K = (int)kb2[r].size();
        for (int p = 0; p < s.np; ++p) {
            const int l = s.rmp_l[(size_t)r * s.np + p];
            if (l < 0) continue;
            for (int ss = 0; ss < s.phys.s[p]; ++ss) {
                slice_of[r][(size_t)p * maxss + ss] = (int)slices[r].size();
                slices[r].push_back({p, l, ss, nt});
                nt += (int64_t)s.lsize(l) * K;
            }
        }
    }
    P.n_t = nt;
    // strips per right charge (host threads), concatenated in charge order
    std::vector<StripList> part(s.nr);
    parallel_for(s.nr, [&](int r) {
        if (s.x_of_r[r] < 0) return;
        const Q rc = s.right.q[r];
        const int jp = spin2(rc);
        StripList &out = part[r];
        std::vector<std::map<int, double>> acc(slices[r].size());
        for (int k = 0; k < (int)kb2[r].size(); ++k) {
            const int b2 = kb2[r][k];
            const int ap = s.W.rspin[b2];
            for (auto &a_ : acc) a_.clear();
            for (int b1 : s.W.col_b1[b2]) {
                const int a = s.W.lspin[b1];
                for (const TermD &term : s.W.terms(b1, b2)) {
                    const int opk = s.W.ops[term.op].spin;
                    for (const Site::OpB &ob : s.opb[term.op]) {
                        if (ob.pin != ob.pout) continue;
                        const int p = ob.pin;
                        const int l = s.rmp_l[(size_t)r * s.np + p];
                        if (l < 0 || dl[(size_t)b1 * s.nl + l] < 0) continue;
                        const int j = spin2(s.left.q[l]), ts = std::abs(j - jp);
                        const std::array<double, 4> &cp = coupling(j, ts, jp, a, opk, ap, j, ts, jp);
                        for (const Ent &e : ob.b->e) {
                            if (e.row != e.col || e.row < 0 || e.row >= s.phys.s[p]) continue;
                            const int si = slice_of[r][(size_t)p * maxss + e.row];
                            if (si < 0) continue;
                            acc[si][b1] += e.coeff * term.scale * cp[e.rspin == 2 ? 3 : 0];
                        }
                    }
                }
            }
            for (size_t si = 0; si < acc.size(); ++si) {
                if (acc[si].empty()) continue;
                const Slice &sl = slices[r][si];
                const int ls = s.lsize(sl.l);
                Strip st{sl.t + (int64_t)k * ls, ls, 1, ls, (int)out.srcs.size(), 0};
                for (auto &kv : acc[si]) out.srcs.push_back({dl[(size_t)kv.first * s.nl + sl.l], ls, kv.second});
                st.s1 = (int)out.srcs.size();
                out.strips.push_back(st);
            }
        }
    });
    double src_el = 0;
    for (int r = 0; r < s.nr; ++r) {
        const int base = (int)P.wapply.srcs.size();
        for (Strip st : part[r].strips) {
            st.s0 += base;
            st.s1 += base;
            src_el += (double)st.rows * (st.s1 - st.s0);
            P.wapply.strips.push_back(st);
        }
        P.wapply.srcs.insert(P.wapply.srcs.end(), part[r].srcs.begin(), part[r].srcs.end());
    }
    // diag slice = T_slice * RB_rc (K = 0 writes zeros)
    for (int r = 0; r < s.nr; ++r) {
        if (s.x_of_r[r] < 0) continue;
        const VBlk &xb = s.X.blk[s.x_of_r[r]];
        const int K = (int)kb2[r].size();
        for (const Slice &sl : slices[r]) {
            const int ls = s.lsize(sl.l);
            GemmGroup g{};
            g.a = sl.t;
            g.b = dr_off[r];
            g.c = xb.off + s.pbl[(size_t)sl.p * s.nl + sl.l] + (int64_t)sl.ss * ls;
            g.M = ls;
            g.N = xb.cols;
            g.K = K;
            g.lda = ls;
            g.ldb = std::max(K, 1);
            g.ldc = xb.rows;
            add_gemm(P.g2, g, tm, tn, kc);
        }
    }
    order_tiles(P.g2);
    make_packs(P.wapply);
    P.exec_flops = P.g2.flops() + 2.0 * src_el;
    return P;
}

}  // namespace mq
