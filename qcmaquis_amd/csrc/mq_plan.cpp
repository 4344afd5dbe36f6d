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

#include <set>
#include <tuple>

namespace mq {

namespace {

struct Site {
    int symm;
    bool su2;
    Idx phys, left, right;
    Mpo W;
    VLayout X;
    PB pbl;  // PB(phys, left)
    PB pbr;  // PB(phys, right, right-fused)
    Site(const mq_site_desc &d)
        : symm(d.symm), su2(d.symm == MQ_SYMM_SU2U1), phys(to_idx(d.phys)), left(to_idx(d.left)),
          right(to_idx(d.right)), W(to_mpo(d.mpo)) {
        if (d.symm != MQ_SYMM_SU2U1 && d.symm != MQ_SYMM_TWOU1) throw std::runtime_error("unknown symmetry");
        X = vec_layout(phys, left, right);
        pbl = PB(phys, left, false);
        pbr = PB(phys, right, true);
    }
};

int ceil_div(int a, int b) { return (a + b - 1) / b; }

void add_gemm(GemmList &gl, GemmGroup g, int tm, int tn, int kchunk_target) {
    g.split = 1;
    g.kchunk = std::max(g.K, 1);
    int mt = ceil_div(std::max(g.M, 1), tm), nt = ceil_div(std::max(g.N, 1), tn);
    if (g.M == 0 || g.N == 0) return;
    // split K when the output is small and K long: aim for chunks of ~kchunk_target
    if (g.K > 2 * kchunk_target && mt * nt < 64) {
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
    for (int ks = 0; ks < g.split; ++ks)
        for (int n = 0; n < nt; ++n)
            for (int m = 0; m < mt; ++m) gl.tiles.push_back({gi, m, n, ks});
}

// ---------------- right-paired psi (shared by all plans) ----------------
struct PRPLayout {
    std::map<Q, int64_t> off;  // ket-left charge -> offset
    std::map<Q, int> ncol;
    int64_t total = 0;
};

PRPLayout prp_layout(const Site &s, std::vector<Copy2D> &tasks) {
    PRPLayout p;
    for (int i = 0; i < s.left.n(); ++i) {
        Q lk = s.left.q[i];
        int n = s.pbr.sz(lk);
        if (n == 0) continue;
        p.off[lk] = p.total;
        p.ncol[lk] = n;
        p.total += (int64_t)s.left.s[i] * n;
    }
    for (auto &xb : s.X.blk) {
        Q r = xb.c;
        for (int si = 0; si < s.phys.n(); ++si) {
            Q sc = s.phys.q[si];
            Q lk = r - sc;
            if (!s.left.has(lk)) continue;
            int ld = s.left.size_of(lk);
            for (int ss = 0; ss < s.phys.s[si]; ++ss) {
                Copy2D c;
                c.dst = p.off.at(lk) + (int64_t)(s.pbr.at(sc, r) + ss * xb.cols) * ld;
                c.src = xb.off + s.pbl.at(sc, lk) + ss * ld;
                c.rows = ld;
                c.cols = xb.cols;
                c.ldd = ld;
                c.lds = xb.rows;
                c.trans = 0;
                c.scale = 1.0;
                tasks.push_back(c);
            }
        }
    }
    return p;
}

// ---------------- T = L^T psi: stacked left-boundary operand -------------
struct TLayout {
    std::map<Q, int> M;                                    // rows of LA_lk / T_lk
    std::map<Q, int64_t> la_off, t_off;                    // per lk
    std::map<std::tuple<int, Q, Q>, int> rowoff;           // (b1, lb, lk) -> row offset in LA_lk
    std::vector<std::tuple<int, Q, Q>> tblocks;            // (b1, lb, lk) in creation order
    int64_t la_total = 0, t_total = 0;
};

// A-operand blocks of BoundaryMPSProduct for b1: (lb, lk, src block, transposed, scale)
struct ABlk {
    Q lb, lk;
    const BBlk *blk;
    bool trans;
    double scale;
};

std::vector<ABlk> a_blocks(const Site &s, const BLayout &L, int b1, const std::map<Q, int64_t> &ref) {
    std::vector<ABlk> out;
    if (s.W.lskip(b1)) {
        int cj = s.W.lherm[b1];
        for (auto &x : L.b[cj]) {
            if (!ref.count(x.lc) || !ref.count(x.rc)) continue;
            double sc = s.su2 ? conj_correction(x.lc, x.rc, s.W.lspin[b1]) * s.W.lphase[b1] : 1.0;
            out.push_back({x.lc, x.rc, &x, false, sc});
        }
    } else {
        for (auto &x : L.b[b1]) {
            if (!ref.count(x.rc) || !ref.count(x.lc)) continue;
            out.push_back({x.rc, x.lc, &x, true, 1.0});
        }
    }
    return out;
}

TLayout t_layout(const Site &s, const BLayout &L, const PRPLayout &prp, std::vector<Copy2D> &arr, GemmList &g1,
                 int tm, int tn, int kc, bool used_rows_only) {
    TLayout t;
    std::vector<std::vector<ABlk>> ab(s.W.rows);
    for (int b1 = 0; b1 < s.W.rows; ++b1) {
        if (used_rows_only && s.W.row_nnz[b1] == 0) continue;
        ab[b1] = a_blocks(s, L, b1, prp.off);
        for (auto &a : ab[b1]) {
            int &M = t.M[a.lk];
            t.rowoff[{b1, a.lb, a.lk}] = M;
            t.tblocks.push_back({b1, a.lb, a.lk});
            M += s.left.size_of(a.lb);
        }
    }
    for (auto &kv : t.M) {
        Q lk = kv.first;
        t.la_off[lk] = t.la_total;
        t.la_total += (int64_t)kv.second * s.left.size_of(lk);
        t.t_off[lk] = t.t_total;
        t.t_total += (int64_t)kv.second * prp.ncol.at(lk);
    }
    for (int b1 = 0; b1 < s.W.rows; ++b1)
        for (auto &a : ab[b1]) {
            int M = t.M[a.lk];
            int lbs = s.left.size_of(a.lb), lks = s.left.size_of(a.lk);
            Copy2D c;
            c.dst = t.la_off[a.lk] + t.rowoff[{b1, a.lb, a.lk}];
            c.src = a.blk->off;
            c.rows = lbs;
            c.cols = lks;
            c.ldd = M;
            c.lds = a.trans ? lks : lbs;
            c.trans = a.trans ? 1 : 0;
            c.scale = a.scale;
            if (a.blk->rows * a.blk->cols != lbs * lks) throw std::runtime_error("boundary block size mismatch");
            arr.push_back(c);
        }
    for (auto &kv : t.M) {
        Q lk = kv.first;
        GemmGroup g{};
        g.a = t.la_off[lk];
        g.b = prp.off.at(lk);
        g.c = t.t_off[lk];
        g.M = kv.second;
        g.N = prp.ncol.at(lk);
        g.K = s.left.size_of(lk);
        g.lda = kv.second;
        g.ldb = g.K;
        g.ldc = kv.second;
        add_gemm(g1, g, tm, tn, kc);
    }
    return t;
}

// contribution of the W application: (al, b2, ar, row0) <- src
struct Contrib {
    Q al, ar;
    int b2;
    int row0;  // row offset of the strip inside the al block
    int rows;  // strip rows (lb size)
    Src src;
};

// lbtm W application (apply_op.hpp:23-101 / abelian apply_op.hpp:24-139)
void lbtm_contribs(const Site &s, const TLayout &t, const PRPLayout &prp, const std::vector<int> &b2_list,
                   std::vector<Contrib> &out, double *axpy_flops) {
    // T blocks by b1
    std::map<int, std::vector<std::pair<Q, Q>>> tb;
    for (auto &x : t.tblocks) tb[std::get<0>(x)].push_back({std::get<1>(x), std::get<2>(x)});
    for (int b2 : b2_list) {
        int ap = s.W.rspin[b2];
        for (int b1 : s.W.col_b1[b2]) {
            auto it = tb.find(b1);
            if (it == tb.end()) continue;
            int a = s.W.lspin[b1];
            for (auto &term : s.W.at.at({b1, b2})) {
                const Op &op = s.W.ops[term.op];
                int k = op.spin;
                for (auto &lblk : it->second) {
                    Q lb = lblk.first, lk = lblk.second;
                    int lsz = s.left.size_of(lb);
                    int M = t.M.at(lk);
                    int64_t toff = t.t_off.at(lk) + t.rowoff.at({b1, lb, lk});
                    for (auto &wb : op.blocks) {
                        Q ar = lk + wb.in;
                        if (!s.right.has(ar)) continue;
                        Q al = lb + wb.out;
                        if (s.su2 && !tri(spin2(ar), ap, spin2(al))) continue;
                        if (!s.X.pos.count(al)) continue;  // out_left_i (phys*left common with right)
                        int rsz = s.right.size_of(ar);
                        double cpl[4] = {term.scale, term.scale, term.scale, term.scale};
                        if (s.su2) {
                            int i = spin2(lb), ip = spin2(al), j = spin2(lk), jp = spin2(ar);
                            set_coupling(j, std::abs(j - jp), jp, a, k, ap, i, std::abs(i - ip), ip, term.scale, cpl);
                        }
                        int iro = s.pbr.at(wb.in, ar);
                        int olo = s.pbl.at(wb.out, lb);
                        for (auto &e : wb.e) {
                            double coef = e.coeff * (s.su2 ? cpl[case_of(e.rspin, e.cspin)] : term.scale);
                            Contrib c;
                            c.al = al;
                            c.ar = ar;
                            c.b2 = b2;
                            c.row0 = olo + e.col * lsz;
                            c.rows = lsz;
                            c.src.off = toff + (int64_t)(iro + e.row * rsz) * M;
                            c.src.ld = M;
                            c.src.coef = coef;
                            out.push_back(c);
                            if (axpy_flops) *axpy_flops += 2.0 * lsz * rsz;
                        }
                    }
                }
            }
        }
    }
}

// the R operand of gemm_trim for (b2, ar, al): non-skip R[b2](ar,al); skip phase*R[conj](al,ar)^T
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

// column segments of C_al and their strips
struct CLayout {
    std::map<Q, int> K;                       // per al
    std::map<Q, int64_t> off;                 // per al
    std::map<std::tuple<Q, int, Q>, int> koff;  // (al, b2, ar)
    std::vector<std::tuple<Q, int, Q>> segs;  // ordered
    int64_t total = 0;
};

void build_strips(const Site &s, const CLayout &cl, const std::vector<Contrib> &cb, StripList &sl) {
    // group contributions by (al, b2, ar, row0)
    std::map<std::tuple<Q, int, Q, int>, std::vector<Src>> by;
    for (auto &c : cb) {
        auto key = std::make_tuple(c.al, c.b2, c.ar, c.row0);
        if (!cl.koff.count(std::make_tuple(c.al, c.b2, c.ar))) continue;
        by[key].push_back(c.src);
    }
    for (auto &seg : cl.segs) {
        Q al = std::get<0>(seg), ar = std::get<2>(seg);
        int b2 = std::get<1>(seg);
        int rows_al = s.pbl.sz(al);
        int k0 = cl.koff.at(seg);
        int rsz = s.right.size_of(ar);
        for (int si = 0; si < s.phys.n(); ++si) {
            Q so = s.phys.q[si];
            Q lb = al - so;
            if (!s.left.has(lb)) continue;
            int lsz = s.left.size_of(lb);
            for (int ss = 0; ss < s.phys.s[si]; ++ss) {
                int row0 = s.pbl.at(so, lb) + ss * lsz;
                Strip st;
                st.dst = cl.off.at(al) + (int64_t)k0 * rows_al + row0;
                st.rows = lsz;
                st.cols = rsz;
                st.ldd = rows_al;
                st.s0 = (int)sl.srcs.size();
                auto it = by.find(std::make_tuple(al, b2, ar, row0));
                if (it != by.end())
                    for (auto &x : it->second) sl.srcs.push_back(x);
                st.s1 = (int)sl.srcs.size();
                sl.strips.push_back(st);
            }
        }
    }
}

}  // namespace

// ============================================================ sigma
HostPlan build_sigma_plan(const mq_site_desc &d, int tm, int tn, int kc) {
    Site s(d);
    BLayout L = to_layout(d.L), R = to_layout(d.R);
    if (L.aux != s.W.rows || R.aux != s.W.cols) throw std::runtime_error("boundary aux_dim does not match the MPO");
    HostPlan P;
    P.kind = PLAN_SIGMA;
    P.symm = s.symm;
    P.vec_len = s.X.total;
    PRPLayout prp = prp_layout(s, P.reshape);
    P.n_prp = prp.total;
    TLayout t = t_layout(s, L, prp, P.arrL, P.g1, tm, tn, kc, true);
    P.n_la = t.la_total;
    P.n_t = t.t_total;
    std::vector<int> b2s;
    for (int b2 = 0; b2 < s.W.cols; ++b2) b2s.push_back(b2);
    std::vector<Contrib> cb;
    lbtm_contribs(s, t, prp, b2s, cb, nullptr);
    // segments: (al, b2, ar) with contributions and an R operand
    std::set<std::tuple<Q, int, Q>> segset;
    for (auto &c : cb) segset.insert(std::make_tuple(c.al, c.b2, c.ar));
    CLayout cl;
    std::map<Q, std::vector<std::tuple<int, Q, const BBlk *, bool, double>>> rops;
    for (auto &sg : segset) {
        Q al = std::get<0>(sg), ar = std::get<2>(sg);
        int b2 = std::get<1>(sg);
        const BBlk *blk;
        bool tr;
        double sc;
        if (!r_operand(s, R, b2, ar, al, &blk, &tr, &sc)) continue;
        int &K = cl.K[al];
        cl.koff[sg] = K;
        cl.segs.push_back(sg);
        rops[al].push_back(std::make_tuple(b2, ar, blk, tr, sc));
        K += s.right.size_of(ar);
    }
    for (auto &xb : s.X.blk) {
        int K = cl.K.count(xb.c) ? cl.K[xb.c] : 0;
        cl.off[xb.c] = cl.total;
        cl.total += (int64_t)xb.rows * K;
    }
    P.n_c = cl.total;
    build_strips(s, cl, cb, P.wapply);
    // RB arrangement and the output GEMMs
    std::map<Q, int64_t> rb_off;
    int64_t rb_total = 0;
    for (auto &xb : s.X.blk) {
        rb_off[xb.c] = rb_total;
        int K = cl.K.count(xb.c) ? cl.K[xb.c] : 0;
        rb_total += (int64_t)K * xb.cols;
    }
    P.n_rb = rb_total;
    for (auto &kv : rops) {
        Q al = kv.first;
        int K = cl.K[al];
        int ral = s.right.size_of(al);
        for (auto &r : kv.second) {
            int b2 = std::get<0>(r);
            Q ar = std::get<1>(r);
            const BBlk *blk = std::get<2>(r);
            bool tr = std::get<3>(r);
            int rar = s.right.size_of(ar);
            if (blk->rows * blk->cols != rar * ral) throw std::runtime_error("right boundary block size mismatch");
            Copy2D c;
            c.dst = rb_off[al] + cl.koff.at(std::make_tuple(al, b2, ar));
            c.src = blk->off;
            c.rows = rar;
            c.cols = ral;
            c.ldd = K;
            c.lds = tr ? ral : rar;
            c.trans = tr ? 1 : 0;
            c.scale = std::get<4>(r);
            P.arrR.push_back(c);
        }
    }
    for (auto &xb : s.X.blk) {
        int K = cl.K.count(xb.c) ? cl.K[xb.c] : 0;
        GemmGroup g{};
        g.a = cl.off[xb.c];
        g.b = rb_off[xb.c];
        g.c = xb.off;
        g.M = xb.rows;
        g.N = xb.cols;
        g.K = K;
        g.lda = xb.rows;
        g.ldb = std::max(K, 1);
        g.ldc = xb.rows;
        add_gemm(P.g2, g, tm, tn, kc);
    }
    P.exec_flops = P.g1.flops() + P.g2.flops();
    for (auto &st : P.wapply.strips) P.exec_flops += 2.0 * st.rows * st.cols * (st.s1 - st.s0);
    P.ref_flops = reference_sigma_flops(d);
    P.min_bytes = 8.0 * (L.total + R.total + 2.0 * s.X.total);
    return P;
}

// ============================================================ left step
HostPlan build_lstep_plan(const mq_site_desc &d, int tm, int tn, int kc) {
    Site s(d);
    BLayout L = to_layout(d.L);
    if (L.aux != s.W.rows) throw std::runtime_error("left boundary aux_dim does not match the MPO");
    HostPlan P;
    P.kind = PLAN_LSTEP;
    P.symm = s.symm;
    P.vec_len = s.X.total;
    PRPLayout prp = prp_layout(s, P.reshape);
    P.n_prp = prp.total;
    TLayout t = t_layout(s, L, prp, P.arrL, P.g1, tm, tn, kc, true);
    P.n_la = t.la_total;
    P.n_t = t.t_total;
    std::vector<int> b2s;
    for (int b2 = 0; b2 < s.W.cols; ++b2)
        if (!s.W.rskip(b2)) b2s.push_back(b2);
    std::vector<Contrib> cb;
    double axf = 0;
    lbtm_contribs(s, t, prp, b2s, cb, &axf);
    std::set<std::tuple<Q, int, Q>> segset;
    for (auto &c : cb) segset.insert(std::make_tuple(c.al, c.b2, c.ar));
    CLayout cl;
    for (auto &sg : segset) {
        Q al = std::get<0>(sg), ar = std::get<2>(sg);
        int b2 = std::get<1>(sg);
        // spin filter of SU2::gemm(transpose(contr), bra_conj, ret[b2], spin(b2)) (move_boundary.hpp:165)
        if (s.su2 && !tri(spin2(ar), s.W.rspin[b2], spin2(al))) continue;
        int &K = cl.K[al];
        cl.koff[sg] = K;
        cl.segs.push_back(sg);
        K += s.right.size_of(ar);
    }
    int64_t tmp_total = 0;
    std::map<Q, int64_t> tmp_off;
    for (auto &xb : s.X.blk) {
        int K = cl.K.count(xb.c) ? cl.K[xb.c] : 0;
        cl.off[xb.c] = cl.total;
        cl.total += (int64_t)xb.rows * K;
        tmp_off[xb.c] = tmp_total;
        tmp_total += (int64_t)K * xb.cols;
    }
    P.n_c = cl.total;
    P.n_tmp = tmp_total;
    build_strips(s, cl, cb, P.wapply);
    for (auto &xb : s.X.blk) {
        int K = cl.K.count(xb.c) ? cl.K[xb.c] : 0;
        if (K == 0) continue;
        GemmGroup g{};
        g.a = cl.off[xb.c];  // C_al stored rows(al) x K -> op(A) = C^T (K x rows)
        g.b = xb.off;        // X_al rows(al) x r_al
        g.c = tmp_off[xb.c];
        g.M = K;
        g.N = xb.cols;
        g.K = xb.rows;
        g.lda = xb.rows;
        g.ldb = xb.rows;
        g.ldc = K;
        g.transA = 1;
        add_gemm(P.g2, g, tm, tn, kc);
    }
    // output boundary: per b2, blocks (ar, al) descending
    P.out_aux = s.W.cols;
    std::map<int, std::vector<std::pair<Q, Q>>> per_b2;
    for (auto &sg : cl.segs) per_b2[std::get<1>(sg)].push_back({std::get<2>(sg), std::get<0>(sg)});
    P.out_bptr.push_back(0);
    int64_t off = 0;
    for (int b2 = 0; b2 < s.W.cols; ++b2) {
        auto it = per_b2.find(b2);
        if (it != per_b2.end()) {
            auto v = it->second;
            std::sort(v.begin(), v.end(), [](auto &x, auto &y) {
                return x.first > y.first || (x.first == y.first && x.second > y.second);
            });
            for (auto &bl : v) {
                Q ar = bl.first, al = bl.second;
                int rar = s.right.size_of(ar), ral = s.right.size_of(al);
                P.out_lc.push_back(ar.a);
                P.out_lc.push_back(ar.b);
                P.out_rc.push_back(al.a);
                P.out_rc.push_back(al.b);
                P.out_rows.push_back(rar);
                P.out_cols.push_back(ral);
                P.out_off.push_back(off);
                Copy2D c;
                c.dst = off;
                c.src = tmp_off[al] + cl.koff.at(std::make_tuple(al, b2, ar));
                c.rows = rar;
                c.cols = ral;
                c.ldd = rar;
                c.lds = cl.K[al];
                c.trans = 0;
                c.scale = 1.0;
                P.scatter.push_back(c);
                off += (int64_t)rar * ral;
            }
        }
        P.out_bptr.push_back((int32_t)P.out_rows.size());
    }
    P.n_out = off;
    P.exec_flops = P.g1.flops() + P.g2.flops();
    for (auto &st : P.wapply.strips) P.exec_flops += 2.0 * st.rows * st.cols * (st.s1 - st.s0);
    // reference: BoundaryMPSProduct GEMMs (all rows), lbtm axpys, SU2::gemm products
    double f = 0;
    {
        std::vector<Copy2D> dummy;
        GemmList gl;
        TLayout tall = t_layout(s, L, prp, dummy, gl, tm, tn, kc, false);
        f += gl.flops();
    }
    f += axf;
    for (auto &sg : cl.segs) {
        Q al = std::get<0>(sg), ar = std::get<2>(sg);
        f += 2.0 * s.right.size_of(ar) * s.pbl.sz(al) * s.right.size_of(al);
    }
    P.ref_flops = f;
    P.min_bytes = 8.0 * (L.total + off + s.X.total);
    return P;
}

// ============================================================ right step
HostPlan build_rstep_plan(const mq_site_desc &d, int tm, int tn, int kc) {
    Site s(d);
    BLayout R = to_layout(d.R);
    if (R.aux != s.W.cols) throw std::runtime_error("right boundary aux_dim does not match the MPO");
    HostPlan P;
    P.kind = PLAN_RSTEP;
    P.symm = s.symm;
    P.vec_len = s.X.total;
    PRPLayout prp = prp_layout(s, P.reshape);
    P.n_prp = prp.total;
    // MPSBoundaryProduct (boundary_times_mps.hpp:226-293): t[b2] = X * Rop[b2], right-trimmed
    // to the right charges of X.  RC_c stacks Rop[b2](c, c') horizontally.
    std::map<Q, int> Nc;
    std::map<std::tuple<int, Q, Q>, int> coff;  // (b2, c, c') -> column offset in t_c
    std::vector<std::tuple<int, Q, Q>> tblk;    // (b2, c, c')
    std::vector<std::tuple<Q, int, const BBlk *, bool, double, Q>> rops;
    for (int b2 = 0; b2 < s.W.cols; ++b2) {
        if (s.W.col_nnz[b2] == 0) continue;
        for (auto &xb : s.X.blk) {
            Q c = xb.c;
            if (s.W.rskip(b2)) {
                int cj = s.W.rherm[b2];
                for (auto &x : R.b[cj]) {
                    if (x.rc != c) continue;  // R[conj](c', c) -> adjoint block (c, c')
                    Q cp = x.lc;
                    if (!s.X.pos.count(cp)) continue;
                    double sc = s.su2 ? conj_correction(c, cp, s.W.rspin[b2]) * s.W.rphase[cj] : 1.0;
                    int &N = Nc[c];
                    coff[{b2, c, cp}] = N;
                    tblk.push_back({b2, c, cp});
                    rops.push_back({c, N, &x, true, sc, cp});
                    N += s.right.size_of(cp);
                }
            } else {
                for (auto &x : R.b[b2]) {
                    if (x.lc != c) continue;
                    Q cp = x.rc;
                    if (!s.X.pos.count(cp)) continue;
                    int &N = Nc[c];
                    coff[{b2, c, cp}] = N;
                    tblk.push_back({b2, c, cp});
                    rops.push_back({c, N, &x, false, 1.0, cp});
                    N += s.right.size_of(cp);
                }
            }
        }
    }
    std::map<Q, int64_t> rc_off, t_off;
    int64_t rc_total = 0, t_total = 0;
    for (auto &xb : s.X.blk) {
        if (!Nc.count(xb.c)) continue;
        rc_off[xb.c] = rc_total;
        rc_total += (int64_t)xb.cols * Nc[xb.c];
        t_off[xb.c] = t_total;
        t_total += (int64_t)xb.rows * Nc[xb.c];
    }
    P.n_rb = rc_total;
    P.n_t = t_total;
    for (auto &r : rops) {
        Q c = std::get<0>(r);
        Q cp = std::get<5>(r);
        const BBlk *blk = std::get<2>(r);
        bool tr = std::get<3>(r);
        int rc_ = s.right.size_of(c), rcp = s.right.size_of(cp);
        if (blk->rows * blk->cols != rc_ * rcp) throw std::runtime_error("right boundary block size mismatch");
        Copy2D cp2;
        cp2.dst = rc_off[c] + (int64_t)std::get<1>(r) * rc_;
        cp2.src = blk->off;
        cp2.rows = rc_;
        cp2.cols = rcp;
        cp2.ldd = rc_;
        cp2.lds = tr ? rcp : rc_;
        cp2.trans = tr ? 1 : 0;
        cp2.scale = std::get<4>(r);
        P.arrR.push_back(cp2);
    }
    for (auto &xb : s.X.blk) {
        if (!Nc.count(xb.c)) continue;
        GemmGroup g{};
        g.a = xb.off;
        g.b = rc_off[xb.c];
        g.c = t_off[xb.c];
        g.M = xb.rows;
        g.N = Nc[xb.c];
        g.K = xb.cols;
        g.lda = xb.rows;
        g.ldb = xb.cols;
        g.ldc = xb.rows;
        add_gemm(P.g1, g, tm, tn, kc);
    }
    // rbtm tasks (apply_op.hpp:114-209): buf[b1](out_l, out_r) += coef * t rows
    struct RC {
        int b1;
        Q ol, orr;
        int col0;  // column offset inside the buf block
        int cols;
        int rows;
        Src src;
    };
    std::vector<RC> cb;
    std::map<int, std::vector<std::tuple<int, Q, Q>>> tb_by_b2;
    for (auto &x : tblk) tb_by_b2[std::get<0>(x)].push_back(x);
    double axf = 0;
    for (int b1 = 0; b1 < s.W.rows; ++b1) {
        if (s.W.lskip(b1)) continue;
        int a = s.W.lspin[b1];
        for (int b2 : s.W.row_b2[b1]) {
            int ap = s.W.rspin[b2];
            auto itb = tb_by_b2.find(b2);
            if (itb == tb_by_b2.end()) continue;
            for (auto &term : s.W.at.at({b1, b2})) {
                const Op &op = s.W.ops[term.op];
                int k = op.spin;
                for (auto &tbk : itb->second) {
                    Q lc = std::get<1>(tbk), rc = std::get<2>(tbk);
                    int rowsc = s.X.blk[s.X.pos.at(lc)].rows;
                    int64_t tbase = t_off.at(lc) + (int64_t)coff.at(tbk) * rowsc;
                    for (auto &wb : op.blocks) {
                        Q ol = lc - wb.in;
                        if (!s.left.has(ol)) continue;
                        Q orr = rc - wb.out;
                        if (s.su2 && !tri(spin2(ol), a, spin2(orr))) continue;
                        if (!s.left.has(orr) || s.pbr.sz(orr) == 0) continue;  // out_right_i
                        int lsz = s.left.size_of(ol);
                        double cpl[4] = {term.scale, term.scale, term.scale, term.scale};
                        if (s.su2) {
                            int i = spin2(orr), ip = spin2(rc), j = spin2(ol), jp = spin2(lc);
                            set_coupling(j, std::abs(j - jp), jp, a, k, ap, i, std::abs(i - ip), ip, term.scale, cpl);
                        }
                        int ilo = s.pbl.at(wb.in, ol);
                        int oro = s.pbr.at(wb.out, rc);
                        int rsz = s.right.size_of(rc);
                        for (auto &e : wb.e) {
                            RC r;
                            r.b1 = b1;
                            r.ol = ol;
                            r.orr = orr;
                            r.col0 = oro + e.col * rsz;
                            r.cols = rsz;
                            r.rows = lsz;
                            r.src.off = tbase + ilo + e.row * lsz;
                            r.src.ld = rowsc;
                            r.src.coef = e.coeff * (s.su2 ? cpl[case_of(e.rspin, e.cspin)] : term.scale);
                            cb.push_back(r);
                            axf += 2.0 * lsz * rsz;
                        }
                    }
                }
            }
        }
    }
    // BUF_r stacks buf[b1](out_l, out_r) vertically over (b1, out_l); spin filter of
    // SU2::gemm(ret[b1], transpose(bra_conj), tmp, spin(b1)) (move_boundary.hpp:224)
    std::map<Q, int> Mr;
    std::map<std::tuple<Q, int, Q>, int> rowoff;  // (out_r, b1, out_l)
    std::vector<std::tuple<Q, int, Q>> rows_order;
    {
        std::set<std::tuple<Q, int, Q>> seen;
        for (auto &c : cb) {
            auto key = std::make_tuple(c.orr, c.b1, c.ol);
            if (seen.count(key)) continue;
            seen.insert(key);
            if (s.su2 && !tri(spin2(c.ol), s.W.lspin[c.b1], spin2(c.orr))) continue;
            if (!prp.off.count(c.orr)) continue;
            rows_order.push_back(key);
        }
        std::sort(rows_order.begin(), rows_order.end(), [](auto &x, auto &y) {
            if (std::get<0>(x) != std::get<0>(y)) return std::get<0>(x) > std::get<0>(y);
            if (std::get<1>(x) != std::get<1>(y)) return std::get<1>(x) < std::get<1>(y);
            return std::get<2>(x) > std::get<2>(y);
        });
        for (auto &key : rows_order) {
            int &M = Mr[std::get<0>(key)];
            rowoff[key] = M;
            M += s.left.size_of(std::get<2>(key));
        }
    }
    std::map<Q, int64_t> buf_off, tmp_off;
    int64_t buf_total = 0, tmp_total = 0;
    for (auto &kv : Mr) {
        Q orr = kv.first;
        buf_off[orr] = buf_total;
        buf_total += (int64_t)kv.second * s.pbr.sz(orr);
        tmp_off[orr] = tmp_total;
        tmp_total += (int64_t)kv.second * s.left.size_of(orr);
    }
    P.n_c = buf_total;
    P.n_tmp = tmp_total;
    {
        std::map<std::tuple<Q, int, Q, int>, std::vector<Src>> by;
        for (auto &c : cb) {
            auto key = std::make_tuple(c.orr, c.b1, c.ol);
            if (!rowoff.count(key)) continue;
            by[std::make_tuple(c.orr, c.b1, c.ol, c.col0)].push_back(c.src);
        }
        for (auto &key : rows_order) {
            Q orr = std::get<0>(key), ol = std::get<2>(key);
            int b1 = std::get<1>(key);
            int M = Mr[orr];
            int lsz = s.left.size_of(ol);
            for (int si = 0; si < s.phys.n(); ++si) {
                Q so = s.phys.q[si];
                Q r = orr + so;
                if (!s.right.has(r)) continue;
                int rsz = s.right.size_of(r);
                for (int ss = 0; ss < s.phys.s[si]; ++ss) {
                    int col0 = s.pbr.at(so, r) + ss * rsz;
                    Strip st;
                    st.dst = buf_off[orr] + rowoff[key] + (int64_t)col0 * M;
                    st.rows = lsz;
                    st.cols = rsz;
                    st.ldd = M;
                    st.s0 = (int)P.wapply.srcs.size();
                    auto it = by.find(std::make_tuple(orr, b1, ol, col0));
                    if (it != by.end())
                        for (auto &x : it->second) P.wapply.srcs.push_back(x);
                    st.s1 = (int)P.wapply.srcs.size();
                    P.wapply.strips.push_back(st);
                }
            }
        }
    }
    for (auto &kv : Mr) {
        Q orr = kv.first;
        GemmGroup g{};
        g.a = buf_off[orr];
        g.b = prp.off.at(orr);  // PRP block bl x Nr -> op(B) = PRP^T
        g.c = tmp_off[orr];
        g.M = kv.second;
        g.N = s.left.size_of(orr);
        g.K = s.pbr.sz(orr);
        g.lda = kv.second;
        g.ldb = s.left.size_of(orr);
        g.ldc = kv.second;
        g.transB = 1;
        add_gemm(P.g2, g, tm, tn, kc);
    }
    P.out_aux = s.W.rows;
    std::map<int, std::vector<std::pair<Q, Q>>> per_b1;
    for (auto &key : rows_order) per_b1[std::get<1>(key)].push_back({std::get<2>(key), std::get<0>(key)});
    P.out_bptr.push_back(0);
    int64_t off = 0;
    for (int b1 = 0; b1 < s.W.rows; ++b1) {
        auto it = per_b1.find(b1);
        if (it != per_b1.end()) {
            auto v = it->second;
            std::sort(v.begin(), v.end(), [](auto &x, auto &y) {
                return x.first > y.first || (x.first == y.first && x.second > y.second);
            });
            for (auto &bl : v) {
                Q ol = bl.first, orr = bl.second;
                int r1 = s.left.size_of(ol), c1 = s.left.size_of(orr);
                P.out_lc.push_back(ol.a);
                P.out_lc.push_back(ol.b);
                P.out_rc.push_back(orr.a);
                P.out_rc.push_back(orr.b);
                P.out_rows.push_back(r1);
                P.out_cols.push_back(c1);
                P.out_off.push_back(off);
                Copy2D c;
                c.dst = off;
                c.src = tmp_off[orr] + rowoff.at(std::make_tuple(orr, b1, ol));
                c.rows = r1;
                c.cols = c1;
                c.ldd = r1;
                c.lds = Mr[orr];
                c.trans = 0;
                c.scale = 1.0;
                P.scatter.push_back(c);
                off += (int64_t)r1 * c1;
            }
        }
        P.out_bptr.push_back((int32_t)P.out_rows.size());
    }
    P.n_out = off;
    P.exec_flops = P.g1.flops() + P.g2.flops();
    for (auto &st : P.wapply.strips) P.exec_flops += 2.0 * st.rows * st.cols * (st.s1 - st.s0);
    double f = P.g1.flops() + axf;
    for (auto &key : rows_order) {
        Q orr = std::get<0>(key), ol = std::get<2>(key);
        f += 2.0 * s.left.size_of(ol) * s.pbr.sz(orr) * s.left.size_of(orr);
    }
    P.ref_flops = f;
    P.min_bytes = 8.0 * (R.total + off + s.X.total);
    return P;
}

// ============================================================ reference FLOPs of sigma
double reference_sigma_flops(const mq_site_desc &d) {
    Site s(d);
    BLayout L = to_layout(d.L), R = to_layout(d.R);
    std::vector<Copy2D> dummy;
    PRPLayout prp = prp_layout(s, dummy);
    bool lbtm = s.symm != MQ_SYMM_SU2U1 || (s.W.rows - s.W.one_rows) < (s.W.cols - s.W.one_cols);
    double f = 0;
    if (lbtm) {
        GemmList gl;
        std::vector<Copy2D> arr;
        TLayout t = t_layout(s, L, prp, arr, gl, 32, 32, 1 << 30, false);
        f += gl.flops();
        std::vector<int> b2s;
        for (int b2 = 0; b2 < s.W.cols; ++b2) b2s.push_back(b2);
        std::vector<Contrib> cb;
        lbtm_contribs(s, t, prp, b2s, cb, &f);
        std::set<std::tuple<Q, int, Q>> segset;
        for (auto &c : cb) segset.insert(std::make_tuple(c.al, c.b2, c.ar));
        for (auto &sg : segset) {
            const BBlk *blk;
            bool tr;
            double sc;
            Q al = std::get<0>(sg), ar = std::get<2>(sg);
            if (!r_operand(s, R, std::get<1>(sg), ar, al, &blk, &tr, &sc)) continue;
            f += 2.0 * s.pbl.sz(al) * s.right.size_of(ar) * s.right.size_of(al);
        }
        return f;
    }
    // rbtm (site_hamil.hpp:149-204)
    // MPSBoundaryProduct over all columns (lazy single-use columns counted when used)
    std::map<std::tuple<int, Q, Q>, bool> tb;  // (b2, c, c') exists
    auto tgemm = [&](int b2, Q c, Q cp) {
        return 2.0 * s.X.blk[s.X.pos.at(c)].rows * s.right.size_of(cp) * s.right.size_of(c);
    };
    std::map<int, std::vector<std::pair<Q, Q>>> tblocks;
    for (int b2 = 0; b2 < s.W.cols; ++b2) {
        for (auto &xb : s.X.blk) {
            Q c = xb.c;
            if (s.W.rskip(b2)) {
                for (auto &x : R.b[s.W.rherm[b2]])
                    if (x.rc == c && s.X.pos.count(x.lc)) tblocks[b2].push_back({c, x.lc});
            } else {
                for (auto &x : R.b[b2])
                    if (x.lc == c && s.X.pos.count(x.rc)) tblocks[b2].push_back({c, x.rc});
            }
        }
        if (s.W.col_nnz[b2] != 1)
            for (auto &p : tblocks[b2]) f += tgemm(b2, p.first, p.second);
    }
    std::set<int> lazy_used;
    for (int b1 = 0; b1 < s.W.rows; ++b1) {
        int a = s.W.lspin[b1];
        // left_b1 blocks (bra, ket): transpose(L[b1]) or L[conj]
        std::set<std::pair<Q, Q>> lblk;
        if (s.W.lskip(b1)) {
            for (auto &x : L.b[s.W.lherm[b1]]) lblk.insert({x.lc, x.rc});
        } else {
            for (auto &x : L.b[b1]) lblk.insert({x.rc, x.lc});
        }
        std::map<std::pair<Q, Q>, std::vector<std::pair<double, int>>> groups;  // (ol, or) -> (flops, b2)
        for (int b2 : s.W.row_b2[b1]) {
            int ap = s.W.rspin[b2];
            for (auto &term : s.W.at.at({b1, b2})) {
                const Op &op = s.W.ops[term.op];
                for (auto &p : tblocks[b2]) {
                    Q lc = p.first, rc = p.second;
                    for (auto &wb : op.blocks) {
                        Q ol = lc - wb.in;
                        if (!s.left.has(ol)) continue;
                        Q orr = rc - wb.out;
                        if (s.su2 && !tri(spin2(ol), a, spin2(orr))) continue;
                        if (!s.left.has(orr) || s.pbr.sz(orr) == 0) continue;
                        (void)ap;
                        double fl = 2.0 * s.left.size_of(ol) * s.right.size_of(rc) * wb.e.size();
                        groups[{ol, orr}].push_back({fl, b2});
                    }
                }
            }
        }
        for (auto &g : groups) {
            Q ol = g.first.first, orr = g.first.second;
            if (!lblk.count({orr, ol})) continue;
            for (auto &x : g.second) {
                f += x.first;
                if (s.W.col_nnz[x.second] == 1) lazy_used.insert(x.second);
            }
            f += 2.0 * s.left.size_of(orr) * s.pbr.sz(orr) * s.left.size_of(ol);
        }
    }
    for (int b2 : lazy_used)
        for (auto &p : tblocks[b2]) f += tgemm(b2, p.first, p.second);
    return f;
}

}  // namespace mq
