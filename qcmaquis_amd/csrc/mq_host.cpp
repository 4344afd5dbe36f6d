// Descriptor parsing and SU(2) coefficients for the MI355X hot path.
#include "mq_host.hpp"

namespace mq {

Mpo to_mpo(const mq_mpo_desc &d) {
    Mpo m;
    m.rows = d.row_dim;
    m.cols = d.col_dim;
    m.col_b1.resize(d.col_dim);
    m.row_b2.resize(d.row_dim);
    m.row_ent.resize(d.row_dim);
    m.term_ptr.push_back(0);
    for (int b2 = 0; b2 < d.col_dim; ++b2)
        for (int e = d.col_ptr[b2]; e < d.col_ptr[b2 + 1]; ++e) {
            int b1 = d.row_idx[e];
            if (b1 < 0 || b1 >= d.row_dim) throw std::runtime_error("mpo: row index out of range");
            // columns are walked in order, so row_b2[b1] stays ascending; a repeat
            // of (b1, b2) would have been added to it just before
            if (!m.row_b2[b1].empty() && m.row_b2[b1].back() == b2)
                throw std::runtime_error("mpo: repeated (row, column) entry");
            m.col_b1[b2].push_back(b1);
            m.row_b2[b1].push_back(b2);
            m.row_ent[b1].push_back((int)m.term_ptr.size() - 1);
            for (int q = d.term_ptr[e]; q < d.term_ptr[e + 1]; ++q) {
                if (d.term_op[q] < 0 || d.term_op[q] >= d.n_ops) throw std::runtime_error("mpo: operator id out of range");
                m.term.push_back({d.term_op[q], d.term_scale[q]});
            }
            m.term_ptr.push_back((int)m.term.size());
        }
    if (d.n_ops < 0 || d.op_bptr[0] != 0) throw std::runtime_error("mpo: bad operator block pointers");
    for (int o = 0; o < d.n_ops; ++o)
        if (d.op_bptr[o + 1] < d.op_bptr[o]) throw std::runtime_error("mpo: bad operator block pointers");
    const int nblk = d.op_bptr[d.n_ops];
    if (d.opb_eptr[0] != 0) throw std::runtime_error("mpo: bad operator entry pointers");
    for (int bb = 0; bb < nblk; ++bb)
        if (d.opb_eptr[bb + 1] < d.opb_eptr[bb]) throw std::runtime_error("mpo: bad operator entry pointers");
    const int nent = d.opb_eptr[nblk];
    m.ent.resize(nent);
    for (int e = 0; e < nent; ++e)
        m.ent[e] = {d.ent_row[e], d.ent_col[e], d.ent_rspin[e], d.ent_cspin[e], d.ent_coeff[e]};
    m.blk.resize(nblk);
    for (int bb = 0; bb < nblk; ++bb) {
        OpBlk &ob = m.blk[bb];
        ob.in = Q(d.opb_in[2 * bb], d.opb_in[2 * bb + 1]);
        ob.out = Q(d.opb_out[2 * bb], d.opb_out[2 * bb + 1]);
        ob.e = {m.ent.data() + d.opb_eptr[bb], m.ent.data() + d.opb_eptr[bb + 1]};
    }
    m.ops.resize(d.n_ops);
    for (int o = 0; o < d.n_ops; ++o) {
        m.ops[o].spin = d.op_spin[o];
        m.ops[o].blocks = {m.blk.data() + d.op_bptr[o], m.blk.data() + d.op_bptr[o + 1]};
    }
    m.lspin.assign(d.left_spin, d.left_spin + d.row_dim);
    m.rspin.assign(d.right_spin, d.right_spin + d.col_dim);
    m.lherm.assign(d.left_herm, d.left_herm + d.row_dim);
    m.rherm.assign(d.right_herm, d.right_herm + d.col_dim);
    m.lphase.assign(d.left_phase, d.left_phase + d.row_dim);
    m.rphase.assign(d.right_phase, d.right_phase + d.col_dim);
    m.row_nnz.assign(d.row_dim, 0);
    m.col_nnz.assign(d.col_dim, 0);
    for (int b = 0; b < m.rows; ++b) m.row_nnz[b] = (int)m.row_b2[b].size();
    for (int b = 0; b < m.cols; ++b) m.col_nnz[b] = (int)m.col_b1[b].size();
    m.one_rows = (int)std::count(m.row_nnz.begin(), m.row_nnz.end(), 1);
    m.one_cols = (int)std::count(m.col_nnz.begin(), m.col_nnz.end(), 1);
    for (int b = 0; b < m.rows; ++b)
        if (m.lherm[b] < 0 || m.lherm[b] >= m.rows) throw std::runtime_error("mpo: left_herm out of range");
    for (int b = 0; b < m.cols; ++b)
        if (m.rherm[b] < 0 || m.rherm[b] >= m.cols) throw std::runtime_error("mpo: right_herm out of range");
    return m;
}

BLayout to_layout(const mq_boundary_desc &d) {
    BLayout L;
    L.aux = d.aux_dim;
    L.total = d.total;
    L.b.resize(d.aux_dim);
    for (int b = 0; b < d.aux_dim; ++b)
        for (int k = d.bptr[b]; k < d.bptr[b + 1]; ++k) {
            BBlk x;
            x.lc = Q(d.blocks.lc[2 * k], d.blocks.lc[2 * k + 1]);
            x.rc = Q(d.blocks.rc[2 * k], d.blocks.rc[2 * k + 1]);
            x.rows = d.blocks.rows[k];
            x.cols = d.blocks.cols[k];
            x.off = d.blocks.offset[k];
            if (x.off < 0 || x.off + (int64_t)x.rows * x.cols > d.total)
                throw std::runtime_error("boundary block outside its arena");
            L.b[b].push_back(x);
        }
    L.desc.assign(d.aux_dim, 1);
    for (int b = 0; b < d.aux_dim; ++b)
        for (size_t k = 1; k < L.b[b].size(); ++k)
            if (!(QPair{L.b[b][k].lc, L.b[b][k].rc} < QPair{L.b[b][k - 1].lc, L.b[b][k - 1].rc})) L.desc[b] = 0;
    return L;
}

VLayout vec_layout(const Idx &phys, const Idx &left, const Idx &right) {
    PB pb(phys, left, false);
    VLayout v;
    for (int i = 0; i < right.n(); ++i) {
        int rows = pb.sz(right.q[i]);
        if (rows == 0) continue;
        VBlk b{right.q[i], rows, right.s[i], v.total};
        v.pos[b.c] = (int)v.blk.size();
        v.blk.push_back(b);
        v.total += (int64_t)rows * right.s[i];
    }
    return v;
}

// Racah single-sum 6j; 9j via Edmonds eq. 6.4.3 (sum over three 6j symbols).
static long double fact(int n) {
    static const std::vector<long double> t = [] {
        std::vector<long double> v(171);
        v[0] = 1.0L;
        for (int i = 1; i < 171; ++i) v[i] = v[i - 1] * i;
        return v;
    }();
    if (n < 0 || n > 170) throw std::runtime_error("factorial argument out of range");
    return t[n];
}
static long double tri_delta(int a, int b, int c) {
    return std::sqrt(fact((a + b - c) / 2) * fact((a - b + c) / 2) * fact((-a + b + c) / 2) / fact((a + b + c) / 2 + 1));
}
double wigner6j(int a, int b, int c, int d, int e, int f) {
    if (!(tri(a, b, c) && tri(a, e, f) && tri(d, b, f) && tri(d, e, c))) return 0.0;
    int s1 = (a + b + c) / 2, s2 = (a + e + f) / 2, s3 = (d + b + f) / 2, s4 = (d + e + c) / 2;
    int p1 = (a + b + d + e) / 2, p2 = (b + c + e + f) / 2, p3 = (c + a + f + d) / 2;
    int lo = std::max(std::max(s1, s2), std::max(s3, s4)), hi = std::min(p1, std::min(p2, p3));
    long double sum = 0;
    for (int t = lo; t <= hi; ++t) {
        long double v = fact(t + 1) / (fact(t - s1) * fact(t - s2) * fact(t - s3) * fact(t - s4) * fact(p1 - t) *
                                      fact(p2 - t) * fact(p3 - t));
        sum += (t & 1) ? -v : v;
    }
    return (double)(sum * tri_delta(a, b, c) * tri_delta(a, e, f) * tri_delta(d, b, f) * tri_delta(d, e, c));
}
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    if (!(tri(a, b, c) && tri(d, e, f) && tri(g, h, i) && tri(a, d, g) && tri(b, e, h) && tri(c, f, i))) return 0.0;
    int lo = std::max(std::abs(a - i), std::max(std::abs(b - f), std::abs(d - h)));
    int hi = std::min(a + i, std::min(b + f, d + h));
    long double t = 0;
    for (int x = lo; x <= hi; ++x) {
        if ((x + a + i) & 1) continue;
        long double w = (long double)(x + 1) * ((x & 1) ? -1.0L : 1.0L);
        t += w * wigner6j(a, b, c, f, i, x) * wigner6j(d, e, f, b, x, h) * wigner6j(g, h, i, x, a, d);
    }
    return (double)t;
}
static double modc(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    return std::sqrt((g + 1.) * (h + 1.) * (c + 1.) * (f + 1.)) * wigner9j(a, b, c, d, e, f, g, h, i);
}
void set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double out[4]) {
    double pre = std::sqrt((i + 1.) * (a + 1.) / ((g + 1.) * (c + 1.))) * init;
    out[0] = out[1] = out[2] = out[3] = 0.0;
    if (tri(a, b, c)) {
        out[0] = pre * modc(a, b, c, d, e, f, g, h, i);
        out[2] = pre * modc(a, b, c, d, e, f, g, 2, i);
    }
    if (tri(a, 2, c)) {
        out[1] = pre * modc(a, 2, c, d, e, f, g, h, i);
        out[3] = pre * modc(a, 2, c, d, e, f, g, 2, i);
    }
}
double conj_correction(Q lc, Q rc, int ts) {
    int S = std::min(spin2(rc), spin2(lc));
    int diff = spin2(rc) - spin2(lc);
    if (ts == 0) return 1.0;
    if (ts == 1) {
        if (diff > 0) return -std::sqrt((S + 1.) / (S + 2.));
        if (diff < 0) return std::sqrt((S + 2.) / (S + 1.));
        return 0.0;
    }
    if (ts == 2) {
        if (diff > 0) return -std::sqrt((S + 1.) / (S + 3.));
        if (diff < 0) return -std::sqrt((S + 3.) / (S + 1.));
        return 1.0;
    }
    throw std::runtime_error("hermitian conjugate only implemented up to rank 1");
}

}  // namespace mq
