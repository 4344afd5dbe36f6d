// Host side of the two-site split: the map from the two-site tensor to the
// blocks of its both-paired matrix, whose SVD gives the two new site tensors.
#include <cmath>
#include <cstdlib>
#include <map>
#include <tuple>

#include "mq_plan.hpp"

namespace mq {

namespace {

// Recoupling coefficient of the SU(2) two-site reduction (ts_reduction.h:
// the 6j symbol {jl jr j; S2 S1 jm} with its phase and dimension factors).
double recouple(int jl, int jr, int j, int S1, int S2, int jm) {
    double c = std::sqrt((double)(j + 1) * (jm + 1)) * wigner6j(jl, jr, j, S2, S1, jm);
    return (((jl + jr + S1 + S2) / 2) % 2) ? -c : c;
}

struct Term {
    int ipo;
    double c;
};

}  // namespace

SplitMap build_split_map(const mq_split_desc &d) {
    const bool su2 = d.symm == MQ_SYMM_SU2U1;
    if (!su2 && d.symm != MQ_SYMM_TWOU1) throw std::runtime_error("split: unknown symmetry");
    const Idx p1 = to_idx(d.phys1), p2 = to_idx(d.phys2), left = to_idx(d.left), right = to_idx(d.right);
    const Idx phys2 = idx_product(p1, p2);
    const PB phys_pb(p1, p2, false), in_left(phys2, left, false), out_left(p1, left, false),
        out_right(p2, right, true);
    const VLayout vl = vec_layout(phys2, left, right);

    // slabs per both-block charge K (the block is created when a slab reaches it)
    struct Pending {
        int olo, oro, ldim, rdim;
        int64_t src0, src1;
        double c0, c1;
        int src_cs;
    };
    std::map<Q, std::vector<Pending>, std::greater<Q>> per_block;
    for (const VBlk &vb : vl.blk) {
        const Q c = vb.c;
        const int rdim = vb.cols;
        for (int is = 0; is < phys2.n(); ++is) {
            const Q s = phys2.q[is];
            const Q lch = c - s;
            if (!left.has(lch)) continue;
            const int ldim = left.size_of(lch);
            const int ilo = in_left.at(s, lch);
            // unreduced slab position -> its terms (input slab position, coefficient)
            std::map<int, std::vector<Term>> outs;
            for (int a = 0; a < p1.n(); ++a)
                for (int b = 0; b < p2.n(); ++b) {
                    const Q s1 = p1.q[a], s2 = p2.q[b];
                    if (s1 + s2 != s) continue;
                    const int sd1 = p1.s[a], sd2 = p2.s[b];
                    const int ipo = phys_pb.at(s1, s2);
                    if (!su2) {
                        for (int x = 0; x < sd1 * sd2; ++x) outs[ipo + x].push_back({ipo + x, 1.0});
                        continue;
                    }
                    if (sd1 != 1 || sd2 != 1) throw std::runtime_error("split: SU2 physical blocks must have size 1");
                    const int S1 = std::abs(s1.b), S2 = std::abs(s2.b), jl = lch.b, jr = c.b;
                    if (jl == jr && jl > 0 && S1 == 1 && S2 == 1) {
                        const int j = s1.b == 1 ? 0 : 2;
                        const int base = s1.b == 1 ? ipo : ipo - 1;
                        for (int jm : {jl - 1, jl + 1}) {
                            const int opo = jm == jl - 1 ? base + 1 : base;
                            outs[opo].push_back({ipo, recouple(jl, jr, j, S1, S2, jm)});
                        }
                    } else {
                        const int j = std::abs(s1.b + s2.b), jm = jl + s1.b;
                        if (jm < 0) continue;
                        outs[ipo].push_back({ipo, recouple(jl, jr, j, S1, S2, jm)});
                    }
                }
            for (int a = 0; a < p1.n(); ++a)
                for (int b = 0; b < p2.n(); ++b) {
                    const Q s1 = p1.q[a], s2 = p2.q[b];
                    if (s1 + s2 != s) continue;
                    const int sd1 = p1.s[a], sd2 = p2.s[b];
                    const Q K = s1 + lch;
                    const int olo = out_left.at(s1, lch), oro = out_right.at(s2, c);
                    auto &vec = per_block[K];
                    for (int x1 = 0; x1 < sd1; ++x1)
                        for (int x2 = 0; x2 < sd2; ++x2) {
                            const int so = phys_pb.at(s1, s2) + x1 * sd2 + x2;
                            Pending pd{olo + x1 * ldim, oro + x2 * rdim, ldim, rdim, -1, -1, 0.0, 0.0, vb.rows};
                            auto it = outs.find(so);
                            if (it != outs.end()) {
                                if (it->second.size() > 2) throw std::runtime_error("split: more than two terms");
                                pd.src0 = vb.off + ilo + (int64_t)it->second[0].ipo * ldim;
                                pd.c0 = it->second[0].c;
                                if (it->second.size() == 2) {
                                    pd.src1 = vb.off + ilo + (int64_t)it->second[1].ipo * ldim;
                                    pd.c1 = it->second[1].c;
                                }
                            }
                            vec.push_back(pd);
                        }
                }
        }
    }

    SplitMap m;
    m.psi_len = vl.total;
    for (auto &kv : per_block) {
        const Q K = kv.first;
        SvdBlk sb{};
        sb.rows = out_left.sz(K);
        sb.cols = out_right.sz(K);
        sb.trans = sb.cols > sb.rows;
        sb.R = std::max(sb.rows, sb.cols);
        sb.C = std::min(sb.rows, sb.cols);
        sb.x_off = m.x_total;
        sb.v_off = m.v_total;
        sb.s_off = m.s_total;
        m.x_total += (int64_t)sb.R * sb.C;
        m.v_total += (int64_t)sb.C * sb.C;
        m.s_total += sb.C;
        m.max_c = std::max(m.max_c, sb.C);
        // X element (i, j) of the block: trans = 0 at j*rows + i, trans = 1 at i*cols + j
        const int drs = sb.trans ? sb.cols : 1, dcs = sb.trans ? 1 : sb.rows;
        for (const Pending &pd : kv.second) {
            if (pd.src0 < 0) continue;  // no term reaches this slab: it stays zero
            if (pd.olo + pd.ldim > sb.rows || pd.oro + pd.rdim > sb.cols)
                throw std::runtime_error("split: slab outside its block");
            SplitSlab sl{};
            sl.dst = sb.x_off + (int64_t)pd.olo * drs + (int64_t)pd.oro * dcs;
            sl.src0 = pd.src0;
            sl.src1 = pd.src1;
            sl.c0 = pd.c0;
            sl.c1 = pd.c1;
            sl.rows = pd.ldim;
            sl.cols = pd.rdim;
            sl.dst_rs = drs;
            sl.dst_cs = dcs;
            sl.src_cs = pd.src_cs;
            if (sl.rows > 0 && sl.cols > 0) m.slabs.push_back(sl);
        }
        m.blk.push_back(sb);
        m.charge.push_back(K.a);
        m.charge.push_back(K.b);
        m.rows.push_back(sb.rows);
        m.cols.push_back(sb.cols);
    }
    m.order.resize(m.blk.size());
    for (size_t i = 0; i < m.order.size(); ++i) m.order[i] = (int32_t)i;
    std::stable_sort(m.order.begin(), m.order.end(), [&](int a, int b) {
        const SvdBlk &x = m.blk[a], &y = m.blk[b];
        return (double)x.R * x.C * x.C > (double)y.R * y.C * y.C;
    });
    return m;
}

void apply_split_map_host(const SplitMap &m, const double *psi, double *both) {
    std::vector<double> X((size_t)m.x_total, 0.0);
    for (const SplitSlab &sl : m.slabs)
        for (int j = 0; j < sl.cols; ++j)
            for (int i = 0; i < sl.rows; ++i) {
                double v = sl.c0 * psi[sl.src0 + i + (int64_t)j * sl.src_cs];
                if (sl.src1 >= 0) v += sl.c1 * psi[sl.src1 + i + (int64_t)j * sl.src_cs];
                X[sl.dst + (int64_t)i * sl.dst_rs + (int64_t)j * sl.dst_cs] = v;
            }
    int64_t o = 0;
    for (const SvdBlk &b : m.blk) {
        for (int i = 0; i < b.rows; ++i)
            for (int j = 0; j < b.cols; ++j)
                both[o + (int64_t)i * b.cols + j] =
                    X[b.x_off + (b.trans ? (int64_t)i * b.cols + j : (int64_t)j * b.rows + i)];
        o += (int64_t)b.rows * b.cols;
    }
}

}  // namespace mq
