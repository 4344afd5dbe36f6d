// Host-side data model of the MI355X hot path: charges, indices, product
// bases, MPO tensors and boundary layouts parsed from the C-ABI descriptors
// (include/mqgpu.h), plus the SU(2) coupling coefficients.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/mqgpu.h"

namespace mq {

struct Q {  // charge: SU2U1 (N, 2S), TwoU1 (Na, Nb)
    int a = 0, b = 0;
    Q() = default;
    Q(int x, int y) : a(x), b(y) {}
    bool operator<(const Q &o) const { return a < o.a || (a == o.a && b < o.b); }
    bool operator>(const Q &o) const { return o < *this; }
    bool operator==(const Q &o) const { return a == o.a && b == o.b; }
    bool operator!=(const Q &o) const { return !(*this == o); }
    Q operator+(const Q &o) const { return Q(a + o.a, b + o.b); }
    Q operator-(const Q &o) const { return Q(a - o.a, b - o.b); }
    Q operator-() const { return Q(-a, -b); }
};
inline int spin2(Q c) { return c.b; }
inline bool tri(int a, int b, int c) { return ((a + b + c) % 2 == 0) && std::abs(a - b) <= c && c <= a + b; }

struct QPair {
    Q x, y;
    bool operator<(const QPair &o) const { return x < o.x || (x == o.x && y < o.y); }
    bool operator==(const QPair &o) const { return x == o.x && y == o.y; }
};

// Index: (charge, size) descending in charge (indexing_stable.hpp:82-130)
struct Idx {
    std::vector<Q> q;
    std::vector<int> s;
    std::map<Q, int> pos;
    void build() {
        std::vector<std::pair<Q, int>> v;
        for (size_t i = 0; i < q.size(); ++i) v.push_back({q[i], s[i]});
        std::sort(v.begin(), v.end(), [](auto &x, auto &y) { return x.first > y.first; });
        q.clear();
        s.clear();
        pos.clear();
        for (auto &p : v) {
            pos[p.first] = (int)q.size();
            q.push_back(p.first);
            s.push_back(p.second);
        }
    }
    int n() const { return (int)q.size(); }
    bool has(Q c) const { return pos.count(c) != 0; }
    int size_of(Q c) const {
        auto it = pos.find(c);
        return it == pos.end() ? 0 : s[it->second];
    }
};

inline Idx to_idx(const mq_index &x) {
    Idx r;
    for (int i = 0; i < x.n; ++i) {
        r.q.push_back(Q(x.charges[2 * i], x.charges[2 * i + 1]));
        r.s.push_back(x.sizes[i]);
    }
    r.build();
    return r;
}

inline Idx idx_product(const Idx &a, const Idx &b) {
    std::map<Q, int> m;
    for (int i = 0; i < a.n(); ++i)
        for (int j = 0; j < b.n(); ++j) m[a.q[i] + b.q[j]] += a.s[i] * b.s[j];
    Idx r;
    for (auto &kv : m) {
        r.q.push_back(kv.first);
        r.s.push_back(kv.second);
    }
    r.build();
    return r;
}

// ProductBasis (indexing_stable.hpp:305-343): offset of (a, b) inside the fused charge.
struct PB {
    std::map<QPair, int> off;
    std::map<Q, int> size;
    PB() = default;
    PB(const Idx &A, const Idx &B, bool right) {
        for (int i = 0; i < A.n(); ++i)
            for (int j = 0; j < B.n(); ++j) {
                Q pc = right ? (B.q[j] - A.q[i]) : (A.q[i] + B.q[j]);
                int &sz = size[pc];
                off[{A.q[i], B.q[j]}] = sz;
                sz += A.s[i] * B.s[j];
            }
    }
    int at(Q a, Q b) const {
        auto it = off.find({a, b});
        if (it == off.end()) throw std::runtime_error("product basis: missing pair");
        return it->second;
    }
    int sz(Q c) const {
        auto it = size.find(c);
        return it == size.end() ? 0 : it->second;
    }
};

struct Ent {
    int row, col, rspin, cspin;
    double coeff;
};
// read-only view of a contiguous range
template <class T>
struct Span {
    const T *b = nullptr, *e_ = nullptr;
    const T *begin() const { return b; }
    const T *end() const { return e_; }
    size_t size() const { return (size_t)(e_ - b); }
    bool empty() const { return b == e_; }
    const T &operator[](size_t i) const { return b[i]; }
};
struct OpBlk {
    Q in, out;
    Span<Ent> e;  // into Mpo::ent
};
struct Op {
    int spin = 0;
    Span<OpBlk> blocks;  // into Mpo::blk
};
struct TermD {
    int op;
    double scale;
};

struct TermSpan {
    const TermD *b, *e;
    const TermD *begin() const { return b; }
    const TermD *end() const { return e; }
};

// Move-only: its operators hold views into its own ent and blk arrays (a move
// keeps the arrays' storage, a copy would not).
struct Mpo {
    Mpo() = default;
    Mpo(const Mpo &) = delete;
    Mpo &operator=(const Mpo &) = delete;
    Mpo(Mpo &&) = default;
    Mpo &operator=(Mpo &&) = default;
    int rows = 0, cols = 0;
    std::vector<std::vector<int>> col_b1, row_b2;  // row_b2[b1] ascending
    std::vector<std::vector<int>> row_ent;         // entry id of (b1, row_b2[b1][k])
    std::vector<int> term_ptr;                      // terms of entry e: term[term_ptr[e] .. term_ptr[e + 1])
    std::vector<TermD> term;
    // terms of the nonzero (b1, b2); throws std::out_of_range for a zero entry
    TermSpan terms(int b1, int b2) const {
        const std::vector<int> &r = row_b2[b1];
        auto it = std::lower_bound(r.begin(), r.end(), b2);
        if (it == r.end() || *it != b2) throw std::out_of_range("mpo: no entry at (b1, b2)");
        const int e = row_ent[b1][it - r.begin()];
        return {term.data() + term_ptr[e], term.data() + term_ptr[e + 1]};
    }
    std::vector<Op> ops;
    std::vector<OpBlk> blk;  // operator blocks, operator after operator
    std::vector<Ent> ent;    // sparse entries, block after block
    std::vector<int> lspin, rspin, lherm, rherm, lphase, rphase, row_nnz, col_nnz;
    int one_rows = 0, one_cols = 0;
    bool lskip(int b) const { return lherm[b] < b; }
    bool rskip(int b) const { return rherm[b] < b; }
};

Mpo to_mpo(const mq_mpo_desc &d);

// Boundary layout: per bond, map (lc, rc) -> (rows, cols, offset)
struct BBlk {
    Q lc, rc;
    int rows, cols;
    int64_t off;
};
struct BLayout {
    int aux = 0;
    std::vector<std::vector<BBlk>> b;  // per bond, descending (lc, rc)
    std::vector<char> desc;            // per bond: blocks strictly descending in (lc, rc)
    int64_t total = 0;
    const BBlk *find(int bond, Q lc, Q rc) const {
        const std::vector<BBlk> &v = b[bond];
        if (!desc.empty() && desc[bond]) {
            const QPair key{lc, rc};
            auto it = std::lower_bound(v.begin(), v.end(), key,
                                       [](const BBlk &x, const QPair &k) { return k < QPair{x.lc, x.rc}; });
            return it != v.end() && it->lc == lc && it->rc == rc ? &*it : nullptr;
        }
        for (auto &x : v)
            if (x.lc == lc && x.rc == rc) return &x;
        return nullptr;
    }
};
BLayout to_layout(const mq_boundary_desc &d);

// canonical left-paired MPS layout: per right charge c present in phys*left
struct VBlk {
    Q c;
    int rows, cols;
    int64_t off;
};
struct VLayout {
    std::vector<VBlk> blk;  // descending c
    std::map<Q, int> pos;
    int64_t total = 0;
};
VLayout vec_layout(const Idx &phys, const Idx &left, const Idx &right);

// SU(2) coupling coefficients (gsl_coupling.h:166-205) and hermitian
// conjugate correction (non-abelian/gemm.hpp:18-46), computed on the host.
double wigner6j(int a, int b, int c, int d, int e, int f);
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i);
void set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double out[4]);
double conj_correction(Q lc, Q rc, int tensor_spin);
inline int case_of(int rs, int cs) {
    if (rs == 2 && cs == 2) return 3;
    if (rs == 2) return 1;
    if (cs == 2) return 2;
    return 0;
}

}  // namespace mq
