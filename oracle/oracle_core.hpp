// FP64 CPU oracle of the DMRG sweep hot path -- TEST INFRASTRUCTURE ONLY.
//
// This is a from-scratch restatement of the reference algorithms, used by
// tests/ and by bench.py's CPU-baseline leg as the checker; it is never part
// of the shipped GPU path.  Each function cites the reference file:line it
// follows (paths relative to the reference's dmrg/framework/dmrg/).
//
// Parity pinning (see the tests that exercise each function):
//  - wigner 9j: the 24 golden values of tests/test_wigner.cpp:21-45.
//  - sigma (lbtm, rbtm): <psi|H_eff psi> + E_core equals the MPS energy
//    (tests/test_mps_mpo_ops/test_siteproblem.cpp:84-94 invariant); lbtm and
//    rbtm schedules agree; golden DMRG energies of H2 (tests/test1.cpp:93),
//    LiH (tests/test2.cpp:57) and brute-force FCI of the seeded systems.
//  - boundary steps: the fully propagated boundary trace equals the energy
//    (tests/SweepOptimizationTools/BoundaryPropagatorElectronic.cpp:56-65),
//    left and right propagation agree.
//  - Jacobi-Davidson: lowest eigenvalue of the explicit H_eff (dense eigh).
//  - lbtm_kernel_rp (site_hamil.hpp:120-125, used when an MPO column holds
//    more than 3 operators) forms the same sums in right-paired order; the
//    oracle evaluates those columns with the left-paired kernel, which gives
//    the same values and the same FLOP count.
//  - Micro-task offsets of rbtm_tasks are 16-bit in the reference
//    (micro_kernels.hpp:433-440); the oracle keeps full-width offsets.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../include/mqgpu.h"

namespace orc {

// ---------------------------------------------------------------- charges
struct Charge {
    int a = 0, b = 0;
    Charge() = default;
    Charge(int x, int y) : a(x), b(y) {}
    bool operator<(const Charge &o) const { return a < o.a || (a == o.a && b < o.b); }
    bool operator>(const Charge &o) const { return o < *this; }
    bool operator==(const Charge &o) const { return a == o.a && b == o.b; }
    bool operator!=(const Charge &o) const { return !(*this == o); }
    Charge operator-() const { return Charge(-a, -b); }
};
inline Charge fuse(Charge x, Charge y) { return Charge(x.a + y.a, x.b + y.b); }
inline int spin(Charge c) { return c.b; }
inline bool triangle(int a, int b, int c) {  // block_matrix/symmetry/gsl_coupling.h:25-28
    return ((a + b + c) % 2 == 0) && std::abs(a - b) <= c && c <= a + b;
}

// ---------------------------------------------------------------- Index
// block_matrix/indexing_stable.hpp:82-130: sorted by descending charge.
struct Index {
    std::vector<std::pair<Charge, int>> d;
    void sort() {
        std::sort(d.begin(), d.end(), [](auto &x, auto &y) { return x.first > y.first; });
    }
    size_t size() const { return d.size(); }
    size_t position(Charge c) const {
        auto it = std::lower_bound(d.begin(), d.end(), c,
                                   [](const std::pair<Charge, int> &x, Charge v) { return x.first > v; });
        if (it != d.end() && it->first == c) return it - d.begin();
        return d.size();
    }
    bool has(Charge c) const { return position(c) != d.size(); }
    int size_of_block(Charge c) const {
        size_t p = position(c);
        if (p == d.size()) throw std::runtime_error("Index: missing charge");
        return d[p].second;
    }
    const std::pair<Charge, int> &operator[](size_t i) const { return d[i]; }
};

inline Index product(const Index &x, const Index &y) {  // indexing_stable.hpp operator*
    std::map<Charge, int> m;
    for (auto &p : x.d)
        for (auto &q : y.d) m[fuse(p.first, q.first)] += p.second * q.second;
    Index r;
    for (auto &kv : m) r.d.push_back(kv);
    r.sort();
    return r;
}
inline Index adjoin(const Index &x) {
    Index r;
    for (auto &p : x.d) r.d.push_back({-p.first, p.second});
    r.sort();
    return r;
}
inline void common_subset(Index &a, Index &b) {  // indexing_stable.hpp common_subset
    Index a2, b2;
    for (auto &p : a.d)
        if (b.has(p.first)) a2.d.push_back(p);
    for (auto &p : b.d)
        if (a2.has(p.first)) b2.d.push_back(p);
    a = a2;
    b = b2;
}

// indexing_stable.hpp:305-375
struct ProductBasis {
    std::map<std::pair<Charge, Charge>, int> off;
    std::map<Charge, int> sz;
    ProductBasis() = default;
    ProductBasis(const Index &a, const Index &b, bool right = false) {
        for (auto &p : a.d)
            for (auto &q : b.d) {
                Charge pc = right ? fuse(-p.first, q.first) : fuse(p.first, q.first);
                int &s = sz[pc];
                off[{p.first, q.first}] = s;
                s += p.second * q.second;
            }
    }
    int operator()(Charge a, Charge b) const {
        auto it = off.find({a, b});
        if (it == off.end()) throw std::runtime_error("ProductBasis: missing pair");
        return it->second;
    }
    int size(Charge pc) const {
        auto it = sz.find(pc);
        if (it == sz.end()) throw std::runtime_error("ProductBasis: missing charge");
        return it->second;
    }
};

// ---------------------------------------------------------------- flops
struct Counter {
    double flops = 0;
};
extern thread_local Counter *g_counter;

// dense kernels (BLAS dgemm via dlopen when present, else a blocked loop)
void dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
           const double *B, int ldb, double beta, double *C, int ldc);
inline void axpy(int n, double a, const double *x, double *y) {
    for (int i = 0; i < n; ++i) y[i] += a * x[i];
    if (g_counter) g_counter->flops += 2.0 * n;
}
const char *blas_name();

// ---------------------------------------------------------------- matrices
struct Matrix {
    int r = 0, c = 0;
    std::vector<double> v;
    Matrix() = default;
    Matrix(int r_, int c_) : r(r_), c(c_), v((size_t)r_ * c_, 0.0) {}
    double &operator()(int i, int j) { return v[(size_t)j * r + i]; }
    double operator()(int i, int j) const { return v[(size_t)j * r + i]; }
    double *ptr(int i, int j) { return v.data() + (size_t)j * r + i; }
    const double *ptr(int i, int j) const { return v.data() + (size_t)j * r + i; }
};

struct Block {
    Charge lc, rc;
    Matrix m;
};

// block_matrix.h:35-178 with DualIndex ordering dual_index.h:39-52.
struct BlockMatrix {
    std::vector<Block> b;
    static bool gt(const Block &x, Charge lc, Charge rc) {
        return x.lc > lc || (x.lc == lc && x.rc > rc);
    }
    size_t find(Charge lc, Charge rc) const {
        auto it = std::lower_bound(b.begin(), b.end(), std::make_pair(lc, rc),
                                   [](const Block &x, const std::pair<Charge, Charge> &v) {
                                       return gt(x, v.first, v.second);
                                   });
        if (it != b.end() && it->lc == lc && it->rc == rc) return it - b.begin();
        return b.size();
    }
    size_t insert(Charge lc, Charge rc, int rows, int cols) {
        auto it = std::lower_bound(b.begin(), b.end(), std::make_pair(lc, rc),
                                   [](const Block &x, const std::pair<Charge, Charge> &v) {
                                       return gt(x, v.first, v.second);
                                   });
        size_t pos = it - b.begin();
        Block nb;
        nb.lc = lc;
        nb.rc = rc;
        nb.m = Matrix(rows, cols);
        b.insert(it, std::move(nb));
        return pos;
    }
    size_t find_or_insert(Charge lc, Charge rc, int rows, int cols) {
        size_t p = find(lc, rc);
        if (p == b.size()) p = insert(lc, rc, rows, cols);
        return p;
    }
    // block_matrix.hpp:401-429 (sizes always agree here)
    void match_and_add(const Block &x) {
        size_t p = find(x.lc, x.rc);
        if (p == b.size()) {
            p = insert(x.lc, x.rc, x.m.r, x.m.c);
        }
        Matrix &t = b[p].m;
        if (t.r != x.m.r || t.c != x.m.c) throw std::runtime_error("match_and_add: size mismatch");
        for (size_t i = 0; i < t.v.size(); ++i) t.v[i] += x.m.v[i];
    }
    size_t n() const { return b.size(); }
    void clear() { b.clear(); }
};

BlockMatrix transpose(const BlockMatrix &x);

// ---------------------------------------------------------------- MPO
struct Entry {
    int row, col, rspin, cspin;
    double coeff;
};
struct OpBlock {
    Charge in, out;
    std::vector<Entry> e;
};
struct Op {
    int spin = 0;
    std::vector<OpBlock> blocks;
};
struct Term {
    int op;
    double scale;
};
struct MPO {
    int row_dim = 0, col_dim = 0;
    std::vector<std::vector<int>> col_rows;            // b2 -> ascending b1
    std::vector<std::vector<int>> row_cols;            // b1 -> ascending b2
    std::map<std::pair<int, int>, std::vector<Term>> at;
    std::vector<Op> ops;
    std::vector<int> left_spin, right_spin, left_herm, right_herm, left_phase, right_phase;
    std::vector<int> row_nnz, col_nnz;
    int num_one_rows = 0, num_one_cols = 0;
    bool left_skip(int b1) const { return left_herm[b1] < b1; }
    bool right_skip(int b2) const { return right_herm[b2] < b2; }
    bool has(int b1, int b2) const { return at.count({b1, b2}) != 0; }
};

using Boundary = std::vector<BlockMatrix>;

// ---------------------------------------------------------------- problem
struct Problem {
    int symm = MQ_SYMM_SU2U1;
    Index phys, left, right;
    MPO mpo;
};

Index to_index(const mq_index &x);
MPO to_mpo(const mq_mpo_desc &d);
Boundary to_boundary(const mq_boundary_desc &d, const double *data);
// canonical left-paired MPS tensor <-> block matrix
BlockMatrix mps_from_flat(const Index &phys, const Index &left, const Index &right, const double *x);
void mps_to_flat(const Index &phys, const Index &left, const Index &right, const BlockMatrix &m, double *x);
int64_t mps_len(const Index &phys, const Index &left, const Index &right);

// reshapes, mp_tensors/reshapes.h:174-336
BlockMatrix reshape_left_to_right(const Index &phys, const Index &left, const Index &right, const BlockMatrix &m1);
BlockMatrix reshape_right_to_left(const Index &phys, const Index &left, const Index &right, const BlockMatrix &m1);

// sigma vector, non-abelian/site_hamil.hpp:28-204 and abelian/site_hamil.hpp:15-87
BlockMatrix site_hamil2(const Problem &p, const BlockMatrix &ket_lp, const Boundary &L, const Boundary &R,
                        int force_schedule /* -1 auto, 0 lbtm, 1 rbtm */);
// boundary steps, common/move_boundary.hpp:130-229
Boundary overlap_mpo_left_step(const Problem &p, const BlockMatrix &ket_lp, const Boundary &L);
Boundary overlap_mpo_right_step(const Problem &p, const BlockMatrix &ket_lp, const Boundary &R);
// diagonal of H_eff for the Davidson preconditioner, non-abelian/h_diag.hpp:20-152 (SU2 only)
BlockMatrix diagonal_hamiltonian(const Problem &p, const Boundary &L, const Boundary &R);

// SU(2) coupling coefficients, block_matrix/symmetry/gsl_coupling.h:166-205
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i);
double wigner6j(int a, int b, int c, int d, int e, int f);
void set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double out[4]);
double conjugate_correction(Charge lc, Charge rc, int tensor_spin);

}  // namespace orc
