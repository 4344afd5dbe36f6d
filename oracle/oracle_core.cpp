// FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_core.hpp"

#include <dlfcn.h>
#include <omp.h>

#include <cstdlib>

namespace orc {

thread_local Counter *g_counter = nullptr;

// ------------------------------------------------------------------ BLAS
typedef void (*dgemm_fn)(const char *, const char *, const int *, const int *, const int *, const double *,
                         const double *, const int *, const double *, const int *, const double *, double *,
                         const int *);
static dgemm_fn g_dgemm = nullptr;
static const char *g_blas = "in-repo blocked dgemm";
static bool g_blas_init = false;

static void blas_init_once();
static void blas_init() {
    // std::call_once-style: the first caller (serial API entry) does the work
    static const bool done = (blas_init_once(), true);
    (void)done;
}
static void blas_init_once() {
    if (g_blas_init) return;
    g_blas_init = true;
    const char *no = getenv("ORACLE_NO_BLAS");
    if (no && no[0] == '1') return;
    // Single-threaded MKL inside the OpenMP loop, as the reference's CPU path.
    setenv("MKL_THREADING_LAYER", "SEQUENTIAL", 0);
    setenv("MKL_INTERFACE_LAYER", "LP64", 0);
    const char *cands[] = {"libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so.2", "libmkl_rt.so.1"};
    for (const char *c : cands) {
        void *h = dlopen(c, RTLD_NOW | RTLD_GLOBAL);
        if (!h) continue;
        void *f = dlsym(h, "dgemm_");
        if (f) {
            g_dgemm = (dgemm_fn)f;
            g_blas = "MKL (libmkl_rt, sequential)";
            return;
        }
    }
}

const char *blas_name() {
    blas_init();
    return g_blas;
}

static void naive_gemm(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                       const double *B, int ldb, double beta, double *C, int ldc) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) C[(size_t)j * ldc + i] *= beta;
    // loop order j, p, i keeps the inner loop contiguous in C and (for 'N') in A
    for (int j = 0; j < n; ++j)
        for (int p = 0; p < k; ++p) {
            double bv = (tb == 'N') ? B[(size_t)j * ldb + p] : B[(size_t)p * ldb + j];
            if (bv == 0.0) continue;
            bv *= alpha;
            double *c = C + (size_t)j * ldc;
            if (ta == 'N') {
                const double *a = A + (size_t)p * lda;
                for (int i = 0; i < m; ++i) c[i] += a[i] * bv;
            } else {
                for (int i = 0; i < m; ++i) c[i] += A[(size_t)i * lda + p] * bv;
            }
        }
}

void dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda, const double *B,
           int ldb, double beta, double *C, int ldc) {
    if (m == 0 || n == 0) return;
    if (g_counter) g_counter->flops += 2.0 * m * n * k;
    if (k == 0) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) C[(size_t)j * ldc + i] *= beta;
        return;
    }
    blas_init();
    if (g_dgemm) {
        char a = ta, b = tb;
        g_dgemm(&a, &b, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
    } else {
        naive_gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    }
}

// C += alpha * A * B (plain block matrices)
static void gemm_acc(Matrix &C, double alpha, const Matrix &A, const Matrix &B) {
    if (A.c != B.r || C.r != A.r || C.c != B.c) throw std::runtime_error("gemm_acc: shape mismatch");
    dgemm('N', 'N', A.r, B.c, A.c, alpha, A.v.data(), std::max(1, A.r), B.v.data(), std::max(1, B.r), 1.0,
          C.v.data(), std::max(1, C.r));
}

BlockMatrix transpose(const BlockMatrix &x) {
    BlockMatrix r;
    for (auto &bl : x.b) {
        size_t p = r.insert(bl.rc, bl.lc, bl.m.c, bl.m.r);
        Matrix &t = r.b[p].m;
        for (int j = 0; j < bl.m.c; ++j)
            for (int i = 0; i < bl.m.r; ++i) t(j, i) = bl.m(i, j);
    }
    return r;
}

// ------------------------------------------------------------------ Wigner
// Racah single-sum 6j; 9j as a sum over three 6j (Edmonds 6.4.3).
static long double lfact(int n) {
    // function-local static: initialised once, thread-safe (called inside OpenMP loops)
    static const std::vector<long double> t = [] {
        std::vector<long double> v(171);
        v[0] = 1.0L;
        for (int i = 1; i < 171; ++i) v[i] = v[i - 1] * i;
        return v;
    }();
    if (n < 0 || n > 170) throw std::runtime_error("factorial out of range");
    return t[n];
}
static long double delta(int a, int b, int c) {
    return std::sqrt(lfact((a + b - c) / 2) * lfact((a - b + c) / 2) * lfact((-a + b + c) / 2) /
                     lfact((a + b + c) / 2 + 1));
}
double wigner6j(int a, int b, int c, int d, int e, int f) {
    if (!(triangle(a, b, c) && triangle(a, e, f) && triangle(d, b, f) && triangle(d, e, c))) return 0.0;
    int s1 = (a + b + c) / 2, s2 = (a + e + f) / 2, s3 = (d + b + f) / 2, s4 = (d + e + c) / 2;
    int p1 = (a + b + d + e) / 2, p2 = (b + c + e + f) / 2, p3 = (c + a + f + d) / 2;
    int tmin = std::max(std::max(s1, s2), std::max(s3, s4));
    int tmax = std::min(p1, std::min(p2, p3));
    long double sum = 0.0L;
    for (int t = tmin; t <= tmax; ++t) {
        long double term = lfact(t + 1) / (lfact(t - s1) * lfact(t - s2) * lfact(t - s3) * lfact(t - s4) *
                                           lfact(p1 - t) * lfact(p2 - t) * lfact(p3 - t));
        sum += (t % 2) ? -term : term;
    }
    return (double)(sum * delta(a, b, c) * delta(a, e, f) * delta(d, b, f) * delta(d, e, c));
}
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    if (!(triangle(a, b, c) && triangle(d, e, f) && triangle(g, h, i) && triangle(a, d, g) && triangle(b, e, h) &&
          triangle(c, f, i)))
        return 0.0;
    int lo = std::max(std::abs(a - i), std::max(std::abs(b - f), std::abs(d - h)));
    int hi = std::min(a + i, std::min(b + f, d + h));
    long double tot = 0.0L;
    for (int x = lo; x <= hi; ++x) {
        if ((x + a + i) % 2) continue;
        long double w = (long double)(x + 1) * ((x % 2) ? -1.0L : 1.0L);
        tot += w * wigner6j(a, b, c, f, i, x) * wigner6j(d, e, f, b, x, h) * wigner6j(g, h, i, x, a, d);
    }
    return (double)tot;
}
static double mod_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    return std::sqrt((g + 1.) * (h + 1.) * (c + 1.) * (f + 1.)) * wigner9j(a, b, c, d, e, f, g, h, i);
}
// gsl_coupling.h:177-204
void set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double out[4]) {
    double pre = std::sqrt((i + 1.) * (a + 1.) / ((g + 1.) * (c + 1.))) * init;
    if (triangle(a, b, c)) {
        out[0] = pre * mod_coupling(a, b, c, d, e, f, g, h, i);
        out[2] = pre * mod_coupling(a, b, c, d, e, f, g, 2, i);
    } else {
        out[0] = out[2] = 0.0;
    }
    if (triangle(a, 2, c)) {
        out[1] = pre * mod_coupling(a, 2, c, d, e, f, g, h, i);
        out[3] = pre * mod_coupling(a, 2, c, d, e, f, g, 2, i);
    } else {
        out[1] = out[3] = 0.0;
    }
}
// non-abelian/gemm.hpp:18-46
double conjugate_correction(Charge lc, Charge rc, int tensor_spin) {
    int S = std::min(spin(rc), spin(lc));
    int diff = spin(rc) - spin(lc);
    double ret = 0.0;
    if (tensor_spin == 0) {
        ret = 1.0;
    } else if (tensor_spin == 1) {
        if (diff > 0) ret = -std::sqrt((S + 1.) / (S + 2.));
        else if (diff < 0) ret = std::sqrt((S + 2.) / (S + 1.));
    } else if (tensor_spin == 2) {
        if (diff > 0) ret = -std::sqrt((S + 1.) / (S + 3.));
        else if (diff < 0) ret = -std::sqrt((S + 3.) / (S + 1.));
        else ret = 1.0;
    } else {
        throw std::runtime_error("hermitian conjugate for reduced tensor operators only implemented up to rank 1");
    }
    return ret;
}

static int case_of(int rspin, int cspin) {  // micro_kernels.hpp:33-38
    if (rspin == 2 && cspin == 2) return 3;
    if (rspin == 2) return 1;
    if (cspin == 2) return 2;
    return 0;
}

// ------------------------------------------------------------------ conversion
Index to_index(const mq_index &x) {
    Index r;
    for (int i = 0; i < x.n; ++i) r.d.push_back({Charge(x.charges[2 * i], x.charges[2 * i + 1]), x.sizes[i]});
    r.sort();
    return r;
}

MPO to_mpo(const mq_mpo_desc &d) {
    MPO m;
    m.row_dim = d.row_dim;
    m.col_dim = d.col_dim;
    m.col_rows.resize(d.col_dim);
    m.row_cols.resize(d.row_dim);
    for (int b2 = 0; b2 < d.col_dim; ++b2)
        for (int e = d.col_ptr[b2]; e < d.col_ptr[b2 + 1]; ++e) {
            int b1 = d.row_idx[e];
            m.col_rows[b2].push_back(b1);
            m.row_cols[b1].push_back(b2);
            std::vector<Term> &t = m.at[{b1, b2}];
            for (int q = d.term_ptr[e]; q < d.term_ptr[e + 1]; ++q) t.push_back({d.term_op[q], d.term_scale[q]});
        }
    for (auto &r : m.row_cols) std::sort(r.begin(), r.end());
    m.ops.resize(d.n_ops);
    for (int o = 0; o < d.n_ops; ++o) {
        m.ops[o].spin = d.op_spin[o];
        for (int bb = d.op_bptr[o]; bb < d.op_bptr[o + 1]; ++bb) {
            OpBlock ob;
            ob.in = Charge(d.opb_in[2 * bb], d.opb_in[2 * bb + 1]);
            ob.out = Charge(d.opb_out[2 * bb], d.opb_out[2 * bb + 1]);
            for (int e = d.opb_eptr[bb]; e < d.opb_eptr[bb + 1]; ++e)
                ob.e.push_back({d.ent_row[e], d.ent_col[e], d.ent_rspin[e], d.ent_cspin[e], d.ent_coeff[e]});
            m.ops[o].blocks.push_back(ob);
        }
    }
    m.left_spin.assign(d.left_spin, d.left_spin + d.row_dim);
    m.right_spin.assign(d.right_spin, d.right_spin + d.col_dim);
    m.left_herm.assign(d.left_herm, d.left_herm + d.row_dim);
    m.right_herm.assign(d.right_herm, d.right_herm + d.col_dim);
    m.left_phase.assign(d.left_phase, d.left_phase + d.row_dim);
    m.right_phase.assign(d.right_phase, d.right_phase + d.col_dim);
    m.row_nnz.assign(d.row_dim, 0);
    m.col_nnz.assign(d.col_dim, 0);
    for (auto &kv : m.at) {
        m.row_nnz[kv.first.first]++;
        m.col_nnz[kv.first.second]++;
    }
    m.num_one_rows = (int)std::count(m.row_nnz.begin(), m.row_nnz.end(), 1);
    m.num_one_cols = (int)std::count(m.col_nnz.begin(), m.col_nnz.end(), 1);
    return m;
}

Boundary to_boundary(const mq_boundary_desc &d, const double *data) {
    Boundary B(d.aux_dim);
    for (int b = 0; b < d.aux_dim; ++b)
        for (int k = d.bptr[b]; k < d.bptr[b + 1]; ++k) {
            Charge lc(d.blocks.lc[2 * k], d.blocks.lc[2 * k + 1]);
            Charge rc(d.blocks.rc[2 * k], d.blocks.rc[2 * k + 1]);
            size_t p = B[b].insert(lc, rc, d.blocks.rows[k], d.blocks.cols[k]);
            Matrix &m = B[b].b[p].m;
            if (data) std::memcpy(m.v.data(), data + d.blocks.offset[k], sizeof(double) * m.v.size());
        }
    return B;
}

int64_t mps_len(const Index &phys, const Index &left, const Index &right) {
    ProductBasis pb(phys, left);
    int64_t n = 0;
    for (auto &rp : right.d) {
        auto it = pb.sz.find(rp.first);
        if (it == pb.sz.end()) continue;
        n += (int64_t)it->second * rp.second;
    }
    return n;
}

BlockMatrix mps_from_flat(const Index &phys, const Index &left, const Index &right, const double *x) {
    ProductBasis pb(phys, left);
    BlockMatrix m;
    int64_t off = 0;
    for (auto &rp : right.d) {
        auto it = pb.sz.find(rp.first);
        if (it == pb.sz.end()) continue;
        size_t p = m.insert(rp.first, rp.first, it->second, rp.second);
        std::memcpy(m.b[p].m.v.data(), x + off, sizeof(double) * m.b[p].m.v.size());
        off += m.b[p].m.v.size();
    }
    return m;
}

void mps_to_flat(const Index &phys, const Index &left, const Index &right, const BlockMatrix &m, double *x) {
    ProductBasis pb(phys, left);
    int64_t off = 0;
    for (auto &rp : right.d) {
        auto it = pb.sz.find(rp.first);
        if (it == pb.sz.end()) continue;
        int64_t n = (int64_t)it->second * rp.second;
        size_t p = m.find(rp.first, rp.first);
        if (p == m.n()) {
            std::memset(x + off, 0, sizeof(double) * n);
        } else {
            const Matrix &mm = m.b[p].m;
            if (mm.r != it->second || mm.c != rp.second) throw std::runtime_error("mps_to_flat: shape mismatch");
            std::memcpy(x + off, mm.v.data(), sizeof(double) * n);
        }
        off += n;
    }
}

// reshapes.h:174-230 [(phys,left), right] -> [left, (-phys,right)]
BlockMatrix reshape_left_to_right(const Index &phys, const Index &left, const Index &right, const BlockMatrix &m1) {
    ProductBasis in_left(phys, left);
    ProductBasis out_right(phys, right, true);
    BlockMatrix m2;
    for (auto &blk : m1.b) {
        size_t r = right.position(blk.rc);
        if (r == right.size()) continue;
        Charge in_r = right[r].first, in_l = blk.lc;
        for (size_t s = 0; s < phys.size(); ++s) {
            Charge sc = phys[s].first;
            size_t l = left.position(fuse(in_l, -sc));
            if (l == left.size()) continue;
            Charge out_l = left[l].first;
            Charge out_r = fuse(-sc, in_r);
            size_t o = m2.find(out_l, out_r);
            if (o == m2.n()) o = m2.insert(out_l, out_r, left[l].second, out_right.size(out_r));
            int in_left_offset = in_left(sc, out_l);
            int out_right_offset = out_right(sc, in_r);
            int sdim = phys[s].second, ldim = left[l].second, rdim = right[r].second;
            Matrix &out = m2.b[o].m;
            for (int ss = 0; ss < sdim; ++ss)
                for (int rr = 0; rr < rdim; ++rr)
                    for (int ll = 0; ll < ldim; ++ll)
                        out(ll, out_right_offset + ss * rdim + rr) = blk.m(in_left_offset + ss * ldim + ll, rr);
        }
    }
    return m2;
}

// reshapes.h:290-336 [left, (-phys,right)] -> [(phys,left), right]
BlockMatrix reshape_right_to_left(const Index &phys, const Index &left, const Index &right, const BlockMatrix &m1) {
    ProductBasis in_right(phys, right, true);
    ProductBasis out_left(phys, left);
    BlockMatrix m2;
    for (auto &blk : m1.b) {
        size_t l = left.position(blk.lc);
        if (l == left.size()) continue;
        Charge in_l = left[l].first;
        Charge in_r = blk.rc;
        for (size_t s = 0; s < phys.size(); ++s) {
            size_t r = right.position(fuse(in_r, phys[s].first));
            if (r == right.size()) continue;
            Charge out_l = fuse(phys[s].first, in_l);
            Charge out_r = right[r].first;
            size_t o = m2.find(out_l, out_r);
            if (o == m2.n()) o = m2.insert(out_l, out_r, out_left.size(out_l), right[r].second);
            int in_right_offset = in_right(phys[s].first, out_r);
            int out_left_offset = out_left(phys[s].first, in_l);
            int sdim = phys[s].second, ldim = left[l].second, rdim = right[r].second;
            Matrix &out = m2.b[o].m;
            for (int ss = 0; ss < sdim; ++ss)
                for (int rr = 0; rr < rdim; ++rr)
                    for (int ll = 0; ll < ldim; ++ll)
                        out(out_left_offset + ss * ldim + ll, rr) = blk.m(ll, in_right_offset + ss * rdim + rr);
        }
    }
    return m2;
}

// ------------------------------------------------------------------ SU2 gemm family
// non-abelian/gemm.hpp:79-115: C(A.lc, B.rc) += s_k A_k B, A.lc restricted to ref.
static void gemm_trim_left(const BlockMatrix &A, const BlockMatrix &B, BlockMatrix &C, const Index &ref,
                           const std::vector<double> *scales) {
    C.clear();
    for (size_t k = 0; k < A.n(); ++k) {
        const Block &a = A.b[k];
        if (!ref.has(a.lc)) continue;
        for (auto &bb : B.b) {
            if (bb.lc != a.rc) continue;
            size_t c = C.find_or_insert(a.lc, bb.rc, a.m.r, bb.m.c);
            gemm_acc(C.b[c].m, scales ? (*scales)[k] : 1.0, a.m, bb.m);
        }
    }
}
// non-abelian/gemm.hpp:145-177: right trim, scales indexed by the B block.
static void gemm_trim_right(const BlockMatrix &A, const BlockMatrix &B, BlockMatrix &C, const Index &ref,
                            const std::vector<double> *scales) {
    C.clear();
    for (auto &a : A.b) {
        for (size_t kb = 0; kb < B.n(); ++kb) {
            const Block &bb = B.b[kb];
            if (bb.lc != a.rc) continue;
            if (!ref.has(bb.rc)) continue;
            size_t c = C.find_or_insert(a.lc, bb.rc, a.m.r, bb.m.c);
            gemm_acc(C.b[c].m, scales ? (*scales)[kb] : 1.0, a.m, bb.m);
        }
    }
}
// non-abelian/gemm.hpp:48-75 (spin = -1: no filter)
static void su2_gemm(const BlockMatrix &A, const BlockMatrix &B, BlockMatrix &C, int spin_) {
    C.clear();
    for (auto &a : A.b)
        for (auto &bb : B.b) {
            if (bb.lc != a.rc) continue;
            if (spin_ != -1 && !triangle(spin(a.lc), spin_, spin(bb.rc))) continue;
            size_t c = C.find_or_insert(a.lc, bb.rc, a.m.r, bb.m.c);
            gemm_acc(C.b[c].m, 1.0, a.m, bb.m);
        }
}
// non-abelian/gemm.hpp:179-204: C(al,al) += s * A(al,ar) B(ar,al)
static void gemm_trim(const BlockMatrix &A, const BlockMatrix &B, BlockMatrix &C, const std::vector<double> &scales,
                      bool conj_a) {
    C.clear();
    for (size_t k = 0; k < A.n(); ++k) {
        const Block &a = A.b[k];
        size_t mb = B.find(a.rc, a.lc);
        if (mb == B.n()) continue;
        size_t c = C.find_or_insert(a.lc, a.lc, a.m.r, B.b[mb].m.c);
        gemm_acc(C.b[c].m, scales[conj_a ? k : mb], a.m, B.b[mb].m);
    }
}

// common/boundary_times_mps.hpp:19-43
static std::vector<double> conjugate_phases(const BlockMatrix &bm, const MPO &mpo, int k, bool left, bool forward,
                                            bool su2) {
    std::vector<double> ret(bm.n(), 1.0);
    if (!su2) return ret;
    int S = left ? mpo.left_spin[k] : mpo.right_spin[k];
    for (size_t b = 0; b < bm.n(); ++b) {
        double s = conjugate_correction(bm.b[b].lc, bm.b[b].rc, S);
        if (forward) s *= left ? mpo.left_phase[mpo.left_herm[k]] : mpo.right_phase[mpo.right_herm[k]];
        else s *= left ? mpo.left_phase[k] : mpo.right_phase[k];
        ret[b] = s;
    }
    return ret;
}

template <class F>
static void omp_loop_counted(int n, F &&f);

// ------------------------------------------------------------------ BoundaryMPSProduct
// common/boundary_times_mps.hpp:84-218 (SU2 gemms).
struct BMP {
    const Boundary &L;
    const MPO &mpo;
    const BlockMatrix &bm;  // ket, right-paired
    Index ref;
    bool su2;
    std::vector<BlockMatrix> data;
    BMP(const BlockMatrix &bm_, const Boundary &L_, const MPO &mpo_, const Index &ref_, bool su2_)
        : L(L_), mpo(mpo_), bm(bm_), ref(ref_), su2(su2_), data(L_.size()) {}
    void one(int b1, BlockMatrix &out) const {
        if (mpo.left_skip(b1)) {
            const BlockMatrix &lc = L[mpo.left_herm[b1]];
            std::vector<double> sc = conjugate_phases(lc, mpo, b1, true, false, su2);
            gemm_trim_left(lc, bm, out, ref, &sc);
        } else {
            BlockMatrix t = transpose(L[b1]);
            gemm_trim_left(t, bm, out, ref, nullptr);
        }
    }
    void populate() {
        omp_loop_counted((int)L.size(), [&](int b1) {
            if (mpo.row_nnz[b1] == 1) return;
            one(b1, data[b1]);
        });
    }
    const BlockMatrix &at(int b1, BlockMatrix &storage) const {
        if (mpo.row_nnz[b1] == 1) {
            one(b1, storage);
            return storage;
        }
        return data[b1];
    }
};

// ------------------------------------------------------------------ lbtm kernel
// non-abelian/apply_op.hpp:23-101 with detail::lbtm micro_kernels.hpp:20-45
static void lbtm_kernel(int b2, BlockMatrix &ret, const BMP &t, const MPO &mpo, const Index &right_i,
                        const Index &out_left_i, const ProductBasis &in_right_pb, const ProductBasis &out_left_pb) {
    int ap = mpo.right_spin[b2];
    for (int b1 : mpo.col_rows[b2]) {
        BlockMatrix local;
        const BlockMatrix &T = t.at(b1, local);
        const std::vector<Term> &access = mpo.at.at({b1, b2});
        int a = mpo.left_spin[b1];
        for (const Term &term : access) {
            const Op &W = mpo.ops[term.op];
            int k = W.spin;
            for (auto &tb : T.b) {
                Charge lc = tb.lc, rc = tb.rc;
                Charge mc = rc;  // ket blocks are diagonal in charge (right-paired)
                for (auto &wb : W.blocks) {
                    Charge out_r = fuse(rc, wb.in);
                    size_t rb = right_i.position(out_r);
                    if (rb == right_i.size()) continue;
                    Charge out_l = fuse(lc, wb.out);
                    if (!triangle(spin(out_r), ap, spin(out_l))) continue;
                    if (!out_left_i.has(out_l)) continue;
                    int r_size = right_i[rb].second;
                    size_t o = ret.find_or_insert(out_l, out_r, out_left_i.size_of_block(out_l), r_size);
                    int i = spin(lc), ip = spin(out_l), j = spin(mc), jp = spin(out_r);
                    int two_sp = std::abs(i - ip), two_s = std::abs(j - jp);
                    double cpl[4];
                    set_coupling(j, two_s, jp, a, k, ap, i, two_sp, ip, term.scale, cpl);
                    int in_right_offset = in_right_pb(wb.in, out_r);
                    int out_left_offset = out_left_pb(wb.out, lc);
                    int l_size = tb.m.r;
                    Matrix &ob = ret.b[o].m;
                    for (int rr = 0; rr < r_size; ++rr)
                        for (const Entry &e : wb.e) {
                            double alfa = e.coeff * cpl[case_of(e.rspin, e.cspin)];
                            axpy(l_size, alfa, tb.m.ptr(0, in_right_offset + e.row * r_size + rr),
                                 ob.ptr(out_left_offset + e.col * l_size, rr));
                        }
                }
            }
        }
    }
}

// abelian/apply_op.hpp:24-139 with lb_tensor_mpo alps_detail.hpp:192-207: dense W blocks.
static void lbtm_kernel_abelian(int b2, BlockMatrix &ret, const BMP &t, const MPO &mpo, const Index &right_i,
                                const Index &out_left_i, const ProductBasis &in_right_pb,
                                const ProductBasis &out_left_pb) {
    for (int b1 : mpo.col_rows[b2]) {
        BlockMatrix local;
        const BlockMatrix &T = t.at(b1, local);
        const std::vector<Term> &access = mpo.at.at({b1, b2});
        for (const Term &term : access) {
            const Op &W = mpo.ops[term.op];
            for (auto &tb : T.b) {
                Charge lc = tb.lc, rc = tb.rc;
                for (auto &wb : W.blocks) {
                    Charge out_r = fuse(rc, wb.in);
                    size_t rb = right_i.position(out_r);
                    if (rb == right_i.size()) continue;
                    Charge out_l = fuse(lc, wb.out);
                    if (!out_left_i.has(out_l)) continue;
                    int r_size = right_i[rb].second;
                    size_t o = ret.find_or_insert(out_l, out_r, out_left_i.size_of_block(out_l), r_size);
                    int in_right_offset = in_right_pb(wb.in, out_r);
                    int out_left_offset = out_left_pb(wb.out, lc);
                    int l_size = tb.m.r;
                    Matrix &ob = ret.b[o].m;
                    for (int rr = 0; rr < r_size; ++rr)
                        for (const Entry &e : wb.e) {
                            double alfa = e.coeff * term.scale;
                            axpy(l_size, alfa, tb.m.ptr(0, in_right_offset + e.row * r_size + rr),
                                 ob.ptr(out_left_offset + e.col * l_size, rr));
                        }
                }
            }
        }
    }
}

// omp_for over [0, n) with schedule(dynamic,1) (utils/parallel/loops.hpp:11-33);
// exceptions are carried out of the parallel region and rethrown.
template <class F>
static void omp_loop_counted(int n, F &&f) {
    Counter *parent = g_counter;
    std::vector<double> fl(omp_get_max_threads(), 0.0);
    std::string err;
#pragma omp parallel
    {
        Counter local;
        if (parent) g_counter = &local;
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < n; ++i) {
            try {
                f(i);
            } catch (std::exception &e) {
#pragma omp critical(orc_err)
                if (err.empty()) err = e.what();
            }
        }
        if (parent) {
            fl[omp_get_thread_num()] += local.flops;
            g_counter = nullptr;
        }
    }
    if (parent) {
        g_counter = parent;
        for (double v : fl) parent->flops += v;
    }
    if (!err.empty()) throw std::runtime_error(err);
}

// non-abelian/site_hamil.hpp:57-147 (and abelian/site_hamil.hpp:15-87)
static BlockMatrix site_hamil_lbtm(const Problem &p, const BlockMatrix &ket_lp, const Boundary &L,
                                   const Boundary &R) {
    const bool su2 = p.symm == MQ_SYMM_SU2U1;
    const MPO &mpo = p.mpo;
    Index out_left_i = product(p.phys, p.left);
    Index right_i_bra = p.right;
    common_subset(out_left_i, right_i_bra);
    ProductBasis out_left_pb(p.phys, p.left);
    ProductBasis in_right_pb(p.phys, p.right, true);
    BlockMatrix ket_rp = reshape_left_to_right(p.phys, p.left, p.right, ket_lp);
    Index ref;  // left basis of the right-paired bra (= ket)
    for (auto &b : ket_rp.b) ref.d.push_back({b.lc, b.m.r});
    ref.sort();
    BMP t(ket_rp, L, mpo, ref, su2);
    t.populate();
    BlockMatrix ret;
    omp_loop_counted(mpo.col_dim, [&](int b2) {
        BlockMatrix contr, tmp;
        if (su2) lbtm_kernel(b2, contr, t, mpo, p.right, out_left_i, in_right_pb, out_left_pb);
        else lbtm_kernel_abelian(b2, contr, t, mpo, p.right, out_left_i, in_right_pb, out_left_pb);
        if (mpo.right_skip(b2)) {
            BlockMatrix adj = transpose(R[mpo.right_herm[b2]]);
            std::vector<double> ph = conjugate_phases(adj, mpo, b2, false, true, su2);
            gemm_trim(contr, adj, tmp, ph, false);
        } else {
            gemm_trim(contr, R[b2], tmp, std::vector<double>(contr.n(), 1.0), true);
        }
#pragma omp critical
        for (auto &blk : tmp.b) ret.match_and_add(blk);
    });
    return ret;
}

// ------------------------------------------------------------------ MPSBoundaryProduct
// common/boundary_times_mps.hpp:226-385
struct MBP {
    const Boundary &R;
    const MPO &mpo;
    const BlockMatrix &bm;  // ket, left-paired
    Index ref;
    bool su2;
    mutable std::vector<BlockMatrix> data;
    mutable std::vector<char> pop;
    // single-use columns: FLOPs of computing each, and whether a task has used it yet
    // (the reference computes such a column on its first use only)
    mutable std::vector<double> lazy_flops;
    mutable std::vector<char> used;
    MBP(const BlockMatrix &bm_, const Boundary &R_, const MPO &mpo_, const Index &ref_, bool su2_)
        : R(R_), mpo(mpo_), bm(bm_), ref(ref_), su2(su2_), data(R_.size()), pop(R_.size(), 0),
          lazy_flops(R_.size(), 0.0), used(R_.size(), 0) {}
    void one(int b2, BlockMatrix &out) const {
        if (mpo.right_skip(b2)) {
            BlockMatrix trv = transpose(R[mpo.right_herm[b2]]);
            std::vector<double> sc = conjugate_phases(trv, mpo, b2, false, true, su2);
            gemm_trim_right(bm, trv, out, ref, &sc);
        } else {
            gemm_trim_right(bm, R[b2], out, ref, nullptr);
        }
    }
    void populate() {
        omp_loop_counted((int)R.size(), [&](int b2) {
            if (mpo.col_nnz[b2] == 1) return;
            one(b2, data[b2]);
        });
    }
    // lazily computed single-use columns (boundary_times_mps.hpp:332-362); computed by the
    // owning thread only: each single-use column belongs to exactly one row b1.
    const BlockMatrix &at(int b2) const {
        if (mpo.col_nnz[b2] == 1 && !pop[b2]) {
            one(b2, data[b2]);
            pop[b2] = 1;
        }
        return data[b2];
    }
    // Computes a single-use column ahead of its use without counting its FLOPs;
    // mark_used then counts them once, at the first task that reads the column,
    // as the reference's lazy at() spends them.
    void prefetch(int b2) const {
        Counter c, *parent = g_counter;
        g_counter = parent ? &c : nullptr;
        at(b2);
        g_counter = parent;
        lazy_flops[b2] = c.flops;
    }
    void mark_used(int b2) const {
        if (mpo.col_nnz[b2] != 1 || used[b2]) return;
        used[b2] = 1;
        if (g_counter) g_counter->flops += lazy_flops[b2];
    }
};

struct MicroTask {
    double scale;
    int b2, k, l_size, r_size, stripe, out_offset;
    int64_t in_offset;
};

using TaskMap = std::map<std::pair<Charge, Charge>, std::vector<MicroTask>,
                         std::function<bool(const std::pair<Charge, Charge> &, const std::pair<Charge, Charge> &)>>;

static bool cmp_pair(const std::pair<Charge, Charge> &x, const std::pair<Charge, Charge> &y) {
    return x.first < y.first || (x.first == y.first && x.second < y.second);
}

// non-abelian/apply_op.hpp:114-209 (op_iterate micro_kernels.hpp:150-180)
static void rbtm_tasks(int b1, const MBP &t, const MPO &mpo, const Index &left_i, const Index &out_right_i,
                       const ProductBasis &in_left_pb, const ProductBasis &out_right_pb, TaskMap &tasks, bool su2) {
    int a = mpo.left_spin[b1];
    for (int b2 : mpo.row_cols[b1]) {
        const BlockMatrix &T = t.at(b2);
        const std::vector<Term> &access = mpo.at.at({b1, b2});
        int ap = mpo.right_spin[b2];
        for (const Term &term : access) {
            const Op &W = mpo.ops[term.op];
            int k = W.spin;
            for (size_t tbk = 0; tbk < T.n(); ++tbk) {
                Charge lc = T.b[tbk].lc, rc = T.b[tbk].rc;
                Charge mc = lc;  // left-paired ket blocks are diagonal in charge
                for (auto &wb : W.blocks) {
                    Charge out_l = fuse(lc, -wb.in);
                    size_t lb = left_i.position(out_l);
                    if (lb == left_i.size()) continue;
                    Charge out_r = fuse(rc, -wb.out);
                    if (su2 && !triangle(spin(out_l), a, spin(out_r))) continue;
                    if (!out_right_i.has(out_r)) continue;
                    int l_size = left_i[lb].second;
                    std::vector<MicroTask> &ot = tasks[{out_l, out_r}];
                    double cpl[4] = {term.scale, term.scale, term.scale, term.scale};
                    if (su2) {
                        int i = spin(out_r), ip = spin(rc), j = spin(out_l), jp = spin(mc);
                        int two_sp = std::abs(i - ip), two_s = std::abs(j - jp);
                        set_coupling(j, two_s, jp, a, k, ap, i, two_sp, ip, term.scale, cpl);
                    }
                    int in_left_offset = in_left_pb(wb.in, out_l);
                    int out_right_offset = out_right_pb(wb.out, rc);
                    int r_size = T.b[tbk].m.c;
                    int stripe = T.b[tbk].m.r;
                    for (const Entry &e : wb.e) {
                        MicroTask mt;
                        mt.scale = e.coeff * (su2 ? cpl[case_of(e.rspin, e.cspin)] : term.scale);
                        mt.b2 = b2;
                        mt.k = (int)tbk;
                        mt.l_size = l_size;
                        mt.r_size = r_size;
                        mt.stripe = stripe;
                        mt.in_offset = in_left_offset + (int64_t)e.row * l_size;
                        mt.out_offset = out_right_offset + e.col * r_size;
                        ot.push_back(mt);
                    }
                }
            }
        }
    }
}

// micro_kernels.hpp:183-199 task_axpy
static void task_axpy(const MicroTask &mt, double *buf, const double *src) {
    for (int rr = 0; rr < mt.r_size; ++rr)
        axpy(mt.l_size, mt.scale, src + (int64_t)mt.stripe * rr, buf + (int64_t)(mt.out_offset + rr) * mt.l_size);
}

// non-abelian/apply_op.hpp:211-234 rbtm_axpy
static void rbtm_axpy(TaskMap &tasks, BlockMatrix &ret, const Index &out_right_i, const MBP &t) {
    for (auto &kv : tasks) {
        auto &ot = kv.second;
        std::stable_sort(ot.begin(), ot.end(),
                         [](const MicroTask &x, const MicroTask &y) { return x.out_offset < y.out_offset; });
        if (ot.empty()) continue;
        size_t o = ret.insert(kv.first.first, kv.first.second, ot[0].l_size,
                              out_right_i.size_of_block(kv.first.second));
        Matrix &buf = ret.b[o].m;
        for (auto &mt : ot) {
            t.mark_used(mt.b2);
            task_axpy(mt, buf.v.data(), t.at(mt.b2).b[mt.k].m.v.data() + mt.in_offset);
        }
    }
}

// non-abelian/apply_op.hpp:268-298 rbtm_axpy_gemm
static void rbtm_axpy_gemm(int b1, TaskMap &tasks, BlockMatrix &prod, const Index &out_right_i,
                           const MPO &mpo, const BlockMatrix &left_b1, const MBP &t, bool su2) {
    std::vector<double> phases = mpo.left_skip(b1) ? conjugate_phases(left_b1, mpo, b1, true, false, su2)
                                                   : std::vector<double>(left_b1.n(), 1.0);
    for (auto &kv : tasks) {
        auto &ot = kv.second;
        std::stable_sort(ot.begin(), ot.end(),
                         [](const MicroTask &x, const MicroTask &y) { return x.out_offset < y.out_offset; });
        if (ot.empty()) continue;
        size_t k = left_b1.find(kv.first.second, kv.first.first);
        if (k == left_b1.n()) continue;
        Matrix buf(ot[0].l_size, out_right_i.size_of_block(kv.first.second));
        for (auto &mt : ot) {
            t.mark_used(mt.b2);
            task_axpy(mt, buf.v.data(), t.at(mt.b2).b[mt.k].m.v.data() + mt.in_offset);
        }
        Charge rc = kv.first.second;
        const Matrix &A = left_b1.b[k].m;
        size_t c = prod.find_or_insert(rc, rc, A.r, buf.c);
        gemm_acc(prod.b[c].m, phases[k], A, buf);
    }
}

// non-abelian/site_hamil.hpp:149-204
static BlockMatrix site_hamil_rbtm(const Problem &p, const BlockMatrix &ket_lp, const Boundary &L,
                                   const Boundary &R) {
    const bool su2 = p.symm == MQ_SYMM_SU2U1;
    const MPO &mpo = p.mpo;
    Index ref;  // right basis of the left-paired bra (= ket)
    for (auto &b : ket_lp.b) ref.d.push_back({b.rc, b.m.c});
    ref.sort();
    MBP t(ket_lp, R, mpo, ref, su2);
    t.populate();
    Index left_i = p.left;
    Index out_right_i = product(adjoin(p.phys), p.right);
    Index left_i_ket = p.left;
    common_subset(out_right_i, left_i_ket);
    ProductBasis in_left_pb(p.phys, p.left);
    ProductBasis out_right_pb(p.phys, p.right, true);
    BlockMatrix collector;
    // Single-use columns are materialised before the parallel loop so that the
    // lazily filled cache is never written concurrently; their FLOPs count at their
    // first use (MBP::mark_used), where the reference computes them.
    for (int b2 = 0; b2 < mpo.col_dim; ++b2)
        if (mpo.col_nnz[b2] == 1) t.prefetch(b2);
    omp_loop_counted(mpo.row_dim, [&](int b1) {
        TaskMap tasks(cmp_pair);
        BlockMatrix tmp2;
        rbtm_tasks(b1, t, mpo, left_i, out_right_i, in_left_pb, out_right_pb, tasks, su2);
        if (mpo.left_skip(b1)) {
            rbtm_axpy_gemm(b1, tasks, tmp2, out_right_i, mpo, L[mpo.left_herm[b1]], t, su2);
        } else {
            BlockMatrix lt = transpose(L[b1]);
            rbtm_axpy_gemm(b1, tasks, tmp2, out_right_i, mpo, lt, t, su2);
        }
#pragma omp critical
        for (auto &blk : tmp2.b) collector.match_and_add(blk);
    });
    return reshape_right_to_left(p.phys, p.left, p.right, collector);
}

// non-abelian/site_hamil.hpp:28-39 (abelian/site_hamil.hpp: lbtm only)
BlockMatrix site_hamil2(const Problem &p, const BlockMatrix &ket_lp, const Boundary &L, const Boundary &R,
                        int force) {
    bool lbtm;
    if (force >= 0) lbtm = (force == 0);
    else if (p.symm != MQ_SYMM_SU2U1) lbtm = true;
    else lbtm = (p.mpo.row_dim - p.mpo.num_one_rows) < (p.mpo.col_dim - p.mpo.num_one_cols);
    return lbtm ? site_hamil_lbtm(p, ket_lp, L, R) : site_hamil_rbtm(p, ket_lp, L, R);
}

// ------------------------------------------------------------------ boundary steps
// common/move_boundary.hpp:130-187
Boundary overlap_mpo_left_step(const Problem &p, const BlockMatrix &ket_lp, const Boundary &L) {
    const bool su2 = p.symm == MQ_SYMM_SU2U1;
    const MPO &mpo = p.mpo;
    BlockMatrix ket_rp = reshape_left_to_right(p.phys, p.left, p.right, ket_lp);
    Index braBasis;
    for (auto &b : ket_rp.b) braBasis.d.push_back({b.lc, b.m.r});
    braBasis.sort();
    BMP t(ket_rp, L, mpo, braBasis, su2);
    t.populate();
    Index out_left_i = product(p.phys, p.left);
    Index bra_right_i = p.right;
    common_subset(out_left_i, bra_right_i);
    ProductBasis out_left_pb(p.phys, p.left);
    ProductBasis in_right_pb(p.phys, p.right, true);
    Boundary ret(mpo.col_dim);
    omp_loop_counted(mpo.col_dim, [&](int b2) {
        if (mpo.right_skip(b2)) return;
        BlockMatrix contr;
        if (su2) lbtm_kernel(b2, contr, t, mpo, p.right, out_left_i, in_right_pb, out_left_pb);
        else lbtm_kernel_abelian(b2, contr, t, mpo, p.right, out_left_i, in_right_pb, out_left_pb);
        BlockMatrix ct = transpose(contr);
        su2_gemm(ct, ket_lp, ret[b2], su2 ? mpo.right_spin[b2] : -1);
    });
    return ret;
}

// common/move_boundary.hpp:191-229
Boundary overlap_mpo_right_step(const Problem &p, const BlockMatrix &ket_lp, const Boundary &R) {
    const bool su2 = p.symm == MQ_SYMM_SU2U1;
    const MPO &mpo = p.mpo;
    Index indexForTrim;
    for (auto &b : ket_lp.b) indexForTrim.d.push_back({b.rc, b.m.c});
    indexForTrim.sort();
    MBP t(ket_lp, R, mpo, indexForTrim, su2);
    t.populate();
    for (int b2 = 0; b2 < mpo.col_dim; ++b2)
        if (mpo.col_nnz[b2] == 1) t.prefetch(b2);
    Index out_right_i = product(adjoin(p.phys), p.right);
    Index bra_left_i = p.left;
    common_subset(out_right_i, bra_left_i);
    ProductBasis in_left_pb(p.phys, p.left);
    ProductBasis out_right_pb(p.phys, p.right, true);
    BlockMatrix bra_rp = reshape_left_to_right(p.phys, p.left, p.right, ket_lp);
    BlockMatrix bra_t = transpose(bra_rp);
    Boundary ret(mpo.row_dim);
    omp_loop_counted(mpo.row_dim, [&](int b1) {
        if (mpo.left_skip(b1)) return;
        TaskMap tasks(cmp_pair);
        BlockMatrix buf;
        rbtm_tasks(b1, t, mpo, p.left, out_right_i, in_left_pb, out_right_pb, tasks, su2);
        rbtm_axpy(tasks, buf, out_right_i, t);
        su2_gemm(buf, bra_t, ret[b1], su2 ? mpo.left_spin[b1] : -1);
    });
    return ret;
}

// non-abelian/h_diag.hpp:20-117 (lbtm_diag_kernel) and :120-152
// (diagonal_hamiltonian): the diagonal of H_eff built from the diagonals of
// the stored boundary blocks L[b1](lc,lc), R[b2](rc,rc) and the diagonal
// entries (row == col) of the phys-diagonal operator blocks.  Boundary slices
// that are not stored (hermitian partners) contribute nothing, as in the
// reference.  The result holds one (rc, rc) block per right charge reached by
// some R[b2].
BlockMatrix diagonal_hamiltonian(const Problem &p, const Boundary &L, const Boundary &R) {
    if (p.symm != MQ_SYMM_SU2U1) throw std::runtime_error("Davidson only implemented for spin-adapted Hamiltonians");
    const MPO &mpo = p.mpo;
    Index right_i = p.right, out_left_i = product(p.phys, p.left);
    common_subset(out_left_i, right_i);
    ProductBasis left_pb(p.phys, p.left);
    BlockMatrix ret;
    for (int b2 = 0; b2 < (int)R.size(); ++b2) {
        BlockMatrix lb2;
        for (int b1 : mpo.col_rows[b2]) {
            const std::vector<Term> &terms = mpo.at.at({b1, b2});
            for (const Term &term : terms) {
                const Op &W = mpo.ops[term.op];
                const int a = mpo.left_spin[b1], k = W.spin, ap = mpo.right_spin[b2];
                for (size_t block = 0; block < right_i.size(); ++block) {
                    const Charge in_charge = right_i[block].first;
                    const size_t o = lb2.find_or_insert(in_charge, in_charge, out_left_i[block].second,
                                                        right_i[block].second);
                    for (size_t s = 0; s < p.phys.size(); ++s) {
                        const Charge phys_charge = p.phys[s].first;
                        const size_t l = p.left.position(fuse(in_charge, -phys_charge));
                        if (l == p.left.size()) continue;
                        const Charge lc = p.left[l].first;
                        const size_t lb = L[b1].find(lc, lc);
                        if (lb == L[b1].n()) continue;
                        const int ls = p.left[l].second;
                        std::vector<double> left_diagonal(ls);
                        for (int i = 0; i < ls; ++i) left_diagonal[i] = L[b1].b[lb].m(i, i);
                        const int left_offset = left_pb(phys_charge, lc);
                        for (const OpBlock &wb : W.blocks) {
                            if (phys_charge != wb.in || wb.in != wb.out) continue;
                            const int i = spin(lc), ip = spin(in_charge), j = spin(lc), jp = spin(in_charge);
                            const int two_sp = std::abs(i - ip), two_s = std::abs(j - jp);
                            const double prefactor = std::sqrt((ip + 1.) * (j + 1.) / ((i + 1.) * (jp + 1.))) * term.scale;
                            const double couplings[2] = {
                                prefactor * mod_coupling(j, two_s, jp, a, k, ap, i, two_sp, ip),
                                prefactor * mod_coupling(j, 2, jp, a, k, ap, i, 2, ip)};
                            for (const Entry &e : wb.e) {
                                if (e.row != e.col) continue;
                                const double alfa_t = e.coeff * couplings[e.rspin == 2 ? 1 : 0];
                                Matrix &out = lb2.b[o].m;
                                for (int col = 0; col < right_i[block].second; ++col)
                                    for (int r = 0; r < ls; ++r)
                                        out(left_offset + e.row * ls + r, col) += left_diagonal[r] * alfa_t;
                            }
                        }
                    }
                }
            }
        }
        for (Block &bk : lb2.b) {
            const size_t rb = R[b2].find(bk.rc, bk.rc);
            if (rb == R[b2].n()) continue;
            const Matrix &Rm = R[b2].b[rb].m;
            for (int c = 0; c < bk.m.c; ++c)
                for (int r = 0; r < bk.m.r; ++r) bk.m(r, c) *= Rm(c, c);
            ret.match_and_add(bk);
        }
    }
    return ret;
}

}  // namespace orc
