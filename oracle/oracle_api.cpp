// extern "C" entry points of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
// Used by tests/ (ctypes) and by the CPU-baseline leg of bench.py.
#include <omp.h>

#include <cstdio>
#include <cstdlib>

#include "oracle_core.hpp"

using namespace orc;

namespace {
thread_local std::string g_err;

struct Result {
    int aux = 0;
    std::vector<int32_t> bptr, lc, rc, rows, cols;
    std::vector<int64_t> offset;
    std::vector<double> data;
};

Result *pack(const Boundary &B) {
    Result *r = new Result();
    r->aux = (int)B.size();
    r->bptr.push_back(0);
    int64_t off = 0;
    for (auto &bm : B) {
        for (auto &blk : bm.b) {
            r->lc.push_back(blk.lc.a);
            r->lc.push_back(blk.lc.b);
            r->rc.push_back(blk.rc.a);
            r->rc.push_back(blk.rc.b);
            r->rows.push_back(blk.m.r);
            r->cols.push_back(blk.m.c);
            r->offset.push_back(off);
            r->data.insert(r->data.end(), blk.m.v.begin(), blk.m.v.end());
            off += blk.m.v.size();
        }
        r->bptr.push_back((int32_t)r->rows.size());
    }
    return r;
}

Problem make_problem(const mq_site_desc *d) {
    blas_name();  // resolve BLAS (and its environment) before any OpenMP region
    Problem p;
    p.symm = d->symm;
    p.phys = to_index(d->phys);
    p.left = to_index(d->left);
    p.right = to_index(d->right);
    p.mpo = to_mpo(d->mpo);
    return p;
}

double dotv(const std::vector<double> &a, const std::vector<double> &b) {
    long double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += (long double)a[i] * b[i];
    return (double)s;
}

// Cyclic Jacobi eigensolver for the small symmetric Rayleigh matrix.
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
                double theta = (aqq - app) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[(size_t)p * n + k], akq = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * akp - s * akq;
                    A[(size_t)q * n + k] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[(size_t)k * n + p], aqk = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * apk - s * aqk;
                    A[(size_t)k * n + q] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[(size_t)p * n + k], vkq = V[(size_t)q * n + k];
                    V[(size_t)p * n + k] = c * vkp - s * vkq;
                    V[(size_t)q * n + k] = s * vkp + c * vkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = A[(size_t)i * n + i];
}
}  // namespace

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }
const char *orc_blas_name() { return blas_name(); }
int orc_num_threads() { return omp_get_max_threads(); }
void orc_set_num_threads(int n) { omp_set_num_threads(n); }

double orc_wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    return wigner9j(a, b, c, d, e, f, g, h, i);
}
double orc_wigner6j(int a, int b, int c, int d, int e, int f) { return wigner6j(a, b, c, d, e, f); }
void orc_set_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double *out) {
    set_coupling(a, b, c, d, e, f, g, h, i, init, out);
}

int64_t orc_vec_len(const mq_site_desc *d) {
    return mps_len(to_index(d->phys), to_index(d->left), to_index(d->right));
}

// y = H_eff x; schedule -1 = reference rule (site_hamil.hpp:35), 0 = lbtm, 1 = rbtm.
int orc_sigma(const mq_site_desc *d, const double *L, const double *R, const double *x, double *y, int schedule,
              double *flops) {
    try {
        Problem p = make_problem(d);
        Boundary Lb = to_boundary(d->L, L), Rb = to_boundary(d->R, R);
        BlockMatrix ket = mps_from_flat(p.phys, p.left, p.right, x);
        Counter c;
        g_counter = flops ? &c : nullptr;
        BlockMatrix s = site_hamil2(p, ket, Lb, Rb, schedule);
        g_counter = nullptr;
        if (flops) *flops = c.flops;
        mps_to_flat(p.phys, p.left, p.right, s, y);
        return 0;
    } catch (std::exception &e) {
        g_counter = nullptr;
        g_err = e.what();
        return -1;
    }
}

int orc_left_step(const mq_site_desc *d, const double *L, const double *A, void **out, double *flops) {
    try {
        Problem p = make_problem(d);
        Boundary Lb = to_boundary(d->L, L);
        BlockMatrix ket = mps_from_flat(p.phys, p.left, p.right, A);
        Counter c;
        g_counter = flops ? &c : nullptr;
        Boundary r = overlap_mpo_left_step(p, ket, Lb);
        g_counter = nullptr;
        if (flops) *flops = c.flops;
        *out = pack(r);
        return 0;
    } catch (std::exception &e) {
        g_counter = nullptr;
        g_err = e.what();
        return -1;
    }
}

int orc_right_step(const mq_site_desc *d, const double *R, const double *A, void **out, double *flops) {
    try {
        Problem p = make_problem(d);
        Boundary Rb = to_boundary(d->R, R);
        BlockMatrix ket = mps_from_flat(p.phys, p.left, p.right, A);
        Counter c;
        g_counter = flops ? &c : nullptr;
        Boundary r = overlap_mpo_right_step(p, ket, Rb);
        g_counter = nullptr;
        if (flops) *flops = c.flops;
        *out = pack(r);
        return 0;
    } catch (std::exception &e) {
        g_counter = nullptr;
        g_err = e.what();
        return -1;
    }
}

void orc_result_info(void *h, int32_t *aux, int32_t *nblocks, int64_t *total) {
    Result *r = (Result *)h;
    *aux = r->aux;
    *nblocks = (int32_t)r->rows.size();
    *total = (int64_t)r->data.size();
}
void orc_result_copy(void *h, int32_t *bptr, int32_t *lc, int32_t *rc, int32_t *rows, int32_t *cols,
                     int64_t *offset, double *data) {
    Result *r = (Result *)h;
    std::copy(r->bptr.begin(), r->bptr.end(), bptr);
    std::copy(r->lc.begin(), r->lc.end(), lc);
    std::copy(r->rc.begin(), r->rc.end(), rc);
    std::copy(r->rows.begin(), r->rows.end(), rows);
    std::copy(r->cols.begin(), r->cols.end(), cols);
    std::copy(r->offset.begin(), r->offset.end(), offset);
    std::copy(r->data.begin(), r->data.end(), data);
}
void orc_result_free(void *h) { delete (Result *)h; }

// Jacobi-Davidson, framework/ietl/jacobi.h:364-493 + optimize/ietl_jacobi_davidson.h:18-77
// (GMRES off: ietl_jcd_gmres = 0, DmrgParameters.h:54).  Ritz vector sign: largest
// component positive.
int orc_jacobi_davidson(const mq_site_desc *d, const double *L, const double *R, const double *x0, double *u_out,
                        double *theta_out, int max_iter, double tol, const double *const *ortho, int n_ortho,
                        int *iters_out, double *flops) {
    try {
        Problem p = make_problem(d);
        Boundary Lb = to_boundary(d->L, L), Rb = to_boundary(d->R, R);
        int64_t n = mps_len(p.phys, p.left, p.right);
        Counter c;
        auto mult = [&](const std::vector<double> &x, std::vector<double> &y) {
            BlockMatrix ket = mps_from_flat(p.phys, p.left, p.right, x.data());
            g_counter = flops ? &c : nullptr;
            BlockMatrix s = site_hamil2(p, ket, Lb, Rb, -1);
            g_counter = nullptr;
            y.assign(n, 0.0);
            mps_to_flat(p.phys, p.left, p.right, s, y.data());
        };
        std::vector<std::vector<double>> O;
        for (int i = 0; i < n_ortho; ++i) O.emplace_back(ortho[i], ortho[i] + n);
        auto project = [&](std::vector<double> &t) {
            for (auto &o : O) {
                double f = dotv(o, t) / dotv(o, o);
                for (int64_t k = 0; k < n; ++k) t[k] -= f * o[k];
            }
        };
        std::vector<std::vector<double>> V(max_iter), VA(max_iter);
        std::vector<double> M((size_t)max_iter * max_iter, 0.0);
        V[0].assign(x0, x0 + n);
        project(V[0]);
        int it = 0;
        double theta = 0;
        std::vector<double> u(n), uA(n);
        while (true) {
            std::vector<double> &t = V[it];
            double tau = std::sqrt(dotv(t, t));
            for (int i = 0; i < it; ++i) {
                double f = dotv(V[i], t);
                for (int64_t k = 0; k < n; ++k) t[k] -= f * V[i][k];
            }
            if (std::sqrt(dotv(t, t)) < 0.25 * tau)
                for (int i = 0; i < it; ++i) {
                    double f = dotv(V[i], t);
                    for (int64_t k = 0; k < n; ++k) t[k] -= f * V[i][k];
                }
            project(t);
            double nt = std::sqrt(dotv(t, t));
            for (auto &v : t) v /= nt;
            mult(t, VA[it]);
            for (int i = 0; i <= it; ++i) M[(size_t)it * max_iter + i] = dotv(V[i], VA[it]);
            int dim = it + 1;
            std::vector<double> A((size_t)dim * dim);
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j <= i; ++j) {
                    double v = M[(size_t)i * max_iter + j];  // M(j,i), j <= i: upper triangle
                    A[(size_t)i * dim + j] = v;
                    A[(size_t)j * dim + i] = v;
                }
            std::vector<double> w, Z;
            sym_eig(dim, A, w, Z);
            int lo = 0;
            for (int i = 1; i < dim; ++i)
                if (w[i] < w[lo]) lo = i;
            theta = w[lo];
            std::vector<double> s(Z.begin() + (size_t)lo * dim, Z.begin() + (size_t)(lo + 1) * dim);
            int big = 0;
            for (int i = 1; i < dim; ++i)
                if (std::fabs(s[i]) > std::fabs(s[big])) big = i;
            if (s[big] < 0)
                for (auto &v : s) v = -v;
            std::fill(u.begin(), u.end(), 0.0);
            std::fill(uA.begin(), uA.end(), 0.0);
            for (int j = 0; j < dim; ++j)
                for (int64_t k = 0; k < n; ++k) {
                    u[k] += s[j] * V[j][k];
                    uA[k] += s[j] * VA[j][k];
                }
            project(uA);
            for (int64_t k = 0; k < n; ++k) uA[k] -= theta * u[k];
            ++it;
            double rn = std::sqrt(dotv(uA, uA));
            if (rn <= tol * std::fabs(theta) || rn < tol || it >= max_iter) break;
            double dru = dotv(uA, u), duu = dotv(u, u);
            V[it].assign(n, 0.0);
            for (int64_t k = 0; k < n; ++k) V[it][k] = -uA[k] + dru / duu * u[k];
        }
        std::copy(u.begin(), u.end(), u_out);
        *theta_out = theta;
        if (iters_out) *iters_out = it;
        if (flops) *flops = c.flops;
        return 0;
    } catch (std::exception &e) {
        g_counter = nullptr;
        g_err = e.what();
        return -1;
    }
}

// Diagonal of H_eff in the flat left-paired layout, non-abelian/h_diag.hpp:120-152.
int orc_diagonal(const mq_site_desc *d, const double *L, const double *R, double *out) {
    try {
        Problem p = make_problem(d);
        Boundary Lb = to_boundary(d->L, L), Rb = to_boundary(d->R, R);
        BlockMatrix dg = diagonal_hamiltonian(p, Lb, Rb);
        mps_to_flat(p.phys, p.left, p.right, dg, out);
        return 0;
    } catch (std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

// Davidson with the diagonal preconditioner: optimize/davidson.h:64-161 driven
// as solve_ietl_davidson does (optimize/ietl_davidson.h:121-160), with
// MultDiagonal::precondition (:96-104).  ortho: vectors projected out as
// SingleSiteVS::project does (the caller applies the Gram-Schmidt of
// ietl_davidson.h:126-129 beforehand).  Eigenvectors of the Rayleigh matrix:
// ascending eigenvalues, each column with its largest component positive.
int orc_davidson(const mq_site_desc *d, const double *L, const double *R, const double *x0, double *u_out,
                 double *theta_out, int max_iter, double tol, const double *const *ortho, int n_ortho,
                 int *iters_out, double *flops) {
    try {
        Problem p = make_problem(d);
        Boundary Lb = to_boundary(d->L, L), Rb = to_boundary(d->R, R);
        int64_t n = mps_len(p.phys, p.left, p.right);
        std::vector<double> D(n);
        mps_to_flat(p.phys, p.left, p.right, diagonal_hamiltonian(p, Lb, Rb), D.data());
        Counter c;
        auto mult = [&](const std::vector<double> &x, std::vector<double> &y) {
            BlockMatrix ket = mps_from_flat(p.phys, p.left, p.right, x.data());
            g_counter = flops ? &c : nullptr;
            BlockMatrix s = site_hamil2(p, ket, Lb, Rb, -1);
            g_counter = nullptr;
            y.assign(n, 0.0);
            mps_to_flat(p.phys, p.left, p.right, s, y.data());
        };
        std::vector<std::vector<double>> O;
        for (int i = 0; i < n_ortho; ++i) O.emplace_back(ortho[i], ortho[i] + n);
        auto project = [&](std::vector<double> &t) {
            for (auto &o : O) {
                double f = dotv(o, t) / dotv(o, o);
                for (int64_t k = 0; k < n; ++k) t[k] -= f * o[k];
            }
        };
        auto norm = [&](const std::vector<double> &v) { return std::sqrt(dotv(v, v)); };
        // MultDiagonal::mult_diag: x /= (theta - D) where that difference is nonzero
        auto mult_diag = [&](double theta, std::vector<double> &x) {
            for (int64_t k = 0; k < n; ++k)
                if (std::fabs(theta - D[k]) != 0.0) x[k] /= (theta - D[k]);
        };
        std::vector<double> t(x0, x0 + n), u, r;
        project(t);
        std::vector<std::vector<double>> V, VA;
        double theta = 0;
        int it = 0;
        const double kappa = 0.25;
        while (true) {
            double tau = norm(t);
            for (size_t i = 0; i < V.size(); ++i) {
                double f = dotv(V[i], t);
                for (int64_t k = 0; k < n; ++k) t[k] -= f * V[i][k];
            }
            if (norm(t) < kappa * tau)
                for (size_t i = 0; i < V.size(); ++i) {
                    double f = dotv(V[i], t);
                    for (int64_t k = 0; k < n; ++k) t[k] -= f * V[i][k];
                }
            project(t);
            double nt = norm(t);
            for (auto &v : t) v /= nt;
            V.push_back(t);
            VA.emplace_back();
            mult(V.back(), VA.back());
            const int dim = (int)V.size();
            std::vector<double> M((size_t)dim * dim);
            for (int i = 0; i < dim; ++i)
                for (int j = i; j < dim; ++j) {
                    M[(size_t)j * dim + i] = dotv(V[i], VA[j]);
                    M[(size_t)i * dim + j] = M[(size_t)j * dim + i];
                }
            std::vector<double> w, Z;
            sym_eig(dim, M, w, Z);
            std::vector<int> order(dim);
            for (int i = 0; i < dim; ++i) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return w[x] < w[y]; });
            // Mevecs(i, j): component i of the j-th lowest eigenvector
            std::vector<double> E((size_t)dim * dim);
            for (int j = 0; j < dim; ++j) {
                const double *z = Z.data() + (size_t)order[j] * dim;
                int big = 0;
                for (int i = 1; i < dim; ++i)
                    if (std::fabs(z[i]) > std::fabs(z[big])) big = i;
                const double sg = z[big] < 0 ? -1.0 : 1.0;
                for (int i = 0; i < dim; ++i) E[(size_t)j * dim + i] = sg * z[i];
            }
            std::vector<std::vector<double>> Vp = V, VAp = VA;
            for (int j = 0; j < dim; ++j) {
                std::fill(V[j].begin(), V[j].end(), 0.0);
                std::fill(VA[j].begin(), VA[j].end(), 0.0);
                for (int i = 0; i < dim; ++i) {
                    const double e = E[(size_t)j * dim + i];
                    for (int64_t k = 0; k < n; ++k) {
                        V[j][k] += Vp[i][k] * e;
                        VA[j][k] += VAp[i][k] * e;
                    }
                }
            }
            theta = w[order[0]];
            r.assign(n, 0.0);
            for (int64_t k = 0; k < n; ++k) r[k] = VA[0][k] - V[0][k] * theta;
            u = V[0];
            ++it;
            const double rn = norm(r);
            if (rn <= tol * std::fabs(theta) || rn < tol || it >= max_iter) break;
            // MultDiagonal::precondition
            std::vector<double> Vcpy = V[0];
            mult_diag(theta, Vcpy);
            const double a = dotv(Vcpy, r), b = dotv(V[0], Vcpy);
            for (int64_t k = 0; k < n; ++k) r[k] -= a / b * V[0][k];
            mult_diag(theta, r);
            std::swap(t, r);
            nt = norm(t);
            for (auto &v : t) v /= nt;
            if (V.size() >= 20) {
                V.resize(2);
                VA.resize(2);
            }
        }
        std::copy(u.begin(), u.end(), u_out);
        *theta_out = theta;
        if (iters_out) *iters_out = it;
        if (flops) *flops = c.flops;
        return 0;
    } catch (std::exception &e) {
        g_counter = nullptr;
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"
