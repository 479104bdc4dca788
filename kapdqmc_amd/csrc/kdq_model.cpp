/*
 * kdq_model.cpp — host construction of the model constants (no GPU work).
 *
 * The keyword constructors tU_Hubbard_Para (src/tU/model.jl:39-96), tV_Hubbard_Para
 * (src/tV/model.jl:30-122) and tUV_Hubbard_Para (src/tUV/model.jl:30-118) and the lattice
 * geometry they call (src/public/Geometry.jl:8-207), so that a C or C++ caller can fill the
 * kdq_model_desc of kdq_create without Julia or Python. The two LAPACK.syevd! calls of each
 * constructor become one symmetric eigensolver written here: Householder reduction to
 * tridiagonal form with the reflectors accumulated, then implicit symmetric QR steps with
 * the Wilkinson shift (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 8.3.1 and 8.3.3).
 * Eigenvector signs and the basis chosen inside a degenerate level differ from LAPACK's;
 * nothing downstream depends on them (matrix functions of K, and the trial state only
 * through the projector onto its span).
 */
#include "../../include/kapdqmc.h"
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

// Sets the message kdq_last_error(NULL) returns (defined beside the handle API).
void kdq_internal_set_error(const std::string &msg);

struct kdq_model_s {
    std::vector<double> K, eK, eKinv, hK, hKinv, Pt, alpha, eta, gamma;
    std::vector<int64_t> nodes;
    std::vector<int32_t> bonds;
};

namespace {

using Mat = std::vector<double>;   // column-major n x n

int mod1(int a, int n) {
    int r = a % n;
    if (r < 0) r += n;
    return r == 0 ? n : r;
}
int cdiv(int a, int b) { return (a + b - 1) / b; }

// Even 1-based index of unit cell (x, y) and its inverse (src/public/Geometry.jl:8-21).
int xy_i(const int *site, int x, int y) { return 2 * (x + (y - 1) * site[0]); }
void i_xy(const int *site, int i, int &x, int &y) {
    int j = cdiv(i, 2);
    x = mod1(j, site[0]);
    y = cdiv(j, site[0]);
}

// 1-based neighbours of site idx, in the reference's order (src/public/Geometry.jl:74-134).
int nn2idx(int lattice, const int *site, int idx, int *nn) {
    int x, y;
    i_xy(site, idx, x, y);
    const int L1 = site[0], L2 = site[1];
    const bool a_site = idx % 2 == 1;
    switch (lattice) {
    case KDQ_LATTICE_SQ90:
        if (a_site) {
            nn[0] = idx + 1;
            nn[1] = xy_i(site, mod1(x - 1, L1), mod1(y - 1, L2));
            nn[2] = xy_i(site, x, mod1(y - 1, L2));
            nn[3] = xy_i(site, mod1(x - 1, L1), y);
        } else {
            nn[0] = xy_i(site, mod1(x + 1, L1), y) - 1;
            nn[1] = xy_i(site, x, mod1(y + 1, L2)) - 1;
            nn[2] = idx - 1;
            nn[3] = xy_i(site, mod1(x + 1, L1), mod1(y + 1, L2)) - 1;
        }
        return 4;
    case KDQ_LATTICE_SQ45:
        if (a_site) {
            nn[0] = xy_i(site, mod1(x - 1, L1), y);
            nn[1] = idx + 1;
            nn[2] = xy_i(site, mod1(x - 1, L1), mod1(y + 1, L2));
            nn[3] = xy_i(site, x, mod1(y - 1, L2));
        } else {
            nn[0] = xy_i(site, x, mod1(y + 1, L2)) - 1;
            nn[1] = xy_i(site, mod1(x + 1, L1), mod1(y - 1, L2)) - 1;
            nn[2] = idx - 1;
            nn[3] = xy_i(site, mod1(x + 1, L1), mod1(y, L2)) - 1;
        }
        return 4;
    case KDQ_LATTICE_HC120:
        if (a_site) {
            nn[0] = idx + 1;
            nn[1] = xy_i(site, mod1(x + 1, L1), y);
            nn[2] = xy_i(site, x, mod1(y - 1, L2));
        } else {
            nn[0] = idx - 1;
            nn[1] = xy_i(site, x, mod1(y + 1, L2)) - 1;
            nn[2] = xy_i(site, mod1(x - 1, L1), y) - 1;
        }
        return 3;
    case KDQ_LATTICE_HC60:
        if (a_site) {
            nn[0] = idx + 1;
            nn[1] = xy_i(site, mod1(x + 1, L1), mod1(y - 1, L2));
            nn[2] = xy_i(site, x, mod1(y - 1, L2));
        } else {
            nn[0] = idx - 1;
            nn[1] = xy_i(site, x, mod1(y + 1, L2)) - 1;
            nn[2] = xy_i(site, mod1(x - 1, L1), mod1(y + 1, L2)) - 1;
        }
        return 3;
    }
    return 0;
}

// Hopping matrix with unit amplitudes and no flux (src/public/Geometry.jl:178-207).
Mat nnK(int lattice, const int *site, int ns) {
    Mat K((size_t)ns * ns, 0.0);
    int nn[4];
    for (int i = 1; i <= ns; ++i) {
        int z = nn2idx(lattice, site, i, nn);
        for (int j = 0; j < z; ++j) K[(size_t)(nn[j] - 1) * ns + (i - 1)] = 1.0;
    }
    return K;
}

bool symmetric(int n, const Mat &K) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i)
            if (K[(size_t)j * n + i] != K[(size_t)i * n + j]) return false;
    return true;
}

// Householder vector of x[0..m): returns beta and v (v[0] = 1) with (I - beta v v^T) x = |x| e1
// (Golub & Van Loan Alg. 5.1.1).
double house(int m, const double *x, double *v) {
    double sigma = 0.0;
    for (int i = 1; i < m; ++i) sigma += x[i] * x[i];
    v[0] = 1.0;
    if (sigma == 0.0) {
        for (int i = 1; i < m; ++i) v[i] = 0.0;
        return 0.0;
    }
    const double nx = std::sqrt(x[0] * x[0] + sigma);
    const double v0 = x[0] <= 0.0 ? x[0] - nx : -sigma / (x[0] + nx);
    for (int i = 1; i < m; ++i) v[i] = x[i] / v0;
    return 2.0 * v0 * v0 / (sigma + v0 * v0);
}

// Eigen-decomposition of the symmetric n x n matrix A (column-major, full storage,
// overwritten): ascending eigenvalues w and orthonormal eigenvectors V (columns).
bool sym_eig(int n, Mat &A, std::vector<double> &w, Mat &V) {
    auto a = [&](int i, int j) -> double & { return A[(size_t)j * n + i]; };
    std::vector<double> hv((size_t)n * n, 0.0), hb(n, 0.0), p(n), q(n);
    // reduction to tridiagonal form: A <- H_k A H_k, k = 0 .. n-3
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;
        double *v = &hv[(size_t)k * n + k + 1];
        const double beta = house(m, &a(k + 1, k), v);
        hb[k] = beta;
        if (beta == 0.0) continue;
        // p = beta A22 v, q = p - (beta/2)(p.v) v, A22 <- A22 - v q^T - q v^T
        for (int i = 0; i < m; ++i) p[i] = 0.0;
        for (int j = 0; j < m; ++j) {
            const double vj = beta * v[j];
            const double *col = &a(k + 1, k + 1 + j);
            for (int i = 0; i < m; ++i) p[i] += col[i] * vj;
        }
        double pv = 0.0;
        for (int i = 0; i < m; ++i) pv += p[i] * v[i];
        for (int i = 0; i < m; ++i) q[i] = p[i] - 0.5 * beta * pv * v[i];
        for (int j = 0; j < m; ++j) {
            double *col = &a(k + 1, k + 1 + j);
            const double vj = v[j], qj = q[j];
            for (int i = 0; i < m; ++i) col[i] -= v[i] * qj + q[i] * vj;
        }
        double nx = 0.0;   // the reflected column is (|x|, 0, ..., 0)
        for (int i = 0; i < m; ++i) nx += a(k + 1 + i, k) * a(k + 1 + i, k);
        nx = std::sqrt(nx);
        a(k + 1, k) = nx;
        a(k, k + 1) = nx;
    }
    std::vector<double> d(n), e(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = a(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
    // Q = H_0 H_1 ... H_{n-3}, accumulated backwards
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int k = n - 3; k >= 0; --k) {
        const double beta = hb[k];
        if (beta == 0.0) continue;
        const int m = n - k - 1;
        const double *v = &hv[(size_t)k * n + k + 1];
        for (int j = k + 1; j < n; ++j) {
            double *col = &V[(size_t)j * n + k + 1];
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += v[i] * col[i];
            s *= beta;
            for (int i = 0; i < m; ++i) col[i] -= s * v[i];
        }
    }
    // implicit symmetric QR with Wilkinson shifts on (d, e), rotations applied to V
    const double eps = 2.220446049250313e-16;
    int iters = 0;
    for (int hi = n - 1; hi > 0;) {
        for (int i = 0; i < hi; ++i)
            if (std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1])) || std::fabs(e[i]) < 1e-300)
                e[i] = 0.0;
        if (e[hi - 1] == 0.0) {
            --hi;
            continue;
        }
        int lo = hi - 1;
        while (lo > 0 && e[lo - 1] != 0.0) --lo;
        if (++iters > 60 * n) return false;
        const double dd = 0.5 * (d[hi - 1] - d[hi]);
        const double eh = e[hi - 1];
        const double mu = d[hi] - eh * eh / (dd + std::copysign(std::hypot(dd, eh), dd == 0.0 ? 1.0 : dd));
        double x = d[lo] - mu, z = e[lo];
        for (int k = lo; k < hi; ++k) {
            const double r = std::hypot(x, z);
            const double c = x / r, s = z / r;
            if (k > lo) e[k - 1] = r;
            const double a0 = d[k], b0 = e[k], c0 = d[k + 1];
            d[k] = c * c * a0 + 2.0 * c * s * b0 + s * s * c0;
            d[k + 1] = s * s * a0 - 2.0 * c * s * b0 + c * c * c0;
            e[k] = c * s * (c0 - a0) + (c * c - s * s) * b0;
            if (k + 1 < hi) {
                z = s * e[k + 1];
                e[k + 1] *= c;
                x = e[k];
            }
            double *vk = &V[(size_t)k * n], *vk1 = &V[(size_t)(k + 1) * n];
            for (int i = 0; i < n; ++i) {
                const double t0 = vk[i], t1 = vk1[i];
                vk[i] = c * t0 + s * t1;
                vk1[i] = c * t1 - s * t0;
            }
        }
    }
    // ascending order
    w.assign(d.begin(), d.end());
    for (int i = 0; i < n; ++i) {
        int m = i;
        for (int j = i + 1; j < n; ++j)
            if (w[j] < w[m]) m = j;
        if (m != i) {
            std::swap(w[i], w[m]);
            std::swap_ranges(V.begin() + (size_t)i * n, V.begin() + (size_t)(i + 1) * n, V.begin() + (size_t)m * n);
        }
    }
    return true;
}

// out = V diag(f) V^T (symmetric), column by column.
Mat vfv(int n, const Mat &V, const std::vector<double> &f) {
    Mat out((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        double *o = &out[(size_t)j * n];
        for (int k = 0; k < n; ++k) {
            const double s = f[k] * V[(size_t)k * n + j];
            if (s == 0.0) continue;
            const double *vk = &V[(size_t)k * n];
            for (int i = 0; i < n; ++i) o[i] += vk[i] * s;
        }
    }
    return out;
}

// C = A B for n x n A and n x m B (column-major).
Mat matmul(int n, int m, const Mat &A, const Mat &B) {
    Mat C((size_t)n * m, 0.0);
    for (int j = 0; j < m; ++j)
        for (int k = 0; k < n; ++k) {
            const double b = B[(size_t)j * n + k];
            if (b == 0.0) continue;
            const double *ak = &A[(size_t)k * n];
            double *c = &C[(size_t)j * n];
            for (int i = 0; i < n; ++i) c[i] += ak[i] * b;
        }
    return C;
}

// Stabilisation nodes (src/tU/model.jl:81-85; the same lines in tV and tUV).
std::vector<int64_t> make_nodes(int nt, int bs) {
    std::vector<int64_t> nodes;
    const int h = nt / 2;
    if (h % bs == 0) {
        for (int t = 0; t <= nt; t += bs) nodes.push_back(t);
        return nodes;
    }
    nodes.push_back(0);
    std::vector<int64_t> left;
    for (int t = h - bs; t >= 1; t -= bs) left.push_back(t);
    nodes.insert(nodes.end(), left.rbegin(), left.rend());
    for (int t = h; t <= nt; t += bs) nodes.push_back(t);
    nodes.push_back(nt);
    return nodes;
}

// Coupling profile of a quench (src/tU/model.jl:41-52; src/tV/model.jl the same).
bool coupling_profile(double h1, double h2, double th_relax, double th_quench, double dt, int nt,
                      std::vector<double> &prof, std::string &err) {
    prof.clear();
    if (th_quench > 0.0 && std::fabs(h1 - h2) > 0.0) {
        const int nq = (int)std::nearbyint(th_quench / dt);
        const int nr = (int)std::nearbyint(th_relax / dt);
        std::vector<double> lin;
        // LinRange(h1, h2, nq + 1)[2:end]: element i is (1 - t) h1 + t h2 with t = i / nq, as
        // Julia's lerp evaluates it (two products and a sum, no fused multiply-add)
        for (int i = 1; i <= nq; ++i) {
            const double t = (double)i / nq;
            volatile double a = (1.0 - t) * h1;
            volatile double b = t * h2;
            lin.push_back(a + b);
        }
        prof.assign(nr, h1);
        prof.insert(prof.end(), lin.begin(), lin.end());
        prof.insert(prof.end(), lin.rbegin(), lin.rend());
        prof.insert(prof.end(), (size_t)nr, h1);
    } else {
        if (!(h1 == h2 && th_quench < 1e-7)) {
            err = "For Θquench=0, the two couplings must be equal";
            return false;
        }
        prof.assign(nt, h1);
    }
    double asym = 0.0;
    for (size_t i = 0; i < prof.size(); ++i) {
        const double t = prof[prof.size() - 1 - i] - prof[i];
        asym += t * t;
    }
    if (std::sqrt(asym) >= 1e-10) {
        err = "coupling profile is not symmetric!";
        return false;
    }
    if ((int)prof.size() != nt) {
        err = "Length of coupling profile does not match Nt!";
        return false;
    }
    return true;
}

int model_fail(const std::string &msg) {
    kdq_internal_set_error("kdq_model_build: " + msg);
    return KDQ_EARG;
}

}  // namespace

extern "C" int kdq_model_build(const kdq_model_params *p, kdq_model *out, kdq_model_desc *desc) {
    if (!p || !out || !desc) return model_fail("null argument");
    *out = nullptr;
    const int kind = p->kind, lattice = p->lattice;
    if (kind < KDQ_MODEL_TU || kind > KDQ_MODEL_TUV) return model_fail("unknown model kind");
    if (lattice < KDQ_LATTICE_HC120 || lattice > KDQ_LATTICE_SQ45) return model_fail("unknown lattice");
    const bool honeycomb = lattice == KDQ_LATTICE_HC120 || lattice == KDQ_LATTICE_HC60;
    if (kind != KDQ_MODEL_TU && !honeycomb) return model_fail("t-V and t-U-V need a honeycomb lattice");
    if (p->site[0] < 1 || p->site[1] < 1 || p->batch_size < 1 || !(p->dt > 0.0))
        return model_fail("site, BatchSize and Δt must be positive");
    if (p->initial != KDQ_INITIAL_H0 && p->initial != KDQ_INITIAL_V) return model_fail("Initial must be H0 or V");
    const int site[2] = {p->site[0], p->site[1]};
    const int ns = 2 * site[0] * site[1];
    if (ns > 8192) return model_fail("lattice too large");
    const int np = ns / 2;

    auto *m = new kdq_model_s();
    std::string err;
    int nt;
    if (kind == KDQ_MODEL_TUV) {
        nt = 2 * (int)std::ceil(p->theta_relax / p->dt);
    } else {
        nt = (int)std::nearbyint(2.0 * (p->theta_relax + p->theta_quench) / p->dt);
    }
    if (nt < 2) {
        delete m;
        return model_fail("Nt < 2");
    }
    static const double s6 = std::sqrt(6.0);
    m->gamma = {1 + s6 / 3, 1 + s6 / 3, 1 - s6 / 3, 1 - s6 / 3};
    double a_tuv = 0.0;
    if (kind == KDQ_MODEL_TUV) {
        // src/tUV/model.jl:78-90
        const double z = 3.0, U = p->c1, Vc = p->c2;
        if (!(U * U > 4 * z * z * Vc * Vc)) {
            delete m;
            return model_fail("For stability, require U^2 > 4 z^2 V^2");
        }
        const double root = std::sqrt(U * U / z / z / 4 - Vc * Vc);
        const double g = U > 0 ? U / z / 2 + root : U / z / 2 - root;
        a_tuv = g == 0.0 ? 0.0 : Vc / g;
        const double sg = std::sqrt(std::fabs(p->dt * g));
        m->eta = {sg * std::sqrt(3 - s6), -sg * std::sqrt(3 - s6), sg * std::sqrt(3 + s6), -sg * std::sqrt(3 + s6)};
    } else {
        // src/tU/model.jl:41-56, src/tV/model.jl:78-92
        std::vector<double> prof;
        if (!coupling_profile(p->c1, p->c2, p->theta_relax, p->theta_quench, p->dt, nt, prof, err)) {
            delete m;
            return model_fail(err);
        }
        m->alpha.resize(nt);
        for (int l = 0; l < nt; ++l) m->alpha[l] = std::sqrt(p->dt * prof[l] / 2);
        m->eta = {std::sqrt(2 * (3 - s6)), -std::sqrt(2 * (3 - s6)), std::sqrt(2 * (3 + s6)), -std::sqrt(2 * (3 + s6))};
    }

    // kinetic propagators: t-U exponentiates Ht K, t-V and t-U-V -Ht K (src/tU/model.jl:61-66,
    // src/tV/model.jl:34-38, src/tUV/model.jl:33-37)
    m->K = nnK(lattice, site, ns);
    if (!symmetric(ns, m->K)) {   // src/public/Geometry.jl:205
        delete m;
        return model_fail("K is not Hermitian!");
    }
    Mat A = m->K;
    const double sgn = kind == KDQ_MODEL_TU ? p->ht : -p->ht;
    for (double &x : A) x *= sgn;
    std::vector<double> w;
    Mat V;
    if (!sym_eig(ns, A, w, V)) {
        delete m;
        return model_fail("eigensolver did not converge");
    }
    std::vector<double> f(ns);
    const double dt = p->dt;
    for (int i = 0; i < ns; ++i) f[i] = std::exp(-dt * w[i] / 2);
    m->hK = vfv(ns, V, f);
    for (int i = 0; i < ns; ++i) f[i] = std::exp(-dt * w[i]);
    m->eK = vfv(ns, V, f);
    for (int i = 0; i < ns; ++i) f[i] = std::exp(dt * w[i] / 2);
    m->hKinv = vfv(ns, V, f);
    for (int i = 0; i < ns; ++i) f[i] = std::exp(dt * w[i]);
    m->eKinv = vfv(ns, V, f);
    if (kind == KDQ_MODEL_TU) {
        // src/tU/model.jl:66
        Mat I = matmul(ns, ns, m->eK, m->eKinv);
        double r = 0.0;
        for (int j = 0; j < ns; ++j)
            for (int i = 0; i < ns; ++i) {
                const double t = I[(size_t)j * ns + i] - (i == j ? 1.0 : 0.0);
                r += t * t;
            }
        if (std::sqrt(r) >= 1e-10) {
            delete m;
            return model_fail("eK*eKinv does not equal identity!");
        }
    }

    // trial state: the lowest (t-U) or highest (t-V, t-U-V) half of the spectrum of K with a
    // staggered potential, or the product state; t-V and t-U-V then apply e^{+Δτ Ht K/2}
    // (src/tU/model.jl:68-79, src/tV/model.jl:40-76, src/tUV/model.jl:39-77)
    m->Pt.assign((size_t)ns * np, 0.0);
    if (p->initial == KDQ_INITIAL_H0) {
        const double mu = kind == KDQ_MODEL_TV ? 1e-3 : 1e-5;
        Mat KK = m->K;
        for (int i = 0; i < ns; ++i) KK[(size_t)i * ns + i] += (i % 2 == 0 ? -mu : mu);
        if (!sym_eig(ns, KK, w, V)) {
            delete m;
            return model_fail("eigensolver did not converge");
        }
        const int c0 = kind == KDQ_MODEL_TU ? 0 : np;
        std::copy(V.begin() + (size_t)c0 * ns, V.begin() + (size_t)(c0 + np) * ns, m->Pt.begin());
    } else {
        const int off = kind == KDQ_MODEL_TU ? 1 : 0;   // B sites (t-U) or A sites
        for (int i = 0; i < np; ++i) m->Pt[(size_t)i * ns + 2 * i + off] = 1.0;
    }
    if (kind != KDQ_MODEL_TU) m->Pt = matmul(ns, np, m->hKinv, m->Pt);

    m->nodes = make_nodes(nt, p->batch_size);
    int z = 0;
    if (kind != KDQ_MODEL_TU) {
        // nnidx_F (src/public/Geometry.jl:24-66): the A sites' bonds, family j = column; t-U-V
        // adds each bond reversed (src/tUV/model.jl:92-100). Column-major (x, y) pairs.
        z = 3;
        const int nb = kind == KDQ_MODEL_TV ? np : ns;
        m->bonds.assign((size_t)nb * z * 2, 0);
        int nn[4];
        for (int b = 0; b < np; ++b) {
            const int i = 2 * b + 1;
            nn2idx(lattice, site, i, nn);
            for (int j = 0; j < z; ++j) {
                if (kind == KDQ_MODEL_TV) {
                    int32_t *pr = &m->bonds[((size_t)j * nb + b) * 2];
                    pr[0] = i;
                    pr[1] = nn[j];
                } else {
                    int32_t *pr = &m->bonds[((size_t)j * nb + 2 * b) * 2];
                    pr[0] = i;
                    pr[1] = nn[j];
                    pr[2] = nn[j];
                    pr[3] = i;
                }
            }
        }
    }

    *desc = kdq_model_desc{};
    desc->kind = kind;
    desc->ns = ns;
    desc->nt = nt;
    desc->n_nodes = (int32_t)m->nodes.size();
    desc->nodes = m->nodes.data();
    desc->lattice = lattice;
    desc->site[0] = site[0];
    desc->site[1] = site[1];
    desc->ht = p->ht;
    desc->alpha = m->alpha.empty() ? nullptr : m->alpha.data();
    desc->eta = m->eta.data();
    desc->gamma = m->gamma.data();
    desc->K = m->K.data();
    desc->eK = m->eK.data();
    desc->eKinv = m->eKinv.data();
    desc->halfeK = m->hK.data();
    desc->halfeKinv = m->hKinv.data();
    desc->Pt = m->Pt.data();
    desc->z = z;
    desc->bonds = m->bonds.empty() ? nullptr : m->bonds.data();
    desc->a = a_tuv;
    desc->is_complex = kind == KDQ_MODEL_TUV && p->c1 > 0;
    desc->U = kind == KDQ_MODEL_TUV ? p->c1 : 0.0;
    desc->V = kind == KDQ_MODEL_TUV ? p->c2 : 0.0;
    *out = m;
    return KDQ_OK;
}

extern "C" int kdq_model_free(kdq_model m) {
    delete m;
    return KDQ_OK;
}
