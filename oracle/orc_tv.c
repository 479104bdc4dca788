/*
 * ORACLE — test infrastructure only (see orc.h).
 * Restatement of the spinless t-V projector DQMC of KAPDQMC (real FP64):
 *   src/tV/GreenMatrix.jl:1-119  (Initial_s, BM_F!, BMinv_F!, get_r!, WrapV!)
 *   src/tV/phy_update.jl:1-314   (phy_update, UpdatePhyLayer!, phy_measure)
 *   src/public/Gupdate.jl:5-11,139-157 (Gupdate! rank-2, get_G!)
 * The dense bond rotations UV[:,:,j] (src/tV/model.jl:103-110) are applied as dense
 * GEMMs exactly as the reference does.
 * phy_measure calls xy_i(Lattice, site, x, y), a method the reference does not define
 * (src/tV/phy_update.jl:274-275); the documented assumption is xy_i(site, x, y).
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* C = op(A) op(B) real, column-major */
static void dgemm(char ta, int m, int n, int k, const double *A, int lda, const double *B, int ldb,
                  double *C, int ldc) {
#pragma omp parallel for schedule(static) if ((long long)m * n * k > 200000)
    for (int j = 0; j < n; ++j) {
        double *cj = C + (size_t)j * ldc;
        for (int i = 0; i < m; ++i) cj[i] = 0.0;
        for (int p = 0; p < k; ++p) {
            double b = B[p + (size_t)j * ldb];
            if (b == 0.0) continue;
            if (ta == 'N') {
                const double *ap = A + (size_t)p * lda;
                for (int i = 0; i < m; ++i) cj[i] += ap[i] * b;
            } else {
                for (int i = 0; i < m; ++i) cj[i] += A[p + (size_t)i * lda] * b;
            }
        }
    }
}

/* real QR with explicit Q via the complex routine (imaginary parts stay zero) */
static void dqr_q(int m, int n, double *A, int lda) {
    zdouble *Z = (zdouble *)malloc(sizeof(zdouble) * (size_t)m * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) Z[IX(i, j, m)] = A[IX(i, j, lda)];
    orc_zqr_q(m, n, Z, m, NULL);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) A[IX(i, j, lda)] = creal(Z[IX(i, j, m)]);
    free(Z);
}

static void dget_G(int ns, int np, const double *BL, const double *BR, double *G) {
    zdouble *zl = (zdouble *)malloc(sizeof(zdouble) * (size_t)np * ns);
    zdouble *zr = (zdouble *)malloc(sizeof(zdouble) * (size_t)ns * np);
    zdouble *zg = (zdouble *)malloc(sizeof(zdouble) * (size_t)ns * ns);
    for (size_t q = 0; q < (size_t)np * ns; ++q) { zl[q] = BL[q]; zr[q] = BR[q]; }
    orc_get_G(ns, np, zl, np, zr, ns, zg, NULL);
    for (size_t q = 0; q < (size_t)ns * ns; ++q) G[q] = creal(zg[q]);
    free(zl); free(zr); free(zg);
}

orc_tv *orc_tv_new(int ns, int nt, int bs, int nn_nodes, const int64_t *nodes, int lattice,
                   const int site[2], double ht, const double *alpha, const double *eta,
                   const double *gamma, const int *bonds, const double *K, const double *eK,
                   const double *eKinv, const double *HalfeK, const double *HalfeKinv,
                   const double *Pt, uint64_t seed, uint32_t chain) {
    orc_tv *m = (orc_tv *)calloc(1, sizeof(orc_tv));
    m->ns = ns; m->np = ns / 2; m->nt = nt; m->bs = bs; m->nn_nodes = nn_nodes;
    m->lattice = lattice; m->site[0] = site[0]; m->site[1] = site[1]; m->ht = ht;
    m->nb = ns / 2;
    m->nodes = (int *)malloc(sizeof(int) * nn_nodes);
    for (int i = 0; i < nn_nodes; ++i) m->nodes[i] = (int)nodes[i];
    m->alpha = dalloc(nt);
    memcpy(m->alpha, alpha, sizeof(double) * nt);
    for (int i = 0; i < 4; ++i) { m->eta[i] = eta[i]; m->gamma[i] = gamma[i]; }
    size_t nn = (size_t)ns * ns, npp = (size_t)m->np;
    m->bx = (int *)malloc(sizeof(int) * 3 * m->nb);
    m->by = (int *)malloc(sizeof(int) * 3 * m->nb);
    for (int q = 0; q < 3 * m->nb; ++q) { m->bx[q] = bonds[2 * q] - 1; m->by[q] = bonds[2 * q + 1] - 1; }
    m->K = dalloc(nn); m->eK = dalloc(nn); m->eKinv = dalloc(nn);
    m->HalfeK = dalloc(nn); m->HalfeKinv = dalloc(nn); m->Pt = dalloc(ns * npp);
    memcpy(m->K, K, nn * 8); memcpy(m->eK, eK, nn * 8); memcpy(m->eKinv, eKinv, nn * 8);
    memcpy(m->HalfeK, HalfeK, nn * 8); memcpy(m->HalfeKinv, HalfeKinv, nn * 8);
    memcpy(m->Pt, Pt, ns * npp * 8);
    m->UV = dalloc(nn * 3);
    const double h = sqrt(2.0) / 2.0;   /* 2^0.5/2 */
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < m->nb; ++i) {
            int x = m->bx[i + j * m->nb], y = m->by[i + j * m->nb];
            double *U = m->UV + nn * j;
            U[IX(x, x, ns)] = U[IX(x, y, ns)] = U[IX(y, x, ns)] = -h;
            U[IX(y, y, ns)] = h;
        }
    m->s = (uint8_t *)calloc((size_t)m->nb * 3 * nt, 1);
    m->G = dalloc(nn);
    m->BLs = dalloc(npp * ns * nn_nodes);
    m->BRs = dalloc(ns * npp * nn_nodes);
    m->T = dalloc(nn); m->T2 = dalloc(nn); m->BM = dalloc(nn);
    m->tNn = dalloc(ns * npp); m->tnN = dalloc(npp * ns); m->tN = dalloc(ns);
    m->seed = seed; m->chain = chain;
    return m;
}

void orc_tv_free(orc_tv *m) {
    if (!m) return;
    free(m->nodes); free(m->alpha); free(m->bx); free(m->by); free(m->K); free(m->eK);
    free(m->eKinv); free(m->HalfeK); free(m->HalfeKinv); free(m->Pt); free(m->UV); free(m->s);
    free(m->G); free(m->BLs); free(m->BRs); free(m->T); free(m->T2); free(m->BM);
    free(m->tNn); free(m->tnN); free(m->tN);
    free(m);
}

#define SIDX(m, i, j, lt) ((size_t)(i) + (size_t)(j) * (m)->nb + (size_t)(lt) * 3 * (m)->nb)

/* Initial_s (src/tV/GreenMatrix.jl:1-10), Philox stream, element index i + j*nb */
void orc_tv_initial_s(orc_tv *m) {
    for (int lt = 0; lt < m->nt; ++lt)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < m->nb; ++i)
                m->s[SIDX(m, i, j, lt)] = (uint8_t)orc_init_value(m->seed, m->chain, 0, (uint32_t)lt,
                                                                 (uint32_t)(i + j * m->nb));
}

/* D of bond family j (0-based) at slice lt (1-based): x -> sign*α η, y -> -sign*α η, exp */
static void bond_diag(orc_tv *m, int j, int lt, double sign, double *D) {
    for (int i = 0; i < m->nb; ++i) {
        int x = m->bx[i + j * m->nb], y = m->by[i + j * m->nb];
        double v = m->alpha[lt - 1] * m->eta[m->s[SIDX(m, i, j, lt - 1)] - 1];
        D[x] = exp(sign * v);
        D[y] = exp(-sign * v);
    }
}

/* WrapV!(…, "B") (src/tV/GreenMatrix.jl:99-119): G <- UV D UV G UV D⁻¹ UV (dense). */
static void wrapV_B(orc_tv *m, double *G, const double *D, int j) {
    int n = m->ns;
    const double *U = m->UV + (size_t)n * n * j;
    dgemm('N', n, n, n, U, n, G, n, m->T, n);
    dgemm('N', n, n, n, m->T, n, U, n, G, n);
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) m->T[IX(r, c, n)] = D[r] * G[IX(r, c, n)] / D[c];
    dgemm('N', n, n, n, U, n, m->T, n, G, n);
    dgemm('N', n, n, n, G, n, U, n, m->T, n);
    memcpy(G, m->T, sizeof(double) * (size_t)n * n);
}

/* G <- eK G eKinv (forward=1) or eKinv G eK (forward=0) */
static void wrapK(orc_tv *m, double *G, int forward) {
    int n = m->ns;
    if (forward) {
        dgemm('N', n, n, n, G, n, m->eKinv, n, m->T, n);
        dgemm('N', n, n, n, m->eK, n, m->T, n, G, n);
    } else {
        dgemm('N', n, n, n, m->eKinv, n, G, n, m->T, n);
        dgemm('N', n, n, n, m->T, n, m->eK, n, G, n);
    }
}

/* BM_F! (src/tV/GreenMatrix.jl:12-38) */
void orc_tv_BM_F(orc_tv *m, int idx, double *BM) {
    int n = m->ns;
    double *X = m->T2;
    memset(X, 0, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) X[IX(i, i, n)] = 1.0;
    for (int lt = m->nodes[idx - 1] + 1; lt <= m->nodes[idx]; ++lt) {
        dgemm('N', n, n, n, m->eK, n, X, n, BM, n);
        for (int j = 2; j >= 0; --j) {
            bond_diag(m, j, lt, 1.0, m->tN);
            const double *U = m->UV + (size_t)n * n * j;
            dgemm('N', n, n, n, U, n, BM, n, X, n);
            for (int c = 0; c < n; ++c)
                for (int r = 0; r < n; ++r) BM[IX(r, c, n)] = m->tN[r] * X[IX(r, c, n)];
            dgemm('N', n, n, n, U, n, BM, n, X, n);
            memcpy(BM, X, sizeof(double) * (size_t)n * n);
        }
    }
}

/* UpdatePhyLayer! (src/tV/phy_update.jl:190-205) for bond family j (0-based) with
 * get_r! (src/tV/GreenMatrix.jl:76-89) and the rank-2 Gupdate! (Gupdate.jl:5-11). */
void orc_tv_update_family(orc_tv *m, int lt, int dir, int j) {
    int n = m->ns;
    double *G = m->G;
    const double h = sqrt(2.0) / 2.0;
    const double uv[2][2] = {{-h, -h}, {-h, h}};
    for (int i = 0; i < m->nb; ++i) {
        int x = m->bx[i + j * m->nb], y = m->by[i + j * m->nb];
        int sub[2] = {x, y};
        size_t si_ = SIDX(m, i, j, lt - 1);
        int si = m->s[si_];
        int choice; double u;
        orc_sweep_draw(&m->rp, m->nt, 3 * m->nb, m->seed, m->chain, m->sweep, lt, dir, (uint32_t)(i + j * m->nb), &choice, &u);
        int allowed[3], c = 0;
        for (int v = 1; v <= 4; ++v) if (v != si) allowed[c++] = v;
        int sx = allowed[choice];
        double ds = m->alpha[lt - 1] * (m->eta[sx - 1] - m->eta[si - 1]);
        double t2[2] = {exp(ds) - 1.0, exp(-ds) - 1.0};
        double Dl[2][2], r[2][2];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) Dl[a][b] = uv[a][0] * t2[0] * uv[0][b] + uv[a][1] * t2[1] * uv[1][b];
        /* r = I + Δ (I - G[sub,sub]) */
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                double acc = Dl[a][b];
                for (int q = 0; q < 2; ++q) acc -= Dl[a][q] * G[IX(sub[q], sub[b], n)];
                r[a][b] = acc + (a == b ? 1.0 : 0.0);
            }
        double p = r[0][0] * r[1][1] - r[0][1] * r[1][0];
        double prob = p * (m->gamma[sx - 1] / m->gamma[si - 1]);
        if (prob < -1e-3) m->n_negative++;
        m->n_propose++;
        if (fabs(u - prob) < ORC_NEAR_TIE) m->n_ties++;
        if (u < prob) {
            m->n_accept++;
            /* r <- inv(r) Δ */
            double ri[2][2] = {{r[1][1] / p, -r[0][1] / p}, {-r[1][0] / p, r[0][0] / p}}, w[2][2];
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) w[a][b] = ri[a][0] * Dl[0][b] + ri[a][1] * Dl[1][b];
            /* zN = w (E_sub - G[sub,:]), G -= G[:,sub] zN */
            double *z0 = m->tnN, *z1 = m->tnN + n;
            for (int col = 0; col < n; ++col) {
                double g0 = G[IX(x, col, n)], g1 = G[IX(y, col, n)];
                double e0 = (col == x) - g0, e1 = (col == y) - g1;
                z0[col] = w[0][0] * e0 + w[0][1] * e1;
                z1[col] = w[1][0] * e0 + w[1][1] * e1;
            }
            double *c0 = m->tNn, *c1 = m->tNn + n;
            for (int k = 0; k < n; ++k) { c0[k] = G[IX(k, x, n)]; c1[k] = G[IX(k, y, n)]; }
            for (int col = 0; col < n; ++col)
                for (int k = 0; k < n; ++k) G[IX(k, col, n)] -= c0[k] * z0[col] + c1[k] * z1[col];
            m->s[si_] = (uint8_t)sx;
        }
    }
}

/* phy_measure (src/tV/phy_update.jl:207-314): out = Ek, Ev, R0[4], R1[4] */
void orc_tv_measure(orc_tv *m, int lt, const double *Gin, double out[10]) {
    int n = m->ns, nt = m->nt;
    double *G0 = dalloc((size_t)n * n), *D = dalloc(n);
    memcpy(G0, Gin, sizeof(double) * (size_t)n * n);
    if (lt > nt / 2.0) {
        for (int t = lt; t >= nt / 2 + 1; --t) {
            for (int j = 0; j < 3; ++j) {
                bond_diag(m, j, t, -1.0, D);
                wrapV_B(m, G0, D, j);
            }
            wrapK(m, G0, 0);
        }
    } else {
        for (int t = lt + 1; t <= nt / 2; ++t) {
            wrapK(m, G0, 1);
            for (int j = 2; j >= 0; --j) {
                bond_diag(m, j, t, 1.0, D);
                wrapV_B(m, G0, D, j);
            }
        }
    }
    dgemm('N', n, n, n, m->HalfeK, n, G0, n, m->T, n);
    dgemm('N', n, n, n, m->T, n, m->HalfeKinv, n, G0, n);
    double ek = 0.0;
    for (size_t q = 0; q < (size_t)n * n; ++q) ek += m->K[q] * G0[q];
    double ev = 0.0;
    for (int q = 0; q < 3 * m->nb; ++q) {
        int x = m->bx[q], y = m->by[q];
        ev += (1 - G0[IX(x, x, n)]) * (1 - G0[IX(y, y, n)]) - G0[IX(x, y, n)] * G0[IX(y, x, n)];
    }
    double R0[4] = {0, 0, 0, 0}, R1[4] = {0, 0, 0, 0};
    int L1 = m->site[0], L2 = m->site[1];
#define g(a, b) G0[IX(a, b, n)]
    for (int rx = 1; rx <= L1; ++rx)
        for (int ry = 1; ry <= L2; ++ry) {
            double tmp[4] = {0, 0, 0, 0};
            for (int ix = 1; ix <= L1; ++ix)
                for (int iy = 1; iy <= L2; ++iy) {
                    int a = orc_xy_i(m->site, ix, iy) - 2;   /* 0-based A site */
                    int b = orc_xy_i(m->site, (ix + rx - 1) % L1 + 1, (iy + ry - 1) % L2 + 1) - 2;
                    if (a == b) {
                        tmp[0] += 1 - g(a, a);
                        tmp[1] += 1 - g(a + 1, a + 1);
                    } else {
                        tmp[0] += (1 - g(a, a)) * (1 - g(b, b)) - g(a, b) * g(b, a);
                        tmp[1] += (1 - g(a + 1, a + 1)) * (1 - g(b + 1, b + 1)) - g(a + 1, b + 1) * g(b + 1, a + 1);
                    }
                    tmp[2] += (1 - g(a + 1, a + 1)) * (1 - g(b, b)) - g(a + 1, b) * g(b, a + 1);
                    tmp[3] += (1 - g(a, a)) * (1 - g(b + 1, b + 1)) - g(a, b + 1) * g(b + 1, a);
                }
            double c = cos(2 * M_PI / L1 * rx + 2 * M_PI / L2 * ry);
            for (int q = 0; q < 4; ++q) { R0[q] += tmp[q]; R1[q] += c * tmp[q]; }
        }
#undef g
    double f = 4.0 / ((double)n * n);
    out[0] = m->ht * ek;
    out[1] = ev;
    for (int q = 0; q < 4; ++q) { out[2 + q] = f * R0[q]; out[6 + q] = f * R1[q]; }
    free(G0); free(D);
}

static int tv_is_node(const orc_tv *m, int t) {
    for (int i = 0; i < m->nn_nodes; ++i) if (m->nodes[i] == t) return 1;
    return 0;
}

#define TBL(k) (m->BLs + (size_t)((k) - 1) * np * n)
#define TBR(k) (m->BRs + (size_t)((k) - 1) * n * np)

static void rq_rows_real(int np, int n, const double *X, double *out) {
    double *XT = dalloc((size_t)n * np);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) XT[IX(j, i, n)] = X[IX(i, j, np)];
    dqr_q(n, np, XT, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) out[IX(i, j, np)] = XT[IX(j, i, n)];
    free(XT);
}

void orc_tv_init_chain(orc_tv *m) {
    int n = m->ns, np = m->np, NN = m->nn_nodes;
    memcpy(TBR(1), m->Pt, sizeof(double) * (size_t)n * np);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) TBL(NN)[IX(i, j, np)] = m->Pt[IX(j, i, n)];
    for (int idx = NN - 1; idx >= 1; --idx) {
        orc_tv_BM_F(m, idx, m->BM);
        dgemm('N', np, n, n, TBL(idx + 1), np, m->BM, n, m->tnN, np);
        rq_rows_real(np, n, m->tnN, TBL(idx));
    }
    m->idx = 1;
    dget_G(n, np, TBL(1), TBR(1), m->G);
}

/* phy_update (src/tV/phy_update.jl:1-188); rows: sweeps x 10 [Ek, Ev, R0[4], R1[4]] */
void orc_tv_phy_update(orc_tv *m, int sweeps, int record, double *rows) {
    int n = m->ns, np = m->np, NN = m->nn_nodes, nt = m->nt;
    int thidx = NN / 2 + 1;
    double *D = dalloc(n);
    orc_tv_init_chain(m);
    for (int sw = 0; sw < sweeps; ++sw) {
        double acc[10] = {0};
        int counter = 0;
        for (int lt = 1; lt <= nt; ++lt) {
            wrapK(m, m->G, 1);
            for (int j = 2; j >= 0; --j) {
                bond_diag(m, j, lt, 1.0, D);
                wrapV_B(m, m->G, D, j);
                orc_tv_update_family(m, lt, 0, j);
            }
            if (record && abs(m->idx - thidx) <= 1) {
                double o[10];
                orc_tv_measure(m, lt, m->G, o);
                for (int q = 0; q < 10; ++q) acc[q] += o[q];
                counter++;
            }
            if (tv_is_node(m, lt)) {
                m->idx += 1;
                int idx = m->idx;
                orc_tv_BM_F(m, idx - 1, m->BM);
                dgemm('N', n, np, n, m->BM, n, TBR(idx - 1), n, m->tNn, n);
                dqr_q(n, np, m->tNn, n);
                memcpy(TBR(idx), m->tNn, sizeof(double) * (size_t)n * np);
                dget_G(n, np, TBL(idx), TBR(idx), m->G);
            }
        }
        for (int lt = nt; lt >= 1; --lt) {
            for (int j = 0; j < 3; ++j) {
                orc_tv_update_family(m, lt, 1, j);
                bond_diag(m, j, lt, -1.0, D);
                wrapV_B(m, m->G, D, j);
            }
            wrapK(m, m->G, 0);
            if (record && abs(m->idx - thidx) <= 1) {
                double o[10];
                orc_tv_measure(m, lt - 1, m->G, o);
                for (int q = 0; q < 10; ++q) acc[q] += o[q];
                counter++;
            }
            if (tv_is_node(m, lt - 1)) {
                m->idx -= 1;
                int idx = m->idx;
                orc_tv_BM_F(m, idx, m->BM);
                dgemm('N', np, n, n, TBL(idx + 1), np, m->BM, n, m->tnN, np);
                rq_rows_real(np, n, m->tnN, TBL(idx));
                dget_G(n, np, TBL(idx), TBR(idx), m->G);
            }
        }
        if (record && rows)
            for (int q = 0; q < 10; ++q) rows[(size_t)sw * 10 + q] = counter ? acc[q] / counter : 0.0;
        m->sweep += 1;
    }
    free(D);
}

/* Equal-time G at slice tau from unstabilised-then-QR products (src/tV/GreenMatrix.jl:124-194
 * debug routine), used to pin the sweep. Valid for small Nt only. */
void orc_tv_Gtau(orc_tv *m, int tau, double *out) {
    int n = m->ns, np = m->np;
    double *R = dalloc((size_t)n * np), *L = dalloc((size_t)np * n), *T = dalloc((size_t)n * n);
    double *D = dalloc(n), *Rt = dalloc((size_t)n * np), *Lt = dalloc((size_t)np * n);
    memcpy(R, m->Pt, sizeof(double) * (size_t)n * np);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) L[IX(i, j, np)] = m->Pt[IX(j, i, n)];
    for (int lt = 1; lt <= tau; ++lt) {           /* R <- B(lt) R, B = Π_j (UV D UV) eK */
        dgemm('N', n, np, n, m->eK, n, R, n, Rt, n);
        for (int j = 2; j >= 0; --j) {
            bond_diag(m, j, lt, 1.0, D);
            const double *U = m->UV + (size_t)n * n * j;
            dgemm('N', n, np, n, U, n, Rt, n, R, n);
            for (int c = 0; c < np; ++c) for (int r = 0; r < n; ++r) R[IX(r, c, n)] *= D[r];
            dgemm('N', n, np, n, U, n, R, n, Rt, n);
        }
        memcpy(R, Rt, sizeof(double) * (size_t)n * np);
        dqr_q(n, np, R, n);
    }
    for (int lt = m->nt; lt > tau; --lt) {        /* L <- L B(lt) */
        for (int j = 0; j < 3; ++j) {
            bond_diag(m, j, lt, 1.0, D);
            const double *U = m->UV + (size_t)n * n * j;
            dgemm('N', np, n, n, L, np, U, n, Lt, np);
            for (int c = 0; c < n; ++c) for (int r = 0; r < np; ++r) Lt[IX(r, c, np)] *= D[c];
            dgemm('N', np, n, n, Lt, np, U, n, L, np);
        }
        dgemm('N', np, n, n, L, np, m->eK, n, Lt, np);
        rq_rows_real(np, n, Lt, L);
    }
    dget_G(n, np, L, R, out);
    free(R); free(L); free(T); free(D); free(Rt); free(Lt);
}

void orc_tv_set_replay(orc_tv *m, const uint8_t *choice, const double *u, uint32_t nsweeps) {
    m->rp.choice = nsweeps ? choice : NULL;
    m->rp.u = nsweeps ? u : NULL;
    m->rp.sweep0 = m->sweep;
    m->rp.nsweeps = nsweeps;
}
