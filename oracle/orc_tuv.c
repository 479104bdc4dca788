/*
 * ORACLE — test infrastructure only (see orc.h).
 * Restatement of the t-U-V projector DQMC of KAPDQMC:
 *   src/tUV/GreenMatrix.jl:1-122 (Initial_s, BM_F!, get_r!, WrapKV!)
 *   src/tUV/phy_update.jl:1-355  (phy_update, UpdatePhyLayer!, phy_measure)
 * The field lives on directed bonds s[Ns*z, Nt]; bond i = (x, y) adds η_s to site x and
 * a·η_s to site y of the HS diagonal; the diagonal enters as exp(i·) when U > 0
 * (model.Type = ComplexF64) and exp(·) when U < 0 (Float64). Both cases use complex
 * storage here (imaginary parts stay zero in the real case).
 * phy_measure calls xy_i(Lattice, site, x, y), a method the reference does not define
 * (src/tUV/phy_update.jl:292-293); the documented assumption is xy_i(site, x, y).
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

static zdouble *zal(size_t n) { return (zdouble *)calloc(n ? n : 1, sizeof(zdouble)); }

orc_tuv *orc_tuv_new(int ns, int nt, int bs, int nn_nodes, const int64_t *nodes, int lattice,
                     const int site[2], double ht, double U, double V, double a, int is_complex,
                     int z, const int *bonds, const double *eta, const double *gamma, const double *K,
                     const double *eK, const double *eKinv, const double *HalfeK,
                     const double *HalfeKinv, const double *Pt, uint64_t seed, uint32_t chain) {
    orc_tuv *m = (orc_tuv *)calloc(1, sizeof(orc_tuv));
    m->ns = ns; m->np = ns / 2; m->nt = nt; m->bs = bs; m->nn_nodes = nn_nodes;
    m->lattice = lattice; m->site[0] = site[0]; m->site[1] = site[1];
    m->ht = ht; m->U = U; m->V = V; m->a = a; m->is_complex = is_complex; m->nbd = ns * z;
    m->nodes = (int *)malloc(sizeof(int) * nn_nodes);
    for (int i = 0; i < nn_nodes; ++i) m->nodes[i] = (int)nodes[i];
    for (int i = 0; i < 4; ++i) { m->eta[i] = eta[i]; m->gamma[i] = gamma[i]; }
    m->bx = (int *)malloc(sizeof(int) * m->nbd);
    m->by = (int *)malloc(sizeof(int) * m->nbd);
    for (int q = 0; q < m->nbd; ++q) { m->bx[q] = bonds[2 * q] - 1; m->by[q] = bonds[2 * q + 1] - 1; }
    size_t nn = (size_t)ns * ns, npp = (size_t)m->np;
    m->K = zal(nn); m->eK = zal(nn); m->eKinv = zal(nn); m->HalfeK = zal(nn); m->HalfeKinv = zal(nn);
    m->Pt = zal(ns * npp);
    for (size_t q = 0; q < nn; ++q) {
        m->K[q] = K[q]; m->eK[q] = eK[q]; m->eKinv[q] = eKinv[q];
        m->HalfeK[q] = HalfeK[q]; m->HalfeKinv[q] = HalfeKinv[q];
    }
    for (size_t q = 0; q < ns * npp; ++q) m->Pt[q] = Pt[q];
    m->s = (uint8_t *)calloc((size_t)m->nbd * nt, 1);
    m->G = zal(nn); m->BLs = zal(npp * ns * nn_nodes); m->BRs = zal(ns * npp * nn_nodes);
    m->T = zal(nn); m->T2 = zal(nn); m->BM = zal(nn); m->tNn = zal(ns * npp); m->tnN = zal(npp * ns);
    m->tN = zal(ns);
    m->seed = seed; m->chain = chain;
    return m;
}

void orc_tuv_free(orc_tuv *m) {
    if (!m) return;
    free(m->nodes); free(m->bx); free(m->by); free(m->K); free(m->eK); free(m->eKinv);
    free(m->HalfeK); free(m->HalfeKinv); free(m->Pt); free(m->s); free(m->G); free(m->BLs);
    free(m->BRs); free(m->T); free(m->T2); free(m->BM); free(m->tNn); free(m->tnN); free(m->tN);
    free(m);
}

void orc_tuv_initial_s(orc_tuv *m) {
    for (int lt = 0; lt < m->nt; ++lt)
        for (int i = 0; i < m->nbd; ++i)
            m->s[IX(i, lt, m->nbd)] = (uint8_t)orc_init_value(m->seed, m->chain, 0, (uint32_t)lt, (uint32_t)i);
}

static zdouble ex(const orc_tuv *m, double x) { return m->is_complex ? cos(x) + I * sin(x) : exp(x); }

/* HS diagonal of slice lt (1-based) with sign ±1 (src/tUV/phy_update.jl:68-78) */
static void tuv_diag(orc_tuv *m, int lt, double sign, zdouble *D) {
    double *acc = (double *)calloc(m->ns, sizeof(double));
    for (int i = 0; i < m->nbd; ++i) {
        double e = m->eta[m->s[IX(i, lt - 1, m->nbd)] - 1];
        acc[m->bx[i]] += e;
        acc[m->by[i]] += m->a * e;
    }
    for (int x = 0; x < m->ns; ++x) D[x] = ex(m, sign * acc[x]);
    free(acc);
}

static void wrap_B(orc_tuv *m, zdouble *G, const zdouble *D, int forward) {
    int n = m->ns;
    zdouble *T = m->T;
    if (forward) {
        orc_zgemm('N', 'N', n, n, n, 1.0, m->eK, n, G, n, 0.0, T, n);
        orc_zgemm('N', 'N', n, n, n, 1.0, T, n, m->eKinv, n, 0.0, G, n);
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) G[IX(r, c, n)] *= D[r] / D[c];
    } else {
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) G[IX(r, c, n)] *= D[r] / D[c];
        orc_zgemm('N', 'N', n, n, n, 1.0, m->eKinv, n, G, n, 0.0, T, n);
        orc_zgemm('N', 'N', n, n, n, 1.0, T, n, m->eK, n, 0.0, G, n);
    }
}

/* BM_F! (src/tUV/GreenMatrix.jl:10-36) */
void orc_tuv_BM_F(orc_tuv *m, int idx, zdouble *BM) {
    int n = m->ns;
    memset(BM, 0, sizeof(zdouble) * (size_t)n * n);
    for (int i = 0; i < n; ++i) BM[IX(i, i, n)] = 1;
    for (int lt = m->nodes[idx - 1] + 1; lt <= m->nodes[idx]; ++lt) {
        orc_zgemm('N', 'N', n, n, n, 1.0, m->eK, n, BM, n, 0.0, m->T2, n);
        tuv_diag(m, lt, 1.0, m->tN);
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) BM[IX(r, c, n)] = m->tN[r] * m->T2[IX(r, c, n)];
    }
}

/* UpdatePhyLayer! (src/tUV/phy_update.jl:188-203) with get_r! (GreenMatrix.jl:73-89) */
void orc_tuv_update_layer(orc_tuv *m, int lt, int dir) {
    int n = m->ns;
    zdouble *G = m->G;
    for (int i = 0; i < m->nbd; ++i) {
        int x = m->bx[i], y = m->by[i];
        int sub[2] = {x, y};
        size_t sidx = IX(i, lt - 1, m->nbd);
        int si = m->s[sidx];
        int choice; double u;
        orc_sweep_draw(&m->rp, m->nt, m->nbd, m->seed, m->chain, m->sweep, lt, dir, (uint32_t)i, &choice, &u);
        int allowed[3], c = 0;
        for (int v = 1; v <= 4; ++v) if (v != si) allowed[c++] = v;
        int sx = allowed[choice];
        double ds = m->eta[sx - 1] - m->eta[si - 1];
        zdouble Dl[2] = {ex(m, ds) - 1.0, ex(m, m->a * ds) - 1.0};
        zdouble r[2][2];
        for (int p = 0; p < 2; ++p)
            for (int q = 0; q < 2; ++q)
                r[p][q] = (p == q ? 1.0 : 0.0) + Dl[p] * ((p == q ? 1.0 : 0.0) - G[IX(sub[p], sub[q], n)]);
        zdouble det = r[0][0] * r[1][1] - r[0][1] * r[1][0];
        double prob = creal(det * conj(det)) * (m->gamma[sx - 1] / m->gamma[si - 1]);
        if (!m->is_complex) prob *= exp((m->eta[si - 1] - m->eta[sx - 1]) * (m->a + 1));
        m->n_propose++;
        if (fabs(u - prob) < ORC_NEAR_TIE) m->n_ties++;
        if (u < prob) {
            m->n_accept++;
            /* w = r^{-1} Δ  (inv22!) */
            zdouble ri[2][2] = {{r[1][1] / det, -r[0][1] / det}, {-r[1][0] / det, r[0][0] / det}}, w[2][2];
            for (int p = 0; p < 2; ++p)
                for (int q = 0; q < 2; ++q) w[p][q] = ri[p][q] * Dl[q];
            zdouble *z0 = m->tnN, *z1 = m->tnN + n;
            for (int col = 0; col < n; ++col) {
                zdouble e0 = (double)(col == x) - G[IX(x, col, n)], e1 = (double)(col == y) - G[IX(y, col, n)];
                z0[col] = w[0][0] * e0 + w[0][1] * e1;
                z1[col] = w[1][0] * e0 + w[1][1] * e1;
            }
            zdouble *c0 = m->tNn, *c1 = m->tNn + n;
            for (int k = 0; k < n; ++k) { c0[k] = G[IX(k, x, n)]; c1[k] = G[IX(k, y, n)]; }
            for (int col = 0; col < n; ++col)
                for (int k = 0; k < n; ++k) G[IX(k, col, n)] -= c0[k] * z0[col] + c1[k] * z1[col];
            m->s[sidx] = (uint8_t)sx;
        }
    }
}

/* phy_measure (src/tUV/phy_update.jl:205-355): Ek, Eu, Ev, CDW0, CDW1, SDW0, SDW1
 * (complex accumulators; the caller keeps real parts as the reference does at :170-173). */
void orc_tuv_measure(orc_tuv *m, int lt, const zdouble *Gin, double out[7]) {
    int n = m->ns, nt = m->nt;
    zdouble *G0 = zal((size_t)n * n), *D = zal(n);
    memcpy(G0, Gin, sizeof(zdouble) * (size_t)n * n);
    if (lt > nt / 2.0) {
        for (int t = lt; t >= nt / 2 + 1; --t) { tuv_diag(m, t, -1.0, D); wrap_B(m, G0, D, 0); }
    } else {
        for (int t = lt + 1; t <= nt / 2; ++t) { tuv_diag(m, t, 1.0, D); wrap_B(m, G0, D, 1); }
    }
    orc_zgemm('N', 'N', n, n, n, 1.0, m->HalfeK, n, G0, n, 0.0, m->T, n);
    orc_zgemm('N', 'N', n, n, n, 1.0, m->T, n, m->HalfeKinv, n, 0.0, G0, n);
#define g(p, q) G0[IX(p, q, n)]
    zdouble ks = 0;
    for (size_t q = 0; q < (size_t)n * n; ++q) ks += m->K[q] * G0[q];
    double Ek = 2 * m->ht * creal(ks);
    zdouble eu = 0;
    for (int i = 0; i < n; ++i)
        eu += m->is_complex ? (1.0 - g(i, i)) * conj(g(i, i)) : (1.0 - g(i, i)) * (1.0 - g(i, i));
    zdouble ev = 0;
    for (int k = 0; k < m->nbd; ++k) {
        int x = m->bx[k], y = m->by[k];
        if (m->is_complex) {
            ev += (1.0 - g(x, x)) * (1.0 - g(y, y)) - g(x, y) * g(y, x);
            ev += conj(g(x, x) * g(y, y) - g(x, y) * g(y, x));
            ev += (1.0 - g(y, y)) * conj(g(x, x)) + (1.0 - g(x, x)) * conj(g(y, y));
        } else {
            ev += 2.0 * ((1.0 - g(x, x)) * (1.0 - g(y, y)) - g(x, y) * g(y, x));
            ev += 2.0 * (1.0 - g(y, y)) * (1.0 - g(x, x));
        }
    }
    zdouble C0 = 0, C1 = 0, S0 = 0, S1 = 0;
    int L1 = m->site[0], L2 = m->site[1];
    for (int rx = 1; rx <= L1; ++rx)
        for (int ry = 1; ry <= L2; ++ry) {
            zdouble t1 = 0, t2 = 0;
            for (int ix = 1; ix <= L1; ++ix)
                for (int iy = 1; iy <= L2; ++iy) {
                    int a = orc_xy_i(m->site, ix, iy) - 2;
                    int b = orc_xy_i(m->site, (ix + rx - 1) % L1 + 1, (iy + ry - 1) % L2 + 1) - 2;
                    double d = (a == b) ? 1.0 : 0.0;
                    if (!m->is_complex) {
                        t1 += 2.0 * ((1.0 - g(a, a)) * (1.0 - g(b, b)) + (d - g(a, b)) * g(b, a));
                        t1 += 2.0 * ((1.0 - g(a + 1, a + 1)) * (1.0 - g(b + 1, b + 1)) + (d - g(a + 1, b + 1)) * g(b + 1, a + 1));
                        t1 -= 2.0 * ((1.0 - g(a + 1, a + 1)) * (1.0 - g(b, b)) - g(a + 1, b) * g(b, a + 1));
                        t1 -= 2.0 * ((1.0 - g(a, a)) * (1.0 - g(b + 1, b + 1)) - g(a, b + 1) * g(b + 1, a));
                        t2 += 2.0 * (1.0 - g(a, a)) * (1.0 - g(b, b));
                        t2 += 2.0 * (1.0 - g(a + 1, a + 1)) * (1.0 - g(b + 1, b + 1));
                        t2 -= 2.0 * (1.0 - g(a + 1, a + 1)) * (1.0 - g(b, b));
                        t2 -= 2.0 * (1.0 - g(a, a)) * (1.0 - g(b + 1, b + 1));
                    } else {
                        /* only the first expression of each `+= ... / +adjoint(...)` pair accumulates */
                        t1 += (1.0 - g(a, a)) * (1.0 - g(b, b)) + (d - g(a, b)) * g(b, a);
                        t1 += (1.0 - g(a + 1, a + 1)) * (1.0 - g(b + 1, b + 1)) + (d - g(a + 1, b + 1)) * g(b + 1, a + 1);
                        t1 -= (1.0 - g(a + 1, a + 1)) * (1.0 - g(b, b)) - g(a + 1, b) * g(b, a + 1);
                        t1 -= (1.0 - g(a, a)) * (1.0 - g(b + 1, b + 1)) - g(a, b + 1) * g(b + 1, a);
                        t2 += (1.0 - g(a, a)) * conj(g(b, b));
                        t2 += (1.0 - g(a + 1, a + 1)) * conj(g(b + 1, b + 1));
                        t2 -= (1.0 - g(a, a)) * conj(g(b + 1, b + 1));
                        t2 -= (1.0 - g(a + 1, a + 1)) * conj(g(b, b));
                    }
                }
            double cs = cos(2 * M_PI / L1 * rx + 2 * M_PI / L2 * ry);
            C0 += t1 + t2; C1 += cs * (t1 + t2); S0 += t1 - t2; S1 += cs * (t1 - t2);
        }
#undef g
    double n2 = (double)n * n;
    out[0] = Ek; out[1] = m->U * creal(eu); out[2] = m->V * creal(ev);
    out[3] = creal(C0) / n2; out[4] = creal(C1) / n2; out[5] = creal(S0) / n2; out[6] = creal(S1) / n2;
    free(G0); free(D);
}

static int tuv_is_node(const orc_tuv *m, int t) {
    for (int i = 0; i < m->nn_nodes; ++i) if (m->nodes[i] == t) return 1;
    return 0;
}

#define UBL(k) (m->BLs + (size_t)((k) - 1) * np * n)
#define UBR(k) (m->BRs + (size_t)((k) - 1) * n * np)

static void rq_rows_c(int np, int n, const zdouble *X, zdouble *out) {
    zdouble *XH = zal((size_t)n * np);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) XH[IX(j, i, n)] = conj(X[IX(i, j, np)]);
    orc_zqr_q(n, np, XH, n, NULL);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) out[IX(i, j, np)] = conj(XH[IX(j, i, n)]);
    free(XH);
}

void orc_tuv_init_chain(orc_tuv *m) {
    int n = m->ns, np = m->np, NN = m->nn_nodes;
    memcpy(UBR(1), m->Pt, sizeof(zdouble) * (size_t)n * np);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) UBL(NN)[IX(i, j, np)] = conj(m->Pt[IX(j, i, n)]);
    for (int idx = NN - 1; idx >= 1; --idx) {
        orc_tuv_BM_F(m, idx, m->BM);
        orc_zgemm('N', 'N', np, n, n, 1.0, UBL(idx + 1), np, m->BM, n, 0.0, m->tnN, np);
        rq_rows_c(np, n, m->tnN, UBL(idx));
    }
    m->idx = 1;
    orc_get_G(n, np, UBL(1), np, UBR(1), n, m->G, NULL);
}

/* phy_update (src/tUV/phy_update.jl:1-186); rows: sweeps x 7 */
void orc_tuv_phy_update(orc_tuv *m, int sweeps, int record, double *rows) {
    int n = m->ns, np = m->np, NN = m->nn_nodes, nt = m->nt;
    int thidx = NN / 2 + 1;
    zdouble *D = zal(n);
    orc_tuv_init_chain(m);
    for (int sw = 0; sw < sweeps; ++sw) {
        double acc[7] = {0};
        int counter = 0;
        for (int lt = 1; lt <= nt; ++lt) {
            tuv_diag(m, lt, 1.0, D);
            wrap_B(m, m->G, D, 1);
            orc_tuv_update_layer(m, lt, 0);
            if (record && abs(m->idx - thidx) <= 1) {
                double o[7];
                orc_tuv_measure(m, lt, m->G, o);
                for (int q = 0; q < 7; ++q) acc[q] += o[q];
                counter++;
            }
            if (tuv_is_node(m, lt)) {
                m->idx += 1;
                int idx = m->idx;
                orc_tuv_BM_F(m, idx - 1, m->BM);
                orc_zgemm('N', 'N', n, np, n, 1.0, m->BM, n, UBR(idx - 1), n, 0.0, m->tNn, n);
                orc_zqr_q(n, np, m->tNn, n, NULL);
                memcpy(UBR(idx), m->tNn, sizeof(zdouble) * (size_t)n * np);
                orc_get_G(n, np, UBL(idx), np, UBR(idx), n, m->G, NULL);
            }
        }
        for (int lt = nt; lt >= 1; --lt) {
            orc_tuv_update_layer(m, lt, 1);
            if (record && abs(m->idx - thidx) <= 1) {
                double o[7];
                orc_tuv_measure(m, lt, m->G, o);
                for (int q = 0; q < 7; ++q) acc[q] += o[q];
                counter++;
            }
            tuv_diag(m, lt, -1.0, D);
            wrap_B(m, m->G, D, 0);
            if (tuv_is_node(m, lt - 1)) {
                m->idx -= 1;
                int idx = m->idx;
                orc_tuv_BM_F(m, idx, m->BM);
                orc_zgemm('N', 'N', np, n, n, 1.0, UBL(idx + 1), np, m->BM, n, 0.0, m->tnN, np);
                rq_rows_c(np, n, m->tnN, UBL(idx));
                orc_get_G(n, np, UBL(idx), np, UBR(idx), n, m->G, NULL);
            }
        }
        if (record && rows)
            for (int q = 0; q < 7; ++q) rows[(size_t)sw * 7 + q] = counter ? acc[q] / counter : 0.0;
        m->sweep += 1;
    }
    free(D);
}

void orc_tuv_set_replay(orc_tuv *m, const uint8_t *choice, const double *u, uint32_t nsweeps) {
    m->rp.choice = nsweeps ? choice : NULL;
    m->rp.u = nsweeps ? u : NULL;
    m->rp.sweep0 = m->sweep;
    m->rp.nsweeps = nsweeps;
}
