/*
 * ORACLE — test infrastructure only (see orc.h).
 * Restatement of the t-U projector DQMC of KAPDQMC:
 *   src/tU/GreenMatrix.jl:1-86   (Initial_s, BM_F!, BMinv_F!, get_r!, WrapKV!)
 *   src/tU/GreenMatrix.jl:91-123 (brute-force Gtau, debug cross-check)
 *   src/tU/phy_update.jl:3-266   (phy_update, phy_measure, UpdatePhyLayer!)
 *   src/public/Gupdate.jl:5-11,139-157 (Gupdate!, get_G!)
 * The model constants (eK, eKinv, HalfeK, HalfeKinv, Pt, alpha, eta, gamma, nodes) are
 * inputs, as in the reference where tU_Hubbard_Para (src/tU/model.jl:39-96) builds them.
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

static zdouble *zalloc(size_t n) { return (zdouble *)calloc(n ? n : 1, sizeof(zdouble)); }

orc_tu *orc_tu_new(int ns, int nt, int bs, int nn_nodes, const int64_t *nodes, int lattice,
                   const int site[2], double ht, double dt, const double *alpha,
                   const double *eta, const double *gamma, const double *K, const double *eK,
                   const double *eKinv, const double *HalfeK, const double *HalfeKinv,
                   const double *Pt, uint64_t seed, uint32_t chain) {
    orc_tu *m = (orc_tu *)calloc(1, sizeof(orc_tu));
    m->ns = ns; m->np = ns / 2; m->nt = nt; m->bs = bs; m->nn_nodes = nn_nodes;
    m->lattice = lattice; m->site[0] = site[0]; m->site[1] = site[1];
    m->ht = ht; m->dt = dt;
    m->nodes = (int *)malloc(sizeof(int) * nn_nodes);
    for (int i = 0; i < nn_nodes; ++i) m->nodes[i] = (int)nodes[i];
    m->alpha = (double *)malloc(sizeof(double) * nt);
    memcpy(m->alpha, alpha, sizeof(double) * nt);
    for (int i = 0; i < 4; ++i) { m->eta[i] = eta[i]; m->gamma[i] = gamma[i]; }
    size_t nn = (size_t)ns * ns, np = (size_t)m->np;
    m->K = zalloc(nn); m->eK = zalloc(nn); m->eKinv = zalloc(nn);
    m->HalfeK = zalloc(nn); m->HalfeKinv = zalloc(nn); m->Pt = zalloc(ns * np);
    for (size_t i = 0; i < nn; ++i) {
        m->K[i] = K[i]; m->eK[i] = eK[i]; m->eKinv[i] = eKinv[i];
        m->HalfeK[i] = HalfeK[i]; m->HalfeKinv[i] = HalfeKinv[i];
    }
    for (size_t i = 0; i < ns * np; ++i) m->Pt[i] = Pt[i];
    m->s = (uint8_t *)calloc((size_t)ns * nt, 1);
    m->G = zalloc(nn);
    m->BLs = zalloc(np * ns * nn_nodes);
    m->BRs = zalloc(ns * np * nn_nodes);
    m->tmpNN = zalloc(nn); m->tmpNN2 = zalloc(nn); m->BM = zalloc(nn);
    m->tmpNn = zalloc(ns * np); m->tmpnN = zalloc(np * ns); m->tmpN = zalloc(ns);
    m->seed = seed; m->chain = chain; m->sweep = 0;
    return m;
}

void orc_tu_free(orc_tu *m) {
    if (!m) return;
    free(m->nodes); free(m->alpha); free(m->K); free(m->eK); free(m->eKinv);
    free(m->HalfeK); free(m->HalfeKinv); free(m->Pt); free(m->s); free(m->G);
    free(m->BLs); free(m->BRs); free(m->tmpNN); free(m->tmpNN2); free(m->BM);
    free(m->tmpNn); free(m->tmpnN); free(m->tmpN);
    free(m);
}

/* Initial_s (GreenMatrix.jl:1-8): i.i.d. uniform {1..4}; Philox stream instead of MT. */
void orc_tu_initial_s(orc_tu *m) {
    for (int lt = 0; lt < m->nt; ++lt)
        for (int i = 0; i < m->ns; ++i)
            m->s[IX(i, lt, m->ns)] = (uint8_t)orc_init_value(m->seed, m->chain, 0, (uint32_t)lt, (uint32_t)i);
}

static inline zdouble cis(double x) { return cos(x) + I * sin(x); }

/* diag(D) * A in place (rows scaled) */
static void zdiag_left(int n, int ncol, const zdouble *D, zdouble *A, int lda) {
    for (int j = 0; j < ncol; ++j)
        for (int i = 0; i < n; ++i) A[IX(i, j, lda)] *= D[i];
}
/* A * diag(D) in place (columns scaled) */
static void zdiag_right(int nrow, int n, const zdouble *D, zdouble *A, int lda) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < nrow; ++i) A[IX(i, j, lda)] *= D[j];
}

/* BM_F! (GreenMatrix.jl:10-27): BM = prod_{lt in block} diag(cis(alpha eta_s)) eK,
 * slices nodes[idx]+1 .. nodes[idx+1] applied left-most last. idx is 1-based. */
void orc_tu_BM_F(orc_tu *m, int idx, zdouble *BM) {
    int ns = m->ns;
    memset(BM, 0, sizeof(zdouble) * (size_t)ns * ns);
    for (int i = 0; i < ns; ++i) BM[IX(i, i, ns)] = 1;
    for (int lt = m->nodes[idx - 1] + 1; lt <= m->nodes[idx]; ++lt) {
        for (int i = 0; i < ns; ++i) m->tmpN[i] = cis(m->alpha[lt - 1] * m->eta[m->s[IX(i, lt - 1, ns)] - 1]);
        orc_zgemm('N', 'N', ns, ns, ns, 1.0, m->eK, ns, BM, ns, 0.0, m->tmpNN2, ns);
        zdiag_left(ns, ns, m->tmpN, m->tmpNN2, ns);
        memcpy(BM, m->tmpNN2, sizeof(zdouble) * (size_t)ns * ns);
    }
}

/* BMinv_F! (GreenMatrix.jl:29-47) */
void orc_tu_BMinv_F(orc_tu *m, int idx, zdouble *BM) {
    int ns = m->ns;
    memset(BM, 0, sizeof(zdouble) * (size_t)ns * ns);
    for (int i = 0; i < ns; ++i) BM[IX(i, i, ns)] = 1;
    for (int lt = m->nodes[idx - 1] + 1; lt <= m->nodes[idx]; ++lt) {
        for (int i = 0; i < ns; ++i) m->tmpN[i] = cis(-m->alpha[lt - 1] * m->eta[m->s[IX(i, lt - 1, ns)] - 1]);
        orc_zgemm('N', 'N', ns, ns, ns, 1.0, BM, ns, m->eKinv, ns, 0.0, m->tmpNN2, ns);
        zdiag_right(ns, ns, m->tmpN, m->tmpNN2, ns);
        memcpy(BM, m->tmpNN2, sizeof(zdouble) * (size_t)ns * ns);
    }
}

/* WrapKV! (GreenMatrix.jl:56-86). D_in is the phase vector as passed by the caller;
 * the "B" variants use D and 1/D (the reference mutates D to 1/D; we do not). */
void orc_tu_WrapKV(orc_tu *m, zdouble *G, const zdouble *D, int forward, char LR) {
    int n = m->ns;
    zdouble *T = m->tmpNN;
    zdouble *Dinv = zalloc(n);
    for (int i = 0; i < n; ++i) Dinv[i] = 1.0 / D[i];
    if (forward) {
        if (LR == 'L') {
            orc_zgemm('N', 'N', n, n, n, 1.0, m->eK, n, G, n, 0.0, T, n);
            zdiag_left(n, n, D, T, n); memcpy(G, T, sizeof(zdouble) * (size_t)n * n);
        } else if (LR == 'R') {
            orc_zgemm('N', 'N', n, n, n, 1.0, G, n, m->eKinv, n, 0.0, T, n);
            zdiag_right(n, n, D, T, n); memcpy(G, T, sizeof(zdouble) * (size_t)n * n);
        } else {
            orc_zgemm('N', 'N', n, n, n, 1.0, m->eK, n, G, n, 0.0, T, n);
            orc_zgemm('N', 'N', n, n, n, 1.0, T, n, m->eKinv, n, 0.0, G, n);
            zdiag_left(n, n, D, G, n);
            zdiag_right(n, n, Dinv, G, n);
        }
    } else {
        if (LR == 'L') {
            memcpy(T, G, sizeof(zdouble) * (size_t)n * n); zdiag_left(n, n, D, T, n);
            orc_zgemm('N', 'N', n, n, n, 1.0, m->eKinv, n, T, n, 0.0, G, n);
        } else if (LR == 'R') {
            memcpy(T, G, sizeof(zdouble) * (size_t)n * n); zdiag_right(n, n, D, T, n);
            orc_zgemm('N', 'N', n, n, n, 1.0, T, n, m->eK, n, 0.0, G, n);
        } else {
            zdiag_left(n, n, D, G, n);
            zdiag_right(n, n, Dinv, G, n);
            orc_zgemm('N', 'N', n, n, n, 1.0, m->eKinv, n, G, n, 0.0, T, n);
            orc_zgemm('N', 'N', n, n, n, 1.0, T, n, m->eK, n, 0.0, G, n);
        }
    }
    free(Dinv);
}

/* UpdatePhyLayer! (phy_update.jl:256-266) with get_r! (GreenMatrix.jl:49-54) and
 * Gupdate! (Gupdate.jl:5-11). Draws come from the shared Philox stream. */
void orc_tu_update_layer(orc_tu *m, int lt, int dir) {
    int n = m->ns;
    zdouble *G = m->G;
    zdouble *zN = m->tmpN;
    for (int i = 0; i < n; ++i) {
        int si = m->s[IX(i, lt - 1, n)];
        int choice; double u;
        orc_sweep_draw(&m->rp, m->nt, m->ns, m->seed, m->chain, m->sweep, lt, dir, (uint32_t)i, &choice, &u);
        /* Random.Sampler over the other three values, increasing order */
        int allowed[3], c = 0;
        for (int v = 1; v <= 4; ++v) if (v != si) allowed[c++] = v;
        int sx = allowed[choice];
        double ds = m->alpha[lt - 1] * (m->eta[sx - 1] - m->eta[si - 1]);
        zdouble Delta = cis(ds) - 1.0;
        zdouble p = 1.0 + Delta * (1.0 - G[IX(i, i, n)]);
        zdouble r = Delta / p;
        double prob = creal(p * conj(p)) * (m->gamma[sx - 1] / m->gamma[si - 1]);
        m->n_propose++;
        if (fabs(u - prob) < ORC_NEAR_TIE) m->n_ties++;
        if (u < prob) {
            m->n_accept++;
            /* zN = r * (e_i - G[i,:]); G -= G[:,i] * zN */
            for (int j = 0; j < n; ++j) zN[j] = -r * G[IX(i, j, n)];
            zN[i] += r;
            /* copy column i first: it is overwritten during the update */
            zdouble *col = m->tmpNn; /* size >= n */
            for (int k = 0; k < n; ++k) col[k] = G[IX(k, i, n)];
            for (int j = 0; j < n; ++j) {
                zdouble z = zN[j];
                if (z == 0) continue;
                for (int k = 0; k < n; ++k) G[IX(k, j, n)] -= col[k] * z;
            }
            m->s[IX(i, lt - 1, n)] = (uint8_t)sx;
        }
    }
}

/* phy_measure (phy_update.jl:154-254), returns (Ek, Eu, CDW0, CDW1, SDW0, SDW1).
 * The lines starting with `+adjoint(` in the reference are separate statements whose
 * values are discarded; only the first expression of each pair accumulates. */
void orc_tu_measure(orc_tu *m, int lt, const zdouble *Gin, double out[6]) {
    int n = m->ns, nt = m->nt;
    zdouble *G0 = zalloc((size_t)n * n);
    zdouble *D = zalloc(n);
    memcpy(G0, Gin, sizeof(zdouble) * (size_t)n * n);
    if (lt > nt / 2.0) {
        for (int t = lt; t >= nt / 2 + 1; --t) {
            for (int i = 0; i < n; ++i) D[i] = cis(-m->alpha[t - 1] * m->eta[m->s[IX(i, t - 1, n)] - 1]);
            orc_tu_WrapKV(m, G0, D, 0, 'B');
        }
    } else {
        for (int t = lt + 1; t <= nt / 2; ++t) {
            for (int i = 0; i < n; ++i) D[i] = cis(m->alpha[t - 1] * m->eta[m->s[IX(i, t - 1, n)] - 1]);
            orc_tu_WrapKV(m, G0, D, 1, 'B');
        }
    }
    zdouble *T = zalloc((size_t)n * n);
    orc_zgemm('N', 'N', n, n, n, 1.0, m->HalfeK, n, G0, n, 0.0, T, n);
    orc_zgemm('N', 'N', n, n, n, 1.0, T, n, m->HalfeKinv, n, 0.0, G0, n);
    free(T);

    zdouble ksum = 0;
    for (size_t q = 0; q < (size_t)n * n; ++q) ksum += m->K[q] * G0[q];
    double Ek = -2.0 * m->ht * creal(ksum);
    zdouble eu = 0;
    for (int i = 0; i < n; ++i) eu += (1.0 - G0[IX(i, i, n)]) * conj(G0[IX(i, i, n)]);
    double Eu = creal(eu);

    zdouble CDW0 = 0, CDW1 = 0, SDW0 = 0, SDW1 = 0;
    if (m->lattice == ORC_HC120 || m->lattice == ORC_HC60) {
        int L1 = m->site[0], L2 = m->site[1];
#define g(a, b) G0[IX((a) - 1, (b) - 1, n)]
        for (int rx = 1; rx <= L1; ++rx)
            for (int ry = 1; ry <= L2; ++ry) {
                zdouble tmp1 = 0, tmp2 = 0;
                for (int ix = 1; ix <= L1; ++ix)
                    for (int iy = 1; iy <= L2; ++iy) {
                        int a = orc_xy_i(m->site, ix, iy) - 1;
                        int b = orc_xy_i(m->site, (ix + rx - 1) % L1 + 1, (iy + ry - 1) % L2 + 1) - 1;
                        double delta = (a == b) ? 1.0 : 0.0;
                        tmp1 += (1.0 - g(a, a)) * (1.0 - g(b, b)) + (delta - g(a, b)) * g(b, a);
                        tmp1 += (1.0 - g(a + 1, a + 1)) * (1.0 - g(b + 1, b + 1)) + (delta - g(a + 1, b + 1)) * g(b + 1, a + 1);
                        tmp1 -= (1.0 - g(a + 1, a + 1)) * (1.0 - g(b, b)) - g(a + 1, b) * g(b, a + 1);
                        tmp1 -= (1.0 - g(a, a)) * (1.0 - g(b + 1, b + 1)) - g(a, b + 1) * g(b + 1, a);
                        tmp2 += (1.0 - g(a, a)) * conj(g(b, b));
                        tmp2 += (1.0 - g(a + 1, a + 1)) * conj(g(b + 1, b + 1));
                        tmp2 -= (1.0 - g(a, a)) * conj(g(b + 1, b + 1));
                        tmp2 -= (1.0 - g(a + 1, a + 1)) * conj(g(b, b));
                    }
                double c = cos(2 * M_PI / L1 * rx + 2 * M_PI / L2 * ry);
                CDW0 += tmp1 + tmp2;
                CDW1 += c * (tmp1 + tmp2);
                SDW0 += tmp1 - tmp2;
                SDW1 += c * (tmp1 - tmp2);
            }
#undef g
        double n2 = (double)n * (double)n;
        out[2] = creal(CDW0) / n2; out[3] = creal(CDW1) / n2;
        out[4] = creal(SDW0) / n2; out[5] = creal(SDW1) / n2;
    } else {
        out[2] = out[3] = out[4] = out[5] = 0;
    }
    out[0] = Ek; out[1] = Eu;
    free(G0); free(D);
}

static int is_node(const orc_tu *m, int t) {
    for (int i = 0; i < m->nn_nodes; ++i) if (m->nodes[i] == t) return 1;
    return 0;
}

#define BL(k) (m->BLs + (size_t)((k) - 1) * np * n)
#define BR(k) (m->BRs + (size_t)((k) - 1) * n * np)

/* U_L <- rows-orthonormal factor of X (np x n): gerqf+orgrq restated as QR of X^H. */
static void rq_rows(int np, int n, const zdouble *X, zdouble *out) {
    zdouble *XH = zalloc((size_t)n * np);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) XH[IX(j, i, n)] = conj(X[IX(i, j, np)]);
    orc_zqr_q(n, np, XH, n, NULL);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) out[IX(i, j, np)] = conj(XH[IX(j, i, n)]);
    free(XH);
}

/* chain initialisation (phy_update.jl:39-50) */
void orc_tu_init_chain(orc_tu *m) {
    int n = m->ns, np = m->np, NN = m->nn_nodes;
    memcpy(BR(1), m->Pt, sizeof(zdouble) * (size_t)n * np);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) BL(NN)[IX(i, j, np)] = conj(m->Pt[IX(j, i, n)]);
    for (int idx = NN - 1; idx >= 1; --idx) {
        orc_tu_BM_F(m, idx, m->BM);
        orc_zgemm('N', 'N', np, n, n, 1.0, BL(idx + 1), np, m->BM, n, 0.0, m->tmpnN, np);
        rq_rows(np, n, m->tmpnN, BL(idx));
    }
    m->idx = 1;
    orc_get_G(n, np, BL(1), np, BR(1), n, m->G, NULL);
}

/* phy_update (phy_update.jl:3-152). rows (sweeps x 6, row-major) receives the per-sweep
 * averages when record != 0. */
void orc_tu_phy_update(orc_tu *m, int sweeps, int record, double *rows) {
    int n = m->ns, np = m->np, NN = m->nn_nodes, nt = m->nt;
    int thidx = NN / 2 + 1;
    zdouble *D = zalloc(n);
    orc_tu_init_chain(m);
    for (int sw = 0; sw < sweeps; ++sw) {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        int counter = 0;
        for (int lt = 1; lt <= nt; ++lt) {
            for (int i = 0; i < n; ++i) D[i] = cis(m->alpha[lt - 1] * m->eta[m->s[IX(i, lt - 1, n)] - 1]);
            orc_tu_WrapKV(m, m->G, D, 1, 'B');
            orc_tu_update_layer(m, lt, 0);
            if (record && thidx - m->idx >= 0 && thidx - m->idx <= 1) {
                double o[6];
                orc_tu_measure(m, lt, m->G, o);
                for (int q = 0; q < 6; ++q) acc[q] += o[q];
                counter++;
            }
            if (is_node(m, lt)) {
                m->idx += 1;
                int idx = m->idx;
                orc_tu_BM_F(m, idx - 1, m->BM);
                orc_zgemm('N', 'N', n, np, n, 1.0, m->BM, n, BR(idx - 1), n, 0.0, m->tmpNn, n);
                orc_zqr_q(n, np, m->tmpNn, n, NULL);
                memcpy(BR(idx), m->tmpNn, sizeof(zdouble) * (size_t)n * np);
                orc_get_G(n, np, BL(idx), np, BR(idx), n, m->G, NULL);
            }
        }
        for (int lt = nt; lt >= 1; --lt) {
            orc_tu_update_layer(m, lt, 1);
            if (record && m->idx - thidx >= 0 && m->idx - thidx <= 1) {
                double o[6];
                orc_tu_measure(m, lt, m->G, o);
                for (int q = 0; q < 6; ++q) acc[q] += o[q];
                counter++;
            }
            for (int i = 0; i < n; ++i) D[i] = cis(-m->alpha[lt - 1] * m->eta[m->s[IX(i, lt - 1, n)] - 1]);
            orc_tu_WrapKV(m, m->G, D, 0, 'B');
            if (is_node(m, lt - 1)) {
                m->idx -= 1;
                int idx = m->idx;
                orc_tu_BM_F(m, idx, m->BM);
                orc_zgemm('N', 'N', np, n, n, 1.0, BL(idx + 1), np, m->BM, n, 0.0, m->tmpnN, np);
                rq_rows(np, n, m->tmpnN, BL(idx));
                orc_get_G(n, np, BL(idx), np, BR(idx), n, m->G, NULL);
            }
        }
        if (record && rows)
            for (int q = 0; q < 6; ++q) rows[(size_t)sw * 6 + q] = counter ? acc[q] / counter : 0.0;
        m->sweep += 1;
    }
    free(D);
}

/* Bounded sample of phy_update for the CPU baseline: chain initialisation, then the
 * forward pass over slices 1..nslices exactly as in the sweep (wrap, local update and
 * the stabilisation at every node), without measurements. */
void orc_tu_forward_slices(orc_tu *m, int nslices) {
    int n = m->ns, np = m->np;
    zdouble *D = zalloc(n);
    orc_tu_init_chain(m);
    for (int lt = 1; lt <= nslices && lt <= m->nt; ++lt) {
        for (int i = 0; i < n; ++i) D[i] = cis(m->alpha[lt - 1] * m->eta[m->s[IX(i, lt - 1, n)] - 1]);
        orc_tu_WrapKV(m, m->G, D, 1, 'B');
        orc_tu_update_layer(m, lt, 0);
        if (is_node(m, lt)) {
            m->idx += 1;
            int idx = m->idx;
            orc_tu_BM_F(m, idx - 1, m->BM);
            orc_zgemm('N', 'N', n, np, n, 1.0, m->BM, n, BR(idx - 1), n, 0.0, m->tmpNn, n);
            orc_zqr_q(n, np, m->tmpNn, n, NULL);
            memcpy(BR(idx), m->tmpNn, sizeof(zdouble) * (size_t)n * np);
            orc_get_G(n, np, BL(idx), np, BR(idx), n, m->G, NULL);
        }
    }
    free(D);
}

/* Gτ (GreenMatrix.jl:91-123): equal-time G at slice tau from products with a QR every
 * BatchSize slices — the reference's own debug cross-check. */
void orc_tu_Gtau(orc_tu *m, int tau, zdouble *out) {
    int n = m->ns, np = m->np;
    zdouble *BLm = zalloc((size_t)np * n), *BRm = zalloc((size_t)n * np);
    zdouble *T1 = zalloc((size_t)np * n), *T2 = zalloc((size_t)n * np);
    zdouble *D = zalloc(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) BLm[IX(i, j, np)] = conj(m->Pt[IX(j, i, n)]);
    memcpy(BRm, m->Pt, sizeof(zdouble) * (size_t)n * np);
    int counter = 0;
    for (int t = m->nt; t >= tau + 1; --t) {
        for (int i = 0; i < n; ++i) D[i] = cis(m->alpha[t - 1] * m->eta[m->s[IX(i, t - 1, n)] - 1]);
        zdiag_right(np, n, D, BLm, np);
        orc_zgemm('N', 'N', np, n, n, 1.0, BLm, np, m->eK, n, 0.0, T1, np);
        memcpy(BLm, T1, sizeof(zdouble) * (size_t)np * n);
        if (++counter == m->bs) { counter = 0; rq_rows(np, n, BLm, BLm); }
    }
    counter = 0;
    for (int t = 1; t <= tau; ++t) {
        for (int i = 0; i < n; ++i) D[i] = cis(m->alpha[t - 1] * m->eta[m->s[IX(i, t - 1, n)] - 1]);
        orc_zgemm('N', 'N', n, np, n, 1.0, m->eK, n, BRm, n, 0.0, T2, n);
        zdiag_left(n, np, D, T2, n);
        memcpy(BRm, T2, sizeof(zdouble) * (size_t)n * np);
        if (++counter == m->bs) { counter = 0; orc_zqr_q(n, np, BRm, n, NULL); }
    }
    rq_rows(np, n, BLm, BLm);
    orc_zqr_q(n, np, BRm, n, NULL);
    orc_get_G(n, np, BLm, np, BRm, n, out, NULL);
    free(BLm); free(BRm); free(T1); free(T2); free(D);
}

void orc_tu_set_replay(orc_tu *m, const uint8_t *choice, const double *u, uint32_t nsweeps) {
    m->rp.choice = nsweeps ? choice : NULL;
    m->rp.u = nsweeps ? u : NULL;
    m->rp.sweep0 = m->sweep;
    m->rp.nsweeps = nsweeps;
}

/* Imaginary parts of the model matrices for complex hopping (flux != 0; the reference keeps
 * K, eK, eKinv, HalfeK, HalfeKinv and Pt as ComplexF64, src/tU/model.jl:24-34); the real
 * parts came with orc_tu_new. */
void orc_tu_set_imag(orc_tu *m, const double *K_im, const double *eK_im, const double *eKinv_im,
                     const double *HalfeK_im, const double *HalfeKinv_im, const double *Pt_im) {
    size_t nn = (size_t)m->ns * m->ns, np = (size_t)m->ns * m->np;
    for (size_t q = 0; q < nn; ++q) {
        m->K[q] += I * K_im[q];
        m->eK[q] += I * eK_im[q];
        m->eKinv[q] += I * eKinv_im[q];
        m->HalfeK[q] += I * HalfeK_im[q];
        m->HalfeKinv[q] += I * HalfeKinv_im[q];
    }
    for (size_t q = 0; q < np; ++q) m->Pt[q] += I * Pt_im[q];
}
