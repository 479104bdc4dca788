/*
 * ORACLE — test infrastructure only (see orc.h).
 * Restatement of the swap-replica Rényi-2 driver of the t-V and t-U-V models, whose local
 * updates are rank 2 (one bond, two sites):
 *   src/tV/SCEE.jl:1-517     (ctrl_SCEEicr, get_ABGM!, UpdateSCEELayer!, WrapK!)
 *   src/tUV/SCEE.jl:1-402    (ctrl_SCEEicr, get_ABGM!, UpdateSCEELayer!)
 *   src/tV/GreenMatrix.jl:12-119, src/tUV/GreenMatrix.jl:10-122 (BM_F!, BMinv_F!, get_r!,
 *                            WrapV!, WrapKV!)
 *   src/public/Gupdate.jl:23-51,91-132,159-212 (G4update!, GMupdate!, get_abTau1!/2!, G4!)
 *   src/public/GF.jl:75-109  (inv22!, GroverMatrix!)
 * Everything is carried in complex arithmetic (t-V matrices have zero imaginary parts).
 * Differences between the two models, as the reference has them: t-V weighs a proposal by
 * det r and det Tau and keeps det g (real, may change sign); t-U-V weighs by |det r|²,
 * |det Tau|² and keeps |det g|², and its SCEE proposal omits the real-case factor
 * e^{(η-η')(a+1)} that its phy_update applies (src/tUV/SCEE.jl:362-366 vs
 * src/tUV/phy_update.jl:195-197).
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

static zdouble *zz(size_t n) { return (zdouble *)calloc(n ? n : 1, sizeof(zdouble)); }

typedef struct rep2 {
    zdouble *Gt, *G0, *Gt0, *G0t;
    zdouble *BLMs, *BRMs, *BMs, *BMinvs;
    uint8_t *s;   /* nel x nt */
} rep2;

typedef struct area2 {
    int nA;
    int *index;          /* 0-based */
    double detg;
    zdouble *gmInv;      /* nA x nA */
    zdouble *a, *b;      /* a: nA x 2 (column-major), b: 2 x nA */
    zdouble Tau[2][2];
} area2;

struct orc_scee2 {
    int kind;            /* 1 t-V, 2 t-U-V */
    orc_tv *tv;
    orc_tuv *tuv;
    int n, np, nt, NN, nel;
    zdouble *eK, *eKinv, *Pt;
    rep2 R[2];
    area2 A, B;
    double lambda;
    int nlam, idx;
    zdouble *tNN, *tNN2, *tNn, *tnN, *tN, *tN2;
};

static const int *nodes_of(const orc_scee2 *e) { return e->kind == 1 ? e->tv->nodes : e->tuv->nodes; }

orc_scee2 *orc_scee2_new(int kind, void *model, int nA, const int *indexA, int nB, const int *indexB,
                         double lambda, int nlam, const uint8_t *s1, const uint8_t *s2) {
    orc_scee2 *e = (orc_scee2 *)calloc(1, sizeof(orc_scee2));
    e->kind = kind;
    if (kind == 1) {
        e->tv = (orc_tv *)model;
        e->n = e->tv->ns; e->np = e->tv->np; e->nt = e->tv->nt; e->NN = e->tv->nn_nodes;
        e->nel = 3 * e->tv->nb;
    } else {
        e->tuv = (orc_tuv *)model;
        e->n = e->tuv->ns; e->np = e->tuv->np; e->nt = e->tuv->nt; e->NN = e->tuv->nn_nodes;
        e->nel = e->tuv->nbd;
    }
    int n = e->n, np = e->np, NN = e->NN;
    size_t nn = (size_t)n * n;
    e->eK = zz(nn); e->eKinv = zz(nn); e->Pt = zz((size_t)n * np);
    for (size_t q = 0; q < nn; ++q) {
        e->eK[q] = kind == 1 ? e->tv->eK[q] : e->tuv->eK[q];
        e->eKinv[q] = kind == 1 ? e->tv->eKinv[q] : e->tuv->eKinv[q];
    }
    for (size_t q = 0; q < (size_t)n * np; ++q) e->Pt[q] = kind == 1 ? e->tv->Pt[q] : e->tuv->Pt[q];
    for (int r = 0; r < 2; ++r) {
        rep2 *p = &e->R[r];
        p->Gt = zz(nn); p->G0 = zz(nn); p->Gt0 = zz(nn); p->G0t = zz(nn);
        p->BLMs = zz((size_t)np * n * NN); p->BRMs = zz((size_t)n * np * NN);
        p->BMs = zz(nn * NN); p->BMinvs = zz(nn * NN);
        p->s = (uint8_t *)malloc((size_t)e->nel * e->nt);
        memcpy(p->s, r == 0 ? s1 : s2, (size_t)e->nel * e->nt);
    }
    area2 *as[2] = {&e->A, &e->B};
    const int na[2] = {nA, nB};
    const int *ix[2] = {indexA, indexB};
    for (int k = 0; k < 2; ++k) {
        area2 *X = as[k];
        X->nA = na[k];
        X->index = (int *)malloc(sizeof(int) * (na[k] ? na[k] : 1));
        for (int i = 0; i < na[k]; ++i) X->index[i] = ix[k][i] - 1;
        X->gmInv = zz((size_t)na[k] * na[k]);
        X->a = zz(2 * (size_t)na[k]);
        X->b = zz(2 * (size_t)na[k]);
    }
    e->lambda = lambda;
    e->nlam = nlam;
    e->tNN = zz(nn); e->tNN2 = zz(nn); e->tNn = zz((size_t)n * (np > 4 ? np : 4));
    e->tnN = zz((size_t)np * n); e->tN = zz(2 * (size_t)n); e->tN2 = zz(2 * (size_t)n);
    return e;
}

void orc_scee2_free(orc_scee2 *e) {
    if (!e) return;
    for (int r = 0; r < 2; ++r) {
        rep2 *p = &e->R[r];
        free(p->Gt); free(p->G0); free(p->Gt0); free(p->G0t); free(p->BLMs); free(p->BRMs);
        free(p->BMs); free(p->BMinvs); free(p->s);
    }
    area2 *as[2] = {&e->A, &e->B};
    for (int k = 0; k < 2; ++k) { free(as[k]->index); free(as[k]->gmInv); free(as[k]->a); free(as[k]->b); }
    free(e->eK); free(e->eKinv); free(e->Pt);
    free(e->tNN); free(e->tNN2); free(e->tNn); free(e->tnN); free(e->tN); free(e->tN2);
    free(e);
}

void orc_scee2_get_fields(orc_scee2 *e, int r, uint8_t *s) { memcpy(s, e->R[r].s, (size_t)e->nel * e->nt); }
void orc_scee2_get_G4(orc_scee2 *e, int r, zdouble *Gt, zdouble *G0, zdouble *Gt0, zdouble *G0t) {
    size_t nn = (size_t)e->n * e->n * sizeof(zdouble);
    memcpy(Gt, e->R[r].Gt, nn); memcpy(G0, e->R[r].G0, nn);
    memcpy(Gt0, e->R[r].Gt0, nn); memcpy(G0t, e->R[r].G0t, nn);
}
double orc_scee2_detg(orc_scee2 *e, int which) { return which ? e->B.detg : e->A.detg; }

/* ---------------------------------------------------------------- per-model slice factors */
static const double UVH = 0.70710678118654752440;   /* √2/2: UV block [-h -h; -h h] */

/* t-V: D of family j (0-based) at slice lt (1-based) for fields s: x -> exp(sgn α η),
 * y -> exp(-sgn α η) (src/tV/SCEE.jl:107-116) */
static void tv_diag(const orc_scee2 *e, const uint8_t *s, int j, int lt, double sgn, zdouble *D) {
    const orc_tv *m = e->tv;
    for (int i = 0; i < m->nb; ++i) {
        const int x = m->bx[i + j * m->nb], y = m->by[i + j * m->nb];
        const double v = m->alpha[lt - 1] * m->eta[s[(size_t)i + (size_t)j * m->nb + (size_t)(lt - 1) * e->nel] - 1];
        D[x] = exp(sgn * v);
        D[y] = exp(-sgn * v);
    }
}

/* t-U-V: site factor of slice lt: exp(sgn·i·Σ) (complex) or exp(sgn·Σ) (real), Σ_x = Σ η over
 * bonds with x first + a Σ η over bonds with x second (src/tUV/SCEE.jl:107-119) */
static void tuv_diag(const orc_scee2 *e, const uint8_t *s, int lt, double sgn, zdouble *D) {
    const orc_tuv *m = e->tuv;
    double *acc = (double *)calloc(e->n, sizeof(double));
    for (int i = 0; i < m->nbd; ++i) {
        const double v = m->eta[s[(size_t)i + (size_t)(lt - 1) * e->nel] - 1];
        acc[m->bx[i]] += v;
        acc[m->by[i]] += m->a * v;
    }
    for (int q = 0; q < e->n; ++q)
        D[q] = m->is_complex ? cos(sgn * acc[q]) + I * sin(sgn * acc[q]) : exp(sgn * acc[q]);
    free(acc);
}

/* apply the 2x2 bond rotation UV_j (symmetric, its own inverse) to the rows (left = 1) or the
 * columns of G: the bond pairs (x, y) of family j mix as [x; y] <- [-h -h; -h h] [x; y] */
static void tv_rot(const orc_scee2 *e, zdouble *G, int j, int left) {
    const orc_tv *m = e->tv;
    const int n = e->n;
    for (int i = 0; i < m->nb; ++i) {
        const int x = m->bx[i + j * m->nb], y = m->by[i + j * m->nb];
        if (left)
            for (int c = 0; c < n; ++c) {
                const zdouble gx = G[IX(x, c, n)], gy = G[IX(y, c, n)];
                G[IX(x, c, n)] = -UVH * gx - UVH * gy;
                G[IX(y, c, n)] = -UVH * gx + UVH * gy;
            }
        else
            for (int rr = 0; rr < n; ++rr) {
                const zdouble gx = G[IX(rr, x, n)], gy = G[IX(rr, y, n)];
                G[IX(rr, x, n)] = -UVH * gx - UVH * gy;
                G[IX(rr, y, n)] = -UVH * gx + UVH * gy;
            }
    }
}

static void dleft(int n, zdouble *G, const zdouble *D) {
    for (int c = 0; c < n; ++c)
        for (int rr = 0; rr < n; ++rr) G[IX(rr, c, n)] *= D[rr];
}
static void dright(int n, zdouble *G, const zdouble *D) {
    for (int c = 0; c < n; ++c)
        for (int rr = 0; rr < n; ++rr) G[IX(rr, c, n)] *= D[c];
}
static void gemm_nn(int n, const zdouble *A, const zdouble *B, zdouble *C) {
    orc_zgemm('N', 'N', n, n, n, 1.0, A, n, B, n, 0.0, C, n);
}

/* t-V WrapV! (src/tV/GreenMatrix.jl:99-119) with the given D: L: UV D UV G; R: G UV D UV;
 * B: UV D UV G UV D⁻¹ UV */
static void tv_wrapV(const orc_scee2 *e, zdouble *G, const zdouble *D, int j, char LR) {
    const int n = e->n;
    if (LR == 'L' || LR == 'B') {
        tv_rot(e, G, j, 1);
        dleft(n, G, D);
        tv_rot(e, G, j, 1);
    }
    if (LR == 'R' || LR == 'B') {
        zdouble *Dr = e->tN2;
        for (int q = 0; q < n; ++q) Dr[q] = LR == 'B' ? 1.0 / D[q] : D[q];
        tv_rot(e, G, j, 0);
        dright(n, G, Dr);
        tv_rot(e, G, j, 0);
    }
}

/* t-V WrapK! (src/tV/SCEE.jl:509-517) of replica r: forward Gt <- eK Gt eKinv, Gt0 <- eK Gt0,
 * G0t <- G0t eKinv; backward with eK and eKinv exchanged */
static void tv_wrapK(orc_scee2 *e, rep2 *p, int forward) {
    const int n = e->n;
    const zdouble *L = forward ? e->eK : e->eKinv, *Rm = forward ? e->eKinv : e->eK;
    size_t nn = (size_t)n * n * sizeof(zdouble);
    gemm_nn(n, p->Gt, Rm, e->tNN); gemm_nn(n, L, e->tNN, p->Gt);
    gemm_nn(n, L, p->Gt0, e->tNN); memcpy(p->Gt0, e->tNN, nn);
    gemm_nn(n, p->G0t, Rm, e->tNN); memcpy(p->G0t, e->tNN, nn);
}

/* t-U-V WrapKV! (src/tUV/GreenMatrix.jl:92-122). Forward: L: D eK G; R: G eKinv D;
 * B: D eK G eKinv D⁻¹. Backward: L: eKinv D G; R: G D eK; B: eKinv D G D⁻¹ eK. */
static void tuv_wrap(orc_scee2 *e, zdouble *G, const zdouble *D, int forward, char LR) {
    const int n = e->n;
    size_t nn = (size_t)n * n * sizeof(zdouble);
    zdouble *T = e->tNN;
    if (forward) {
        if (LR == 'L' || LR == 'B') { gemm_nn(n, e->eK, G, T); memcpy(G, T, nn); dleft(n, G, D); }
        if (LR == 'R') { gemm_nn(n, G, e->eKinv, T); memcpy(G, T, nn); dright(n, G, D); }
        if (LR == 'B') {
            gemm_nn(n, G, e->eKinv, T); memcpy(G, T, nn);
            for (int q = 0; q < n; ++q) e->tN2[q] = 1.0 / D[q];
            dright(n, G, e->tN2);
        }
    } else {
        if (LR == 'L') { dleft(n, G, D); gemm_nn(n, e->eKinv, G, T); memcpy(G, T, nn); }
        if (LR == 'R') { dright(n, G, D); gemm_nn(n, G, e->eK, T); memcpy(G, T, nn); }
        if (LR == 'B') {
            dleft(n, G, D);
            for (int q = 0; q < n; ++q) e->tN2[q] = 1.0 / D[q];
            dright(n, G, e->tN2);
            gemm_nn(n, e->eKinv, G, T);
            gemm_nn(n, T, e->eK, G);
        }
    }
}

/* BM_F! / BMinv_F! of replica r, block idx (1-based) */
static void blocks2(orc_scee2 *e, int r, int idx) {
    const int n = e->n;
    const size_t nn = (size_t)n * n;
    const int *nodes = nodes_of(e);
    zdouble *BM = e->R[r].BMs + nn * (idx - 1), *BI = e->R[r].BMinvs + nn * (idx - 1);
    zdouble *D = e->tN, *T = e->tNN2;
    memset(BM, 0, nn * sizeof(zdouble));
    memset(BI, 0, nn * sizeof(zdouble));
    for (int q = 0; q < n; ++q) { BM[IX(q, q, n)] = 1.0; BI[IX(q, q, n)] = 1.0; }
    for (int lt = nodes[idx - 1] + 1; lt <= nodes[idx]; ++lt) {
        if (e->kind == 1) {
            /* src/tV/GreenMatrix.jl:12-38: BM <- Π_{j=3..1} (UV_j D_j UV_j) eK BM */
            gemm_nn(n, e->eK, BM, T); memcpy(BM, T, nn * sizeof(zdouble));
            for (int j = 2; j >= 0; --j) {
                tv_diag(e, e->R[r].s, j, lt, 1.0, D);
                tv_rot(e, BM, j, 1); dleft(n, BM, D); tv_rot(e, BM, j, 1);
            }
            /* src/tV/GreenMatrix.jl:40-67: BMinv <- BMinv eKinv Π_{j=3..1} (UV_j D_j⁻¹ UV_j) */
            gemm_nn(n, BI, e->eKinv, T); memcpy(BI, T, nn * sizeof(zdouble));
            for (int j = 2; j >= 0; --j) {
                tv_diag(e, e->R[r].s, j, lt, -1.0, D);
                tv_rot(e, BI, j, 0); dright(n, BI, D); tv_rot(e, BI, j, 0);
            }
        } else {
            /* src/tUV/GreenMatrix.jl:10-64: BM <- D eK BM; BMinv <- BMinv eKinv D⁻¹ */
            tuv_diag(e, e->R[r].s, lt, 1.0, D);
            gemm_nn(n, e->eK, BM, T); memcpy(BM, T, nn * sizeof(zdouble)); dleft(n, BM, D);
            tuv_diag(e, e->R[r].s, lt, -1.0, D);
            gemm_nn(n, BI, e->eKinv, T); memcpy(BI, T, nn * sizeof(zdouble)); dright(n, BI, D);
        }
    }
}

static void rq_rows2(int np, int n, const zdouble *X, zdouble *out) {
    zdouble *XH = zz((size_t)n * np);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) XH[IX(j, i, n)] = conj(X[IX(i, j, np)]);
    orc_zqr_q(n, np, XH, n, NULL);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < np; ++i) out[IX(i, j, np)] = conj(XH[IX(j, i, n)]);
    free(XH);
}

#define RBL(p, k) ((p)->BLMs + (size_t)((k) - 1) * np * n)
#define RBR(p, k) ((p)->BRMs + (size_t)((k) - 1) * n * np)
#define RBM(p, k) ((p)->BMs + (size_t)((k) - 1) * n * n)
#define RBMI(p, k) ((p)->BMinvs + (size_t)((k) - 1) * n * n)

static void update_BR2(orc_scee2 *e, int r, int i) {
    int n = e->n, np = e->np;
    rep2 *p = &e->R[r];
    orc_zgemm('N', 'N', n, np, n, 1.0, RBM(p, i - 1), n, RBR(p, i - 1), n, 0.0, e->tNn, n);
    orc_zqr_q(n, np, e->tNn, n, NULL);
    memcpy(RBR(p, i), e->tNn, sizeof(zdouble) * (size_t)n * np);
}

static void update_BL2(orc_scee2 *e, int r, int i) {
    int n = e->n, np = e->np;
    rep2 *p = &e->R[r];
    orc_zgemm('N', 'N', np, n, n, 1.0, RBL(p, i + 1), np, RBM(p, i), n, 0.0, e->tnN, np);
    rq_rows2(np, n, e->tnN, RBL(p, i));
}

/* G4! (src/public/Gupdate.jl:159-212) */
static void G4_2(orc_scee2 *e, int r, int idx, int forward) {
    int n = e->n, np = e->np, NN = e->NN;
    int thidx = NN / 2 + 1;
    rep2 *p = &e->R[r];
    size_t nn = (size_t)n * n;
    orc_get_G(n, np, RBL(p, idx), np, RBR(p, idx), n, p->Gt, NULL);
    if (idx == thidx) {
        memcpy(p->G0, p->Gt, nn * sizeof(zdouble));
        memcpy(p->Gt0, p->Gt, nn * sizeof(zdouble));
        memcpy(p->G0t, p->Gt, nn * sizeof(zdouble));
        zdouble *shift = forward ? p->G0t : p->Gt0;
        for (int i = 0; i < n; ++i) shift[IX(i, i, n)] -= 1.0;
        return;
    }
    orc_get_G(n, np, RBL(p, thidx), np, RBR(p, thidx), n, p->G0, NULL);
    memset(p->Gt0, 0, nn * sizeof(zdouble));
    memset(p->G0t, 0, nn * sizeof(zdouble));
    for (int i = 0; i < n; ++i) { p->Gt0[IX(i, i, n)] = 1; p->G0t[IX(i, i, n)] = 1; }
    zdouble *T = e->tNN, *T2 = e->tNN2;
    if (idx < thidx) {
        for (int j = idx; j <= thidx - 1; ++j) {
            if (j == idx) memcpy(T2, p->Gt, nn * sizeof(zdouble));
            else orc_get_G(n, np, RBL(p, j), np, RBR(p, j), n, T2, NULL);
            gemm_nn(n, T2, p->G0t, T);
            gemm_nn(n, RBM(p, j), T, p->G0t);
            for (size_t q = 0; q < nn; ++q) T[q] = -T2[q];
            for (int i = 0; i < n; ++i) T[IX(i, i, n)] += 1.0;
            gemm_nn(n, p->Gt0, T, T2);
            gemm_nn(n, T2, RBMI(p, j), p->Gt0);
        }
        for (size_t q = 0; q < nn; ++q) p->Gt0[q] = -p->Gt0[q];
    } else {
        for (int j = thidx; j <= idx - 1; ++j) {
            if (j == thidx) memcpy(T2, p->G0, nn * sizeof(zdouble));
            else orc_get_G(n, np, RBL(p, j), np, RBR(p, j), n, T2, NULL);
            gemm_nn(n, T2, p->Gt0, T);
            gemm_nn(n, RBM(p, j), T, p->Gt0);
            for (size_t q = 0; q < nn; ++q) T[q] = -T2[q];
            for (int i = 0; i < n; ++i) T[IX(i, i, n)] += 1.0;
            gemm_nn(n, p->G0t, T, T2);
            gemm_nn(n, T2, RBMI(p, j), p->G0t);
        }
        for (size_t q = 0; q < nn; ++q) p->G0t[q] = -p->G0t[q];
    }
}

/* GroverMatrix! (GF.jl:101-109) of region X, det g kept as det (t-V, src/tV/SCEE.jl:436) or
 * |det|² (t-U-V, src/tUV/SCEE.jl:345), then inverted */
static void grover2(orc_scee2 *e, area2 *X) {
    int n = e->n, nA = X->nA;
    if (nA == 0) { X->detg = 1.0; return; }
    zdouble *G1 = e->R[0].G0, *G2 = e->R[1].G0;
    zdouble *GM = X->gmInv;
    for (int j = 0; j < nA; ++j)
        for (int i = 0; i < nA; ++i) {
            zdouble s = 0;
            for (int k = 0; k < nA; ++k) s += G1[IX(X->index[i], X->index[k], n)] * G2[IX(X->index[k], X->index[j], n)];
            GM[IX(i, j, nA)] = 2.0 * s - G1[IX(X->index[i], X->index[j], n)] - G2[IX(X->index[i], X->index[j], n)] + (i == j ? 1.0 : 0.0);
        }
    zdouble d = orc_zdet(nA, GM, nA);
    X->detg = e->kind == 1 ? creal(d) : creal(d * conj(d));
    orc_zinv(nA, GM, nA);
}

static void get_ABGM2(orc_scee2 *e, int idx, int forward) {
    G4_2(e, 0, idx, forward);
    G4_2(e, 1, idx, forward);
    grover2(e, &e->A);
    grover2(e, &e->B);
}

static zdouble det22(zdouble m[2][2]) { return m[0][0] * m[1][1] - m[0][1] * m[1][0]; }

/* get_abTau1! / get_abTau2! (Gupdate.jl:91-132) for the bond sub = (x, y) with the 2x2 r;
 * returns det(Tau) */
static zdouble abtau2(orc_scee2 *e, area2 *X, int which, const int sub[2], zdouble r[2][2], const zdouble *G0o,
                      const zdouble *Gt0, const zdouble *G0t) {
    int n = e->n, nA = X->nA;
    X->Tau[0][0] = 1.0; X->Tau[0][1] = 0.0; X->Tau[1][0] = 0.0; X->Tau[1][1] = 1.0;
    if (nA == 0) return 1.0;
    const int *ix = X->index;
    zdouble *zN = e->tN;   /* 2 x nA, [row + 2 q] */
    if (which == 1) {
        /* zN = Gt0[sub, A] (2 G0o[A,A] - I); b = zN gmInv; a = G0t[A, sub] r */
        for (int t = 0; t < 2; ++t)
            for (int q = 0; q < nA; ++q) {
                zdouble s = 0;
                for (int k = 0; k < nA; ++k)
                    s += Gt0[IX(sub[t], ix[k], n)] * (2.0 * G0o[IX(ix[k], ix[q], n)] - (k == q ? 1.0 : 0.0));
                zN[t + 2 * q] = s;
            }
        for (int t = 0; t < 2; ++t)
            for (int q = 0; q < nA; ++q) {
                zdouble s = 0;
                for (int k = 0; k < nA; ++k) s += zN[t + 2 * k] * X->gmInv[IX(k, q, nA)];
                X->b[t + 2 * q] = s;
            }
        for (int q = 0; q < nA; ++q)
            for (int u = 0; u < 2; ++u)
                X->a[q + (size_t)nA * u] = G0t[IX(ix[q], sub[0], n)] * r[0][u] + G0t[IX(ix[q], sub[1], n)] * r[1][u];
    } else {
        /* a = (2 G0o[A,A] - I) G0t[A, sub]; zN = r Gt0[sub, A]; b = zN gmInv */
        for (int q = 0; q < nA; ++q)
            for (int u = 0; u < 2; ++u) {
                zdouble s = 0;
                for (int k = 0; k < nA; ++k)
                    s += (2.0 * G0o[IX(ix[q], ix[k], n)] - (k == q ? 1.0 : 0.0)) * G0t[IX(ix[k], sub[u], n)];
                X->a[q + (size_t)nA * u] = s;
            }
        for (int t = 0; t < 2; ++t)
            for (int q = 0; q < nA; ++q)
                zN[t + 2 * q] = r[t][0] * Gt0[IX(sub[0], ix[q], n)] + r[t][1] * Gt0[IX(sub[1], ix[q], n)];
        for (int t = 0; t < 2; ++t)
            for (int q = 0; q < nA; ++q) {
                zdouble s = 0;
                for (int k = 0; k < nA; ++k) s += zN[t + 2 * k] * X->gmInv[IX(k, q, nA)];
                X->b[t + 2 * q] = s;
            }
    }
    for (int t = 0; t < 2; ++t)
        for (int u = 0; u < 2; ++u) {
            zdouble s = (t == u) ? 1.0 : 0.0;
            for (int q = 0; q < nA; ++q) s += X->b[t + 2 * q] * X->a[q + (size_t)nA * u];
            X->Tau[t][u] = s;
        }
    return det22(X->Tau);
}

/* GMupdate! (Gupdate.jl:45-51): gmInv -= (gmInv a) Tau⁻¹ b, Tau⁻¹ by inv22! (GF.jl:75-94) */
static void gm_update2(area2 *X) {
    int nA = X->nA;
    if (nA == 0) return;
    zdouble dt = det22(X->Tau);
    zdouble ti[2][2] = {{X->Tau[1][1] / dt, -X->Tau[0][1] / dt}, {-X->Tau[1][0] / dt, X->Tau[0][0] / dt}};
    zdouble *Nz = (zdouble *)malloc(sizeof(zdouble) * 2 * nA), *zb = (zdouble *)malloc(sizeof(zdouble) * 2 * nA);
    for (int q = 0; q < nA; ++q)
        for (int u = 0; u < 2; ++u) {
            zdouble s = 0;
            for (int k = 0; k < nA; ++k) s += X->gmInv[IX(q, k, nA)] * X->a[k + (size_t)nA * u];
            Nz[q + (size_t)nA * u] = s;
        }
    for (int t = 0; t < 2; ++t)
        for (int c = 0; c < nA; ++c) zb[t + 2 * c] = ti[t][0] * X->b[0 + 2 * c] + ti[t][1] * X->b[1 + 2 * c];
    for (int c = 0; c < nA; ++c)
        for (int q = 0; q < nA; ++q)
            X->gmInv[IX(q, c, nA)] -= Nz[q] * zb[0 + 2 * c] + Nz[q + (size_t)nA] * zb[1 + 2 * c];
    free(Nz);
    free(zb);
}

/* G4update! (Gupdate.jl:23-37), rank 2 at sub with the 2x2 r */
static void g4_update2(orc_scee2 *e, rep2 *p, const int sub[2], zdouble r[2][2]) {
    int n = e->n;
    zdouble *zN = e->tN, *cG0t = e->tNn, *cGt = e->tNn + 2 * (size_t)n;   /* columns sub of G0t, Gt */
    for (int u = 0; u < 2; ++u)
        for (int k = 0; k < n; ++k) {
            cG0t[k + (size_t)n * u] = p->G0t[IX(k, sub[u], n)];
            cGt[k + (size_t)n * u] = p->Gt[IX(k, sub[u], n)];
        }
    /* zN = r Gt0[sub, :]; G0 += G0t[:, sub] zN; Gt0 += Gt[:, sub] zN */
    for (int c = 0; c < n; ++c)
        for (int t = 0; t < 2; ++t) zN[t + 2 * c] = r[t][0] * p->Gt0[IX(sub[0], c, n)] + r[t][1] * p->Gt0[IX(sub[1], c, n)];
    for (int c = 0; c < n; ++c)
        for (int k = 0; k < n; ++k) {
            p->G0[IX(k, c, n)] += cG0t[k] * zN[2 * c] + cG0t[k + n] * zN[1 + 2 * c];
            p->Gt0[IX(k, c, n)] += cGt[k] * zN[2 * c] + cGt[k + n] * zN[1 + 2 * c];
        }
    /* zN = -r Gt[sub, :] + r at columns sub; G0t -= G0t[:, sub] zN; Gt -= Gt[:, sub] zN */
    for (int c = 0; c < n; ++c)
        for (int t = 0; t < 2; ++t)
            zN[t + 2 * c] = -(r[t][0] * p->Gt[IX(sub[0], c, n)] + r[t][1] * p->Gt[IX(sub[1], c, n)]);
    for (int u = 0; u < 2; ++u)
        for (int t = 0; t < 2; ++t) zN[t + 2 * sub[u]] += r[t][u];
    for (int c = 0; c < n; ++c)
        for (int k = 0; k < n; ++k) {
            p->G0t[IX(k, c, n)] -= cG0t[k] * zN[2 * c] + cG0t[k + n] * zN[1 + 2 * c];
            p->Gt[IX(k, c, n)] -= cGt[k] * zN[2 * c] + cGt[k + n] * zN[1 + 2 * c];
        }
}

/* get_r! of the model (src/tV/GreenMatrix.jl:76-89, src/tUV/GreenMatrix.jl:73-89): Δ of the
 * proposal, r = I + Δ (I - Gt[sub, sub]); weight det r (t-V) or |det r|² (t-U-V); r <- r⁻¹ Δ */
static double get_r2(orc_scee2 *e, int lt, int si, int sx, const int sub[2], const zdouble *Gt, zdouble r[2][2]) {
    const int n = e->n;
    zdouble Dl[2][2];
    if (e->kind == 1) {
        const orc_tv *m = e->tv;
        const double ds = m->alpha[lt - 1] * (m->eta[sx - 1] - m->eta[si - 1]);
        const double t2[2] = {exp(ds) - 1.0, exp(-ds) - 1.0};
        const double uv[2][2] = {{-UVH, -UVH}, {-UVH, UVH}};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) Dl[a][b] = uv[a][0] * t2[0] * uv[0][b] + uv[a][1] * t2[1] * uv[1][b];
    } else {
        const orc_tuv *m = e->tuv;
        const double ds = m->eta[sx - 1] - m->eta[si - 1];
        const double v[2] = {ds, m->a * ds};
        Dl[0][1] = 0.0; Dl[1][0] = 0.0;
        for (int q = 0; q < 2; ++q)
            Dl[q][q] = m->is_complex ? (cos(v[q]) + I * sin(v[q])) - 1.0 : exp(v[q]) - 1.0;
    }
    zdouble rr[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            zdouble acc = Dl[a][b];
            for (int q = 0; q < 2; ++q) acc -= Dl[a][q] * Gt[IX(sub[q], sub[b], n)];
            rr[a][b] = acc + (a == b ? 1.0 : 0.0);
        }
    zdouble p = det22(rr);
    zdouble ri[2][2] = {{rr[1][1] / p, -rr[0][1] / p}, {-rr[1][0] / p, rr[0][0] / p}};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) r[a][b] = ri[a][0] * Dl[0][b] + ri[a][1] * Dl[1][b];
    return e->kind == 1 ? creal(p) : creal(p * conj(p));
}

/* UpdateSCEELayer! over the elements [e0, e0 + ne) of slice lt: per element replica 1 then
 * replica 2 (src/tV/SCEE.jl:447-494 for one family, src/tUV/SCEE.jl:355-402). Replica r
 * draws from Philox stream 1 + r with the element's linear index. */
static void update_elems(orc_scee2 *e, int lt, int dir, uint32_t sweep, int e0, int ne) {
    const uint32_t tag = (uint32_t)lt | ((uint32_t)dir << 31);
    const uint64_t seed = e->kind == 1 ? e->tv->seed : e->tuv->seed;
    const uint32_t chain = e->kind == 1 ? e->tv->chain : e->tuv->chain;
    const double *gamma = e->kind == 1 ? e->tv->gamma : e->tuv->gamma;
    for (int el = e0; el < e0 + ne; ++el) {
        int sub[2];
        if (e->kind == 1) { sub[0] = e->tv->bx[el]; sub[1] = e->tv->by[el]; }
        else { sub[0] = e->tuv->bx[el]; sub[1] = e->tuv->by[el]; }
        for (int r = 0; r < 2; ++r) {
            rep2 *p = &e->R[r], *o = &e->R[1 - r];
            uint8_t *sv = p->s + (size_t)el + (size_t)(lt - 1) * e->nel;
            int si = *sv, choice;
            double u;
            orc_draw(seed, chain, 1 + r, sweep, tag, (uint32_t)el, &choice, &u);
            int allowed[3], c = 0;
            for (int v = 1; v <= 4; ++v) if (v != si) allowed[c++] = v;
            const int sx = allowed[choice];
            zdouble rm[2][2];
            double prob = get_r2(e, lt, si, sx, sub, p->Gt, rm) * (gamma[sx - 1] / gamma[si - 1]);
            zdouble tA = abtau2(e, &e->A, r + 1, sub, rm, o->G0, p->Gt0, p->G0t);
            zdouble tB = abtau2(e, &e->B, r + 1, sub, rm, o->G0, p->Gt0, p->G0t);
            const double dA = e->kind == 1 ? creal(tA) : creal(tA * conj(tA));
            const double dB = e->kind == 1 ? creal(tB) : creal(tB * conj(tB));
            prob *= pow(dA, e->lambda) * pow(dB, 1.0 - e->lambda);
            if (fabs(u - prob) < ORC_NEAR_TIE) ++*(e->kind == 1 ? &e->tv->n_ties : &e->tuv->n_ties);
            if (u < prob) {
                e->A.detg *= dA;
                e->B.detg *= dB;
                gm_update2(&e->A);
                gm_update2(&e->B);
                g4_update2(e, p, sub, rm);
                *sv = (uint8_t)sx;
            }
        }
    }
}

/* chain set-up of ctrl_SCEEicr (src/tV/SCEE.jl:47-86, src/tUV/SCEE.jl:47-86) */
void orc_scee2_init(orc_scee2 *e) {
    int n = e->n, np = e->np, NN = e->NN;
    for (int idx = 1; idx <= NN - 1; ++idx) { blocks2(e, 0, idx); blocks2(e, 1, idx); }
    for (int r = 0; r < 2; ++r) {
        rep2 *p = &e->R[r];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < np; ++i) RBL(p, NN)[IX(i, j, np)] = conj(e->Pt[IX(j, i, n)]);
        memcpy(RBR(p, 1), e->Pt, sizeof(zdouble) * (size_t)n * np);
    }
    for (int i = 1; i <= NN - 1; ++i)
        for (int r = 0; r < 2; ++r) {
            update_BL2(e, r, NN - i);
            update_BR2(e, r, i + 1);
        }
    e->idx = 1;
    get_ABGM2(e, 1, 1);
}

static void slice_wrap(orc_scee2 *e, int lt, int forward, int j) {
    /* one wrap step of both replicas: t-V family j (forward: D, backward: D⁻¹), t-U-V the
     * whole slice; the reference order is L on Gt0, B on Gt, R on G0t with D inverted by B */
    zdouble *D = e->tN2 + 0;   /* tN2 is overwritten inside the wraps: keep D in its own buffer */
    zdouble *Dv = (zdouble *)malloc(sizeof(zdouble) * e->n), *Di = (zdouble *)malloc(sizeof(zdouble) * e->n);
    (void)D;
    for (int r = 0; r < 2; ++r) {
        rep2 *p = &e->R[r];
        if (e->kind == 1) tv_diag(e, p->s, j, lt, forward ? 1.0 : -1.0, Dv);
        else tuv_diag(e, p->s, lt, forward ? 1.0 : -1.0, Dv);
        for (int q = 0; q < e->n; ++q) Di[q] = 1.0 / Dv[q];
        if (e->kind == 1) {
            tv_wrapV(e, p->Gt0, Dv, j, 'L');
            tv_wrapV(e, p->Gt, Dv, j, 'B');
            tv_wrapV(e, p->G0t, Di, j, 'R');
        } else {
            tuv_wrap(e, p->Gt0, Dv, forward, 'L');
            tuv_wrap(e, p->Gt, Dv, forward, 'B');
            tuv_wrap(e, p->G0t, Di, forward, 'R');
        }
    }
    free(Dv);
    free(Di);
}

static int is_node(const orc_scee2 *e, int t) {
    const int *nodes = nodes_of(e);
    for (int q = 0; q < e->NN; ++q)
        if (nodes[q] == t) return 1;
    return 0;
}

/* One forward slice without node handling (wraps + updates), for decision replays. */
void orc_scee2_forward_slice(orc_scee2 *e, int lt, uint32_t sweep) {
    if (e->kind == 1) {
        tv_wrapK(e, &e->R[0], 1);
        tv_wrapK(e, &e->R[1], 1);
        for (int j = 2; j >= 0; --j) {
            slice_wrap(e, lt, 1, j);
            update_elems(e, lt, 0, sweep, j * e->tv->nb, e->tv->nb);
        }
    } else {
        slice_wrap(e, lt, 1, 0);
        update_elems(e, lt, 0, sweep, 0, e->nel);
    }
}

/* ctrl_SCEEicr sweeps (src/tV/SCEE.jl:87-431, src/tUV/SCEE.jl:87-339). */
void orc_scee2_sweeps(orc_scee2 *e, int sweeps, double *O) {
    int nt = e->nt, NN = e->NN, thidx = NN / 2 + 1;
    uint32_t *sweepc = e->kind == 1 ? &e->tv->sweep : &e->tuv->sweep;
    orc_scee2_init(e);
    for (int sw = 0; sw < sweeps; ++sw) {
        uint32_t sweep = *sweepc;
        double tmpO = 0.0;
        int counter = 0;
        for (int lt = 1; lt <= nt; ++lt) {
            orc_scee2_forward_slice(e, lt, sweep);
            tmpO += pow(e->A.detg / e->B.detg, 1.0 / e->nlam);
            counter++;
            if (is_node(e, lt)) {
                e->idx += 1;
                int idx = e->idx;
                blocks2(e, 0, idx - 1);
                blocks2(e, 1, idx - 1);
                for (int i = idx; i <= (thidx > idx ? thidx : idx); ++i) { update_BR2(e, 0, i); update_BR2(e, 1, i); }
                for (int i = idx - 1; i >= (thidx < idx ? thidx : idx); --i) { update_BL2(e, 0, i); update_BL2(e, 1, i); }
                get_ABGM2(e, idx, 1);
            }
        }
        for (int lt = nt; lt >= 1; --lt) {
            if (e->kind == 1) {
                for (int j = 0; j < 3; ++j) {
                    update_elems(e, lt, 1, sweep, j * e->tv->nb, e->tv->nb);
                    slice_wrap(e, lt, 0, j);
                }
                tv_wrapK(e, &e->R[0], 0);
                tv_wrapK(e, &e->R[1], 0);
            } else {
                update_elems(e, lt, 1, sweep, 0, e->nel);
            }
            tmpO += pow(e->A.detg / e->B.detg, 1.0 / e->nlam);
            counter++;
            if (e->kind == 2) slice_wrap(e, lt, 0, 0);
            if (is_node(e, lt - 1)) {
                e->idx -= 1;
                int idx = e->idx;
                blocks2(e, 0, idx);
                blocks2(e, 1, idx);
                for (int i = idx; i >= (thidx < idx ? thidx : idx); --i) { update_BL2(e, 0, i); update_BL2(e, 1, i); }
                for (int i = idx + 1; i <= (thidx > idx ? thidx : idx); ++i) { update_BR2(e, 0, i); update_BR2(e, 1, i); }
                get_ABGM2(e, idx, 0);
            }
        }
        O[sw] = tmpO / counter;
        *sweepc += 1;
    }
}
