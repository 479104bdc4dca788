/*
 * ORACLE — test infrastructure only (see orc.h).
 * Restatement of the swap-replica Rényi-2 driver of the t-U model:
 *   src/tU/SCEE.jl:1-324            (ctrl_SCEEicr, get_ABGM!, UpdateSCEELayer!)
 *   src/public/Gupdate.jl:23-51,91-132,159-212 (G4update!, GMupdate!, get_abTau1!/2!, G4!)
 *   src/public/GF.jl:101-109        (GroverMatrix!)
 * Two replicas (fields ss[1], ss[2]) share the orc_tu model constants; each replica holds
 * G(τ), G(Θ), G(τ,Θ), G(Θ,τ) and its block products, and the Grover matrices of regions A
 * and B are tracked through their inverses and |det|.
 * The disorder operator of the t-U model (orc_dop_new) runs the same driver on one replica:
 *   src/tU/disorder_operate.jl:1-247 (ctrl_SCDOPicr, get_ABX!, UpdateSCDOPLayer!)
 *   src/public/Gupdate.jl:52-82     (GMupdate! and get_abTau! of DOPBuffer_)
 * with X_A = e^{iα} I + (1 - e^{iα}) G0[A,A] in place of the Grover matrix.
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

static zdouble *zz(size_t n) { return (zdouble *)calloc(n ? n : 1, sizeof(zdouble)); }

typedef struct g4rep {
    zdouble *Gt, *G0, *Gt0, *G0t;
    zdouble *BLMs, *BRMs, *BMs, *BMinvs;   /* np x n x NN, n x np x NN, n x n x NN (x2) */
    uint8_t *s;                           /* n x nt */
} g4rep;

typedef struct area {
    int nA;
    int *index;        /* 0-based */
    double detg;
    zdouble *gmInv;    /* nA x nA */
    zdouble *a, *b;    /* nA */
    zdouble Tau;
} area;

struct orc_scee {
    orc_tu *m;         /* model constants and scratch; its s/G are not used */
    int nrep, dop;     /* 2 replicas (Rényi-2) or 1 (disorder operator) */
    zdouble ea;        /* e^{iα} of the disorder operator */
    g4rep R[2];
    area A, B;
    double lambda;
    int nlam;
    int idx;
    zdouble *tNN, *tNN2, *tNn, *tnN, *tN;
};

static void alloc_rep(orc_tu *m, g4rep *r) {
    int n = m->ns, np = m->np, NN = m->nn_nodes;
    r->Gt = zz((size_t)n * n); r->G0 = zz((size_t)n * n); r->Gt0 = zz((size_t)n * n); r->G0t = zz((size_t)n * n);
    r->BLMs = zz((size_t)np * n * NN); r->BRMs = zz((size_t)n * np * NN);
    r->BMs = zz((size_t)n * n * NN); r->BMinvs = zz((size_t)n * n * NN);
    r->s = (uint8_t *)calloc((size_t)n * m->nt, 1);
}

static void alloc_area(area *A, int nA, const int *index1) {
    A->nA = nA;
    A->index = (int *)malloc(sizeof(int) * (nA ? nA : 1));
    for (int i = 0; i < nA; ++i) A->index[i] = index1[i] - 1;
    A->gmInv = zz((size_t)nA * nA); A->a = zz(nA); A->b = zz(nA);
}

orc_scee *orc_scee_new(orc_tu *m, int nA, const int *indexA, int nB, const int *indexB, double lambda,
                       int nlam, const uint8_t *s1, const uint8_t *s2) {
    orc_scee *e = (orc_scee *)calloc(1, sizeof(orc_scee));
    e->m = m;
    e->nrep = 2;
    alloc_rep(m, &e->R[0]);
    alloc_rep(m, &e->R[1]);
    memcpy(e->R[0].s, s1, (size_t)m->ns * m->nt);
    memcpy(e->R[1].s, s2, (size_t)m->ns * m->nt);
    alloc_area(&e->A, nA, indexA);
    alloc_area(&e->B, nB, indexB);
    e->lambda = lambda; e->nlam = nlam;
    int n = m->ns, np = m->np;
    e->tNN = zz((size_t)n * n); e->tNN2 = zz((size_t)n * n); e->tNn = zz((size_t)n * np);
    e->tnN = zz((size_t)np * n); e->tN = zz(n);
    return e;
}

/* ctrl_SCDOPicr (src/tU/disorder_operate.jl:1): one field s, DOPBuffer(alpha, index) of both
 * regions */
orc_scee *orc_dop_new(orc_tu *m, double alpha, int nA, const int *indexA, int nB, const int *indexB,
                      double lambda, int nlam, const uint8_t *s) {
    orc_scee *e = orc_scee_new(m, nA, indexA, nB, indexB, lambda, nlam, s, s);
    e->nrep = 1;
    e->dop = 1;
    e->ea = cos(alpha) + I * sin(alpha);
    return e;
}

void orc_scee_free(orc_scee *e) {
    if (!e) return;
    for (int r = 0; r < 2; ++r) {
        g4rep *p = &e->R[r];
        free(p->Gt); free(p->G0); free(p->Gt0); free(p->G0t); free(p->BLMs); free(p->BRMs);
        free(p->BMs); free(p->BMinvs); free(p->s);
    }
    area *as[2] = {&e->A, &e->B};
    for (int k = 0; k < 2; ++k) { free(as[k]->index); free(as[k]->gmInv); free(as[k]->a); free(as[k]->b); }
    free(e->tNN); free(e->tNN2); free(e->tNn); free(e->tnN); free(e->tN);
    free(e);
}

void orc_scee_get_fields(orc_scee *e, int r, uint8_t *s) { memcpy(s, e->R[r].s, (size_t)e->m->ns * e->m->nt); }
void orc_scee_get_G4(orc_scee *e, int r, zdouble *Gt, zdouble *G0, zdouble *Gt0, zdouble *G0t) {
    size_t nn = (size_t)e->m->ns * e->m->ns * sizeof(zdouble);
    memcpy(Gt, e->R[r].Gt, nn); memcpy(G0, e->R[r].G0, nn);
    memcpy(Gt0, e->R[r].Gt0, nn); memcpy(G0t, e->R[r].G0t, nn);
}
double orc_scee_detg(orc_scee *e, int which) { return which ? e->B.detg : e->A.detg; }

static inline zdouble cisd(double x) { return cos(x) + I * sin(x); }

/* BM_F!/BMinv_F! of replica r, with the orc_tu machinery pointed at the replica fields */
static void blocks(orc_scee *e, int r, int idx) {
    orc_tu *m = e->m;
    size_t nn = (size_t)m->ns * m->ns;
    uint8_t *keep = m->s;
    m->s = e->R[r].s;
    orc_tu_BM_F(m, idx, e->R[r].BMs + nn * (idx - 1));
    orc_tu_BMinv_F(m, idx, e->R[r].BMinvs + nn * (idx - 1));
    m->s = keep;
}

static void rq_rows(int np, int n, const zdouble *X, zdouble *out) {
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

static void update_BR(orc_scee *e, int r, int i) {
    orc_tu *m = e->m;
    int n = m->ns, np = m->np;
    g4rep *p = &e->R[r];
    orc_zgemm('N', 'N', n, np, n, 1.0, RBM(p, i - 1), n, RBR(p, i - 1), n, 0.0, e->tNn, n);
    orc_zqr_q(n, np, e->tNn, n, NULL);
    memcpy(RBR(p, i), e->tNn, sizeof(zdouble) * (size_t)n * np);
}

static void update_BL(orc_scee *e, int r, int i) {
    orc_tu *m = e->m;
    int n = m->ns, np = m->np;
    g4rep *p = &e->R[r];
    orc_zgemm('N', 'N', np, n, n, 1.0, RBL(p, i + 1), np, RBM(p, i), n, 0.0, e->tnN, np);
    rq_rows(np, n, e->tnN, RBL(p, i));
}

/* G4! (src/public/Gupdate.jl:159-212) */
static void G4(orc_scee *e, int r, int idx, int forward) {
    orc_tu *m = e->m;
    int n = m->ns, np = m->np, NN = m->nn_nodes;
    int thidx = NN / 2 + 1;
    g4rep *p = &e->R[r];
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
            orc_zgemm('N', 'N', n, n, n, 1.0, T2, n, p->G0t, n, 0.0, T, n);
            orc_zgemm('N', 'N', n, n, n, 1.0, RBM(p, j), n, T, n, 0.0, p->G0t, n);
            for (size_t q = 0; q < nn; ++q) T[q] = -T2[q];
            for (int i = 0; i < n; ++i) T[IX(i, i, n)] += 1.0;
            orc_zgemm('N', 'N', n, n, n, 1.0, p->Gt0, n, T, n, 0.0, T2, n);
            orc_zgemm('N', 'N', n, n, n, 1.0, T2, n, RBMI(p, j), n, 0.0, p->Gt0, n);
        }
        for (size_t q = 0; q < nn; ++q) p->Gt0[q] = -p->Gt0[q];
    } else {
        for (int j = thidx; j <= idx - 1; ++j) {
            if (j == thidx) memcpy(T2, p->G0, nn * sizeof(zdouble));
            else orc_get_G(n, np, RBL(p, j), np, RBR(p, j), n, T2, NULL);
            orc_zgemm('N', 'N', n, n, n, 1.0, T2, n, p->Gt0, n, 0.0, T, n);
            orc_zgemm('N', 'N', n, n, n, 1.0, RBM(p, j), n, T, n, 0.0, p->Gt0, n);
            for (size_t q = 0; q < nn; ++q) T[q] = -T2[q];
            for (int i = 0; i < n; ++i) T[IX(i, i, n)] += 1.0;
            orc_zgemm('N', 'N', n, n, n, 1.0, p->G0t, n, T, n, 0.0, T2, n);
            orc_zgemm('N', 'N', n, n, n, 1.0, T2, n, RBMI(p, j), n, 0.0, p->G0t, n);
        }
        for (size_t q = 0; q < nn; ++q) p->G0t[q] = -p->G0t[q];
    }
}

/* GroverMatrix! (GF.jl:101-109) of region X from G0 of both replicas, |det|² and inverse */
static void grover(orc_scee *e, area *X) {
    int n = e->m->ns, nA = X->nA;
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
    X->detg = creal(d * conj(d));
    orc_zinv(nA, GM, nA);
}

/* get_ABX! (src/tU/disorder_operate.jl:202-223): X = e^{iα} I + (1 - e^{iα}) G0[A,A],
 * detX = |det X|², and X⁻¹ (kept in gmInv) */
static void xform(orc_scee *e, area *X) {
    int n = e->m->ns, nA = X->nA;
    if (nA == 0) { X->detg = 1.0; return; }
    zdouble *G0 = e->R[0].G0, *XM = X->gmInv;
    for (int j = 0; j < nA; ++j)
        for (int i = 0; i < nA; ++i)
            XM[IX(i, j, nA)] = (1.0 - e->ea) * G0[IX(X->index[i], X->index[j], n)] + (i == j ? e->ea : 0.0);
    zdouble d = orc_zdet(nA, XM, nA);
    X->detg = creal(d * conj(d));
    orc_zinv(nA, XM, nA);
}

/* get_ABGM! (src/tU/SCEE.jl:264-276), or get_ABX! of the disorder operator */
static void get_ABGM(orc_scee *e, int idx, int forward) {
    for (int r = 0; r < e->nrep; ++r) G4(e, r, idx, forward);
    if (e->dop) {
        xform(e, &e->A);
        xform(e, &e->B);
        return;
    }
    grover(e, &e->A);
    grover(e, &e->B);
}

/* get_abTau! of DOPBuffer_ (Gupdate.jl:70-82): a = (1 - e^{iα}) G0t[A, i] r,
 * b = Gt0[i, A] X⁻¹, Tau = b a + 1 */
static zdouble abtau_dop(orc_scee *e, area *X, int i, zdouble r, const zdouble *Gt0, const zdouble *G0t) {
    int n = e->m->ns, nA = X->nA;
    if (nA == 0) { X->Tau = 1.0; return 1.0; }
    const int *ix = X->index;
    for (int q = 0; q < nA; ++q) X->a[q] = (1.0 - e->ea) * (G0t[IX(ix[q], i, n)] * r);
    for (int q = 0; q < nA; ++q) {
        zdouble s = 0;
        for (int k = 0; k < nA; ++k) s += Gt0[IX(i, ix[k], n)] * X->gmInv[IX(k, q, nA)];
        X->b[q] = s;
    }
    zdouble tau = 1.0;
    for (int q = 0; q < nA; ++q) tau += X->b[q] * X->a[q];
    X->Tau = tau;
    return tau;
}

/* get_abTau1! / get_abTau2! (Gupdate.jl:91-132); returns Tau (1 x 1 for t-U) */
static zdouble abtau(orc_scee *e, area *X, int which, int i, zdouble r, const zdouble *G0o, const zdouble *Gt0,
                     const zdouble *G0t) {
    int n = e->m->ns, nA = X->nA;
    if (nA == 0) { X->Tau = 1.0; return 1.0; }
    zdouble *zN = e->tN;   /* nA <= n */
    const int *ix = X->index;
    if (which == 1) {
        /* zN = Gt0[i, A] (2 G0o[A,A] - I); b = zN gmInv; a = G0t[A, i] r */
        for (int q = 0; q < nA; ++q) {
            zdouble s = 0;
            for (int k = 0; k < nA; ++k) s += Gt0[IX(i, ix[k], n)] * (2.0 * G0o[IX(ix[k], ix[q], n)] - (k == q ? 1.0 : 0.0));
            zN[q] = s;
        }
        for (int q = 0; q < nA; ++q) {
            zdouble s = 0;
            for (int k = 0; k < nA; ++k) s += zN[k] * X->gmInv[IX(k, q, nA)];
            X->b[q] = s;
        }
        for (int q = 0; q < nA; ++q) X->a[q] = G0t[IX(ix[q], i, n)] * r;
    } else {
        /* a = (2 G0o[A,A] - I) G0t[A, i]; zN = r Gt0[i, A]; b = zN gmInv */
        for (int q = 0; q < nA; ++q) {
            zdouble s = 0;
            for (int k = 0; k < nA; ++k) s += (2.0 * G0o[IX(ix[q], ix[k], n)] - (k == q ? 1.0 : 0.0)) * G0t[IX(ix[k], i, n)];
            X->a[q] = s;
        }
        for (int q = 0; q < nA; ++q) zN[q] = r * Gt0[IX(i, ix[q], n)];
        for (int q = 0; q < nA; ++q) {
            zdouble s = 0;
            for (int k = 0; k < nA; ++k) s += zN[k] * X->gmInv[IX(k, q, nA)];
            X->b[q] = s;
        }
    }
    zdouble tau = 1.0;
    for (int q = 0; q < nA; ++q) tau += X->b[q] * X->a[q];
    X->Tau = tau;
    return tau;
}

/* GMupdate! (Gupdate.jl:45-51): gmInv -= (gmInv a)(Tau⁻¹ b) */
static void gm_update(area *X) {
    int nA = X->nA;
    if (nA == 0) return;
    zdouble *Nz = (zdouble *)malloc(sizeof(zdouble) * nA);
    for (int q = 0; q < nA; ++q) {
        zdouble s = 0;
        for (int k = 0; k < nA; ++k) s += X->gmInv[IX(q, k, nA)] * X->a[k];
        Nz[q] = s;
    }
    zdouble ti = 1.0 / X->Tau;
    for (int c = 0; c < nA; ++c) {
        zdouble zc = ti * X->b[c];
        for (int q = 0; q < nA; ++q) X->gmInv[IX(q, c, nA)] -= Nz[q] * zc;
    }
    free(Nz);
}

/* G4update! (Gupdate.jl:23-37) with r = Δ / (1 + Δ (1 - Gt_ii)) at site i */
static void g4_update(orc_scee *e, g4rep *p, int i, zdouble r) {
    int n = e->m->ns;
    zdouble *zN = e->tN, *cG0t = e->tNn, *cGt = e->tNn + n;   /* tNn has n*np >= 2n entries */
    for (int k = 0; k < n; ++k) { cG0t[k] = p->G0t[IX(k, i, n)]; cGt[k] = p->Gt[IX(k, i, n)]; }
    for (int c = 0; c < n; ++c) zN[c] = r * p->Gt0[IX(i, c, n)];
    for (int c = 0; c < n; ++c)
        for (int k = 0; k < n; ++k) {
            p->G0[IX(k, c, n)] += cG0t[k] * zN[c];
            p->Gt0[IX(k, c, n)] += cGt[k] * zN[c];
        }
    for (int c = 0; c < n; ++c) zN[c] = -r * p->Gt[IX(i, c, n)];
    zN[i] += r;
    for (int c = 0; c < n; ++c)
        for (int k = 0; k < n; ++k) {
            p->G0t[IX(k, c, n)] -= cG0t[k] * zN[c];
            p->Gt[IX(k, c, n)] -= cGt[k] * zN[c];
        }
}

/* UpdateSCEELayer! (src/tU/SCEE.jl:278-324). Replica r draws from Philox stream 1 + r.
 * Disorder operator: UpdateSCDOPLayer! (src/tU/disorder_operate.jl:225-247), one replica. */
static void update_layer(orc_scee *e, int lt, int dir, uint32_t sweep) {
    orc_tu *m = e->m;
    int n = m->ns;
    uint32_t tag = (uint32_t)lt | ((uint32_t)dir << 31);
    for (int i = 0; i < n; ++i) {
        for (int r = 0; r < e->nrep; ++r) {
            g4rep *p = &e->R[r], *o = &e->R[1 - r];
            uint8_t *sv = p->s + IX(i, lt - 1, n);
            int si = *sv, choice;
            double u;
            orc_draw(m->seed, m->chain, 1 + r, sweep, tag, (uint32_t)i, &choice, &u);
            int allowed[3], c = 0;
            for (int v = 1; v <= 4; ++v) if (v != si) allowed[c++] = v;
            int sx = allowed[choice];
            zdouble Delta = cisd(m->alpha[lt - 1] * (m->eta[sx - 1] - m->eta[si - 1])) - 1.0;
            zdouble pc = 1.0 + Delta * (1.0 - p->Gt[IX(i, i, n)]);
            zdouble rr = Delta / pc;
            double prob = creal(pc * conj(pc)) * (m->gamma[sx - 1] / m->gamma[si - 1]);
            zdouble tA = e->dop ? abtau_dop(e, &e->A, i, rr, p->Gt0, p->G0t)
                                : abtau(e, &e->A, r + 1, i, rr, o->G0, p->Gt0, p->G0t);
            zdouble tB = e->dop ? abtau_dop(e, &e->B, i, rr, p->Gt0, p->G0t)
                                : abtau(e, &e->B, r + 1, i, rr, o->G0, p->Gt0, p->G0t);
            double dA = creal(tA * conj(tA)), dB = creal(tB * conj(tB));
            prob *= pow(dA, e->lambda) * pow(dB, 1.0 - e->lambda);
            m->n_propose++;
            if (fabs(u - prob) < ORC_NEAR_TIE) m->n_ties++;
            if (u < prob) {
                m->n_accept++;
                e->A.detg *= dA;
                e->B.detg *= dB;
                gm_update(&e->A);
                gm_update(&e->B);
                g4_update(e, p, i, rr);
                *sv = (uint8_t)sx;
            }
        }
    }
}

/* One forward slice of ctrl_SCEEicr (wraps of src/tU/SCEE.jl:90-100 and the layer update
 * :128) without node handling; for checks that replay the decisions. */
void orc_scee_forward_slice(orc_scee *e, int lt, uint32_t sweep) {
    orc_tu *m = e->m;
    int n = m->ns;
    zdouble *D[2] = {zz(n), zz(n)}, *Di[2] = {zz(n), zz(n)};
    for (int r = 0; r < e->nrep; ++r)
        for (int i = 0; i < n; ++i) {
            D[r][i] = cisd(m->alpha[lt - 1] * m->eta[e->R[r].s[IX(i, lt - 1, n)] - 1]);
            Di[r][i] = 1.0 / D[r][i];
        }
    for (int r = 0; r < e->nrep; ++r) orc_tu_WrapKV(m, e->R[r].Gt0, D[r], 1, 'L');
    for (int r = 0; r < e->nrep; ++r) orc_tu_WrapKV(m, e->R[r].Gt, D[r], 1, 'B');
    for (int r = 0; r < e->nrep; ++r) orc_tu_WrapKV(m, e->R[r].G0t, Di[r], 1, 'R');
    update_layer(e, lt, 0, sweep);
    for (int r = 0; r < 2; ++r) { free(D[r]); free(Di[r]); }
}

/* chain set-up of ctrl_SCEEicr (src/tU/SCEE.jl:49-86) */
void orc_scee_init(orc_scee *e) {
    orc_tu *m = e->m;
    int n = m->ns, np = m->np, NN = m->nn_nodes;
    for (int idx = 1; idx <= NN - 1; ++idx)
        for (int r = 0; r < e->nrep; ++r) blocks(e, r, idx);
    for (int r = 0; r < e->nrep; ++r) {
        g4rep *p = &e->R[r];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < np; ++i) RBL(p, NN)[IX(i, j, np)] = conj(m->Pt[IX(j, i, n)]);
        memcpy(RBR(p, 1), m->Pt, sizeof(zdouble) * (size_t)n * np);
    }
    for (int i = 1; i <= NN - 1; ++i)
        for (int r = 0; r < e->nrep; ++r) {
            update_BL(e, r, NN - i);
            update_BR(e, r, i + 1);
        }
    e->idx = 1;
    get_ABGM(e, 1, 1);
}

/* ctrl_SCEEicr sweeps (src/tU/SCEE.jl:87-251), or ctrl_SCDOPicr sweeps
 * (src/tU/disorder_operate.jl:63-188). O receives one value per sweep. */
static void wraps(orc_scee *e, int lt, int forward, zdouble *D, zdouble *Di) {
    orc_tu *m = e->m;
    int n = m->ns;
    const double sg = forward ? 1.0 : -1.0;
    /* "L" with D, "B" with D (which the reference then mutates to 1/D), "R" with 1/D, each
     * matrix for every replica before the next */
    for (int r = 0; r < e->nrep; ++r)
        for (int i = 0; i < n; ++i) {
            D[r * n + i] = cisd(sg * m->alpha[lt - 1] * m->eta[e->R[r].s[IX(i, lt - 1, n)] - 1]);
            Di[r * n + i] = 1.0 / D[r * n + i];
        }
    for (int r = 0; r < e->nrep; ++r) orc_tu_WrapKV(m, e->R[r].Gt0, D + r * n, forward, 'L');
    for (int r = 0; r < e->nrep; ++r) orc_tu_WrapKV(m, e->R[r].Gt, D + r * n, forward, 'B');
    for (int r = 0; r < e->nrep; ++r) orc_tu_WrapKV(m, e->R[r].G0t, Di + r * n, forward, 'R');
}

void orc_scee_sweeps(orc_scee *e, int sweeps, double *O) {
    orc_tu *m = e->m;
    int n = m->ns, nt = m->nt, NN = m->nn_nodes, thidx = NN / 2 + 1;
    zdouble *D = zz(2 * (size_t)n), *Di = zz(2 * (size_t)n);
    orc_scee_init(e);
    for (int sw = 0; sw < sweeps; ++sw) {
        uint32_t sweep = m->sweep;
        double tmpO = 0.0;
        int counter = 0;
        for (int lt = 1; lt <= nt; ++lt) {
            wraps(e, lt, 1, D, Di);
            update_layer(e, lt, 0, sweep);
            tmpO += pow(e->A.detg / e->B.detg, 1.0 / e->nlam);
            counter++;
            int node = 0;
            for (int q = 0; q < NN; ++q) node |= (m->nodes[q] == lt);
            if (node) {
                e->idx += 1;
                int idx = e->idx;
                for (int r = 0; r < e->nrep; ++r) blocks(e, r, idx - 1);
                for (int i = idx; i <= (thidx > idx ? thidx : idx); ++i)
                    for (int r = 0; r < e->nrep; ++r) update_BR(e, r, i);
                for (int i = idx - 1; i >= (thidx < idx ? thidx : idx); --i)
                    for (int r = 0; r < e->nrep; ++r) update_BL(e, r, i);
                get_ABGM(e, idx, 1);
            }
        }
        for (int lt = nt; lt >= 1; --lt) {
            update_layer(e, lt, 1, sweep);
            tmpO += pow(e->A.detg / e->B.detg, 1.0 / e->nlam);
            counter++;
            int node = 0;
            for (int q = 0; q < NN; ++q) node |= (m->nodes[q] == lt - 1);
            if (node) {
                e->idx -= 1;
                int idx = e->idx;
                for (int r = 0; r < e->nrep; ++r) blocks(e, r, idx);
                for (int i = idx; i >= (thidx < idx ? thidx : idx); --i)
                    for (int r = 0; r < e->nrep; ++r) update_BL(e, r, i);
                for (int i = idx + 1; i <= (thidx > idx ? thidx : idx); ++i)
                    for (int r = 0; r < e->nrep; ++r) update_BR(e, r, i);
                get_ABGM(e, idx, 0);
            } else {
                wraps(e, lt, 0, D, Di);
            }
        }
        O[sw] = tmpO / counter;
        m->sweep += 1;
    }
    free(D); free(Di);
}
