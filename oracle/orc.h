/*
 * ORACLE — test infrastructure only.
 *
 * FP64 CPU restatement of the KAPDQMC (Julia) projector-DQMC algorithm, used as the
 * parity checker for the HIP library. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product path never calls it.
 *
 * Conventions: all matrices are column-major (as in Julia), complex matrices are
 * interleaved `double complex` (as Julia's ComplexF64). Field values are 1-based
 * (1..4) exactly as the reference's UInt8 arrays. Every function cites the
 * reference file:line it restates.
 *
 * Parity status per function (checks live in tests/test_oracle_*.py):
 *   geometry            pinned: K symmetric, coordination, sublattice structure,
 *                       hand-checked 3x3 neighbour tables
 *   jacobi eigensolver  pinned: numpy.linalg.eigh
 *   householder QR, LU  pinned: numpy/scipy (QR span, LU solve residual, det)
 *   philox4x32-10       pinned: Random123 known-answer vectors
 *   model constants     pinned: closed forms (HS identity, eK*eKinv = I)
 *   wrap / B blocks     pinned: brute-force unstabilised products (Gtau)
 *   get_r / Gupdate     pinned: brute-force determinant ratio + recomputed G
 *   get_G               pinned: projector identities, free-fermion closed form
 *   phy_measure         pinned: hand-evaluated formulas on a fixed G (incl. the
 *                       reference's discarded `+adjoint(...)` statements)
 *   phy_update          pinned: composition of the above + G vs Gtau after sweeps
 */
#ifndef KDQ_ORACLE_H
#define KDQ_ORACLE_H
#include <complex.h>
#include <stdint.h>

/* A decision whose uniform draw u and acceptance probability p differ by less than this is a
 * near tie: rounding differences between two correct implementations can flip it. */
#define ORC_NEAR_TIE 1e-12

typedef double complex zdouble;

/* ---------------- random numbers (shared specification with the HIP library) ------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* one proposal draw: choice in {0,1,2} and uniform in [0,1) */
void orc_draw(uint64_t seed, uint32_t chain, uint32_t stream, uint32_t sweep, uint32_t slice_tag,
              uint32_t elem, int *choice, double *u);
/* initial-field value in {1,2,3,4} */
int orc_init_value(uint64_t seed, uint32_t chain, uint32_t stream, uint32_t slice, uint32_t elem);

/* ---------------- geometry (src/public/Geometry.jl) ---------------------------------- */
enum { ORC_HC120 = 0, ORC_HC60 = 1, ORC_SQ90 = 2, ORC_SQ45 = 3 };
int orc_xy_i(const int site[2], int x, int y);                 /* Geometry.jl:8-13 */
void orc_i_xy(const int site[2], int i, int *x, int *y);       /* Geometry.jl:18-21 */
int orc_nn2idx(int lattice, const int site[2], int idx, int nn[4]); /* Geometry.jl:74-134 */
void orc_nnK(int lattice, const int site[2], double *K);        /* Geometry.jl:178-207 */
void orc_nnidx_F(int lattice, const int site[2], int *nnidx);  /* Geometry.jl:24-66 (HC) */
int orc_area_index(int lattice, const int site[2], const int a0[2], const int a1[2], int *index);

/* ---------------- dense linear algebra ---------------------------------------------- */
void orc_zgemm(char ta, char tb, int m, int n, int k, zdouble alpha, const zdouble *A, int lda,
               const zdouble *B, int ldb, zdouble beta, zdouble *C, int ldc);
void orc_jacobi_eigh(int n, const double *A, double *w, double *V);
void orc_zqr_q(int m, int n, zdouble *A, int lda, double *logabsr);   /* geqrf+orgqr */
int orc_zgetrf(int n, zdouble *A, int lda, int *ipiv);                  /* getrf */
void orc_zgetrs(int n, int nrhs, const zdouble *A, int lda, const int *ipiv, zdouble *B, int ldb);
zdouble orc_zdet(int n, const zdouble *A, int lda);
int orc_zinv(int n, zdouble *A, int lda);                               /* getrf+getri */
void orc_get_G(int ns, int np, const zdouble *BL, int ldbl, const zdouble *BR, int ldbr,
               zdouble *G, double *logabsdet);

/* Replayed proposal draws: entry ((s * 2 + dir) * nt + lt - 1) * n_elem + e is the draw of
 * field element e at slice lt of pass dir (0 forward, 1 backward) of sweep sweep0 + s; the
 * caller fills it in the reference's visiting order (UpdatePhyLayer!: sampler, then uniform).
 * u == NULL: the Philox stream. */
typedef struct orc_replay {
    const uint8_t *choice;
    const double *u;
    uint32_t sweep0, nsweeps;
} orc_replay;
void orc_sweep_draw(const orc_replay *r, int nt, int n_elem, uint64_t seed, uint32_t chain, uint32_t sweep,
                    int lt, int dir, uint32_t elem, int *choice, double *u);

/* ---------------- t-U model (src/tU) ------------------------------------------------- */
typedef struct orc_tu {
    int ns, np, nt, bs, nn_nodes, lattice, site[2];
    double ht, dt;
    int *nodes;             /* [nn_nodes] */
    double *alpha;          /* [nt] */
    double eta[4], gamma[4];
    zdouble *K, *eK, *eKinv, *HalfeK, *HalfeKinv, *Pt;   /* ns x ns, Pt ns x np */
    uint8_t *s;             /* ns x nt, values 1..4 */
    zdouble *G, *BLs, *BRs; /* G ns x ns; BLs np x ns x NN; BRs ns x np x NN */
    zdouble *tmpNN, *tmpNN2, *BM, *tmpNn, *tmpnN, *tmpN;
    int idx;                /* current node index (1-based, as the reference) */
    uint64_t seed;
    uint32_t chain, sweep;  /* RNG stream position */
    long long n_accept, n_propose;
    long long n_ties;       /* decisions with |u - p| < ORC_NEAR_TIE */
    orc_replay rp;
} orc_tu;

orc_tu *orc_tu_new(int ns, int nt, int bs, int nn_nodes, const int64_t *nodes, int lattice,
                   const int site[2], double ht, double dt, const double *alpha,
                   const double *eta, const double *gamma, const double *K, const double *eK,
                   const double *eKinv, const double *HalfeK, const double *HalfeKinv,
                   const double *Pt, uint64_t seed, uint32_t chain);
void orc_tu_set_imag(orc_tu *m, const double *K_im, const double *eK_im, const double *eKinv_im,
                     const double *HalfeK_im, const double *HalfeKinv_im, const double *Pt_im);
void orc_tu_free(orc_tu *m);
void orc_tu_initial_s(orc_tu *m);
void orc_tu_BM_F(orc_tu *m, int idx, zdouble *BM);
void orc_tu_BMinv_F(orc_tu *m, int idx, zdouble *BM);
void orc_tu_WrapKV(orc_tu *m, zdouble *G, const zdouble *D_in, int forward, char LR);
void orc_tu_update_layer(orc_tu *m, int lt, int dir);
void orc_tu_set_replay(orc_tu *m, const uint8_t *choice, const double *u, uint32_t nsweeps);
void orc_tu_measure(orc_tu *m, int lt, const zdouble *G, double out[6]);
void orc_tu_init_chain(orc_tu *m);
void orc_tu_phy_update(orc_tu *m, int sweeps, int record, double *rows);
void orc_tu_Gtau(orc_tu *m, int tau, zdouble *out);
void orc_tu_forward_slices(orc_tu *m, int nslices);

/* ---------------- spinless t-V model (src/tV) ---------------------------------------- */
typedef struct orc_tv {
    int ns, np, nt, bs, nn_nodes, lattice, site[2], nb;  /* nb = ns/2 bonds per family */
    double ht;
    int *nodes;
    double *alpha;
    double eta[4], gamma[4];
    int *bx, *by;           /* [nb*3] 0-based sites of bond (i, family j) at i + j*nb */
    double *K, *eK, *eKinv, *HalfeK, *HalfeKinv, *Pt, *UV;  /* UV: ns x ns x 3 */
    uint8_t *s;             /* [nb, 3, nt] column-major */
    double *G, *BLs, *BRs, *T, *T2, *BM, *tNn, *tnN, *tN;
    int idx;
    uint64_t seed;
    uint32_t chain, sweep;
    long long n_accept, n_propose, n_negative;
    long long n_ties;
    orc_replay rp;
} orc_tv;

orc_tv *orc_tv_new(int ns, int nt, int bs, int nn_nodes, const int64_t *nodes, int lattice,
                   const int site[2], double ht, const double *alpha, const double *eta,
                   const double *gamma, const int *bonds, const double *K, const double *eK,
                   const double *eKinv, const double *HalfeK, const double *HalfeKinv,
                   const double *Pt, uint64_t seed, uint32_t chain);
void orc_tv_free(orc_tv *m);
void orc_tv_initial_s(orc_tv *m);
void orc_tv_BM_F(orc_tv *m, int idx, double *BM);
void orc_tv_update_family(orc_tv *m, int lt, int dir, int j);
void orc_tv_set_replay(orc_tv *m, const uint8_t *choice, const double *u, uint32_t nsweeps);
void orc_tv_measure(orc_tv *m, int lt, const double *G, double out[10]);
void orc_tv_init_chain(orc_tv *m);
void orc_tv_phy_update(orc_tv *m, int sweeps, int record, double *rows);
void orc_tv_Gtau(orc_tv *m, int tau, double *out);

/* ---------------- t-U-V model (src/tUV) ---------------------------------------------- */
typedef struct orc_tuv {
    int ns, np, nt, bs, nn_nodes, lattice, site[2], nbd;   /* nbd = ns*z directed bonds */
    double ht, U, V, a;
    int is_complex;
    int *nodes;
    double eta[4], gamma[4];
    int *bx, *by;           /* [nbd] 0-based (x, y) of directed bond i (linear index of nnidx) */
    zdouble *K, *eK, *eKinv, *HalfeK, *HalfeKinv, *Pt;
    uint8_t *s;             /* [nbd, nt] */
    zdouble *G, *BLs, *BRs, *T, *T2, *BM, *tNn, *tnN, *tN;
    int idx;
    uint64_t seed;
    uint32_t chain, sweep;
    long long n_accept, n_propose;
    long long n_ties;
    orc_replay rp;
} orc_tuv;

orc_tuv *orc_tuv_new(int ns, int nt, int bs, int nn_nodes, const int64_t *nodes, int lattice,
                     const int site[2], double ht, double U, double V, double a, int is_complex,
                     int z, const int *bonds, const double *eta, const double *gamma, const double *K,
                     const double *eK, const double *eKinv, const double *HalfeK,
                     const double *HalfeKinv, const double *Pt, uint64_t seed, uint32_t chain);
void orc_tuv_free(orc_tuv *m);
void orc_tuv_initial_s(orc_tuv *m);
void orc_tuv_BM_F(orc_tuv *m, int idx, zdouble *BM);
void orc_tuv_update_layer(orc_tuv *m, int lt, int dir);
void orc_tuv_set_replay(orc_tuv *m, const uint8_t *choice, const double *u, uint32_t nsweeps);
void orc_tuv_measure(orc_tuv *m, int lt, const zdouble *G, double out[7]);
void orc_tuv_init_chain(orc_tuv *m);
void orc_tuv_phy_update(orc_tuv *m, int sweeps, int record, double *rows);

/* ---------------- t-U swap-replica Rényi-2 (src/tU/SCEE.jl) --------------------------- */
typedef struct orc_scee orc_scee;
orc_scee *orc_scee_new(orc_tu *m, int nA, const int *indexA, int nB, const int *indexB, double lambda,
                       int nlam, const uint8_t *s1, const uint8_t *s2);
void orc_scee_free(orc_scee *e);
void orc_scee_init(orc_scee *e);
void orc_scee_sweeps(orc_scee *e, int sweeps, double *O);
void orc_scee_forward_slice(orc_scee *e, int lt, uint32_t sweep);
void orc_scee_get_fields(orc_scee *e, int r, uint8_t *s);
void orc_scee_get_G4(orc_scee *e, int r, zdouble *Gt, zdouble *G0, zdouble *Gt0, zdouble *G0t);
double orc_scee_detg(orc_scee *e, int which);
/* Disorder operator ctrl_SCDOPicr of the t-U model (src/tU/disorder_operate.jl:1-247): the
 * orc_scee driver on one replica with X_A = e^{iα}I + (1-e^{iα})G0[A,A]; the functions above
 * apply (replica 0; orc_scee_detg returns |det X_A|², |det X_B|²). */
orc_scee *orc_dop_new(orc_tu *m, double alpha, int nA, const int *indexA, int nB, const int *indexB,
                      double lambda, int nlam, const uint8_t *s);

/* ---------------- t-V / t-U-V swap-replica Rényi-2 (src/tV/SCEE.jl, src/tUV/SCEE.jl) ---
 * kind 1: t-V (model = orc_tv *), kind 2: t-U-V (model = orc_tuv *); fields in the model's
 * layout ([nb, 3, nt] or [nbd, nt]). Parity: G4 and the Grover determinants against
 * brute-force products; every decision of a forward slice replayed with the full weight. */
typedef struct orc_scee2 orc_scee2;
orc_scee2 *orc_scee2_new(int kind, void *model, int nA, const int *indexA, int nB, const int *indexB,
                         double lambda, int nlam, const uint8_t *s1, const uint8_t *s2);
void orc_scee2_free(orc_scee2 *e);
void orc_scee2_init(orc_scee2 *e);
void orc_scee2_sweeps(orc_scee2 *e, int sweeps, double *O);
void orc_scee2_forward_slice(orc_scee2 *e, int lt, uint32_t sweep);
void orc_scee2_get_fields(orc_scee2 *e, int r, uint8_t *s);
void orc_scee2_get_G4(orc_scee2 *e, int r, zdouble *Gt, zdouble *G0, zdouble *Gt0, zdouble *G0t);
double orc_scee2_detg(orc_scee2 *e, int which);

#endif
