/*
 * kapdqmc.h — C ABI of the MI355X projector-DQMC library (libkapdqmc_hip.so).
 *
 * The library replaces the numeric bodies of the reference's Julia entry points for the
 * projector-DQMC inner loop. Every call returns an int status (KDQ_OK or a negative
 * code); kdq_last_error() gives the message. A handle owns one HIP stream and a batch of
 * independent Markov chains that advance together; calls on distinct handles may run on
 * different host threads, a handle itself is not re-entrant.
 *
 * Data crossing the boundary uses the reference's layouts: column-major matrices, field
 * arrays as UInt8 with values 1..4, complex numbers interleaved (ComplexF64).
 */
#ifndef KAPDQMC_H
#define KAPDQMC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KDQ_OK 0
#define KDQ_EARG (-1)    /* invalid argument or shape */
#define KDQ_EHIP (-2)    /* HIP runtime error (allocation, launch, device) */
#define KDQ_ENUM (-3)    /* numerical failure (singular overlap, non-finite G) */
#define KDQ_ESTATE (-4)  /* call not valid in the handle's current state */

enum { KDQ_MODEL_TU = 0, KDQ_MODEL_TV = 1, KDQ_MODEL_TUV = 2 };
enum { KDQ_LATTICE_HC120 = 0, KDQ_LATTICE_HC60 = 1, KDQ_LATTICE_SQ90 = 2, KDQ_LATTICE_SQ45 = 3 };
enum { KDQ_PREC_FP64 = 0, KDQ_PREC_MIXED = 1 };

/* Model constants of one chain: the fields of tU_Hubbard_Para_ (src/tU/model.jl:14-37),
 * tV_Hubbard_Para_ (src/tV/model.jl:3-28) or tUV_Hubbard_Para_ (src/tUV/model.jl:3-27)
 * that the sweep reads. All matrices are column-major, ns x ns (Pt: ns x ns/2). */
typedef struct kdq_model_desc {
    int32_t kind;            /* KDQ_MODEL_* */
    int32_t ns;              /* model.Ns */
    int32_t nt;              /* model.Nt */
    int32_t n_nodes;         /* length(model.nodes) */
    const int64_t *nodes;    /* model.nodes */
    int32_t lattice;         /* KDQ_LATTICE_* (model.Lattice) */
    int32_t site[2];         /* model.site */
    double ht;               /* model.Ht */
    const double *alpha;     /* model.α, length nt */
    const double *eta;       /* model.η, length 4 */
    const double *gamma;     /* model.γ, length 4 */
    const double *K;         /* model.K (real part) */
    const double *eK;        /* model.eK */
    const double *eKinv;     /* model.eKinv */
    const double *halfeK;    /* model.HalfeK */
    const double *halfeKinv; /* model.HalfeKinv */
    const double *Pt;        /* model.Pt (real part) */
    /* t-V and t-U-V only (ignored for t-U) */
    int32_t z;               /* coordination number (3 on the honeycomb lattice) */
    const int32_t *bonds;    /* model.nnidx as 1-based (x, y) pairs in column-major order:
                                t-V: ns/2 x z bonds (family j = column), t-U-V: ns x z directed bonds */
    double a;                /* t-U-V: model.a */
    int32_t is_complex;      /* t-U-V: model.Type == ComplexF64 */
    double U, V;             /* t-U-V: model.U, model.V (observable prefactors) */
    /* t-U with complex hopping (flux != 0, src/public/Geometry.jl:178-198, stored ComplexF64 in
     * tU_Hubbard_Para_, src/tU/model.jl:24-34): imaginary parts of K, eK, eKinv, HalfeK,
     * HalfeKinv (ns x ns) and Pt (ns x ns/2), the real parts above. All six NULL: real
     * matrices (flux = 0). All chains of a batch are real or all complex; complex matrices
     * need FP64 precision. */
    const double *K_im, *eK_im, *eKinv_im, *halfeK_im, *halfeKinv_im, *Pt_im;
} kdq_model_desc;

/* ---- host model construction (no GPU work) ----
 * The keyword constructors tU_Hubbard_Para (src/tU/model.jl:39-96), tV_Hubbard_Para
 * (src/tV/model.jl:30-122) and tUV_Hubbard_Para (src/tUV/model.jl:30-118) with the lattice
 * geometry they call (nn2idx, nnK_Matrix, nnidx_F: src/public/Geometry.jl:8-207), so that a
 * C or C++ caller fills the kdq_model_desc of kdq_create without Julia. flux = 0. */
enum { KDQ_INITIAL_H0 = 0, KDQ_INITIAL_V = 1 };
typedef struct kdq_model_params {
    int32_t kind;            /* KDQ_MODEL_* */
    int32_t lattice;         /* KDQ_LATTICE_* (t-V, t-U-V: honeycomb only) */
    int32_t site[2];         /* site */
    int32_t batch_size;      /* BatchSize */
    int32_t initial;         /* Initial: KDQ_INITIAL_H0 ("H0") or KDQ_INITIAL_V ("V") */
    double ht;               /* Ht */
    double c1, c2;           /* t-U: Hu1, Hu2; t-V: Hv1, Hv2; t-U-V: Hu, Hv */
    double dt;               /* Δt */
    double theta_relax;      /* t-U, t-V: Θrelax; t-U-V: Θ */
    double theta_quench;     /* t-U, t-V: Θquench; t-U-V: ignored */
} kdq_model_params;
typedef struct kdq_model_s *kdq_model;
/* Builds the constants on the host and fills *desc with pointers into storage that *out
 * owns until kdq_model_free. The two LAPACK.syevd! calls of each constructor are one
 * symmetric eigensolver here (Householder tridiagonalisation, implicit-shift QR);
 * eigenvector signs and bases inside degenerate levels may differ from LAPACK's, which
 * changes no matrix function of K and no projector onto the trial state's span. The
 * reference's @asserts (equal couplings without a quench, symmetric coupling profile,
 * eK·eKinv = I, U² > 4z²V²) return KDQ_EARG with their message in kdq_last_error(NULL). */
int kdq_model_build(const kdq_model_params *p, kdq_model *out, kdq_model_desc *desc);
int kdq_model_free(kdq_model m);

/* Per-chain run settings. The random stream is Philox4x32-10 keyed by `seed`, with the
 * chain id and the sweep counter in the counter words, so a run is bit-reproducible
 * and resumable from (fields, sweep counter). */
typedef struct kdq_run_desc {
    uint64_t seed;
    uint32_t chain;          /* stream id (< 2^24) */
    uint32_t sweep0;         /* sweep counter of the first sweep */
    int32_t precision;       /* KDQ_PREC_FP64, or KDQ_PREC_MIXED (t-U, see kdq_get_precision_error) */
    int32_t delay;           /* local-update block size k: sites (t-U, 1..64) or field elements
                                (t-V / t-U-V, 1..32); 1 = one Sherman-Morrison update per flip */
} kdq_run_desc;

typedef struct kdq_handle_s *kdq_handle;

/* Creates a batch of `nchains` chains on the current HIP device. All chains of a batch
 * must share the model kind, precision and delay; ns, couplings, seeds, nt and nodes may
 * differ. Chains of equal (nt, nodes) advance together in the same launches; a batch with
 * several (nt, nodes) groups keeps one stream per group, and kdq_phy_update sweeps the
 * groups concurrently (one host thread each). Chain-indexed calls and per-chain outputs
 * keep the batch's chain order; the per-slice calls (kdq_wrap, kdq_update_layer,
 * kdq_measure, kdq_measure_noalloc) and kdq_slmc_move need one nt and return KDQ_EARG on
 * such a batch. Swap-replica handles (kdq_scee_create, kdq_dop_create) need one nt.
 * Replaces: the buffer set-up at the top of phy_update (src/tU/phy_update.jl:7-37). */
int kdq_create(const kdq_model_desc *models, const kdq_run_desc *runs, int32_t nchains,
               kdq_handle *out);
int kdq_destroy(kdq_handle h);
const char *kdq_last_error(kdq_handle h);   /* h may be NULL: last global error */
int kdq_num_chains(kdq_handle h);

/* Number of observables per measurement row: 6 (t-U), 10 (t-V), 7 (t-U-V). */
int kdq_num_obs(kdq_handle h);

/* Field arrays (UInt8, values 1..4, column-major) of chain `c`: t-U s[ns, nt],
 * t-V s[ns/2, 3, nt], t-U-V s[ns*z, nt].
 * set/get correspond to passing `s` into and out of phy_update (phy_update.jl:3,151). */
int kdq_set_fields(kdq_handle h, int32_t c, const uint8_t *s);
int kdq_get_fields(kdq_handle h, int32_t c, uint8_t *s);
/* Initial_s (src/tU/GreenMatrix.jl:1-8) from the chain's Philox stream. */
int kdq_initial_s(kdq_handle h);

/* Replayed random stream of chain c: the local-update draws of the next `nsweeps` sweeps
 * (from the handle's current sweep counter) come from the caller instead of Philox, so that
 * the reference's own generator (the MersenneTwister of phy_update, src/tU/phy_update.jl:31)
 * can drive the decisions. One draw per proposal, as UpdatePhyLayer! takes them
 * (src/tU/phy_update.jl:259-262, src/tV/phy_update.jl:194-200, src/tUV/phy_update.jl:192-198):
 * choice = the index (0, 1, 2) of `rand(rng, model.samplers_dict[s[i]])` among the other three
 * field values in increasing order, u = the `rand(rng)` the acceptance compares with.
 * Layout: entry ((s * 2 + pass) * nt + lt - 1) * n_elem + e, s = sweep since this call,
 * pass 0 forward / 1 backward, e = the field element's linear index within the slice (the
 * kdq_set_fields layout), n_elem = ns (t-U), 3 ns/2 (t-V), ns z (t-U-V). The reference visits
 * the forward pass lt = 1..nt and the backward pass lt = nt..1, elements in order, except t-V
 * (families 3, 2, 1 forward and 1, 2, 3 backward). The arrays are copied. nsweeps = 0 returns
 * the chain to its Philox stream. kdq_phy_update / kdq_update_layer calls that would reach
 * beyond the replayed sweeps return KDQ_EARG; choices above 2 are KDQ_EARG. Sweep handles
 * only. */
int kdq_set_replay(kdq_handle h, int32_t c, const uint8_t *choice, const double *u, int64_t nsweeps);

/* phy_update(path, model, s, Sweeps, record) for every chain (t-U: src/tU/phy_update.jl:3-152,
 * t-V: src/tV/phy_update.jl:1-188, t-U-V: src/tUV/phy_update.jl:1-186).
 * rows (may be NULL when record == 0) receives nchains x sweeps x kdq_num_obs doubles averaged
 * per sweep, the row the reference appends to its CSV: t-U [Ek Eu CDW0 CDW1 SDW0 SDW1]
 * (phy_update.jl:144), t-V [Ek Ev R0[4] R1[4]] (:178), t-U-V [Ek Eu Ev CDW0 CDW1 SDW0 SDW1]
 * (:178). The caller keeps the CSV writing. */
int kdq_phy_update(kdq_handle h, int32_t sweeps, int32_t record, double *rows);

/* ---- fine-grained entry points (parity checks, custom drivers) ---- */
/* Current equal-time G of chain c as ns x ns ComplexF64 (interleaved, column-major). */
int kdq_get_G(kdq_handle h, int32_t c, double *G);
int kdq_set_G(kdq_handle h, int32_t c, const double *G);
/* The full wrap of one slice as the sweep applies it: t-U / t-U-V WrapKV!(…, "B")
 * (src/tU/GreenMatrix.jl:56-86, src/tUV/GreenMatrix.jl:92-122); t-V the kinetic wrap and the
 * three WrapV!(…, "B") bond rotations (src/tV/phy_update.jl:60-71 and :131-142). */
int kdq_wrap(kdq_handle h, int32_t lt, int32_t forward);
/* UpdatePhyLayer! at slice lt (src/tU/phy_update.jl:256-266, src/tUV/phy_update.jl:188-203;
 * t-V: all three bond families in sweep order, src/tV/phy_update.jl:190-205) with draws from
 * the Philox stream of sweep `sweep` and pass direction dir (0 forward, 1 backward). */
int kdq_update_layer(kdq_handle h, int32_t lt, int32_t dir, uint32_t sweep);
/* Rebuilds the U_L chain and sets G = get_G!(BL[1], BR[1]) (phy_update.jl:39-50). */
int kdq_init_chain(kdq_handle h);
/* phy_measure(model, lt, s, G) of the current G (src/tU/phy_update.jl:154-254,
 * src/tV/phy_update.jl:207-314, src/tUV/phy_update.jl:205-355); out receives
 * nchains x kdq_num_obs doubles. */
int kdq_measure(kdq_handle h, int32_t lt, double *out);
/* kdq_measure without any device allocation. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The per-chain working matrices of the measurement (the
 * copy of G wrapped to Nt/2 and its wrap temporary, 4 ns² doubles per chain), their operand
 * tables and the accumulators live in a static device array of the library, one per GPU, of
 * at most 32 MiB; a batch that needs more returns KDQ_EARG (use kdq_measure). The handle's
 * device must be current. Every call on one GPU shares that GPU's array (calls from several
 * threads or handles that overlap on one GPU race); each call clears the accumulators it
 * reads. Same observables and layout as kdq_measure (sweep handles only). */
int kdq_measure_noalloc(kdq_handle h, int32_t lt, double *out);

/* Time-displaced Green's functions of chain c from its current fields (t-U), replacing
 * G4(model, s, τ1, τ2) (src/tU/GreenMatrix.jl:126-226): G1 = G(τ1), G2 = G(τ2),
 * G12 = G(τ1, τ2), G21 = G(τ2, τ1), each ns x ns ComplexF64 (interleaved, column-major),
 * 0 <= τ1, τ2 <= Nt; τ1 == τ2 gives G, G, -(I - G), G as the reference does. Blocks of
 * BatchSize slices between the two times with the equal-time G at every block boundary,
 * orthonormalised chains; the sweep state (chains, G) is not changed. */
int kdq_G4(kdq_handle h, int32_t c, int32_t tau1, int32_t tau2, double *G1, double *G2, double *G12, double *G21);

/* ---- failure detection (no counterpart in the reference beyond its commented checks) ----
 * At every stabilisation node of a sweep the G recomputed from the chain (get_G!) is compared
 * with the wrapped and updated G it replaces: the reference's commented wrap-error check
 * (src/tU/phy_update.jl:54-58,92-97). out receives nchains x 4 doubles: [max |G_wrapped -
 * G_recomputed| over the nodes since the last call (FP64 handles; mixed precision reports its
 * error through kdq_get_precision_error), non-finite G (1/0), singular overlap (1/0),
 * propagator lost rank (1/0)]; the maximum is reset by the call. A chain whose G turns
 * non-finite (a non-finite model constant or field value) is flagged and the call that saw it
 * returns KDQ_ENUM naming it, after the other chains finished their sweeps unaffected (rows of
 * the others are valid); kdq_set_fields clears the chain's flags. Sweep handles. */
int kdq_get_health(kdq_handle h, double *out);
/* Projector check of every chain's current G (any slice): out receives nchains x 2 doubles,
 * [max |G² - G|, |tr G - (ns - ns/2)|] (G is the projector I - U_R (U_L U_R)⁻¹ U_L). */
int kdq_check_G(kdq_handle h, double *out);

/* Counters: out[0] = proposals, out[1] = accepted, per chain (nchains x 2). */
int kdq_get_counters(kdq_handle h, double *out);

/* Near ties, per chain (nchains doubles; swap-replica handles: one per replica entry, chain c
 * of replica r at c + r * nchains): Metropolis decisions whose uniform draw u and acceptance
 * probability p differ by less than 1e-12. Those are the decisions that rounding differences
 * between two correct FP64 implementations (this library, the Julia reference on OpenBLAS)
 * may flip; all other decisions are compared one by one. Reference: the accept test
 * `rand(rng) < p` of UpdatePhyLayer! (src/tU/phy_update.jl:262, src/tV/phy_update.jl:200,
 * src/tUV/phy_update.jl:198) and UpdateSCEELayer! (src/tU/SCEE.jl:292,313). */
int kdq_get_near_ties(kdq_handle h, double *out);

/* Run options (no counterpart in the reference):
 *  KDQ_OPT_GRAPH      1 (default): kdq_phy_update captures one sweep into a HIP graph (on
 *                     the first call, or when `record` changes) and replays it; 0: launches
 *                     every kernel from the host each sweep. While kernel timing is on
 *                     (kdq_profile mask non-zero) sweeps run with host launches.
 *                     Sweeps of more than 20000 launches (small blocks on large lattices)
 *                     also run with host launches.
 *  KDQ_OPT_DELAY      decision block size k of t-U chains (1..64; 1..16 with replicas); t-V
 *                     replica chains: bonds per delayed block (8 or 4, 1..3: element by element).
 *  KDQ_OPT_GEMM_TILE  0 (default: by occupancy), 32 or 64: GEMM output tile edge.
 *  KDQ_OPT_DEFER      t-U chains. 0 (default): every block of k sites rewrites G (rank-|A|
 *                     flush). 1: a slice's local updates keep G0 and low-rank factors U, V
 *                     (the current G is G0 - U V); each block reads its corrected k x k
 *                     submatrix and appends its rank-|A| update to U and V, and G is rewritten
 *                     once per slice by one GEMM of the accumulated rank. Same Markov chain;
 *                     G agrees to rounding.
 *  KDQ_OPT_PIPE       t-U chains with the per-block flush, in captured sweeps. 1 (default):
 *                     the decisions of block b+1 run on a second stream beside block b's
 *                     flush; they read their k x k submatrix of the updated G from the G
 *                     before that flush, block b's accepted sites and its W, so they do not
 *                     wait for the flush to rewrite G. 0: decisions and flushes alternate on
 *                     one stream. Same Markov chain; G agrees to rounding. Sweeps with host
 *                     launches (graph option 0, kernel timing on, kdq_update_layer) always
 *                     use the one-stream order.
 *  KDQ_OPT_FLUSH      form of the rank-|A| rewrite of G after a block of decisions (every
 *                     model, swap-replica four-matrix flushes included). 0 (default): two
 *                     launches (the product W (E_A - G[A,:]) and the gathered columns into
 *                     scratch, then the rewrite). 1: one launch, a workgroup per strip of 32
 *                     columns forms its part of the product and streams the strip's rows
 *                     (faster alone, slower beside the pipelined decisions). Identical results.
 *  KDQ_OPT_ASIDE      t-U chains, captured sweeps. 0 (default): the stabilisation node forms
 *                     its block's chain products. 1: the product of each updated slice into
 *                     the stabilised chain (U_R forward, U_L backward) runs on a third stream
 *                     as soon as the slice's fields are final, beside the next slices' wraps
 *                     and updates, and the node waits for it. Identical results.
 *  KDQ_OPT_CIRC       1 (default): when every chain's eK and eKinv are block-circulant over
 *                     its L x L periodic cells (ns = 2 L², checked at kdq_create to 1e-12 of
 *                     the largest entry; true for every lattice the model constructors build),
 *                     the kinetic products of the wraps and the chain propagation run through
 *                     cell momenta (length-L DFTs and 2 x 2 blocks per momentum) instead of
 *                     dense GEMMs. 0: dense GEMMs. Results agree to rounding.
 *  KDQ_OPT_QSPLIT     t-U sweeps. 1 (default): at a stabilisation node the
 *                     recomputed G is formed from the chain after the first (shifted)
 *                     Cholesky-QR pass, already well conditioned and of the same span, while
 *                     the second pass, which restores orthonormality for the next block's
 *                     propagation, runs beside it on a second stream. 0: both passes, then G.
 *                     G agrees to rounding; kdq_get_logdet then refers to the first-pass
 *                     factors.
 *  KDQ_OPT_MASIDE     t-U sweeps (not while kernel timing is on). 1 (default): a measurement's
 *                     products and observable tiles run on a second stream beside the next
 *                     slices (its copy of G and any wrap reading fields the sweep may still
 *                     change stay in line); the next measurement and the end of the sweep wait
 *                     for it. 0: in line. Identical results.
 *  KDQ_OPT_GINV       The overlap solve of get_G! (X = (U_L U_R)⁻¹ U_L). 1 (default): the inverse
 *                     by Gauss-Jordan elimination with partial pivoting held in registers of
 *                     one workgroup per chain, then a GEMM (Np <= 144; larger Np take the LU).
 *                     0: blocked LU with partial pivoting and its triangular solves. G agrees
 *                     to rounding.
 *  KDQ_OPT_QINV       Each Cholesky-QR pass of the orthonormalisation. 1 (default): the Cholesky
 *                     factor of UᴴU and its inverse in one register pass of one workgroup per
 *                     chain, then U L⁻ᴴ as a GEMM into a second buffer (Np <= 144). 0: the
 *                     blocked Cholesky and the in-place triangular solve. Same span; the
 *                     orthonormal basis agrees to rounding.
 *  KDQ_OPT_PREWRAP    Captured t-U FP64 sweeps with pipelined blocks. 1 (default): the wrap
 *                     before a layer runs as its two products and the layer's first block is
 *                     decided between them, from the first product and the dense kinetic factor
 *                     (its G[S,S] rounds differently from the second product's), beside the
 *                     second product. 0: the first block after the whole wrap.
 *  KDQ_OPT_MCHAIN     t-U sweep measurements with block-circulant factors. 1 (default): G_sym =
 *                     HalfeK B(c, lt) G B(c, lt)⁻¹ HalfeK⁻¹ as two chained cell-momentum passes
 *                     (every slice factor in LDS, one launch per side). 0: products with the
 *                     slot matrices built slice by slice. The rows agree to rounding. */
enum { KDQ_OPT_GRAPH = 0, KDQ_OPT_DELAY = 1, KDQ_OPT_GEMM_TILE = 2, KDQ_OPT_DEFER = 3, KDQ_OPT_PIPE = 4,
       KDQ_OPT_FLUSH = 5, KDQ_OPT_ASIDE = 6, KDQ_OPT_CIRC = 7, KDQ_OPT_QSPLIT = 8, KDQ_OPT_MASIDE = 9,
       KDQ_OPT_GINV = 10, KDQ_OPT_QINV = 11, KDQ_OPT_PREWRAP = 12, KDQ_OPT_MCHAIN = 13 };
int kdq_set_option(kdq_handle h, int32_t key, int32_t value);

/* Mixed precision (kdq_run_desc.precision = KDQ_PREC_MIXED, t-U): G is carried in FP32
 * between stabilisations (FP32 MFMA wraps, FP32 rank-k flushes; the Metropolis ratio, W,
 * the chains, the orthonormalisation and the overlap solve stay FP64) and replaced by the
 * FP64 recompute at every node. out[c] = max |G_FP32 - G_FP64| over the recomputes since the
 * last call (0 for FP64 handles); the call resets it. No counterpart in the reference. */
int kdq_get_precision_error(kdq_handle h, double *out);

/* log W(s) of every chain's current fields (t-U): 2 log|det(P† B(2Θ,0) P)| + Σ log γ(s),
 * the weight whose ratios the sweep samples (nchains doubles). Used by the self-learning
 * global moves; no counterpart in the reference. Leaves G to be rebuilt by the next sweep. */
int kdq_log_weight(kdq_handle h, double *out);

/* log|det(U_L U_R)| of the overlap factorised by the last get_G! of every chain (nchains
 * doubles). No counterpart in the reference, whose get_G! (src/public/Gupdate.jl:139-157)
 * discards the LU diagonal; the weight of a configuration up to the chain normalisation. */
int kdq_get_logdet(kdq_handle h, double *out);

/* ---- swap-replica Rényi-2 entropy (SCEE) of the t-U model ----
 * Regions A and B and the λ interpolation of ctrl_SCEEicr (src/tU/SCEE.jl:1):
 * indexA / indexB are the reference's 1-based site lists (area_index), λ and Nλ its
 * arguments. */
typedef struct kdq_scee_desc {
    int32_t nA;
    const int64_t *indexA;
    int32_t nB;
    const int64_t *indexB;
    double lambda;
    int32_t nlam;
    int32_t pad;
} kdq_scee_desc;
/* Creates `nchains` two-replica chains; run settings as in kdq_create (delay: sites per
 * decision block, 1..16). Replaces the buffer set-up of ctrl_SCEEicr (src/tU/SCEE.jl:1-48). */
int kdq_scee_create(const kdq_model_desc *models, const kdq_run_desc *runs, const kdq_scee_desc *regions,
                    int32_t nchains, kdq_handle *out);
/* Fields of replica 0 / 1 of chain c (ss[1], ss[2] of ctrl_SCEEicr), layout as kdq_set_fields. */
int kdq_scee_set_fields(kdq_handle h, int32_t c, int32_t replica, const uint8_t *s);
int kdq_scee_get_fields(kdq_handle h, int32_t c, int32_t replica, uint8_t *s);
/* The chain set-up of ctrl_SCEEicr (src/tU/SCEE.jl:49-86): block products, both chains,
 * get_ABGM! at the first node. */
int kdq_scee_init(kdq_handle h);
/* ctrl_SCEEicr sweeps (src/tU/SCEE.jl:87-251, set-up included): O receives nchains x sweeps
 * values, per sweep the mean over slices of (det g_A / det g_B)^(1/Nλ), the quantity the
 * reference writes per sweep. */
int kdq_scee_sweeps(kdq_handle h, int32_t sweeps, double *O);
/* G(τ), G(Θ), G(τ,Θ), G(Θ,τ) of replica `replica` of chain c (ns x ns ComplexF64 each). */
int kdq_scee_get_G4(kdq_handle h, int32_t c, int32_t replica, double *Gt, double *G0, double *Gt0, double *G0t);
/* |det g_A|² and |det g_B|² (nchains x 2; 1 for an empty region). */
int kdq_scee_detg(kdq_handle h, double *out);

/* Disorder operator of the t-U model (ctrl_SCDOPicr, src/tU/disorder_operate.jl:1): `nchains`
 * one-replica chains weighted by W · |det X_A|^{2λ} |det X_B|^{2(1-λ)} with
 * X = e^{iα} I + (1 - e^{iα}) G(Θ)[region] (get_ABX!, :202-223; DOPBuffer(alpha, index) of
 * both regions). regions: as kdq_scee_create (λ, Nλ, 1-based site lists). The kdq_scee_*
 * calls apply with replica 0 only: kdq_scee_sweeps returns per sweep the mean over slices of
 * (det X_A / det X_B)^(1/Nλ) (:96-98), kdq_scee_detg |det X_A|², |det X_B|². The field draws
 * from Philox stream 1. Replaces the buffer set-up of ctrl_SCDOPicr (:1-48). */
int kdq_dop_create(const kdq_model_desc *models, const kdq_run_desc *runs, const kdq_scee_desc *regions,
                   double alpha, int32_t nchains, kdq_handle *out);

/* ---- kernel timing (no counterpart in the reference; used by the benchmark) ----
 * Launches of the kernel families whose bit is set in `mask` are bracketed by HIP events
 * on the handle's stream; kdq_profile_read returns the summed device time (ms) and the
 * launch count per family (arrays of KDQ_PROF_FAMILIES). kdq_profile resets both. */
enum {
    KDQ_PROF_WRAP = 0,       /* the two GEMMs of each sweep WrapKV! */
    KDQ_PROF_CHAIN = 1,      /* per-slice propagation of U_R / U_L between nodes */
    KDQ_PROF_GETG = 2,       /* GEMMs of get_G! */
    KDQ_PROF_DECIDE = 3,     /* block Metropolis decisions */
    KDQ_PROF_FLUSH = 4,      /* rank-k flush of accepted flips */
    KDQ_PROF_QR = 5,         /* orthonormalisation of U_R / U_L (Gram GEMM, Cholesky, solve) */
    KDQ_PROF_INV = 6,        /* overlap LU + solve + log|det| */
    KDQ_PROF_MEAS_WRAP = 7,  /* wraps and HalfeK sandwich of phy_measure */
    KDQ_PROF_MEASURE = 8,    /* observable reduction */
    KDQ_PROF_GEMM_OTHER = 9,
    KDQ_PROF_FAMILIES = 10
};
int kdq_profile(kdq_handle h, uint32_t mask);
int kdq_profile_read(kdq_handle h, double *ms, int64_t *launches);

/* ---- self-learning global moves (t-U; no counterpart in the reference) ----
 * A translation-equivariant convolutional effective model E(s) of the HS field on an
 * L1 x L2 honeycomb lattice with nt slices (3 x 3 x 3 convolutions over (slice, y, x) with
 * 8 -> C -> C -> 1 channels, tanh; periodic in x and y, zero-padded in the slice direction),
 * trained on (field, log W) pairs so that E + log W is as constant as possible.
 * Precision: the parameters (kdq_net_get_params / set_params, Adam) are FP32 masters; the
 * kernels read the weights W1, W2, W3 rounded to BF16 (nearest even) and accumulate in FP32
 * with FP32 biases, so E and its gradient are those of the network with the rounded weights. */
typedef struct kdq_net_s *kdq_net;
int kdq_net_create(int32_t L1, int32_t L2, int32_t nt, int32_t channels, uint64_t seed, int32_t max_batch,
                   kdq_net *out);
int kdq_net_destroy(kdq_net n);
const char *kdq_net_last_error(kdq_net n);
int kdq_net_num_params(kdq_net n);
int kdq_net_get_params(kdq_net n, float *out);      /* [W1 C x 8 x 27, b1 C, W2 C x C x 27, b2 C, W3 C x 27] */
int kdq_net_set_params(kdq_net n, const float *in);
/* E(s) of `batch` field arrays (batch x ns x nt, each as kdq_set_fields takes it). */
int kdq_net_energy(kdq_net n, const uint8_t *fields, int32_t batch, double *E);
/* loss = Var_b(E(s_b) + logw_b) and its gradient (grad may be NULL); batch >= 2. */
int kdq_net_loss_grad(kdq_net n, const uint8_t *fields, const double *logw, int32_t batch, double *loss,
                      float *grad);
/* One Adam step (β = 0.9, 0.999) on that loss. */
int kdq_net_train_step(kdq_net n, const uint8_t *fields, const double *logw, int32_t batch, float lr,
                       double *loss);
/* Inner chain alone: nsteps local steps on exp(-E) from each of the `batch` configurations in
 * `fields` (changed in place; Philox stream 3, key = seed, chain = batch index). One step
 * picks a slice class r in [0, 7) and proposes one single-field change (uniform site, uniform
 * new value among the other three) in every slice t = r + 7 j. E receives E of the final
 * fields, accepted the number of accepted changes (either may be NULL). */
int kdq_net_sample(kdq_net n, uint8_t *fields, int32_t batch, int32_t nsteps, uint64_t seed, double *E,
                   int32_t *accepted);
/* One global move per chain: nsteps inner-chain steps on exp(-E) from the current fields,
 * then acceptance with probability min(1, W(s')/W(s) e^{E(s')-E(s)}) using the exact FP64
 * weights of kdq_log_weight (stream 4); accepted[c] = 1 or 0. The local sweep's G is rebuilt
 * by the next kdq_phy_update. Chains must fit the network's batch and lattice. */
int kdq_slmc_move(kdq_handle h, kdq_net net, int32_t nsteps, double *accepted);

#ifdef __cplusplus
}
#endif
#endif
