// Host side of the library: device memory layout per chain, the sweep driver and the
// C ABI declared in include/kapdqmc.h.
//
// The driver follows phy_update (src/tU/phy_update.jl:3-152) step for step; every step is
// a batched launch over all chains of the handle. Device memory per chain (FP64):
//   eK, eKinv, HalfeK, HalfeKinv, K : N x N real
//   G[2], T, Gm, Tm                 : N x N complex planar (G ping-pongs across flushes)
//   UR[NN], ULH[NN], P0, P1         : N x Np complex planar (U_L is stored as U_Lᴴ)
//   M (Np x Np), X (Np x N)         : complex planar, get_G! workspace
//   s                               : N x Nt UInt8
#include "../../include/kapdqmc.h"
#include "kdq_kernels.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

std::mutex g_err_mu;
std::string g_err;

void set_global_error(const std::string &s) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = s;
}

struct ChainHost {
    int ns = 0, np = 0;
    std::vector<void *> allocs;
    double *eK, *eKinv, *hK, *hKinv, *K, *Pt, *alpha;
    double *G[2], *T, *Gm, *Tm, *P0, *P1, *P2, *M, *X, *S, *QT;
    float *G32[2] = {nullptr, nullptr}, *T32 = nullptr, *eK32 = nullptr, *eKinv32 = nullptr;   // mixed precision
    double *FZ = nullptr, *FC = nullptr;     // flush scratch: W·Y (64 x N) and gathered columns (N x 64)
    float *FZ32 = nullptr, *FC32 = nullptr;
    double *DU = nullptr, *DV = nullptr;     // deferred-flush factors of a t-U slice (N x N each)
    double *DI = nullptr;                    // inverses of the 16 x 16 diagonal blocks (Np x 32)
    double *S2 = nullptr, *DI2 = nullptr;    // Gram matrix and block inverses of a concurrent second pass
    double *circ = nullptr;                  // Λ(k) of eK and eK⁻¹ (block-circulant kinetic factors)
    int cL = 0;                              // their cell edge (0: dense products)
    float *DU32 = nullptr, *DV32 = nullptr;
    int *rk = nullptr, *rank = nullptr;
    int *ex = nullptr, *ey = nullptr, *inc = nullptr;
    std::vector<double *> UR, ULH;
    uint8_t *s;
    int *acc_cnt, *acc_idx, *info, *pw_fld;
    double *W, *counters, *meas, *logdet;
};

struct Table {  // one logical operand across chains
    Opnd *d = nullptr;
};

}  // namespace

// The host model builder (kdq_model.cpp) reports through the same global message.
void kdq_internal_set_error(const std::string &msg) { set_global_error(msg); }

struct kdq_handle_s {
    int nch = 0, nt = 0, nn = 0, max_ns = 0, max_np = 0, delay = 16;
    int kind = KDQ_MODEL_TU, nobs = 6, max_elem = 0, max_nb = 0;
    bool cK = false;   // complex model matrices (t-U with flux): eK, eK⁻¹, HalfeK, HalfeK⁻¹, K, P
    int circ_L = 0;    // largest cell edge when every chain's eK, eK⁻¹ are block-circulant (else 0)
    bool circ = true;  // kinetic products through cell momenta where circ_L > 0 (KDQ_OPT_CIRC)
    // rank-2 block tables (t-V: per family, t-U-V: one group)
    BlockDesc *d_blocks = nullptr;
    int *d_block_offset = nullptr, *d_nblocks = nullptr;
    int max_blocks = 0;
    std::vector<int64_t> nodes;
    std::vector<ChainHost> ch;
    std::vector<kdq_model_desc> md;
    std::vector<kdq_run_desc> rd;
    std::vector<ChainDev> cd_host;   // host copy of d_chains
    struct SceeState *scee = nullptr;   // swap-replica state (kdq_scee_create handles only)
    hipStream_t stream = nullptr;
    ChainDev *d_chains = nullptr;
    // operand tables
    Opnd *t_eK, *t_eKinv, *t_hK, *t_hKinv, *t_Pt, *t_G[2], *t_T, *t_Gm, *t_Tm, *t_P0, *t_P1, *t_P2, *t_M, *t_X, *t_S;
    Opnd *t_FZ = nullptr, *t_FC = nullptr;     // flush scratch operands
    OpndF *t_FZ32 = nullptr, *t_FC32 = nullptr;
    Opnd *t_DU = nullptr, *t_DV = nullptr;     // deferred-flush factors (t-U)
    Opnd *t_DI = nullptr;                      // diagonal-block inverses of the Np x Np factors
    Opnd *t_S2 = nullptr, *t_DI2 = nullptr;    // the same for a second Cholesky-QR pass on the aux stream
    Opnd *t_QT = nullptr;                      // N x Np: the first Cholesky-QR pass's result (register factor path)
    bool qinv = true;                          // Cholesky-QR by the register factor-inverse + GEMM (Np <= 144)
    bool qsplit = true;                        // t-U captured sweeps: that second pass beside get_G!
    bool maside = true;                        // t-U sweeps: a measurement's products and tiles on aux
    bool ginv = true;                          // get_G!: register Gauss-Jordan inverse + GEMM (Np <= 144)
    bool prewrap = true;                       // pipelined t-U layers: block 0 decided beside the wrap's 2nd product
    bool mchain = true;                        // t-U sweep measurements as chained cell-momentum passes
    bool mchain_ok = true;                     // every chain's HalfeK, HalfeK⁻¹ block-circulant too
    bool pw_next = false;                      // the next layer's block 0 was launched with the wrap before it
    double *d_pwpart = nullptr;                // prewrap partial products: [chain][N/32][2][32][32]
    bool mpending = false;                     // such a measurement not yet joined
    hipEvent_t ev_mfork = nullptr, ev_mdone = nullptr;
    hipEvent_t ev_q1 = nullptr, ev_q2 = nullptr;
    OpndF *t_DU32 = nullptr, *t_DV32 = nullptr;
    bool defer = false;                        // t-U slices flush once (U, V factors)
    std::vector<Opnd *> t_UR, t_ULH;
    std::vector<void *> dev_misc;
    double *d_logdet = nullptr;   // [nch] log|det(U_L U_R)| of the last get_G
    double *d_lw = nullptr;       // [nch] accumulator of log|det R| for kdq_log_weight
    // self-learning global moves: trial / backup fields, device pointer table, RNG ids,
    // inner-chain scratch, cached log W of the current fields
    std::vector<uint8_t *> sl_trial, sl_backup;
    uint8_t **d_sl_ptrs = nullptr;
    uint64_t *d_sl_seed = nullptr;
    uint32_t *d_sl_chain = nullptr;
    double *d_sl_E = nullptr;
    std::vector<double> sl_logw;
    bool sl_logw_valid = false;
    int sl_move = 0;
    int *d_info = nullptr;        // [nch] sticky: nonzero after a singular overlap
    int *d_qinfo = nullptr;       // [nch] sticky: nonzero after a rank-deficient propagator
    int *d_poison = nullptr;      // [nch] sticky: nonzero after a non-finite G at a stabilisation node
    double *d_wraperr = nullptr;  // [nch] max |G_wrapped - G_recomputed| at the nodes since the last read
    double *d_mpart = nullptr;    // tile partial sums of the measurement
    int *d_piv = nullptr;         // [nch][max_np] LU pivots of the overlap
    int cur = 0;           // which G buffer is current
    int idx = 1;           // current node index (1-based, as the reference)
    uint32_t sweep = 0;    // sweep counter of the next sweep
    std::string err;
    double *h_meas = nullptr;  // pinned host staging
    // per-family HIP-event timing (kdq_profile): launches of enabled families are
    // bracketed by events on the handle's stream and harvested in batches.
    uint32_t prof_mask = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double prof_ms[KDQ_PROF_FAMILIES] = {};
    int64_t prof_n[KDQ_PROF_FAMILIES] = {};
    // mixed precision: FP32 copy of G between stabilisations (t-U)
    int mixed = 0, cur32 = 0;
    OpndF *t_G32[2] = {nullptr, nullptr}, *t_T32 = nullptr, *t_eK32 = nullptr, *t_eKinv32 = nullptr;
    double *d_maxerr = nullptr;   // [nch] max |G_FP32 - G_FP64| at the recomputes since the last read
    // options (kdq_set_option) and the captured sweep
    int use_graph = 1, gemm_tile = 0;
    bool pipe = true;                         // t-U decisions of block b+1 beside block b's flush
    int flush1 = 0;                           // 1: the one-launch (strip) flush
    // time-displaced G (kdq_G4) scratch, allocated on first use: left chain per block
    // boundary, and N x N work matrices
    std::vector<Opnd *> g4_ul;
    Opnd *g4_m[6] = {};
    // t-U measurements of a sweep: slot pairs (X_k, X̃_k) of the products that take G to the
    // symmetric slice (slot 0: HalfeK, HalfeK⁻¹), which recursion they hold and how far
    std::vector<Opnd *> mq, mqt;
    int mq_dir = 0, mq_have = 0;
    hipStream_t side = nullptr;               // their stream
    hipEvent_t ev_app = nullptr, ev_dec[2] = {nullptr, nullptr};   // their dependencies
    bool aside = false;                       // t-U chain products beside the sweep (captured sweeps)
    hipStream_t aux = nullptr;                // their stream
    hipEvent_t ev_fields = nullptr, ev_chain = nullptr;   // a slice's fields final / a block's products done
    Opnd *chain_cur = nullptr;                // the last product of the block in progress
    uint32_t *d_sweep = nullptr;              // device sweep counter read by graph replays
    const uint32_t *sweep_dev = nullptr;      // non-null while a sweep is being captured
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_record = -1;
    std::vector<void *> rp_buf;               // [nch] replayed draws (kdq_set_replay), nullptr: Philox
    // A batch whose chains differ in nt or nodes: one sub-handle per (nt, nodes) group, each
    // with its own stream; this handle only routes calls. gch[g] lists the batch indices of
    // group g's chains, gmap[c] = (group, index within the group).
    std::vector<kdq_handle> grp;
    std::vector<std::vector<int>> gch;
    std::vector<std::pair<int, int>> gmap;

    ~kdq_handle_s() {
        if (stream) hipStreamSynchronize(stream);
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        for (auto &c : ch)
            for (void *p : c.allocs) hipFree(p);
        for (void *p : dev_misc) hipFree(p);
        for (void *p : rp_buf)
            if (p) hipFree(p);
        if (h_meas) hipHostFree(h_meas);
        for (auto &p : ev_pending) {
            hipEventDestroy(p.second.first);
            hipEventDestroy(p.second.second);
        }
        for (hipEvent_t e : ev_pool) hipEventDestroy(e);
        if (side) hipStreamSynchronize(side);
        if (aux) hipStreamSynchronize(aux);
        if (ev_app) hipEventDestroy(ev_app);
        for (hipEvent_t e : ev_dec)
            if (e) hipEventDestroy(e);
        if (ev_fields) hipEventDestroy(ev_fields);
        if (ev_chain) hipEventDestroy(ev_chain);
        if (ev_q1) hipEventDestroy(ev_q1);
        if (ev_q2) hipEventDestroy(ev_q2);
        if (ev_mfork) hipEventDestroy(ev_mfork);
        if (ev_mdone) hipEventDestroy(ev_mdone);
        if (side) hipStreamDestroy(side);
        if (aux) hipStreamDestroy(aux);
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            (void)hipGetLastError(); /* do not leave the error sticky for later checks */ \
            h->err = std::string(#expr) + ": " + hipGetErrorString(e_);                  \
            set_global_error(h->err);                                                    \
            return KDQ_EHIP;                                                             \
        }                                                                                \
    } while (0)

int fail(kdq_handle h, int code, const std::string &msg) {
    if (h) h->err = msg;
    set_global_error(msg);
    return code;
}

hipEvent_t take_event(kdq_handle h) {
    if (!h->ev_pool.empty()) {
        hipEvent_t e = h->ev_pool.back();
        h->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

// Moves finished event pairs into the accumulators (synchronises on the last one).
void harvest(kdq_handle h) {
    if (h->ev_pending.empty()) return;
    hipEventSynchronize(h->ev_pending.back().second.second);
    for (auto &p : h->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            h->prof_ms[p.first] += ms;
            h->prof_n[p.first] += 1;
        }
        h->ev_pool.push_back(p.second.first);
        h->ev_pool.push_back(p.second.second);
    }
    h->ev_pending.clear();
}

// Brackets the launches of one kernel family with events on the handle's stream (eager
// launches only: a sweep is not captured into a graph while a family is timed).
struct ProfScope {
    kdq_handle h;
    int fam;
    hipEvent_t a = nullptr;
    ProfScope(kdq_handle h_, int fam_) : h(h_), fam(fam_) {
        if (h->prof_mask & (1u << fam)) {
            a = take_event(h);
            hipEventRecord(a, h->stream);
        }
    }
    ~ProfScope() {
        if (!a) return;
        hipEvent_t b = take_event(h);
        hipEventRecord(b, h->stream);
        h->ev_pending.push_back({fam, {a, b}});
        if (h->ev_pending.size() >= 4096) harvest(h);
    }
};

template <class T>
hipError_t dalloc(ChainHost &c, T **p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(T) * (count ? count : 1));
    if (e == hipSuccess) {
        c.allocs.push_back(q);
        hipMemset(q, 0, sizeof(T) * (count ? count : 1));
    }
    *p = (T *)q;
    return e;
}

int make_table_f(kdq_handle h, OpndF **out, const std::vector<OpndF> &v) {
    void *q = nullptr;
    HIPCHK(hipMalloc(&q, sizeof(OpndF) * v.size()));
    h->dev_misc.push_back(q);
    HIPCHK(hipMemcpy(q, v.data(), sizeof(OpndF) * v.size(), hipMemcpyHostToDevice));
    *out = (OpndF *)q;
    return KDQ_OK;
}

int make_table(kdq_handle h, Opnd **out, const std::vector<Opnd> &v) {
    void *q = nullptr;
    HIPCHK(hipMalloc(&q, sizeof(Opnd) * v.size()));
    h->dev_misc.push_back(q);
    HIPCHK(hipMemcpy(q, v.data(), sizeof(Opnd) * v.size(), hipMemcpyHostToDevice));
    *out = (Opnd *)q;
    return KDQ_OK;
}

Opnd cplx(double *p, int ld, size_t plane) { return Opnd{p, p + plane, ld, 0}; }
Opnd real(double *p, int ld) { return Opnd{p, nullptr, ld, 0}; }

// ----------------------------------------------------------------- building blocks
GemmArgs gargs(kdq_handle h, const Opnd *A, const Opnd *B, const Opnd *C, int md, int nd, int kd) {
    GemmArgs g{};
    g.chains = h->d_chains;
    g.A = A; g.B = B; g.C = C; g.Cin = nullptr;
    g.mdim = md; g.ndim = nd; g.kdim = kd; g.mode = 0;
    return g;
}

int maxdim(kdq_handle h, int code) { return code == DIM_NS ? h->max_ns : h->max_np; }

int gemm(kdq_handle h, bool ac, bool bc, int opa, int opb, const GemmArgs &g, int fam = KDQ_PROF_GEMM_OTHER) {
    ProfScope ps(h, fam);
    HIPCHK(launch_gemm(h->stream, h->nch, maxdim(h, g.mdim), maxdim(h, g.ndim), maxdim(h, g.kdim), ac, bc, opa, opb, g,
                       h->gemm_tile));
    return KDQ_OK;
}

// dst (N x Np, complex) <- the trial state P (real, or complex for complex hopping)
hipError_t load_P(kdq_handle h, Opnd *dst) {
    if (h->cK) return launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, h->t_Pt, dst, DIM_NS, DIM_NP, true);
    return launch_real_to_cplx(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, h->t_Pt, dst, DIM_NS, DIM_NP);
}

// A GEMM with a model constant (eK, eK⁻¹, HalfeK, HalfeK⁻¹, P) as operand A (amodel) or B
// (bmodel): real for flux = 0, complex for complex hopping (the kinetic factors are then
// Hermitian), where a transposed model operand becomes its conjugate transpose.
int gemm_m(kdq_handle h, bool amodel, bool ac, bool bmodel, bool bc, int opa, int opb, const GemmArgs &g,
           int fam = KDQ_PROF_GEMM_OTHER) {
    if (amodel && h->cK) {
        ac = true;
        if (opa == 1) opa = 2;
    }
    if (bmodel && h->cK) {
        bc = true;
        if (opb == 1) opb = 2;
    }
    return gemm(h, ac, bc, opa, opb, g, fam);
}

int gemm_f(kdq_handle h, bool ac, bool bc, int opa, int opb, const GemmArgsF &g, int fam) {
    ProfScope ps(h, fam);
    HIPCHK(launch_gemm_f32(h->stream, h->nch, maxdim(h, g.mdim), maxdim(h, g.ndim), maxdim(h, g.kdim), ac, bc, opa, opb,
                           g, h->gemm_tile));
    return KDQ_OK;
}

// ---- kinetic products: through cell momenta (k_circ) when eK, eK⁻¹ are block-circulant ----
// t-U only. The t-V and t-U-V measurements wrap a copy of G over up to a block of slices
// without stabilisation (src/tV/phy_update.jl:207-314), which amplifies any difference in the
// last bits of the kinetic factors by ~1e5 on some observables; through cell momenta they act
// as exactly circulant matrices, which the dense eK from the eigen-decomposition is only to
// ~1e-15, so those models keep the dense products that round like the reference's.
bool use_circ(kdq_handle h) { return h->circ && h->circ_L > 0 && h->kind == KDQ_MODEL_TU; }

// grid width and LDS of a k_circ launch over every chain's lines (ldim: DIM_NS or DIM_NP)
void circ_launch_shape(kdq_handle h, int ldim, int &gx, size_t &bytes, int &lmax) {
    gx = 1;
    bytes = 0;
    lmax = 1;
    for (int c = 0; c < h->nch; ++c) {
        const int L = h->ch[c].cL, lines = ldim == DIM_NP ? h->ch[c].np : h->ch[c].ns, lpw = circ_lines_per_wg(L);
        gx = std::max(gx, (lines + lpw - 1) / lpw);
        bytes = std::max(bytes, circ_lds_bytes(L));
        lmax = std::max(lmax, L);
    }
}

// Y = line ⊙ (out ⊙ op(E) (in ⊙ X)) over the ldim columns of X, E = eK (mat 0) or eK⁻¹ (mat 1);
// opa as gemm_m (1: transposed, the conjugate transpose for complex hopping). The dense path:
// gemm_m with A = E, rowS = out, inS = in, colS = line.
int kin_left(kdq_handle h, int mat, int opa, const Opnd *X, const Opnd *Y, int ldim, DiagSpec in, DiagSpec out,
             DiagSpec line, int fam) {
    if (use_circ(h)) {
        ProfScope ps(h, fam);
        CircArgs a{h->d_chains, X, Y, mat, opa == 0 ? 0 : (h->cK ? 2 : 1), 0, ldim, in, out, line};
        int gx = 1;
        size_t bytes = 0;
        int lmax = 1;
        circ_launch_shape(h, ldim, gx, bytes, lmax);
        HIPCHK(launch_circ(h->stream, h->nch, gx, bytes, lmax, a));
        return KDQ_OK;
    }
    GemmArgs g = gargs(h, mat ? h->t_eKinv : h->t_eK, X, Y, DIM_NS, ldim, DIM_NS);
    g.rowS = out;
    g.inS = in;
    g.colS = line;
    return gemm_m(h, true, false, false, true, opa, 0, g, fam);
}

// The product of nsteps kin_left factors over the columns of X (cell-momentum path only): step
// s with the diagonal factors of slice lt + s·dlt (in, out, line as given for s = 0), the columns
// kept in LDS between the steps; the same arithmetic as nsteps kin_left launches.
int kin_left_chain(kdq_handle h, int mat, int opa, const Opnd *X, const Opnd *Y, int ldim, DiagSpec in, DiagSpec out,
                   DiagSpec line, int nsteps, int dlt, int fam) {
    ProfScope ps(h, fam);
    CircArgs a{h->d_chains, X, Y, mat, opa == 0 ? 0 : (h->cK ? 2 : 1), 0, ldim, in, out, line, nsteps, dlt};
    int gx = 1;
    size_t bytes = 0;
    int lmax = 1;
    circ_launch_shape(h, ldim, gx, bytes, lmax);
    HIPCHK(launch_circ(h->stream, h->nch, gx, bytes, lmax, a));
    return KDQ_OK;
}

// Y = line ⊙ ((X ⊙ in) E ⊙ out) over the rows of the ns x ns matrix X (line: per row, in / out:
// per column). The dense path: gemm_m with B = E, rowS = line, inS = in, colS = out.
int kin_right(kdq_handle h, int mat, const Opnd *X, const Opnd *Y, DiagSpec in, DiagSpec out, DiagSpec line,
              int fam) {
    if (use_circ(h)) {
        ProfScope ps(h, fam);
        CircArgs a{h->d_chains, X, Y, mat, 0, 1, DIM_NS, in, out, line};
        int gx = 1;
        size_t bytes = 0;
        int lmax = 1;
        circ_launch_shape(h, DIM_NS, gx, bytes, lmax);
        HIPCHK(launch_circ(h->stream, h->nch, gx, bytes, lmax, a));
        return KDQ_OK;
    }
    GemmArgs g = gargs(h, X, mat ? h->t_eKinv : h->t_eK, Y, DIM_NS, DIM_NS, DIM_NS);
    g.rowS = line;
    g.inS = in;
    g.colS = out;
    return gemm_m(h, false, true, true, false, 0, 0, g, fam);
}

GemmArgsF gargsf(kdq_handle h, const OpndF *A, const OpndF *B, const OpndF *C) {
    GemmArgsF g{};
    g.chains = h->d_chains;
    g.A = A; g.B = B; g.C = C; g.Cin = nullptr;
    g.mdim = g.ndim = g.kdim = DIM_NS;
    g.mode = 0;
    return g;
}

// HS diagonal of slice lt for the site-diagonal models (t-U: kind 1, t-U-V: kind 2)
DiagSpec hs_diag(kdq_handle h, int lt, int sign, int conj = 0) {
    return DiagSpec{h->kind == KDQ_MODEL_TU ? 1 : 2, lt, sign, conj};
}

int bondrot(kdq_handle h, Opnd *X, int family, int lt, int sign, int mode, int ncol_dim, int fam_prof) {
    ProfScope ps(h, fam_prof);
    BondRotArgs b{h->d_chains, X, family, lt, sign, mode, ncol_dim, 0};
    HIPCHK(launch_bondrot(h->stream, h->nch, h->max_nb, mode == 0 ? h->max_nb : maxdim(h, ncol_dim), b));
    return KDQ_OK;
}

// t-V kinetic wrap: G <- eK G eKinv (forward) or eKinv G eK (backward)
int wrap_K(kdq_handle h, Opnd *Gt, Opnd *Tt, bool forward, int fam) {
    const DiagSpec none{};
    if (int r = kin_left(h, forward ? 0 : 1, 0, Gt, Tt, DIM_NS, none, none, none, fam)) return r;
    return kin_right(h, forward ? 1 : 0, Tt, Gt, none, none, none, fam);
}

// G <- D eK G eKinv D⁻¹ (forward) or eKinv D⁻¹ G D eK (backward), D = cis(α η_s) at slice lt.
int wrap_B(kdq_handle h, Opnd *Gt, Opnd *Tt, int lt, bool forward, int fam = KDQ_PROF_WRAP) {
    const DiagSpec none{};
    if (forward) {   // T = eK G, then G = D T eKinv D⁻¹
        if (int r = kin_left(h, 0, 0, Gt, Tt, DIM_NS, none, none, none, fam)) return r;
        return kin_right(h, 1, Tt, Gt, none, hs_diag(h, lt, -1), hs_diag(h, lt, +1), fam);
    }
    // D' = D(-): T = eKinv · (D' G) · D'⁻¹, then G = T eK
    if (int r = kin_left(h, 1, 0, Gt, Tt, DIM_NS, hs_diag(h, lt, -1), none, hs_diag(h, lt, +1), fam)) return r;
    return kin_right(h, 0, Tt, Gt, none, none, none, fam);
}

// WrapKV!(…, "B") on the FP32 copy of G (mixed precision): the same two GEMMs as wrap_B in
// FP32 MFMA with FP32 accumulation, HS phases evaluated in FP64 and rounded.
int wrap_B_f32(kdq_handle h, int lt, bool forward, int fam = KDQ_PROF_WRAP) {
    OpndF *Gt = h->t_G32[h->cur32], *Tt = h->t_T32;
    if (forward) {
        GemmArgsF g1 = gargsf(h, h->t_eK32, Gt, Tt);
        if (int r = gemm_f(h, false, true, 0, 0, g1, fam)) return r;
        GemmArgsF g2 = gargsf(h, Tt, h->t_eKinv32, Gt);
        g2.rowS = hs_diag(h, lt, +1);
        g2.colS = hs_diag(h, lt, -1);
        return gemm_f(h, true, false, 0, 0, g2, fam);
    }
    GemmArgsF g1 = gargsf(h, h->t_eKinv32, Gt, Tt);
    g1.inS = hs_diag(h, lt, -1);
    g1.colS = hs_diag(h, lt, +1);
    if (int r = gemm_f(h, false, true, 0, 0, g1, fam)) return r;
    GemmArgsF g2 = gargsf(h, Tt, h->t_eK32, Gt);
    return gemm_f(h, true, false, 0, 0, g2, fam);
}

// Mixed precision: the FP64 G just recomputed from the chain replaces the FP32 copy; with
// track set, max |G_FP32 - G_FP64| is recorded first.
int refresh_g32(kdq_handle h, bool track) {
    if (!h->mixed) return KDQ_OK;
    HIPCHK(launch_to_f32(h->stream, h->nch, h->max_ns, h->d_chains, h->t_G[h->cur], h->t_G32[h->cur32], DIM_NS,
                         track ? h->d_maxerr : nullptr));
    return KDQ_OK;
}

// U <- orthonormal basis of span(U) for the N x Np propagator, in place: two Cholesky-QR
// passes (S = UᴴU, S + σI = L Lᴴ, U <- U L⁻ᴴ), the first shifted by σ = c·u·‖U‖_F².
// one pass on the handle's stream with the Gram / block-inverse buffers S, DI
int ortho_pass(kdq_handle h, Opnd *U, int pass, double *logacc, Opnd *S, Opnd *DI) {
    const double u = 1.1102230246251565e-16;
    const double n = h->max_ns, k = h->max_np;
    const double shift = 11.0 * (n * k + k * (k + 1.0)) * u;
    GemmArgs g = gargs(h, U, U, S, DIM_NP, DIM_NP, DIM_NS);
    if (int r = gemm(h, true, true, 2, 0, g, KDQ_PROF_QR)) return r;
    ProfScope ps(h, KDQ_PROF_QR);
    CholArgs ca{h->d_chains, S, DIM_NP, 0, pass == 0 ? shift : 0.0, h->d_qinfo, logacc, DI};
    HIPCHK(launch_chol(h->stream, h->nch, h->max_np, ca));
    TrsmArgs ta{h->d_chains, U, S, DIM_NS, DIM_NP, DI};
    HIPCHK(launch_trsm_rlh(h->stream, h->nch, h->max_ns, h->max_np, ta));
    return KDQ_OK;
}

// The same pass out of place, dst = src L⁻ᴴ: the register Cholesky factor-inverse (X = L⁻¹
// into S) and one GEMM src · Xᴴ (Np <= 144)
bool qinv_on(kdq_handle h) { return h->qinv && inv_supported(h->max_np); }

int ortho_pass_to(kdq_handle h, Opnd *src, Opnd *dst, int pass, double *logacc, Opnd *S) {
    const double u = 1.1102230246251565e-16;
    const double n = h->max_ns, k = h->max_np;
    const double shift = 11.0 * (n * k + k * (k + 1.0)) * u;
    GemmArgs g = gargs(h, src, src, S, DIM_NP, DIM_NP, DIM_NS);
    if (int r = gemm(h, true, true, 2, 0, g, KDQ_PROF_QR)) return r;
    {
        ProfScope ps(h, KDQ_PROF_QR);
        CholInvArgs ca{h->d_chains, S, S, DIM_NP, 0, pass == 0 ? shift : 0.0, h->d_qinfo, logacc};
        HIPCHK(launch_cholinv(h->stream, h->nch, h->max_np, ca));
    }
    GemmArgs g2 = gargs(h, src, S, dst, DIM_NS, DIM_NP, DIM_NP);
    return gemm(h, true, true, 0, 2, g2, KDQ_PROF_QR);
}

int orthonormalise(kdq_handle h, Opnd *U, double *logacc = nullptr) {
    if (qinv_on(h)) {
        if (int r = ortho_pass_to(h, U, h->t_QT, 0, logacc, h->t_S)) return r;
        return ortho_pass_to(h, h->t_QT, U, 1, logacc, h->t_S);
    }
    for (int pass = 0; pass < 2; ++pass)
        if (int r = ortho_pass(h, U, pass, logacc, h->t_S, h->t_DI)) return r;
    return KDQ_OK;
}

// The second pass beside get_G! (t-U sweeps, captured or not): G = I - U_R (U_L U_R)⁻¹ U_L depends
// only on span(U), and after the shifted first pass U is already well conditioned (κ = O(1)
// for κ(U) well inside u^(-1/2)·10^4, Fukaya et al. 2020), so the node's G is formed from a copy
// of the first-pass factor while the second pass, which only restores orthonormality for the
// next block's propagation, runs on the aux stream with its own Gram and block-inverse
// buffers. Returns the copy (t_P2; with the register factor path the first pass writes t_QT,
// which the second pass maps back to U, so no copy is made); the caller waits for ev_q2
// before U or that buffer is used again.
bool qsplit_on(kdq_handle h) { return h->kind == KDQ_MODEL_TU && h->qsplit && !h->scee && h->aux && h->t_S2; }

Opnd *orthonormalise_split(kdq_handle h, Opnd *U, int &rc) {
    if (qinv_on(h)) {   // first pass U -> t_QT, which get_G! reads while the second maps it back to U
        rc = ortho_pass_to(h, U, h->t_QT, 0, nullptr, h->t_S);
        if (rc) return nullptr;
        hipError_t e = hipEventRecord(h->ev_q1, h->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(h->aux, h->ev_q1, 0);
        if (e != hipSuccess) {
            rc = fail(h, KDQ_EHIP, std::string("orthonormalise_split: ") + hipGetErrorString(e));
            return nullptr;
        }
        std::swap(h->stream, h->aux);
        rc = ortho_pass_to(h, h->t_QT, U, 1, nullptr, h->t_S2);
        std::swap(h->stream, h->aux);
        if (rc) return nullptr;
        if ((e = hipEventRecord(h->ev_q2, h->aux)) != hipSuccess) {
            rc = fail(h, KDQ_EHIP, std::string("orthonormalise_split: ") + hipGetErrorString(e));
            return nullptr;
        }
        return h->t_QT;
    }
    rc = ortho_pass(h, U, 0, nullptr, h->t_S, h->t_DI);
    if (rc) return nullptr;
    hipError_t e = launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, U, h->t_P2, DIM_NS, DIM_NP, true);
    if (e == hipSuccess) e = hipEventRecord(h->ev_q1, h->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(h->aux, h->ev_q1, 0);
    if (e != hipSuccess) {
        rc = fail(h, KDQ_EHIP, std::string("orthonormalise_split: ") + hipGetErrorString(e));
        return nullptr;
    }
    std::swap(h->stream, h->aux);   // the second pass goes to the aux stream
    rc = ortho_pass(h, U, 1, nullptr, h->t_S2, h->t_DI2);
    std::swap(h->stream, h->aux);
    if (rc) return nullptr;
    if ((e = hipEventRecord(h->ev_q2, h->aux)) != hipSuccess) {
        rc = fail(h, KDQ_EHIP, std::string("orthonormalise_split: ") + hipGetErrorString(e));
        return nullptr;
    }
    return h->t_P2;
}

// dst = B(hi) ... B(lo) · src for U_R (N x Np), not orthonormalised (t-U); src may be dst
int prop_R_range(kdq_handle h, Opnd *src, Opnd *dst, int lo, int hi) {
    Opnd *cur = src;
    for (int lt = lo; lt <= hi; ++lt) {
        Opnd *out = (lt == hi && cur != dst) ? dst : ((cur == h->t_P0) ? h->t_P1 : h->t_P0);
        if (int r = kin_left(h, 0, 0, cur, out, DIM_NP, DiagSpec{}, hs_diag(h, lt, +1), DiagSpec{}, KDQ_PROF_CHAIN))
            return r;
        cur = out;
    }
    if (cur != dst) HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, cur, dst, DIM_NS, DIM_NP, true));
    return KDQ_OK;
}

// U_Lᴴ <- (U_L · B(hi) ... B(lo))ᴴ with U_L stored as U_Lᴴ (N x Np), not orthonormalised (t-U);
// src may be dst
int prop_L_range(kdq_handle h, Opnd *src, Opnd *dst, int lo, int hi) {
    Opnd *cur = src;
    for (int lt = hi; lt >= lo; --lt) {
        Opnd *out = (lt == lo && cur != dst) ? dst : ((cur == h->t_P0) ? h->t_P1 : h->t_P0);
        if (int r = kin_left(h, 0, 1, cur, out, DIM_NP, hs_diag(h, lt, +1, 1), DiagSpec{}, DiagSpec{}, KDQ_PROF_CHAIN))
            return r;
        cur = out;
    }
    if (cur != dst) HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, cur, dst, DIM_NS, DIM_NP, true));
    return KDQ_OK;
}

// dst = QR_Q(B_block(idx) · src) for U_R: slices nodes[idx-1]+1 .. nodes[idx] (1-based idx)
int propagate_R(kdq_handle h, Opnd *src, Opnd *dst, int idx, double *logacc = nullptr, bool ortho = true) {
    const int lo = (int)h->nodes[idx - 1] + 1, hi = (int)h->nodes[idx];
    if (h->kind == KDQ_MODEL_TU && use_circ(h) && hi > lo) {   // the block's factors in one launch
        if (int r = kin_left_chain(h, 0, 0, src, dst, DIM_NP, DiagSpec{}, hs_diag(h, lo, +1), DiagSpec{}, hi - lo + 1,
                                   +1, KDQ_PROF_CHAIN))
            return r;
        return ortho ? orthonormalise(h, dst, logacc) : KDQ_OK;
    }
    Opnd *cur = src;
    for (int lt = lo; lt <= hi; ++lt) {
        Opnd *out = (lt == hi) ? dst : ((cur == h->t_P0) ? h->t_P1 : h->t_P0);
        const DiagSpec rs = h->kind != KDQ_MODEL_TV ? hs_diag(h, lt, +1) : DiagSpec{};
        if (int r = kin_left(h, 0, 0, cur, out, DIM_NP, DiagSpec{}, rs, DiagSpec{}, KDQ_PROF_CHAIN)) return r;
        if (h->kind == KDQ_MODEL_TV)   // B(lt) = V1 V2 V3 eK: V3 first
            for (int j = 2; j >= 0; --j)
                if (int r = bondrot(h, out, j, lt, +1, 1, DIM_NP, KDQ_PROF_CHAIN)) return r;
        cur = out;
    }
    if (lo > hi) HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, src, dst, DIM_NS, DIM_NP, true));
    return ortho ? orthonormalise(h, dst, logacc) : KDQ_OK;
}

// dst = QR_Q((U_L · B_block(idx))ᴴ) with U_L stored as U_Lᴴ: U_Lᴴ <- eKᵀ (conj(D_l) ⊙ U_Lᴴ)
// (t-V: U_Lᴴ <- eKᵀ V3 V2 V1 U_Lᴴ)
// for l = nodes[idx] down to nodes[idx-1]+1.
int propagate_L(kdq_handle h, Opnd *src, Opnd *dst, int idx, bool ortho = true) {
    const int lo = (int)h->nodes[idx - 1] + 1, hi = (int)h->nodes[idx];
    if (h->kind == KDQ_MODEL_TU && use_circ(h) && hi > lo) {   // the block's factors in one launch
        if (int r = kin_left_chain(h, 0, 1, src, dst, DIM_NP, hs_diag(h, hi, +1, 1), DiagSpec{}, DiagSpec{}, hi - lo + 1,
                                   -1, KDQ_PROF_CHAIN))
            return r;
        return ortho ? orthonormalise(h, dst) : KDQ_OK;
    }
    Opnd *cur = src;
    for (int lt = hi; lt >= lo; --lt) {
        Opnd *out = (lt == lo) ? dst : ((cur == h->t_P0) ? h->t_P1 : h->t_P0);
        if (h->kind == KDQ_MODEL_TV) {
            // B(lt)ᴴ = eKᵀ V3 V2 V1 (V symmetric, real): rotate rows of a copy, then eKᵀ
            Opnd *tmp = h->t_P2;
            HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, cur, tmp, DIM_NS, DIM_NP, true));
            for (int j = 0; j < 3; ++j)
                if (int r = bondrot(h, tmp, j, lt, +1, 1, DIM_NP, KDQ_PROF_CHAIN)) return r;
            if (int r = kin_left(h, 0, 1, tmp, out, DIM_NP, DiagSpec{}, DiagSpec{}, DiagSpec{}, KDQ_PROF_CHAIN))
                return r;
        } else {
            // conj(D): D⁻¹ for unimodular phases, D for real ones
            if (int r = kin_left(h, 0, 1, cur, out, DIM_NP, hs_diag(h, lt, +1, 1), DiagSpec{}, DiagSpec{},
                                 KDQ_PROF_CHAIN))
                return r;
        }
        cur = out;
    }
    if (lo > hi) HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, src, dst, DIM_NS, DIM_NP, true));
    return ortho ? orthonormalise(h, dst) : KDQ_OK;
}

// Chain products beside the sweep (captured t-U sweeps with the pipelined blocks): the product
// of slice lt into the block's chain — U_R <- B(lt) U_R going forward, U_Lᴴ <- (U_L B(lt))ᴴ
// going backward, the operands and order of propagate_R / propagate_L — is issued on the aux
// stream once the slice's fields are final (after its local updates, so after every decision
// of the slice), and overlaps the wraps and updates of the block's later slices. The node
// waits for the block's last product (chain_join) and orthonormalises on the handle's stream.
// Same products in the same order: identical results to the in-line form.
bool chain_aside(kdq_handle h) {
    return h->kind == KDQ_MODEL_TU && h->aside && h->sweep_dev && !h->prof_mask && !h->scee && h->aux;
}

int chain_step(kdq_handle h, int lt, bool forward) {
    const int b = forward ? h->idx : h->idx - 1;   // the node block holding slice lt
    const int lo = (int)h->nodes[b - 1] + 1, hi = (int)h->nodes[b];
    Opnd *src = forward ? h->t_UR[b - 1] : h->t_ULH[b];
    Opnd *dst = forward ? h->t_UR[b] : h->t_ULH[b - 1];
    const bool first = forward ? lt == lo : lt == hi, last = forward ? lt == hi : lt == lo;
    Opnd *cur = first ? src : h->chain_cur;
    Opnd *out = last ? dst : ((cur == h->t_P0) ? h->t_P1 : h->t_P0);
    HIPCHK(hipEventRecord(h->ev_fields, h->stream));
    HIPCHK(hipStreamWaitEvent(h->aux, h->ev_fields, 0));
    const DiagSpec none{};
    std::swap(h->stream, h->aux);   // the product goes to the aux stream
    const int r = forward ? kin_left(h, 0, 0, cur, out, DIM_NP, none, hs_diag(h, lt, +1), none, KDQ_PROF_CHAIN)
                          : kin_left(h, 0, 1, cur, out, DIM_NP, hs_diag(h, lt, +1, 1), none, none, KDQ_PROF_CHAIN);
    std::swap(h->stream, h->aux);
    if (r) return r;
    h->chain_cur = out;
    if (last) HIPCHK(hipEventRecord(h->ev_chain, h->aux));
    return KDQ_OK;
}

// the node after a block of chain_step products: dst = QR_Q of the block's product
int chain_join(kdq_handle h, Opnd *dst, bool ortho = true) {
    HIPCHK(hipStreamWaitEvent(h->stream, h->ev_chain, 0));
    return ortho ? orthonormalise(h, dst) : KDQ_OK;
}

// get_G!: G = I - U_R (U_L U_R)⁻¹ U_L into the current G buffer
// G = I - U_R (U_L U_R)⁻¹ U_L for explicit chain operands (log|det| not recorded)
// X = M⁻¹ U_L for M in t_M (overwritten): the register Gauss-Jordan inverse and one GEMM
// (Np <= 144, option on), else the blocked LU and its solve
int overlap_solve(kdq_handle h, Opnd *ULH, double *logdet) {
    if (h->ginv && inv_supported(h->max_np)) {
        InvArgs ia{h->d_chains, h->t_M, h->t_M, DIM_NP, 0, logdet, h->d_info};
        HIPCHK(launch_inv(h->stream, h->nch, h->max_np, ia));
        GemmArgs g2 = gargs(h, h->t_M, ULH, h->t_X, DIM_NP, DIM_NS, DIM_NP);
        return gemm(h, true, true, 0, 2, g2, KDQ_PROF_GETG);
    }
    LuArgs la{h->d_chains, h->t_M, DIM_NP, h->max_np, h->d_piv, logdet, h->d_info, nullptr, h->t_DI};
    HIPCHK(launch_lu(h->stream, h->nch, h->max_np, la));
    LuSolveArgs sa{h->d_chains, h->t_M, ULH, h->t_X, h->d_piv, DIM_NP, DIM_NS, h->max_np, 0, h->t_DI};
    HIPCHK(launch_lu_solve(h->stream, h->nch, h->max_np, h->max_ns, sa));
    return KDQ_OK;
}

int get_G_from(kdq_handle h, Opnd *UR, Opnd *ULH, Opnd *dst) {
    GemmArgs g1 = gargs(h, ULH, UR, h->t_M, DIM_NP, DIM_NP, DIM_NS);
    if (int r = gemm(h, true, true, 2, 0, g1, KDQ_PROF_GETG)) return r;
    if (int r = overlap_solve(h, ULH, nullptr)) return r;
    GemmArgs g3 = gargs(h, UR, h->t_X, dst, DIM_NS, DIM_NS, DIM_NP);
    g3.mode = 1;
    return gemm(h, true, true, 0, 0, g3, KDQ_PROF_GETG);
}

int get_G(kdq_handle h, int idx, Opnd *dst = nullptr, Opnd *UR_ = nullptr, Opnd *ULH_ = nullptr) {
    Opnd *ULH = ULH_ ? ULH_ : h->t_ULH[idx - 1], *UR = UR_ ? UR_ : h->t_UR[idx - 1];
    GemmArgs g1 = gargs(h, ULH, UR, h->t_M, DIM_NP, DIM_NP, DIM_NS);
    if (int r = gemm(h, true, true, 2, 0, g1, KDQ_PROF_GETG)) return r;
    {
        ProfScope ps(h, KDQ_PROF_INV);
        if (int r = overlap_solve(h, ULH, h->d_logdet)) return r;
    }
    GemmArgs g3 = gargs(h, UR, h->t_X, dst ? dst : h->t_G[h->cur], DIM_NS, DIM_NS, DIM_NP);
    g3.mode = 1;
    return gemm(h, true, true, 0, 0, g3, KDQ_PROF_GETG);
}

// log W(s) of every chain's current fields (t-U): 2 log|det(P† B(2Θ,0) P)| + Σ log γ(s).
// U_R is rebuilt from P through every node block with the orthonormalisation's discarded
// triangular factors summed (U = Q L₂ᴴ L₁ᴴ, so log|det R| = Σ log L₁,jj + Σ log L₂,jj), and
// the last factor is the LU of P† U_R. The chain buffers are left for the next
// phy_update, which rebuilds them. No counterpart in the reference (its sweep only needs
// ratios).
int log_weight(kdq_handle h, double *out) {
    const int NN = h->nn;
    HIPCHK(hipMemsetAsync(h->d_lw, 0, sizeof(double) * h->nch, h->stream));
    HIPCHK(load_P(h, h->t_UR[0]));
    for (int idx = 1; idx <= NN - 1; ++idx)
        if (int r = propagate_R(h, h->t_UR[idx - 1], h->t_UR[idx], idx, h->d_lw)) return r;
    GemmArgs g = gargs(h, h->t_Pt, h->t_UR[NN - 1], h->t_M, DIM_NP, DIM_NP, DIM_NS);   // Pᵀ U_R
    if (int r = gemm_m(h, true, false, false, true, 1, 0, g, KDQ_PROF_GETG)) return r;
    LuArgs la{h->d_chains, h->t_M, DIM_NP, h->max_np, h->d_piv, h->d_logdet, h->d_info};
    HIPCHK(launch_lu(h->stream, h->nch, h->max_np, la));
    std::vector<double> lr(h->nch), lm(h->nch);
    HIPCHK(hipMemcpyAsync(lr.data(), h->d_lw, sizeof(double) * h->nch, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(lm.data(), h->d_logdet, sizeof(double) * h->nch, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < h->nch; ++c) {
        const size_t ne = (size_t)h->ch[c].ns * h->nt;
        std::vector<uint8_t> sv(ne);
        HIPCHK(hipMemcpy(sv.data(), h->ch[c].s, ne, hipMemcpyDeviceToHost));
        double lg = 0.0;
        for (size_t q = 0; q < ne; ++q) lg += std::log(h->md[c].gamma[sv[q] - 1]);
        out[c] = 2.0 * (lr[c] + lm[c]) + lg;
    }
    return KDQ_OK;
}

// ---- time-displaced Green's functions: G4(model, s, τ1, τ2) of src/tU/GreenMatrix.jl:126-226
// for every chain from its current fields (t-U). Blocks of BatchSize slices between τ2 and τ1
// (the last one shorter); U_R from P through slices 1..τ2 and U_L from Pᴴ through Nt..τ1+1,
// orthonormalised after every block; the left chain at each block boundary is stored, the
// right chain is advanced block by block, and G at each boundary gives
//   G(τ1,τ2) = BB[nb] G[nb] ... BB[1] G[1],  G(τ2,τ1) = -(I - G[1]) BB[1]⁻¹ ... (I - G[nb]) BB[nb]⁻¹
// (G[1] = G(τ2), G[nb+1] = G(τ1)). Works on its own scratch; the chain state and G of the
// sweep are not touched.
int g4_alloc(kdq_handle h) {
    if (!h->g4_ul.empty()) return KDQ_OK;
    std::vector<std::vector<Opnd>> vul(h->nn + 1), vm(6);   // NN left chains and the right one
    for (int c = 0; c < h->nch; ++c) {
        ChainHost &C = h->ch[c];
        const int n = C.ns, np = C.np;
        for (int k = 0; k <= h->nn; ++k) {
            double *p = nullptr;
            if (dalloc(C, &p, 2 * (size_t)n * np) != hipSuccess) return fail(h, KDQ_EHIP, "kdq_G4: device allocation failed");
            vul[k].push_back(cplx(p, n, (size_t)n * np));
        }
        for (int k = 0; k < 6; ++k) {
            double *p = nullptr;
            if (dalloc(C, &p, 2 * (size_t)n * n) != hipSuccess) return fail(h, KDQ_EHIP, "kdq_G4: device allocation failed");
            vm[k].push_back(cplx(p, n, (size_t)n * n));
        }
    }
    h->g4_ul.resize(h->nn + 1);
    for (int k = 0; k <= h->nn; ++k)
        if (int r = make_table(h, &h->g4_ul[k], vul[k])) return r;
    for (int k = 0; k < 6; ++k)
        if (int r = make_table(h, &h->g4_m[k], vm[k])) return r;
    return KDQ_OK;
}

// X <- B(hi) ... B(lo) X (N x N), through tmp
int g4_left(kdq_handle h, Opnd *X, Opnd *tmp, int lo, int hi) {
    Opnd *cur = X, *oth = tmp;
    for (int lt = lo; lt <= hi; ++lt) {
        GemmArgs g = gargs(h, h->t_eK, cur, oth, DIM_NS, DIM_NS, DIM_NS);   // D eK X
        g.rowS = hs_diag(h, lt, +1);
        if (int r = gemm_m(h, true, false, false, true, 0, 0, g)) return r;
        std::swap(cur, oth);
    }
    if (cur != X) HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, cur, X, DIM_NS, DIM_NS, true));
    return KDQ_OK;
}

// Y <- Y B(lo)⁻¹ ... B(hi)⁻¹ (N x N), B⁻¹ = eK⁻¹ D⁻¹, through tmp
int g4_right_inv(kdq_handle h, Opnd *Y, Opnd *tmp, int lo, int hi) {
    Opnd *cur = Y, *oth = tmp;
    for (int lt = lo; lt <= hi; ++lt) {
        GemmArgs g = gargs(h, cur, h->t_eKinv, oth, DIM_NS, DIM_NS, DIM_NS);   // Y eK⁻¹ D⁻¹
        g.colS = hs_diag(h, lt, -1);
        if (int r = gemm_m(h, false, true, true, false, 0, 0, g)) return r;
        std::swap(cur, oth);
    }
    if (cur != Y) HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, cur, Y, DIM_NS, DIM_NS, true));
    return KDQ_OK;
}

// out[0..3] = G(τ1), G(τ2), G(τ1,τ2), G(τ2,τ1) (tables of the scratch matrices), τ1 > τ2
int g4_compute(kdq_handle h, int t1, int t2, Opnd **out) {
    if (int r = g4_alloc(h)) return r;
    int bs = 1;   // BatchSize: the largest gap between stabilisation nodes
    for (size_t k = 1; k < h->nodes.size(); ++k) bs = std::max(bs, (int)(h->nodes[k] - h->nodes[k - 1]));
    const int nb = (t1 - t2 + bs - 1) / bs;
    if (nb + 1 > h->nn) return fail(h, KDQ_EARG, "kdq_G4: more blocks between the slices than the chain holds");
    Opnd **UL = h->g4_ul.data(), **M = h->g4_m;
    Opnd *Gi = M[0], *X12 = M[1], *Y21 = M[2], *T = M[3], *T2 = M[4], *G2 = M[5];
    Opnd *UR = UL[h->nn];   // the right chain (N x Np), advanced block by block
    // U_L at τ1: from Pᴴ through slices Nt .. τ1+1 (orthonormalised every BatchSize slices)
    HIPCHK(load_P(h, UL[nb]));
    for (int hi = h->nt; hi > t1; hi -= bs) {
        const int lo = std::max(t1 + 1, hi - bs + 1);
        if (int r = prop_L_range(h, UL[nb], UL[nb], lo, hi)) return r;
        if (int r = orthonormalise(h, UL[nb])) return r;
    }
    // U_L at the block boundaries τ2 + i·BatchSize, i = nb-1 .. 0
    for (int i = nb - 1; i >= 0; --i) {
        const int lo = t2 + i * bs + 1, hi = std::min(t2 + (i + 1) * bs, t1);
        if (int r = prop_L_range(h, UL[i + 1], UL[i], lo, hi)) return r;
        if (int r = orthonormalise(h, UL[i])) return r;
    }
    // U_R at τ2: from P through slices 1 .. τ2
    HIPCHK(load_P(h, UR));
    for (int lo = 1; lo <= t2; lo += bs) {
        const int hi = std::min(t2, lo + bs - 1);
        if (int r = prop_R_range(h, UR, UR, lo, hi)) return r;
        if (int r = orthonormalise(h, UR)) return r;
    }
    for (int i = 0; i < nb; ++i) {
        const int lo = t2 + i * bs + 1, hi = std::min(t2 + (i + 1) * bs, t1);
        if (int r = get_G_from(h, UR, UL[i], Gi)) return r;   // G at the block's lower edge
        if (i == 0) {
            HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, Gi, G2, DIM_NS, DIM_NS, true));
            HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, Gi, X12, DIM_NS, DIM_NS, true));
            HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, Gi, Y21, DIM_NS, DIM_NS, true));
            HIPCHK(launch_diag(h->stream, h->nch, h->max_ns, h->d_chains, Y21, DIM_NS, -1.0, 0));   // G - I
        } else {
            GemmArgs g = gargs(h, Gi, X12, T, DIM_NS, DIM_NS, DIM_NS);   // G[i] X12
            if (int r = gemm(h, true, true, 0, 0, g)) return r;
            std::swap(h->g4_m[1], h->g4_m[3]);
            X12 = h->g4_m[1];
            T = h->g4_m[3];
            GemmArgs gy = gargs(h, Y21, Gi, T, DIM_NS, DIM_NS, DIM_NS);   // Y21 - Y21 G[i]
            gy.mode = 2;
            gy.Cin = Y21;
            if (int r = gemm(h, true, true, 0, 0, gy)) return r;
            std::swap(h->g4_m[2], h->g4_m[3]);
            Y21 = h->g4_m[2];
            T = h->g4_m[3];
        }
        if (int r = g4_left(h, X12, T2, lo, hi)) return r;          // BB[i] ...
        if (int r = g4_right_inv(h, Y21, T2, lo, hi)) return r;     // ... BB[i]⁻¹
        if (int r = prop_R_range(h, UR, UR, lo, hi)) return r;
        if (int r = orthonormalise(h, UR)) return r;
    }
    if (int r = get_G_from(h, UR, UL[nb], Gi)) return r;            // G(τ1)
    out[0] = Gi;
    out[1] = G2;
    out[2] = X12;
    out[3] = Y21;
    return KDQ_OK;
}

int init_chain(kdq_handle h) {
    const int NN = h->nn;
    HIPCHK(load_P(h, h->t_UR[0]));
    HIPCHK(load_P(h, h->t_ULH[NN - 1]));
    for (int idx = NN - 1; idx >= 1; --idx)
        if (int r = propagate_L(h, h->t_ULH[idx], h->t_ULH[idx - 1], idx)) return r;
    h->idx = 1;
    if (int r = get_G(h, 1)) return r;
    return refresh_g32(h, false);
}

// Rank-2 blocks of one family (t-V) or of the whole slice (t-U-V, family 0).
int update_family_r2(kdq_handle h, int lt, int dir, uint32_t sweep, int family) {
    for (int b = 0; b < h->max_blocks; ++b) {
        DecideR2Args a{};
        a.chains = h->d_chains;
        a.G = h->t_G[h->cur];
        a.blocks = h->d_blocks;
        a.block_offset = h->d_block_offset;
        a.nblocks = h->d_nblocks;
        a.lt = lt; a.dir = dir; a.family = family; a.block = b; a.sweep = sweep;
        a.sweep_dev = h->sweep_dev;
        {
            ProfScope ps(h, KDQ_PROF_DECIDE);
            HIPCHK(launch_decide_r2(h->stream, h->nch, a));
        }
        FlushArgs f{h->d_chains, h->t_G[h->cur], h->t_G[1 - h->cur], 64, 0, nullptr, nullptr, h->t_FZ, h->t_FC};   // two slots per element
        {
            ProfScope ps(h, KDQ_PROF_FLUSH);
            f.one_launch = h->flush1;
            HIPCHK(launch_flush(h->stream, h->nch, h->max_ns, f));
        }
        h->cur ^= 1;
    }
    return KDQ_OK;
}

// The rank-R flush at the end of a deferred slice: G_new = G0 - U[:, :R] V[:R, :] (V stored
// transposed, so the GEMM reads it as Vᵀ).
int defer_flush(kdq_handle h) {
    ProfScope ps(h, KDQ_PROF_FLUSH);
    if (h->mixed) {
        GemmArgsF g{};
        g.chains = h->d_chains;
        g.A = h->t_DU32; g.B = h->t_DV32; g.C = h->t_G32[1 - h->cur32]; g.Cin = h->t_G32[h->cur32];
        g.mdim = DIM_NS; g.ndim = DIM_NS; g.kdim = DIM_RANK; g.mode = 2;
        HIPCHK(launch_gemm_f32(h->stream, h->nch, h->max_ns, h->max_ns, h->max_ns, true, true, 0, 1, g, h->gemm_tile));
        h->cur32 ^= 1;
    } else {
        GemmArgs g{};
        g.chains = h->d_chains;
        g.A = h->t_DU; g.B = h->t_DV; g.C = h->t_G[1 - h->cur]; g.Cin = h->t_G[h->cur];
        g.mdim = DIM_NS; g.ndim = DIM_NS; g.kdim = DIM_RANK; g.mode = 2;
        HIPCHK(launch_gemm(h->stream, h->nch, h->max_ns, h->max_ns, h->max_ns, true, true, 0, 1, g, h->gemm_tile));
        h->cur ^= 1;
    }
    return KDQ_OK;
}

// Pipelined t-U layer (captured sweeps). The decisions of block b+1 need only their k x k
// submatrix of G_{b+1} = G_b - G_b[:,A_b] W_b (E_A - G_b[A_b,:]); they read it from G_b and
// block b's accepted sites and W, so they depend on the flush of block b-1 (which wrote G_b)
// and on the decisions of block b, not on block b's flush. They run on the side stream, the
// flushes (prep, apply) on the handle's stream:
//   side:   decide_0  decide_1  decide_2 ...      decide_{b+1} waits for apply_{b-1}
//   main:       prep_0 apply_0 prep_1 apply_1 ... prep_b waits for decide_b
// so decide_{b+1} is issued while prep_b runs and shares the device with apply_b. Accepted
// counts, sites and W alternate between two slots by block parity.
int update_layer_pipe(kdq_handle h, int lt, int dir, uint32_t sweep, bool prewrapped = false) {
    const int kb = h->delay;
    const bool mx = h->mixed;
    auto decide = [&](int s0, int slot, bool pre, const Opnd *Gb, const OpndF *Gbf) -> hipError_t {
        DecideArgs a{};
        a.chains = h->d_chains;
        a.G = Gb;
        a.Gf = Gbf;
        a.lt = lt; a.dir = dir; a.s0 = s0; a.kb = kb; a.sweep = sweep;
        a.sweep_dev = h->sweep_dev;
        a.slot = slot;
        a.pre = pre ? 1 : 0;
        return launch_decide_tu(h->side, h->nch, a);
    };
    const int nb = (h->max_ns + kb - 1) / kb;
    // G_0 (this layer's G after the wrap) is ready; block 0 was already decided beside the
    // wrap's second product when prewrapped (wrap_prewrapped)
    HIPCHK(hipEventRecord(h->ev_app, h->stream));
    HIPCHK(hipStreamWaitEvent(h->side, h->ev_app, 0));
    if (!prewrapped) {
        HIPCHK(decide(0, 0, false, h->t_G[h->cur], mx ? h->t_G32[h->cur32] : nullptr));
        HIPCHK(hipEventRecord(h->ev_dec[0], h->side));
    }
    for (int b = 0; b < nb; ++b) {
        const int slot = b & 1;
        // G_b, the input of this block's flush
        const Opnd *Gb = h->t_G[h->cur];
        const OpndF *Gbf = mx ? h->t_G32[h->cur32] : nullptr;
        if (b + 1 < nb) {
            if (b >= 1) HIPCHK(hipStreamWaitEvent(h->side, h->ev_app, 0));   // apply_{b-1} wrote G_b
            HIPCHK(decide((b + 1) * kb, slot ^ 1, true, Gb, Gbf));
            HIPCHK(hipEventRecord(h->ev_dec[slot ^ 1], h->side));
        }
        HIPCHK(hipStreamWaitEvent(h->stream, h->ev_dec[slot], 0));
        if (mx) {
            FlushArgsF f{h->d_chains, h->t_G32[h->cur32], h->t_G32[1 - h->cur32], kb, 0, nullptr, nullptr,
                         h->t_FZ32, h->t_FC32, slot};
            f.one_launch = h->flush1;
            HIPCHK(launch_flush_f32(h->stream, h->nch, h->max_ns, f));
            h->cur32 ^= 1;
        } else {
            FlushArgs f{h->d_chains, h->t_G[h->cur], h->t_G[1 - h->cur], kb, 0, nullptr, nullptr, h->t_FZ, h->t_FC,
                        slot};
            f.one_launch = h->flush1;
            if (prewrapped && b == 0) {   // block 0's accepted fields, held back while the wrap read them
                f.commit_lt = lt;
                f.commit_s0 = 0;
            }
            HIPCHK(launch_flush(h->stream, h->nch, h->max_ns, f));
            h->cur ^= 1;
        }
        if (b + 2 < nb) HIPCHK(hipEventRecord(h->ev_app, h->stream));
    }
    return KDQ_OK;
}

// pipelined only in captured sweeps: with host launches the cross-stream events cost more
// than the overlap gains (measured: 705 against 672 ms per sweep, 64 chains at N = 288)
bool pipe_on(kdq_handle h) {
    const bool defer = h->defer && !h->scee;
    return h->kind == KDQ_MODEL_TU && h->pipe && h->sweep_dev && !defer && !h->scee && !h->prof_mask &&
           h->max_ns > h->delay;
}

bool prewrap_on(kdq_handle h) {
    return pipe_on(h) && h->prewrap && !h->mixed && !h->flush1 && h->delay <= 32 && h->d_pwpart;
}

// The wrap of slice lw (WrapKV!, "B"; forward before layer lt = lw, backward after layer lw and
// before layer lt = lw - 1) as its two products with block 0 of layer lt decided in between, on
// the side stream: its G[S,S] of the wrapped G comes from the first product T and the dense
// factor of the second (k_decide_tu, prewrap), so the decisions overlap the second product
// instead of following the whole wrap. The layer is then run with update_layer_pipe(...,
// prewrapped = true).
int wrap_prewrapped(kdq_handle h, int lw, bool forward, int lt, int dir, uint32_t sweep) {
    Opnd *Gt = h->t_G[h->cur], *Tt = h->t_T;
    const DiagSpec none{};
    DecideArgs a{};
    a.chains = h->d_chains;
    a.G = Gt;
    a.lt = lt; a.dir = dir; a.s0 = 0; a.kb = h->delay; a.sweep = sweep;
    a.sweep_dev = h->sweep_dev;
    a.slot = 0;
    a.prewrap = 1;
    a.T = Tt;
    if (forward) {   // T = eK G, then G = D T eK⁻¹ D⁻¹
        if (int r = kin_left(h, 0, 0, Gt, Tt, DIM_NS, none, none, none, KDQ_PROF_WRAP)) return r;
        a.E = h->t_eKinv;
        a.pw_line = hs_diag(h, lw, +1);
        a.pw_out = hs_diag(h, lw, -1);
    } else {         // T = eK⁻¹ (D' G) D'⁻¹, then G = T eK
        if (int r = kin_left(h, 1, 0, Gt, Tt, DIM_NS, hs_diag(h, lw, -1), none, hs_diag(h, lw, +1), KDQ_PROF_WRAP))
            return r;
        a.E = h->t_eK;
    }
    a.pw_part = h->d_pwpart;
    a.pw_nkc = (h->max_ns + 31) / 32;
    HIPCHK(hipEventRecord(h->ev_app, h->stream));
    HIPCHK(hipStreamWaitEvent(h->side, h->ev_app, 0));
    HIPCHK(launch_prewrap_part(h->side, h->nch, a));
    HIPCHK(launch_decide_tu(h->side, h->nch, a));
    HIPCHK(hipEventRecord(h->ev_dec[0], h->side));
    if (forward) return kin_right(h, 1, Tt, Gt, none, hs_diag(h, lw, -1), hs_diag(h, lw, +1), KDQ_PROF_WRAP);
    return kin_right(h, 0, Tt, Gt, none, none, none, KDQ_PROF_WRAP);
}

int update_layer(kdq_handle h, int lt, int dir, uint32_t sweep) {
    if (h->kind == KDQ_MODEL_TUV) return update_family_r2(h, lt, dir, sweep, 0);
    const int kb = h->delay;
    const bool defer = h->defer && !h->scee;
    if (pipe_on(h)) return update_layer_pipe(h, lt, dir, sweep);
    int blk = 0;
    for (int s0 = 0; s0 < h->max_ns; s0 += kb, ++blk) {
        DecideArgs a{};
        a.chains = h->d_chains;
        a.G = h->t_G[h->cur];
        a.lt = lt; a.dir = dir; a.s0 = s0; a.kb = kb; a.sweep = sweep;
        a.sweep_dev = h->sweep_dev;
        if (h->mixed) a.Gf = h->t_G32[h->cur32];
        if (defer) {
            a.deferred = 1;
            a.blk = blk;
            a.U = h->t_DU; a.V = h->t_DV;
            if (h->mixed) { a.Uf = h->t_DU32; a.Vf = h->t_DV32; }
        }
        // t-U only; t-V layers go through slice_tv
        {
            ProfScope ps(h, KDQ_PROF_DECIDE);
            HIPCHK(launch_decide_tu(h->stream, h->nch, a));
        }
        if (defer) {
            ProfScope ps(h, KDQ_PROF_FLUSH);
            if (h->mixed) {
                DeferArgsT<float> d{h->d_chains, h->t_G32[h->cur32], h->t_DU32, h->t_DV32, blk, kb};
                HIPCHK(launch_defer_xy(h->stream, h->nch, h->max_ns, d));
            } else {
                DeferArgsT<double> d{h->d_chains, h->t_G[h->cur], h->t_DU, h->t_DV, blk, kb};
                HIPCHK(launch_defer_xy(h->stream, h->nch, h->max_ns, d));
            }
            continue;
        }
        if (h->mixed) {
            FlushArgsF f{h->d_chains, h->t_G32[h->cur32], h->t_G32[1 - h->cur32], kb, 0, nullptr, nullptr, h->t_FZ32, h->t_FC32};
            ProfScope ps(h, KDQ_PROF_FLUSH);
            f.one_launch = h->flush1;
            HIPCHK(launch_flush_f32(h->stream, h->nch, h->max_ns, f));
            h->cur32 ^= 1;
            continue;
        }
        FlushArgs f{h->d_chains, h->t_G[h->cur], h->t_G[1 - h->cur], kb, 0, nullptr, nullptr, h->t_FZ, h->t_FC};
        {
            ProfScope ps(h, KDQ_PROF_FLUSH);
            f.one_launch = h->flush1;
            HIPCHK(launch_flush(h->stream, h->nch, h->max_ns, f));
        }
        h->cur ^= 1;
    }
    if (defer) return defer_flush(h);
    return KDQ_OK;
}

// t-V forward slice (src/tV/phy_update.jl:60-73): kinetic wrap, then for families 3,2,1
// the bond rotation of the family and its rank-2 updates.
int slice_tv_forward(kdq_handle h, Opnd *Gt, Opnd *Tt, int lt, bool update, uint32_t sweep, int fam) {
    if (int r = wrap_K(h, Gt, Tt, true, fam)) return r;
    for (int j = 2; j >= 0; --j) {
        Opnd *cur = update ? h->t_G[h->cur] : Gt;
        if (int r = bondrot(h, cur, j, lt, +1, 0, DIM_NS, fam)) return r;
        if (update)
            if (int r = update_family_r2(h, lt, 0, sweep, j)) return r;
    }
    return KDQ_OK;
}

// t-V backward slice (src/tV/phy_update.jl:131-142): for families 1,2,3 the rank-2 updates
// and the inverse bond rotation, then the inverse kinetic wrap.
int slice_tv_backward(kdq_handle h, Opnd *Gt, Opnd *Tt, int lt, bool update, uint32_t sweep, int fam) {
    for (int j = 0; j < 3; ++j) {
        if (update)
            if (int r = update_family_r2(h, lt, 1, sweep, j)) return r;
        Opnd *cur = update ? h->t_G[h->cur] : Gt;
        if (int r = bondrot(h, cur, j, lt, -1, 0, DIM_NS, fam)) return r;
    }
    return wrap_K(h, update ? h->t_G[h->cur] : Gt, Tt, false, fam);
}

// One slice wrap of a G copy (no updates) in either direction, any model.
int wrap_slice(kdq_handle h, Opnd *Gt, Opnd *Tt, int lt, bool forward, int fam) {
    if (h->mixed && Gt == h->t_G[h->cur]) return wrap_B_f32(h, lt, forward, fam);   // the live G is the FP32 copy
    if (h->kind == KDQ_MODEL_TV)
        return forward ? slice_tv_forward(h, Gt, Tt, lt, false, 0, fam) : slice_tv_backward(h, Gt, Tt, lt, false, 0, fam);
    return wrap_B(h, Gt, Tt, lt, forward, fam);
}

// phy_measure: wrap a copy of G to τ = Nt/2, HalfeK sandwich, observables (accumulated).
// Gm, Tm: per-chain N x N complex working matrices; meas: accumulators (nullptr: the chains').
// Measurements beside the sweep (t-U): the front of a measurement — the copy of G and any
// wrap that reads fields the sweep may still change — runs on the handle's stream; its back —
// the products with HalfeK or the slot products and the observable tiles, which read only the
// copy, the slot matrices and the model — goes to the aux stream (meas_fork / meas_unfork).
// The next measurement, which rewrites the copy and may rewrite slot matrices, and the end of
// the sweep wait for it (meas_join). Same kernels on the same operands: identical results.
bool meas_aside(kdq_handle h) {
    return h->kind == KDQ_MODEL_TU && h->maside && !h->prof_mask && !h->scee && h->aux;
}

int meas_join(kdq_handle h) {
    if (h->mpending) {
        HIPCHK(hipStreamWaitEvent(h->stream, h->ev_mdone, 0));
        h->mpending = false;
    }
    return KDQ_OK;
}

int meas_fork(kdq_handle h) {
    HIPCHK(hipEventRecord(h->ev_mfork, h->stream));
    HIPCHK(hipStreamWaitEvent(h->aux, h->ev_mfork, 0));
    std::swap(h->stream, h->aux);
    return KDQ_OK;
}

int meas_unfork(kdq_handle h) {
    std::swap(h->stream, h->aux);
    HIPCHK(hipEventRecord(h->ev_mdone, h->aux));
    h->mpending = true;
    return KDQ_OK;
}

int measure_into(kdq_handle h, int lt, Opnd *Gm, Opnd *Tm, double *meas, double *part, bool fork = false) {
    if (h->mixed)
        HIPCHK(launch_to_f64(h->stream, h->nch, h->max_ns, h->d_chains, h->t_G32[h->cur32], Gm, DIM_NS, 1));
    else
        HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, h->t_G[h->cur], Gm, DIM_NS, DIM_NS, true));
    const int nt = h->nt;
    if (lt > nt / 2.0) {
        for (int t = lt; t >= nt / 2 + 1; --t)
            if (int r = wrap_slice(h, Gm, Tm, t, false, KDQ_PROF_MEAS_WRAP)) return r;
    } else {
        for (int t = lt + 1; t <= nt / 2; ++t)
            if (int r = wrap_slice(h, Gm, Tm, t, true, KDQ_PROF_MEAS_WRAP)) return r;
    }
    if (fork)
        if (int r = meas_fork(h)) return r;
    GemmArgs g1 = gargs(h, h->t_hK, Gm, Tm, DIM_NS, DIM_NS, DIM_NS);
    int r = gemm_m(h, true, false, false, true, 0, 0, g1, KDQ_PROF_MEAS_WRAP);
    GemmArgs g2 = gargs(h, Tm, h->t_hKinv, Gm, DIM_NS, DIM_NS, DIM_NS);
    if (!r) r = gemm_m(h, false, true, true, false, 0, 0, g2, KDQ_PROF_MEAS_WRAP);
    if (!r) {
        MeasArgs m{h->d_chains, Gm, meas, part, meas_tiles(h->max_ns)};
        ProfScope ps(h, KDQ_PROF_MEASURE);
        if (hipError_t e = launch_measure(h->stream, h->nch, m)) r = fail(h, KDQ_EHIP, hipGetErrorString(e));
    }
    if (fork) {
        if (int r2 = meas_unfork(h)) return r2;
    }
    return r;
}

int measure(kdq_handle h, int lt) { return measure_into(h, lt, h->t_Gm, h->t_Tm, nullptr, h->d_mpart); }

// ---- t-U measurements inside a sweep (phy_measure at the slices of the two windows).
// measure_into wraps a copy of G slice by slice to c = Nt/2: |lt − c| wrap pairs per
// measurement, 100 per pass of the N = 288 bench. The same G_sym = HalfeK B(c, lt) G B(c, lt)⁻¹
// HalfeK⁻¹ comes from two complex products with X = HalfeK B(c, lt) and X̃ = its inverse
// times HalfeK⁻¹, and those follow one slice at a time (one pair of real-by-complex products
// per slice) by two recursions over slots k = |lt − c|:
//   down: X_k = X_(k−1) D_s eK,  X̃_k = eK⁻¹ D_s⁻¹ X̃_(k−1)  with s = c − k + 1  (slices below c)
//   up:   Y_k = Y_(k−1) eK⁻¹ D_s⁻¹,  Ỹ_k = D_s eK Ỹ_(k−1)  with s = c + k    (slices above c)
// (B_s = D_s eK as in wrap_B), X_0 = Y_0 = HalfeK, X̃_0 = Ỹ_0 = HalfeK⁻¹. A slot must be built from
// the fields the measurement sees: going forward, the slices between lt and c are not yet
// updated below c (built at the window's first slice) and already final above c (built
// slice by slice); going backward the other way round, and the measured slice itself, just
// updated, is wrapped on the copy first. Products of at most one block of slices (κ ≤ e^{0.6 k}
// at Δτ = 0.1) keep the rounding near that of the stepwise wraps.
bool is_node(kdq_handle h, int t);

int mq_alloc(kdq_handle h) {
    if (!h->mq.empty()) return KDQ_OK;
    // the largest |lt − c| a window reaches (the windows as sweep_body visits them)
    const int c = h->nt / 2, thidx = h->nn / 2 + 1;
    int kmax = 1, idx = 1;
    for (int lt = 1; lt <= h->nt; ++lt) {
        const int d = idx - thidx;
        if (d <= 0 && d >= -1) kmax = std::max(kmax, std::abs(lt - c));
        if (is_node(h, lt)) ++idx;
    }
    for (int lt = h->nt; lt >= 1; --lt) {
        const int d = idx - thidx;
        if (d >= 0 && d <= 1) kmax = std::max(kmax, std::abs(lt - c));
        if (is_node(h, lt - 1)) --idx;
    }
    const int slots = kmax + 1;
    std::vector<std::vector<Opnd>> vq(slots), vqt(slots);
    for (int c2 = 0; c2 < h->nch; ++c2) {
        ChainHost &C = h->ch[c2];
        const int n = C.ns;
        for (int k = 0; k < slots; ++k)
            for (auto *v : {&vq, &vqt}) {
                double *p = nullptr;
                if (dalloc(C, &p, 2 * (size_t)n * n) != hipSuccess)
                    return fail(h, KDQ_EHIP, "measurement propagators: device allocation failed");
                (*v)[k].push_back(cplx(p, n, (size_t)n * n));
            }
    }
    h->mq.resize(slots);
    h->mqt.resize(slots);
    for (int k = 0; k < slots; ++k) {
        if (int r = make_table(h, &h->mq[k], vq[k])) return r;
        if (int r = make_table(h, &h->mqt[k], vqt[k])) return r;
    }
    // slot 0: HalfeK and HalfeK⁻¹ as complex matrices
    if (h->cK) {
        HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, h->t_hK, h->mq[0], DIM_NS, DIM_NS, true));
        HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, h->t_hKinv, h->mqt[0], DIM_NS, DIM_NS, true));
    } else {
        HIPCHK(launch_real_to_cplx(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, h->t_hK, h->mq[0], DIM_NS, DIM_NS));
        HIPCHK(launch_real_to_cplx(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, h->t_hKinv, h->mqt[0], DIM_NS, DIM_NS));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

// slots 1..k of recursion dir (1: down, 2: up) from the current fields
int mq_ensure(kdq_handle h, int dir, int k) {
    const int c = h->nt / 2;
    if (h->mq_dir != dir) {
        h->mq_dir = dir;
        h->mq_have = 0;
    }
    for (int j = h->mq_have + 1; j <= k; ++j) {
        const DiagSpec none{};
        if (dir == 1) {   // X_j = X_{j-1} D eK, X̃_j = eK⁻¹ D⁻¹ X̃_{j-1} at slice c - j + 1
            const int sl = c - j + 1;
            if (int r = kin_right(h, 0, h->mq[j - 1], h->mq[j], hs_diag(h, sl, +1), none, none, KDQ_PROF_MEAS_WRAP))
                return r;
            if (int r = kin_left(h, 1, 0, h->mqt[j - 1], h->mqt[j], DIM_NS, hs_diag(h, sl, -1), none, none,
                                 KDQ_PROF_MEAS_WRAP))
                return r;
        } else {          // X_j = X_{j-1} eK⁻¹ D⁻¹, X̃_j = D eK X̃_{j-1} at slice c + j
            const int sl = c + j;
            if (int r = kin_right(h, 1, h->mq[j - 1], h->mq[j], none, hs_diag(h, sl, -1), none, KDQ_PROF_MEAS_WRAP))
                return r;
            if (int r = kin_left(h, 0, 0, h->mqt[j - 1], h->mqt[j], DIM_NS, none, hs_diag(h, sl, +1), none,
                                 KDQ_PROF_MEAS_WRAP))
                return r;
        }
    }
    h->mq_have = std::max(h->mq_have, k);
    return KDQ_OK;
}

// observables of X G X̃ (wrap_first: G wrapped backward through slice lt on the copy first)
int measure_q(kdq_handle h, int lt, int k, bool wrap_first, bool fork = false) {
    Opnd *Gm = h->t_Gm, *Tm = h->t_Tm;
    if (h->mixed)
        HIPCHK(launch_to_f64(h->stream, h->nch, h->max_ns, h->d_chains, h->t_G32[h->cur32], Gm, DIM_NS, 1));
    else
        HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, h->t_G[h->cur], Gm, DIM_NS, DIM_NS, true));
    if (wrap_first)
        if (int r = wrap_slice(h, Gm, Tm, lt, false, KDQ_PROF_MEAS_WRAP)) return r;
    if (fork)
        if (int r = meas_fork(h)) return r;
    GemmArgs g1 = gargs(h, h->mq[k], Gm, Tm, DIM_NS, DIM_NS, DIM_NS);
    int r = gemm(h, true, true, 0, 0, g1, KDQ_PROF_MEAS_WRAP);
    GemmArgs g2 = gargs(h, Tm, h->mqt[k], Gm, DIM_NS, DIM_NS, DIM_NS);
    if (!r) r = gemm(h, true, true, 0, 0, g2, KDQ_PROF_MEAS_WRAP);
    if (!r) {
        MeasArgs m{h->d_chains, Gm, nullptr, h->d_mpart, meas_tiles(h->max_ns)};
        ProfScope ps(h, KDQ_PROF_MEASURE);
        if (hipError_t e = launch_measure(h->stream, h->nch, m)) r = fail(h, KDQ_EHIP, hipGetErrorString(e));
    }
    if (fork) {
        if (int r2 = meas_unfork(h)) return r2;
    }
    return r;
}

// phy_measure at slice lt of a t-U sweep with the cell-momentum products: G_sym =
// HalfeK B(c, lt) G B(c, lt)⁻¹ HalfeK⁻¹ (c = Nt/2; B(c, lt) = B_c ... B_(lt+1) below c and
// B_(c+1)⁻¹ ... B_lt⁻¹ above, with the fields the measurement sees) as one chained launch over
// the columns of G (the slices' factors, then HalfeK) and one over the rows of the result (their
// inverses in the same order, then HalfeK⁻¹) — the stepwise wraps of measure_into with every
// left factor first. Both passes read fields of slices the sweep updates next, so they run in
// line; the tile reduction of G_sym runs beside the sweep (option maside).
int measure_chain(kdq_handle h, int lt) {
    const bool fk = meas_aside(h);
    if (fk)   // the previous measurement's reduction still reads Gm
        if (int r = meas_join(h)) return r;
    Opnd *Gm = h->t_Gm, *Tm = h->t_Tm, *src = h->t_G[h->cur];
    if (h->mixed) {
        HIPCHK(launch_to_f64(h->stream, h->nch, h->max_ns, h->d_chains, h->t_G32[h->cur32], Gm, DIM_NS, 1));
        src = Gm;
    }
    const int c = h->nt / 2, k = lt < c ? c - lt : lt - c;
    const DiagSpec none{};
    CircArgs l{h->d_chains, src, Tm, 0, 0, 0, DIM_NS, none, none, none, k + 1, 0, 1, 2};
    CircArgs r{h->d_chains, Tm, Gm, 0, 0, 1, DIM_NS, none, none, none, k + 1, 0, 1, 3};
    if (lt < c) {          // B_(lt+1) .. B_c, B_s = D_s eK; inverses eK⁻¹ D_s⁻¹ from the right
        l.mat = 0; l.out = hs_diag(h, lt + 1, +1); l.dlt = +1;
        r.mat = 1; r.out = hs_diag(h, lt + 1, -1); r.dlt = +1;
    } else if (lt > c) {   // B_lt⁻¹ .. B_(c+1)⁻¹, B_s⁻¹ = eK⁻¹ D_s⁻¹; then D_s eK from the right
        l.mat = 1; l.in = hs_diag(h, lt, -1); l.dlt = -1;
        r.mat = 0; r.in = hs_diag(h, lt, +1); r.dlt = -1;
    }
    {
        ProfScope ps(h, KDQ_PROF_MEAS_WRAP);
        int gx = 1;
        size_t bytes = 0;
        int lmax = 1;
        circ_launch_shape(h, DIM_NS, gx, bytes, lmax);
        HIPCHK(launch_circ(h->stream, h->nch, gx, bytes, lmax, l));
        HIPCHK(launch_circ(h->stream, h->nch, gx, bytes, lmax, r));
    }
    if (fk)
        if (int r2 = meas_fork(h)) return r2;
    int rc = KDQ_OK;
    {
        MeasArgs m{h->d_chains, Gm, nullptr, h->d_mpart, meas_tiles(h->max_ns)};
        ProfScope ps(h, KDQ_PROF_MEASURE);
        if (hipError_t e = launch_measure(h->stream, h->nch, m)) rc = fail(h, KDQ_EHIP, hipGetErrorString(e));
    }
    if (fk)
        if (int r2 = meas_unfork(h)) return r2;
    return rc;
}

// phy_measure at slice lt of a t-U sweep (after the slice's update): chained cell-momentum
// passes when the factors are block-circulant; otherwise stepwise wraps where they cost no more
// (|lt − c| ≤ 1, or 2 going backward above c), the slot products elsewhere
int measure_sweep(kdq_handle h, int lt, bool forward) {
    if (use_circ(h) && h->mchain && h->mchain_ok) return measure_chain(h, lt);
    if (h->kind != KDQ_MODEL_TU || h->mq.empty()) return measure(h, lt);
    const bool fk = meas_aside(h);
    if (fk)   // the previous measurement's back half still reads the copy and a slot matrix
        if (int r = meas_join(h)) return r;
    auto near = [&]() { return measure_into(h, lt, h->t_Gm, h->t_Tm, nullptr, h->d_mpart, fk); };
    const int c = h->nt / 2;
    if (lt <= c) {
        const int k = c - lt;
        if (!forward || k >= 2)   // backward: slot k - 1 .. k follow slice by slice
            if (int r = mq_ensure(h, 1, k)) return r;
        return k <= 1 ? near() : measure_q(h, lt, k, false, fk);
    }
    if (forward) {
        const int k = lt - c;
        if (int r = mq_ensure(h, 2, k)) return r;
        return k <= 1 ? near() : measure_q(h, lt, k, false, fk);
    }
    const int k = lt - 1 - c;   // slices c+1 .. lt-1, not yet updated in this pass
    if (k <= 1) return near();
    if (int r = mq_ensure(h, 2, k)) return r;
    return measure_q(h, lt, k, true, fk);
}

int elems(kdq_handle h, int c) {
    const int n = h->ch[c].ns;
    return h->kind == KDQ_MODEL_TU ? n : (h->kind == KDQ_MODEL_TV ? 3 * (n / 2) : 3 * n);
}

bool is_node(kdq_handle h, int t) {
    for (int64_t v : h->nodes)
        if (v == t) return true;
    return false;
}

int read_meas(kdq_handle h, double *dst /* nch x nobs */) {
    const int S = KDQ_MEAS_SLOTS, no = h->nobs;
    for (int c = 0; c < h->nch; ++c)
        HIPCHK(hipMemcpyAsync(h->h_meas + S * c, h->ch[c].meas, S * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < h->nch; ++c) {
        const double cnt = h->h_meas[S * c + S - 1];
        for (int k = 0; k < no; ++k) dst[no * c + k] = cnt > 0 ? h->h_meas[S * c + k] / cnt : 0.0;
    }
    return KDQ_OK;
}

// Reads the sticky singular-overlap flags; KDQ_ENUM names the first failing chain.
int check_health(kdq_handle h) {
    std::vector<int> info(h->nch), qinfo(h->nch);
    HIPCHK(hipMemcpyAsync(info.data(), h->d_info, sizeof(int) * h->nch, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(qinfo.data(), h->d_qinfo, sizeof(int) * h->nch, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<int> pois(h->nch, 0);
    if (h->d_poison)
        HIPCHK(hipMemcpyAsync(pois.data(), h->d_poison, sizeof(int) * h->nch, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < h->nch; ++c) {
        if (pois[c])
            return fail(h, KDQ_ENUM, "non-finite Green's function in chain " + std::to_string(c) +
                                         " (the other chains are unaffected; kdq_set_fields resets the chain)");
        if (info[c]) return fail(h, KDQ_ENUM, "singular overlap U_L U_R in chain " + std::to_string(c));
        if (qinfo[c])
            return fail(h, KDQ_ENUM, "propagator lost rank between stabilisations in chain " + std::to_string(c) +
                                         " (reduce BatchSize)");
    }
    return KDQ_OK;
}

// Measurement windows: t-U measures 0 <= Θidx - idx <= 1 going forward and
// 0 <= idx - Θidx <= 1 going backward (src/tU/phy_update.jl:69,112); t-V and t-U-V measure
// |idx - Θidx| <= 1 in both passes (src/tV/phy_update.jl:93,144; src/tUV/phy_update.jl:82,127).
bool in_window(kdq_handle h, int thidx, bool forward) {
    const int d = h->idx - thidx;
    if (h->kind == KDQ_MODEL_TU) return forward ? (d <= 0 && d >= -1) : (d >= 0 && d <= 1);
    return d >= -1 && d <= 1;
}

// Stabilisation at node idx: G recomputed from the chain (get_G!). FP64: into the spare
// buffer, compared with the wrapped and updated G it replaces (max |difference| per chain, the
// reference's commented wrap-error check, src/tU/phy_update.jl:54-58,92-97; a non-finite entry
// in either marks the chain), then made current. Mixed precision: the FP32 copy is compared and
// replaced by refresh_g32.
int node_refresh(kdq_handle h, int idx, Opnd *UR = nullptr, Opnd *ULH = nullptr) {
    if (h->mixed) {
        if (int r = get_G(h, idx, nullptr, UR, ULH)) return r;
        HIPCHK(launch_node_check(h->stream, h->nch, h->max_ns, h->d_chains, h->t_G[h->cur], h->t_G[h->cur], DIM_NS,
                                 h->d_wraperr, h->d_poison));
        return refresh_g32(h, true);
    }
    if (int r = get_G(h, idx, h->t_G[1 - h->cur], UR, ULH)) return r;
    HIPCHK(launch_node_check(h->stream, h->nch, h->max_ns, h->d_chains, h->t_G[h->cur], h->t_G[1 - h->cur], DIM_NS,
                             h->d_wraperr, h->d_poison));
    h->cur ^= 1;
    return KDQ_OK;
}

// One sweep (forward and backward over all slices, stabilisation at the nodes, measurements
// in the reference's windows) for every chain. Only launches: no host reads, no
// synchronisation, no memset, so the sequence can be captured into a graph.
int sweep_body(kdq_handle h, int record, uint32_t sweep) {
    h->pw_next = false;
    const int NN = h->nn, nt = h->nt, thidx = NN / 2 + 1;
    const bool tv = h->kind == KDQ_MODEL_TV;
    (void)NN;
    if (record) HIPCHK(launch_zero_meas(h->stream, h->nch, h->d_chains));
    const bool aside = chain_aside(h), split = qsplit_on(h);
    h->mq_dir = 0;
    for (int lt = 1; lt <= nt; ++lt) {
        if (tv) {
            if (int r = slice_tv_forward(h, h->t_G[h->cur], h->t_T, lt, true, sweep, KDQ_PROF_WRAP)) return r;
        } else {
            if (prewrap_on(h)) {
                if (int r = wrap_prewrapped(h, lt, true, lt, 0, sweep)) return r;
                if (int r = update_layer_pipe(h, lt, 0, sweep, true)) return r;
            } else {
                if (int r = h->mixed ? wrap_B_f32(h, lt, true) : wrap_B(h, h->t_G[h->cur], h->t_T, lt, true)) return r;
                if (int r = update_layer(h, lt, 0, sweep)) return r;
            }
            if (aside)
                if (int r = chain_step(h, lt, true)) return r;
        }
        if (record && in_window(h, thidx, true))
            if (int r = tv ? measure(h, lt) : measure_sweep(h, lt, true)) return r;
        if (is_node(h, lt)) {
            h->idx += 1;
            Opnd *U = h->t_UR[h->idx - 1];
            if (aside) {
                if (int r = chain_join(h, U, !split)) return r;
            } else if (int r = propagate_R(h, h->t_UR[h->idx - 2], U, h->idx - 1, nullptr, !split)) {
                return r;
            }
            Opnd *Uq = nullptr;   // the first-pass copy the node's G is formed from (split)
            if (split) {
                int rc = 0;
                Uq = orthonormalise_split(h, U, rc);
                if (rc) return rc;
            }
            if (int r = node_refresh(h, h->idx, Uq, nullptr)) return r;
            if (split) HIPCHK(hipStreamWaitEvent(h->stream, h->ev_q2, 0));
        }
    }
    h->mq_dir = 0;
    for (int lt = nt; lt >= 1; --lt) {
        if (tv) {
            if (int r = slice_tv_backward(h, h->t_G[h->cur], h->t_T, lt, true, sweep, KDQ_PROF_WRAP)) return r;
            if (record && in_window(h, thidx, false))
                if (int r = measure(h, lt - 1)) return r;
        } else {
            if (h->pw_next) {
                h->pw_next = false;
                if (int r = update_layer_pipe(h, lt, 1, sweep, true)) return r;
            } else if (int r = update_layer(h, lt, 1, sweep)) {
                return r;
            }
            if (aside)
                if (int r = chain_step(h, lt, false)) return r;
            if (record && in_window(h, thidx, false))
                if (int r = measure_sweep(h, lt, false)) return r;
            // the next layer's block 0 beside this wrap unless a node replaces G in between
            if (prewrap_on(h) && lt > 1 && !is_node(h, lt - 1)) {
                if (int r = wrap_prewrapped(h, lt, false, lt - 1, 1, sweep)) return r;
                h->pw_next = true;
            } else if (int r = h->mixed ? wrap_B_f32(h, lt, false) : wrap_B(h, h->t_G[h->cur], h->t_T, lt, false)) {
                return r;
            }
        }
        if (is_node(h, lt - 1)) {
            h->idx -= 1;
            Opnd *U = h->t_ULH[h->idx - 1];
            if (aside) {
                if (int r = chain_join(h, U, !split)) return r;
            } else if (int r = propagate_L(h, h->t_ULH[h->idx], U, h->idx, !split)) {
                return r;
            }
            Opnd *Uq = nullptr;
            if (split) {
                int rc = 0;
                Uq = orthonormalise_split(h, U, rc);
                if (rc) return rc;
            }
            if (int r = node_refresh(h, h->idx, nullptr, Uq)) return r;
            if (split) HIPCHK(hipStreamWaitEvent(h->stream, h->ev_q2, 0));
        }
    }
    if (int r = meas_join(h)) return r;
    if (h->sweep_dev) HIPCHK(launch_bump(h->stream, h->d_sweep));
    return KDQ_OK;
}

void drop_graph(kdq_handle h) {
    if (h->gexec) hipGraphExecDestroy(h->gexec);
    if (h->graph) hipGraphDestroy(h->graph);
    h->gexec = nullptr;
    h->graph = nullptr;
    h->g_record = -1;
}

// Captures sweep_body into h->gexec (the sweep counter comes from h->d_sweep, which the
// graph increments at its end). The host-side state a sweep mutates (current G buffer, node
// index) must come back to its start value, or the graph would not be replayable.
int capture_sweep(kdq_handle h, int record) {
    drop_graph(h);
    const int cur0 = h->cur, idx0 = h->idx, cur32_0 = h->cur32;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
    h->sweep_dev = h->d_sweep;
    int r = sweep_body(h, record, 0);
    h->sweep_dev = nullptr;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (r) {
        if (g) hipGraphDestroy(g);
        return r;
    }
    HIPCHK(e);
    if (h->cur != cur0 || h->idx != idx0 || h->cur32 != cur32_0) {
        hipGraphDestroy(g);
        return fail(h, KDQ_ESTATE, "sweep is not replayable as a graph (state parity)");
    }
    h->graph = g;
    HIPCHK(hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0));
    h->g_record = record;
    return KDQ_OK;
}

// Upper estimate of the kernel launches of one sweep. A captured sweep is one linear chain
// of graph nodes; past 20000 nodes (small decision blocks on large
// lattices) the sweep runs with host launches instead of being captured.
long sweep_launch_estimate(kdq_handle h) {
    // t-U: decide + factor append per block and one flush GEMM (deferred), or decide + two
    // flush launches per block; rank-2 models: decide + two flush launches per block
    const long blocks = h->kind == KDQ_MODEL_TU ? (h->max_ns + h->delay - 1) / h->delay : 3L * h->max_blocks;
    const bool defer = h->kind == KDQ_MODEL_TU && h->defer && !h->scee;
    const long per_slice = (defer ? 2 * blocks + 1 : 3 * blocks) + 8 + (h->kind == KDQ_MODEL_TV ? 12 : 0);
    return 2L * h->nt * per_slice + 40L * h->nn + 2L * h->nt;
}
constexpr long kMaxGraphLaunches = 20000;

int phy_update(kdq_handle h, int sweeps, int record, double *rows) {
    const int no = h->nobs;
    h->sl_logw_valid = false;
    if (record && h->kind == KDQ_MODEL_TU && !h->scee)
        if (int r = mq_alloc(h)) return r;
    if (int r = init_chain(h)) return r;
    std::vector<double> tmp(no * h->nch);
    const bool graph = h->use_graph && h->prof_mask == 0 && sweep_launch_estimate(h) <= kMaxGraphLaunches;
    if (graph) {
        HIPCHK(hipMemcpyAsync(h->d_sweep, &h->sweep, sizeof(uint32_t), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (!h->gexec || h->g_record != record)
            if (int r = capture_sweep(h, record)) return r;
    }
    for (int sw = 0; sw < sweeps; ++sw) {
        if (graph) HIPCHK(hipGraphLaunch(h->gexec, h->stream));
        else if (int r = sweep_body(h, record, h->sweep)) return r;
        if (record && rows) {
            if (int r = read_meas(h, tmp.data())) return r;
            for (int c = 0; c < h->nch; ++c)
                for (int k = 0; k < no; ++k) rows[((size_t)c * sweeps + sw) * no + k] = tmp[no * c + k];
        }
        h->sweep += 1;
    }
    return check_health(h);
}

int create_impl(const kdq_model_desc *models, const kdq_run_desc *runs, int32_t nchains,
                const uint32_t *streams, kdq_handle *out);
void scee_free(kdq_handle h);
void scee_set_kb(kdq_handle h, int kb);
}  // namespace

// =================================================================== C ABI
extern "C" {

const char *kdq_last_error(kdq_handle h) {
    if (h) return h->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err.c_str();
}

// Block-circulant kinetic factors: with ns = 2 L² (site 2·(x + L·y) + b on an L x L periodic
// cell lattice), checks that eK and eK⁻¹ satisfy E[(c,a),(c',b)] = E_ab(c' - c) to 1e-12 of
// their largest entry (the model constructors' eigen-decompositions leave ~1e-15) and returns L
// with tab = [eK, eK⁻¹][k][a][b] complex Λ(k) = (1/L²) Σ_d E_ab(d) e^{+2πi (kx dx + ky dy)/L};
// 0 when the matrices are not block-circulant or L has no kernel instantiation.
int circ_table(const kdq_model_desc &m, bool cplx, std::vector<double> &tab, bool &half) {
    const int L = m.site[0];
    if (m.site[1] != L || L < 1 || m.ns != 2 * L * L || !circ_supported(L)) return 0;
    const int ns = m.ns, nc = L * L;
    const size_t nn = (size_t)ns * ns;
    // eK, eK⁻¹ and the measurement's HalfeK, HalfeK⁻¹ (mats 0..3), each checked block-circulant;
    // HalfeK's pair only serves the chained measurement (half = false when it is not circulant)
    tab.assign((size_t)4 * nc * 8, 0.0);
    half = true;
    for (int mat = 0; mat < 4; ++mat) {
        const double *res[4] = {m.eK, m.eKinv, m.halfeK, m.halfeKinv};
        const double *ims[4] = {m.eK_im, m.eKinv_im, m.halfeK_im, m.halfeKinv_im};
        const double *re = res[mat];
        const double *im = cplx ? ims[mat] : nullptr;
        if (!re) {
            if (mat < 2) return 0;
            half = false;
            break;
        }
        double big = 0.0;
        for (size_t q = 0; q < nn; ++q) big = std::max(big, std::abs(re[q]) + (im ? std::abs(im[q]) : 0.0));
        const double tol = 1e-12 * big;
        bool circulant = true;
        for (int j = 0; j < ns && circulant; ++j)
            for (int i = 0; i < ns; ++i) {
                const int ci = i / 2, a = i % 2, cj = j / 2, b = j % 2;
                const int dx = ((cj % L) - (ci % L) + L) % L, dy = ((cj / L) - (ci / L) + L) % L;
                const size_t o = (size_t)i + (size_t)j * ns, o0 = (size_t)a + (size_t)(2 * (dx + L * dy) + b) * ns;
                if (std::abs(re[o] - re[o0]) > tol || (im && std::abs(im[o] - im[o0]) > tol)) {
                    circulant = false;
                    break;
                }
            }
        if (!circulant) {
            if (mat < 2) return 0;
            half = false;
            break;
        }
        for (int k = 0; k < nc; ++k) {
            const int kx = k % L, ky = k / L;
            for (int ab = 0; ab < 4; ++ab) {
                const int a = ab >> 1, b = ab & 1;
                double sr = 0.0, si = 0.0;
                for (int d = 0; d < nc; ++d) {
                    const int dx = d % L, dy = d / L;
                    const size_t o = (size_t)a + (size_t)(2 * d + b) * ns;
                    const double er = re[o], ei = im ? im[o] : 0.0;
                    const double ph = 2.0 * M_PI * (double)((kx * dx + ky * dy) % L) / L;
                    const double cr = std::cos(ph), cs = std::sin(ph);
                    sr += er * cr - ei * cs;
                    si += er * cs + ei * cr;
                }
                double *t = tab.data() + ((size_t)mat * nc + k) * 8 + 2 * ab;
                t[0] = sr / nc;
                t[1] = si / nc;
            }
        }
    }
    return L;
}

int kdq_create(const kdq_model_desc *models, const kdq_run_desc *runs, int32_t nchains, kdq_handle *out) {
    if (!models || !runs || !out || nchains < 1) return create_impl(models, runs, nchains, nullptr, out);
    // group the chains by (nt, nodes); one group is the plain batch
    std::vector<std::vector<int>> groups;
    for (int c = 0; c < nchains; ++c) {
        const kdq_model_desc &m = models[c];
        if (!m.nodes || m.n_nodes < 1) return create_impl(models, runs, nchains, nullptr, out);   // its own message
        int g = 0;
        for (; g < (int)groups.size(); ++g) {
            const kdq_model_desc &f = models[groups[g][0]];
            if (f.nt == m.nt && f.n_nodes == m.n_nodes && std::equal(m.nodes, m.nodes + m.n_nodes, f.nodes)) break;
        }
        if (g == (int)groups.size()) groups.emplace_back();
        groups[g].push_back(c);
    }
    if (groups.size() == 1) return create_impl(models, runs, nchains, nullptr, out);
    *out = nullptr;
    for (int c = 1; c < nchains; ++c) {   // what the groups must still share
        if (models[c].kind != models[0].kind)
            return fail(nullptr, KDQ_EARG, "kdq_create: all chains of a batch must share the model kind");
        if (runs[c].delay != runs[0].delay) return fail(nullptr, KDQ_EARG, "kdq_create: delay must be equal across chains");
        if (runs[c].precision != runs[0].precision)
            return fail(nullptr, KDQ_EARG, "kdq_create: precision must be equal across chains");
    }
    kdq_handle h = new kdq_handle_s();
    h->nch = nchains;
    h->kind = models[0].kind;
    h->gmap.resize(nchains);
    for (const auto &grp : groups) {
        std::vector<kdq_model_desc> gm;
        std::vector<kdq_run_desc> gr;
        for (int c : grp) {
            gm.push_back(models[c]);
            gr.push_back(runs[c]);
        }
        kdq_handle sub = nullptr;
        if (int r = create_impl(gm.data(), gr.data(), (int)grp.size(), nullptr, &sub)) {
            kdq_destroy(h);
            return r;
        }
        for (int l = 0; l < (int)grp.size(); ++l) h->gmap[grp[l]] = {(int)h->grp.size(), l};
        h->grp.push_back(sub);
        h->gch.push_back(grp);
    }
    h->nobs = h->grp[0]->nobs;
    h->delay = h->grp[0]->delay;
    *out = h;
    return KDQ_OK;
}
}  // extern "C"

namespace {
// ---- routing for batches of unequal nt (kdq_create with more than one (nt, nodes) group)
int g_chain(kdq_handle h, int c, kdq_handle &sub, int &lc, const char *who) {
    if (c < 0 || c >= h->nch) return fail(h, KDQ_EARG, std::string(who) + ": bad argument");
    sub = h->grp[h->gmap[c].first];
    lc = h->gmap[c].second;
    return KDQ_OK;
}

template <class F>
int g_each(kdq_handle h, F f) {
    for (kdq_handle sub : h->grp)
        if (int r = f(sub)) return fail(h, r, sub->err);
    return KDQ_OK;
}

// per-chain outputs of width w: each group's rows land at its chains' batch indices
template <class F>
int g_gather(kdq_handle h, double *out, int w, F f) {
    for (size_t g = 0; g < h->grp.size(); ++g) {
        kdq_handle sub = h->grp[g];
        std::vector<double> buf((size_t)sub->nch * w);
        if (int r = f(sub, buf.data())) return fail(h, r, sub->err);
        for (int l = 0; l < sub->nch; ++l)
            for (int k = 0; k < w; ++k) out[(size_t)h->gch[g][l] * w + k] = buf[(size_t)l * w + k];
    }
    return KDQ_OK;
}

// The groups sweep concurrently, one host thread each (each group replays its own captured
// sweep on its own stream); rows come back in the batch's chain order.
int g_phy_update(kdq_handle h, int sweeps, int record, double *rows) {
    const size_t G = h->grp.size();
    const int no = h->nobs;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::vector<std::vector<double>> buf(G);
    std::vector<int> rc(G, KDQ_OK);
    std::vector<std::thread> th;
    // device allocations before the threads: one group's allocation must not fall into
    // another group's graph capture
    if (record)
        for (kdq_handle sub : h->grp)
            if (sub->kind == KDQ_MODEL_TU && !sub->scee)
                if (int r = mq_alloc(sub)) return fail(h, r, sub->err);
    for (size_t g = 0; g < G; ++g) {
        if (record) buf[g].resize(std::max<size_t>(1, (size_t)h->grp[g]->nch * sweeps * no));   // non-null at 0 sweeps
        th.emplace_back([&, g] {
            if (hipSetDevice(dev) != hipSuccess) {
                rc[g] = fail(h->grp[g], KDQ_EHIP, "kdq_phy_update: hipSetDevice failed in a group's thread");
                return;
            }
            rc[g] = kdq_phy_update(h->grp[g], sweeps, record, record ? buf[g].data() : nullptr);
        });
    }
    for (auto &t : th) t.join();
    for (size_t g = 0; g < G; ++g)
        if (rc[g]) return fail(h, rc[g], h->grp[g]->err);
    if (record)
        for (size_t g = 0; g < G; ++g)
            for (int l = 0; l < h->grp[g]->nch; ++l)
                std::memcpy(rows + (size_t)h->gch[g][l] * sweeps * no, buf[g].data() + (size_t)l * sweeps * no,
                            sizeof(double) * sweeps * no);
    return KDQ_OK;
}

int g_per_slice(kdq_handle h, const char *who) {
    return fail(h, KDQ_EARG, std::string(who) + ": per-slice calls need one nt; this batch has chains of " +
                                 std::to_string(h->grp.size()) + " different nt (use one handle per nt)");
}
}  // namespace

namespace {
// Builds a handle over `nchains` entries; streams[c] (null: all 0) is the Philox stream of
// entry c (the SCEE handle holds replica r of chain c as entry c + r·nch with stream 1 + r).
int create_impl(const kdq_model_desc *models, const kdq_run_desc *runs, int32_t nchains,
                const uint32_t *streams, kdq_handle *out) {
    if (!models || !runs || !out || nchains < 1) return fail(nullptr, KDQ_EARG, "kdq_create: null argument or nchains < 1");
    *out = nullptr;
    int ns_all = 0;
    for (int c = 0; c < nchains; ++c) ns_all = std::max(ns_all, (int)models[c].ns);
    for (int c = 0; c < nchains; ++c) {
        const kdq_model_desc &m = models[c];
        if (m.kind < KDQ_MODEL_TU || m.kind > KDQ_MODEL_TUV) return fail(nullptr, KDQ_EARG, "kdq_create: unknown model kind");
        if (m.kind != models[0].kind) return fail(nullptr, KDQ_EARG, "kdq_create: all chains of a batch must share the model kind");
        if (m.ns < 2 || m.ns % 2 || m.nt < 1 || m.n_nodes < 2 || !m.nodes || (!m.alpha && m.kind != KDQ_MODEL_TUV) ||
            !m.eta || !m.gamma || !m.K || !m.eK || !m.eKinv || !m.halfeK || !m.halfeKinv || !m.Pt)
            return fail(nullptr, KDQ_EARG, "kdq_create: invalid model description for chain " + std::to_string(c));
        if (m.kind != KDQ_MODEL_TU && (m.z != 3 || !m.bonds || m.lattice > KDQ_LATTICE_HC60))
            return fail(nullptr, KDQ_EARG, "kdq_create: t-V and t-U-V need honeycomb bonds (z = 3)");
        const int kmax = m.kind == KDQ_MODEL_TU ? 64 : 32;
        if (runs[c].delay > kmax) return fail(nullptr, KDQ_EARG, "kdq_create: delay above the block limit of the model");
        if (m.ns / 2 > 256) return fail(nullptr, KDQ_EARG, "kdq_create: ns/2 above 256 is not supported");
        if (m.nt != models[0].nt || m.n_nodes != models[0].n_nodes)
            return fail(nullptr, KDQ_EARG, "kdq_create: all chains of a batch must share nt and nodes");
        for (int k = 0; k < m.n_nodes; ++k)
            if (m.nodes[k] != models[0].nodes[k]) return fail(nullptr, KDQ_EARG, "kdq_create: nodes differ between chains");
        if (m.nodes[0] != 0 || m.nodes[m.n_nodes - 1] != m.nt) return fail(nullptr, KDQ_EARG, "kdq_create: nodes must start at 0 and end at nt");
        if (runs[c].delay < 1 || runs[c].delay > 64) return fail(nullptr, KDQ_EARG, "kdq_create: delay must be in 1..64");
        if (runs[c].delay != runs[0].delay) return fail(nullptr, KDQ_EARG, "kdq_create: delay must be equal across chains");
        if (runs[c].precision != KDQ_PREC_FP64 && runs[c].precision != KDQ_PREC_MIXED)
            return fail(nullptr, KDQ_EARG, "kdq_create: unknown precision");
        if (runs[c].precision != runs[0].precision) return fail(nullptr, KDQ_EARG, "kdq_create: precision must be equal across chains");
        if (runs[c].precision == KDQ_PREC_MIXED && (m.kind != KDQ_MODEL_TU || streams))
            return fail(nullptr, KDQ_EARG, "kdq_create: mixed precision is built for the plain t-U sweep only");
        if (runs[c].chain >= (1u << 24)) return fail(nullptr, KDQ_EARG, "kdq_create: chain id must be < 2^24");
        const int nim = (m.K_im != nullptr) + (m.eK_im != nullptr) + (m.eKinv_im != nullptr) +
                        (m.halfeK_im != nullptr) + (m.halfeKinv_im != nullptr) + (m.Pt_im != nullptr);
        if (nim != 0 && nim != 6)
            return fail(nullptr, KDQ_EARG, "kdq_create: give all six imaginary parts of the model matrices or none");
        if (nim && m.kind != KDQ_MODEL_TU)
            return fail(nullptr, KDQ_EARG, "kdq_create: complex model matrices are for the t-U model only");
        if ((nim != 0) != (models[0].K_im != nullptr))
            return fail(nullptr, KDQ_EARG, "kdq_create: the chains of a batch must all be real or all complex");
        if (nim && runs[c].precision != KDQ_PREC_FP64)
            return fail(nullptr, KDQ_EARG, "kdq_create: complex model matrices need FP64 precision");
    }
    kdq_handle h = new kdq_handle_s();
    h->nch = nchains;
    h->nt = models[0].nt;
    h->nn = models[0].n_nodes;
    h->nodes.assign(models[0].nodes, models[0].nodes + h->nn);
    h->delay = runs[0].delay;
    h->kind = models[0].kind;
    h->mixed = runs[0].precision == KDQ_PREC_MIXED;
    h->cK = models[0].K_im != nullptr;
    h->defer = false;   // per-block flush by default (faster on the measured workloads)
    h->nobs = h->kind == KDQ_MODEL_TU ? 6 : (h->kind == KDQ_MODEL_TV ? 10 : 7);
    h->md.assign(models, models + nchains);
    h->rd.assign(runs, runs + nchains);
    h->sweep = runs[0].sweep0;
    h->ch.resize(nchains);
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_app, hipEventDisableTiming);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&h->ev_dec[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_fields, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_chain, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_q1, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_q2, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_mfork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_mdone, hipEventDisableTiming);
    if (e != hipSuccess) {
        int r = fail(h, KDQ_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        delete h;
        return r;
    }
    std::vector<ChainDev> cd(nchains);
    std::vector<Opnd> v_eK, v_eKinv, v_hK, v_hKinv, v_Pt, v_G0, v_G1, v_T, v_Gm, v_Tm, v_P0, v_P1, v_P2, v_QT, v_M, v_X, v_S;
    std::vector<OpndF> v_G32a, v_G32b, v_T32, v_eK32, v_eKinv32, v_FZ32, v_FC32;
    std::vector<Opnd> v_FZ, v_FC, v_DU, v_DV, v_DI, v_S2, v_DI2;
    std::vector<OpndF> v_DU32, v_DV32;
    std::vector<BlockDesc> blk_all;
    std::vector<int> blk_offset, blk_count;
    std::vector<std::vector<Opnd>> v_UR(h->nn), v_ULH(h->nn);
    std::vector<int> circ_L(nchains, 0);
    for (int c = 0; c < nchains; ++c) {
        const kdq_model_desc &m = models[c];
        ChainHost &C = h->ch[c];
        const int n = m.ns, np = m.ns / 2;
        C.ns = n; C.np = np;
        h->max_ns = std::max(h->max_ns, n);
        h->max_np = std::max(h->max_np, np);
        const size_t nn = (size_t)n * n, nP = (size_t)n * np, pp = (size_t)np * np;
        bool ok = true;
        const size_t planes = h->cK ? 2 : 1;   // complex model matrices: real plane, then imaginary
        ok &= dalloc(C, &C.eK, planes * nn) == hipSuccess;
        ok &= dalloc(C, &C.eKinv, planes * nn) == hipSuccess;
        ok &= dalloc(C, &C.hK, planes * nn) == hipSuccess;
        ok &= dalloc(C, &C.hKinv, planes * nn) == hipSuccess;
        ok &= dalloc(C, &C.K, planes * nn) == hipSuccess;
        ok &= dalloc(C, &C.Pt, planes * nP) == hipSuccess;
        ok &= dalloc(C, &C.alpha, (size_t)m.nt) == hipSuccess;
        ok &= dalloc(C, &C.G[0], 2 * nn) == hipSuccess;
        ok &= dalloc(C, &C.G[1], 2 * nn) == hipSuccess;
        ok &= dalloc(C, &C.T, 2 * nn) == hipSuccess;
        ok &= dalloc(C, &C.Gm, 2 * nn) == hipSuccess;
        ok &= dalloc(C, &C.Tm, 2 * nn) == hipSuccess;
        ok &= dalloc(C, &C.P0, 2 * nP) == hipSuccess;
        ok &= dalloc(C, &C.P1, 2 * nP) == hipSuccess;
        ok &= dalloc(C, &C.P2, 2 * nP) == hipSuccess;
        ok &= dalloc(C, &C.QT, 2 * nP) == hipSuccess;
        ok &= dalloc(C, &C.M, 2 * pp) == hipSuccess;
        ok &= dalloc(C, &C.S, 2 * pp) == hipSuccess;
        ok &= dalloc(C, &C.X, 2 * nP) == hipSuccess;
        C.UR.resize(h->nn);
        C.ULH.resize(h->nn);
        for (int k = 0; k < h->nn; ++k) {
            ok &= dalloc(C, &C.UR[k], 2 * nP) == hipSuccess;
            ok &= dalloc(C, &C.ULH[k], 2 * nP) == hipSuccess;
        }
        const int n_elem = m.kind == KDQ_MODEL_TU ? n : (m.kind == KDQ_MODEL_TV ? 3 * (n / 2) : n * m.z);
        h->max_elem = std::max(h->max_elem, n_elem);
        h->max_nb = std::max(h->max_nb, n / 2);
        ok &= dalloc(C, &C.s, (size_t)n_elem * m.nt) == hipSuccess;
        std::vector<int> hx, hy, hinc;
        if (m.kind != KDQ_MODEL_TU) {
            hx.resize(n_elem);
            hy.resize(n_elem);
            for (int e = 0; e < n_elem; ++e) {
                hx[e] = m.bonds[2 * e] - 1;
                hy[e] = m.bonds[2 * e + 1] - 1;
                if (hx[e] < 0 || hx[e] >= n || hy[e] < 0 || hy[e] >= n || hx[e] == hy[e]) {
                    int r = fail(h, KDQ_EARG, "kdq_create: bond list out of range");
                    delete h;
                    return r;
                }
            }
            ok &= dalloc(C, &C.ex, n_elem) == hipSuccess;
            ok &= dalloc(C, &C.ey, n_elem) == hipSuccess;
            if (m.kind == KDQ_MODEL_TUV) {
                std::vector<int> cnt(n, 0);
                hinc.assign((size_t)n * 6, 0);
                for (int e = 0; e < n_elem; ++e) {
                    for (int w = 0; w < 2; ++w) {
                        const int x = w ? hy[e] : hx[e];
                        if (cnt[x] >= 6) { ok = false; break; }
                        hinc[(size_t)x * 6 + cnt[x]++] = e * 2 + w;
                    }
                }
                for (int x = 0; x < n; ++x) ok &= cnt[x] == 6;
                if (!ok) {
                    int r = fail(h, KDQ_EARG, "kdq_create: every site must be on 3 + 3 directed bonds");
                    delete h;
                    return r;
                }
                ok &= dalloc(C, &C.inc, (size_t)n * 6) == hipSuccess;
            }
        }
        ok &= dalloc(C, &C.acc_cnt, 2) == hipSuccess;
        ok &= dalloc(C, &C.acc_idx, 2 * 64) == hipSuccess;
        ok &= dalloc(C, &C.pw_fld, 64) == hipSuccess;
        ok &= dalloc(C, &C.info, 1) == hipSuccess;
        ok &= dalloc(C, &C.W, 2 * 2 * 64 * 64) == hipSuccess;
        ok &= dalloc(C, &C.FZ, 2 * 64 * (size_t)n) == hipSuccess;
        ok &= dalloc(C, &C.FC, 2 * 64 * (size_t)n) == hipSuccess;
        ok &= dalloc(C, &C.rk, (size_t)ns_all + 2) == hipSuccess;
        ok &= dalloc(C, &C.DI, 2 * 32 * (size_t)np) == hipSuccess;   // blocks run over the largest chain
        if (m.kind == KDQ_MODEL_TU) {
            ok &= dalloc(C, &C.S2, 2 * pp) == hipSuccess;
            ok &= dalloc(C, &C.DI2, 2 * 32 * (size_t)np) == hipSuccess;
        }
        ok &= dalloc(C, &C.rank, 1) == hipSuccess;
        if (m.kind == KDQ_MODEL_TU && !h->mixed) {
            ok &= dalloc(C, &C.DU, 2 * nn) == hipSuccess;
            ok &= dalloc(C, &C.DV, 2 * nn) == hipSuccess;
        }
        ok &= dalloc(C, &C.counters, 3) == hipSuccess;
        ok &= dalloc(C, &C.meas, KDQ_MEAS_SLOTS) == hipSuccess;
        ok &= dalloc(C, &C.logdet, 1) == hipSuccess;
        if (h->mixed) {
            ok &= dalloc(C, &C.G32[0], 2 * nn) == hipSuccess;
            ok &= dalloc(C, &C.G32[1], 2 * nn) == hipSuccess;
            ok &= dalloc(C, &C.T32, 2 * nn) == hipSuccess;
            ok &= dalloc(C, &C.eK32, nn) == hipSuccess;
            ok &= dalloc(C, &C.eKinv32, nn) == hipSuccess;
            ok &= dalloc(C, &C.FZ32, 2 * 64 * (size_t)n) == hipSuccess;
            ok &= dalloc(C, &C.FC32, 2 * 64 * (size_t)n) == hipSuccess;
            if (m.kind == KDQ_MODEL_TU) {
                ok &= dalloc(C, &C.DU32, 2 * nn) == hipSuccess;
                ok &= dalloc(C, &C.DV32, 2 * nn) == hipSuccess;
            }
        }
        if (!ok) {
            int r = fail(h, KDQ_EHIP, "kdq_create: device allocation failed");
            delete h;
            return r;
        }
        hipMemcpy(C.eK, m.eK, nn * 8, hipMemcpyHostToDevice);
        hipMemcpy(C.eKinv, m.eKinv, nn * 8, hipMemcpyHostToDevice);
        if (h->mixed) {
            std::vector<float> f1(nn), f2(nn);
            for (size_t q = 0; q < nn; ++q) {
                f1[q] = (float)m.eK[q];
                f2[q] = (float)m.eKinv[q];
            }
            hipMemcpy(C.eK32, f1.data(), nn * 4, hipMemcpyHostToDevice);
            hipMemcpy(C.eKinv32, f2.data(), nn * 4, hipMemcpyHostToDevice);
        }
        hipMemcpy(C.hK, m.halfeK, nn * 8, hipMemcpyHostToDevice);
        hipMemcpy(C.hKinv, m.halfeKinv, nn * 8, hipMemcpyHostToDevice);
        hipMemcpy(C.K, m.K, nn * 8, hipMemcpyHostToDevice);
        hipMemcpy(C.Pt, m.Pt, nP * 8, hipMemcpyHostToDevice);
        if (h->cK) {
            hipMemcpy(C.eK + nn, m.eK_im, nn * 8, hipMemcpyHostToDevice);
            hipMemcpy(C.eKinv + nn, m.eKinv_im, nn * 8, hipMemcpyHostToDevice);
            hipMemcpy(C.hK + nn, m.halfeK_im, nn * 8, hipMemcpyHostToDevice);
            hipMemcpy(C.hKinv + nn, m.halfeKinv_im, nn * 8, hipMemcpyHostToDevice);
            hipMemcpy(C.K + nn, m.K_im, nn * 8, hipMemcpyHostToDevice);
            hipMemcpy(C.Pt + nP, m.Pt_im, nP * 8, hipMemcpyHostToDevice);
        }
        {
            std::vector<double> tab;
            bool half = false;
            circ_L[c] = circ_table(m, h->cK, tab, half);
            if (!half) h->mchain_ok = false;
            if (circ_L[c]) {
                if (dalloc(C, &C.circ, tab.size()) != hipSuccess) {
                    int r = fail(h, KDQ_EHIP, "kdq_create: device allocation failed");
                    delete h;
                    return r;
                }
                hipMemcpy(C.circ, tab.data(), tab.size() * 8, hipMemcpyHostToDevice);
            }
        }
        if (m.alpha) hipMemcpy(C.alpha, m.alpha, (size_t)m.nt * 8, hipMemcpyHostToDevice);
        else {
            std::vector<double> ones((size_t)m.nt, 1.0);
            hipMemcpy(C.alpha, ones.data(), (size_t)m.nt * 8, hipMemcpyHostToDevice);
        }
        if (C.ex) {
            hipMemcpy(C.ex, hx.data(), sizeof(int) * hx.size(), hipMemcpyHostToDevice);
            hipMemcpy(C.ey, hy.data(), sizeof(int) * hy.size(), hipMemcpyHostToDevice);
        }
        if (C.inc) hipMemcpy(C.inc, hinc.data(), sizeof(int) * hinc.size(), hipMemcpyHostToDevice);
        // rank-2 blocks: t-V per family (bonds i of family j are elements j*nb + i), t-U-V over the slice
        if (m.kind != KDQ_MODEL_TU) {
            const int kb = runs[c].delay, nb = n / 2;
            const int nfam = m.kind == KDQ_MODEL_TV ? 3 : 1;
            const int per = m.kind == KDQ_MODEL_TV ? nb : n_elem;
            const int nblk = (per + kb - 1) / kb;
            blk_offset.push_back((int)blk_all.size());
            blk_count.push_back(nblk);
            h->max_blocks = std::max(h->max_blocks, nblk);
            for (int f = 0; f < nfam; ++f)
                for (int b = 0; b < nblk; ++b) {
                    BlockDesc bd;
                    memset(&bd, 0, sizeof(bd));
                    bd.e0 = f * per + b * kb;
                    bd.ne = std::min(kb, per - b * kb);
                    for (int q = 0; q < bd.ne; ++q) {
                        const int e = bd.e0 + q;
                        const int xs[2] = {hx[e], hy[e]};
                        uint8_t sl[2];
                        for (int w = 0; w < 2; ++w) {
                            int k = 0;
                            while (k < bd.nslot && bd.sites[k] != xs[w]) ++k;
                            if (k == bd.nslot) bd.sites[bd.nslot++] = xs[w];
                            sl[w] = (uint8_t)k;
                        }
                        bd.sl0[q] = sl[0];
                        bd.sl1[q] = sl[1];
                    }
                    blk_all.push_back(bd);
                }
        }

        ChainDev &d = cd[c];
        memset(&d, 0, sizeof(d));
        d.ns = n; d.np = np; d.nt = m.nt;
        d.alpha = C.alpha;
        for (int k = 0; k < 4; ++k) { d.eta[k] = m.eta[k]; d.gamma[k] = m.gamma[k]; }
        d.s = C.s;
        d.ht = m.ht;
        d.lattice = m.lattice; d.L1 = m.site[0]; d.L2 = m.site[1];
        d.K = C.K;
        d.Kim = h->cK ? C.K + nn : nullptr;
        d.seed = runs[c].seed;
        d.chain_id = runs[c].chain;
        d.stream = streams ? streams[c] : 0;
        d.acc_cnt = C.acc_cnt; d.acc_idx = C.acc_idx; d.pw_fld = C.pw_fld; d.W = C.W;
        d.rk = C.rk; d.rank = C.rank;
        d.counters = C.counters; d.meas = C.meas;
        d.kind = m.kind;
        d.n_elem = n_elem;
        d.is_complex = m.kind == KDQ_MODEL_TUV ? (m.is_complex ? 1 : 0) : 1;
        d.nb = n / 2;
        d.ex = C.ex; d.ey = C.ey; d.inc = C.inc;
        d.a = m.kind == KDQ_MODEL_TUV ? m.a : 0.0;
        d.U = m.U; d.V = m.V;

        if (h->cK) {
            v_eK.push_back(cplx(C.eK, n, nn)); v_eKinv.push_back(cplx(C.eKinv, n, nn));
            v_hK.push_back(cplx(C.hK, n, nn)); v_hKinv.push_back(cplx(C.hKinv, n, nn));
            v_Pt.push_back(cplx(C.Pt, n, nP));
        } else {
            v_eK.push_back(real(C.eK, n)); v_eKinv.push_back(real(C.eKinv, n));
            v_hK.push_back(real(C.hK, n)); v_hKinv.push_back(real(C.hKinv, n));
            v_Pt.push_back(real(C.Pt, n));
        }
        v_G0.push_back(cplx(C.G[0], n, nn)); v_G1.push_back(cplx(C.G[1], n, nn));
        v_T.push_back(cplx(C.T, n, nn)); v_Gm.push_back(cplx(C.Gm, n, nn)); v_Tm.push_back(cplx(C.Tm, n, nn));
        v_P0.push_back(cplx(C.P0, n, nP)); v_P1.push_back(cplx(C.P1, n, nP)); v_P2.push_back(cplx(C.P2, n, nP));
        v_QT.push_back(cplx(C.QT, n, nP));
        v_M.push_back(cplx(C.M, np, pp)); v_X.push_back(cplx(C.X, np, nP));
        v_S.push_back(cplx(C.S, np, pp));
        v_FZ.push_back(Opnd{C.FZ, C.FZ + 64 * (size_t)n, 64, 0});
        v_DI.push_back(Opnd{C.DI, C.DI + 32 * (size_t)np, np, 0});
        v_S2.push_back(Opnd{C.S2, C.S2 ? C.S2 + pp : nullptr, np, 0});
        v_DI2.push_back(Opnd{C.DI2, C.DI2 ? C.DI2 + 32 * (size_t)np : nullptr, np, 0});
        v_DU.push_back(Opnd{C.DU, C.DU ? C.DU + nn : nullptr, n, 0});
        v_DV.push_back(Opnd{C.DV, C.DV ? C.DV + nn : nullptr, n, 0});
        v_DU32.push_back(OpndF{C.DU32, C.DU32 ? C.DU32 + nn : nullptr, n, 0});
        v_DV32.push_back(OpndF{C.DV32, C.DV32 ? C.DV32 + nn : nullptr, n, 0});
        v_FC.push_back(Opnd{C.FC, C.FC + 64 * (size_t)n, n, 0});
        if (h->mixed) {
            v_FZ32.push_back(OpndF{C.FZ32, C.FZ32 + 64 * (size_t)n, 64, 0});
            v_FC32.push_back(OpndF{C.FC32, C.FC32 + 64 * (size_t)n, n, 0});
            v_G32a.push_back(OpndF{C.G32[0], C.G32[0] + nn, n, 0});
            v_G32b.push_back(OpndF{C.G32[1], C.G32[1] + nn, n, 0});
            v_T32.push_back(OpndF{C.T32, C.T32 + nn, n, 0});
            v_eK32.push_back(OpndF{C.eK32, nullptr, n, 0});
            v_eKinv32.push_back(OpndF{C.eKinv32, nullptr, n, 0});
        }
        for (int k = 0; k < h->nn; ++k) {
            v_UR[k].push_back(cplx(C.UR[k], n, nP));
            v_ULH[k].push_back(cplx(C.ULH[k], n, nP));
        }
    }
    // kinetic products through cell momenta when every chain's factors are block-circulant
    h->circ_L = *std::min_element(circ_L.begin(), circ_L.end()) > 0 ? *std::max_element(circ_L.begin(), circ_L.end()) : 0;
    for (int c = 0; c < nchains; ++c) {
        cd[c].cL = h->ch[c].cL = h->circ_L ? circ_L[c] : 0;
        cd[c].circ = h->circ_L ? h->ch[c].circ : nullptr;
    }
    {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(ChainDev) * nchains) != hipSuccess) {
            int r = fail(h, KDQ_EHIP, "kdq_create: device allocation failed");
            delete h;
            return r;
        }
        h->dev_misc.push_back(q);
        hipMemcpy(q, cd.data(), sizeof(ChainDev) * nchains, hipMemcpyHostToDevice);
        h->d_chains = (ChainDev *)q;
        h->cd_host = cd;
    }
    int r = 0;
    r |= make_table(h, &h->t_eK, v_eK); r |= make_table(h, &h->t_eKinv, v_eKinv);
    r |= make_table(h, &h->t_hK, v_hK); r |= make_table(h, &h->t_hKinv, v_hKinv);
    r |= make_table(h, &h->t_Pt, v_Pt);
    r |= make_table(h, &h->t_G[0], v_G0); r |= make_table(h, &h->t_G[1], v_G1);
    r |= make_table(h, &h->t_T, v_T); r |= make_table(h, &h->t_Gm, v_Gm); r |= make_table(h, &h->t_Tm, v_Tm);
    r |= make_table(h, &h->t_P0, v_P0); r |= make_table(h, &h->t_P1, v_P1); r |= make_table(h, &h->t_P2, v_P2);
    r |= make_table(h, &h->t_QT, v_QT);
    if (!blk_all.empty()) {
        void *q1 = nullptr, *q2 = nullptr, *q3 = nullptr;
        if (hipMalloc(&q1, sizeof(BlockDesc) * blk_all.size()) == hipSuccess) h->dev_misc.push_back(q1); else r = 1;
        if (hipMalloc(&q2, sizeof(int) * nchains) == hipSuccess) h->dev_misc.push_back(q2); else r = 1;
        if (hipMalloc(&q3, sizeof(int) * nchains) == hipSuccess) h->dev_misc.push_back(q3); else r = 1;
        if (!r) {
            hipMemcpy(q1, blk_all.data(), sizeof(BlockDesc) * blk_all.size(), hipMemcpyHostToDevice);
            hipMemcpy(q2, blk_offset.data(), sizeof(int) * nchains, hipMemcpyHostToDevice);
            hipMemcpy(q3, blk_count.data(), sizeof(int) * nchains, hipMemcpyHostToDevice);
        }
        h->d_blocks = (BlockDesc *)q1;
        h->d_block_offset = (int *)q2;
        h->d_nblocks = (int *)q3;
    }
    r |= make_table(h, &h->t_M, v_M); r |= make_table(h, &h->t_X, v_X); r |= make_table(h, &h->t_S, v_S);
    r |= make_table(h, &h->t_FZ, v_FZ); r |= make_table(h, &h->t_FC, v_FC);
    r |= make_table(h, &h->t_DU, v_DU); r |= make_table(h, &h->t_DV, v_DV);
    r |= make_table(h, &h->t_DI, v_DI);
    r |= make_table(h, &h->t_S2, v_S2); r |= make_table(h, &h->t_DI2, v_DI2);
    r |= make_table_f(h, &h->t_DU32, v_DU32); r |= make_table_f(h, &h->t_DV32, v_DV32);
    if (h->mixed) {
        r |= make_table_f(h, &h->t_G32[0], v_G32a); r |= make_table_f(h, &h->t_G32[1], v_G32b);
        r |= make_table_f(h, &h->t_T32, v_T32);
        r |= make_table_f(h, &h->t_eK32, v_eK32); r |= make_table_f(h, &h->t_eKinv32, v_eKinv32);
        r |= make_table_f(h, &h->t_FZ32, v_FZ32); r |= make_table_f(h, &h->t_FC32, v_FC32);
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(double) * nchains) == hipSuccess) {
            h->dev_misc.push_back(q);
            hipMemset(q, 0, sizeof(double) * nchains);
            h->d_maxerr = (double *)q;
        } else {
            r = 1;
        }
    }
    h->t_UR.resize(h->nn);
    h->t_ULH.resize(h->nn);
    for (int k = 0; k < h->nn; ++k) {
        r |= make_table(h, &h->t_UR[k], v_UR[k]);
        r |= make_table(h, &h->t_ULH[k], v_ULH[k]);
    }
    {
        void *q1 = nullptr, *q2 = nullptr, *q3 = nullptr, *q4 = nullptr;
        if (hipMalloc(&q1, sizeof(double) * nchains) == hipSuccess) h->dev_misc.push_back(q1); else r = 1;
        if (hipMalloc(&q2, sizeof(int) * nchains) == hipSuccess) h->dev_misc.push_back(q2); else r = 1;
        if (hipMalloc(&q3, sizeof(int) * nchains) == hipSuccess) h->dev_misc.push_back(q3); else r = 1;
        if (hipMalloc(&q4, sizeof(int) * nchains * (size_t)h->max_np) == hipSuccess) h->dev_misc.push_back(q4); else r = 1;
        h->d_logdet = (double *)q1;
        h->d_info = (int *)q2;
        h->d_qinfo = (int *)q3;
        h->d_piv = (int *)q4;
        void *q5 = nullptr, *q6 = nullptr, *q7 = nullptr, *q8 = nullptr, *q9 = nullptr;
        if (hipMalloc(&q9, sizeof(double) * meas_part_doubles(nchains, h->max_ns)) == hipSuccess) h->dev_misc.push_back(q9); else r = 1;
        h->d_mpart = (double *)q9;
        if (hipMalloc(&q5, sizeof(uint32_t)) == hipSuccess) h->dev_misc.push_back(q5); else r = 1;
        if (hipMalloc(&q6, sizeof(double) * nchains) == hipSuccess) h->dev_misc.push_back(q6); else r = 1;
        if (hipMalloc(&q7, sizeof(int) * nchains) == hipSuccess) h->dev_misc.push_back(q7); else r = 1;
        if (hipMalloc(&q8, sizeof(double) * nchains) == hipSuccess) h->dev_misc.push_back(q8); else r = 1;
        h->d_sweep = (uint32_t *)q5;
        h->d_lw = (double *)q6;
        h->d_poison = (int *)q7;
        h->d_wraperr = (double *)q8;
        if (h->kind == KDQ_MODEL_TU) {   // prewrap partial products (t-U pipelined layers)
            void *q10 = nullptr;
            const size_t pw = (size_t)nchains * ((h->max_ns + 31) / 32) * 2048;
            if (hipMalloc(&q10, sizeof(double) * pw) == hipSuccess) h->dev_misc.push_back(q10); else r = 1;
            h->d_pwpart = (double *)q10;
        }
        if (!r) {
            hipMemset(q2, 0, sizeof(int) * nchains);
            hipMemset(q3, 0, sizeof(int) * nchains);
            hipMemset(q7, 0, sizeof(int) * nchains);
            hipMemset(q8, 0, sizeof(double) * nchains);
        }
    }
    if (r || hipHostMalloc((void **)&h->h_meas, sizeof(double) * KDQ_MEAS_SLOTS * nchains, hipHostMallocDefault) != hipSuccess) {
        int rr = fail(h, KDQ_EHIP, "kdq_create: table allocation failed");
        delete h;
        return rr;
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        int rr = fail(h, KDQ_EHIP, "kdq_create: device error during set-up");
        delete h;
        return rr;
    }
    *out = h;
    return KDQ_OK;
}
}  // namespace

extern "C" {
int kdq_destroy(kdq_handle h) {
    if (h)
        for (kdq_handle sub : h->grp) kdq_destroy(sub);
    if (h && h->scee) scee_free(h);
    delete h;
    return KDQ_OK;
}

int kdq_num_chains(kdq_handle h) { return h ? h->nch : 0; }

int kdq_num_obs(kdq_handle h) { return h ? h->nobs : 0; }

int kdq_set_fields(kdq_handle h, int32_t c, const uint8_t *s) {
    if (!h || !s || c < 0 || c >= h->nch) return fail(h, KDQ_EARG, "kdq_set_fields: bad argument");
    if (!h->grp.empty()) {
        kdq_handle sub;
        int l;
        g_chain(h, c, sub, l, "kdq_set_fields");
        if (int r = kdq_set_fields(sub, l, s)) return fail(h, r, sub->err);
        return KDQ_OK;
    }
    const size_t n = (size_t)elems(h, c) * h->nt;
    for (size_t i = 0; i < n; ++i)
        if (s[i] < 1 || s[i] > 4) return fail(h, KDQ_EARG, "kdq_set_fields: field values must be 1..4");
    HIPCHK(hipMemcpyAsync(h->ch[c].s, s, n, hipMemcpyHostToDevice, h->stream));
    static const int zero = 0;   // new fields: the chain's sticky failure flags start over
    HIPCHK(hipMemcpyAsync(h->d_info + c, &zero, sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_qinfo + c, &zero, sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_poison + c, &zero, sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->sl_logw_valid = false;
    return KDQ_OK;
}

int kdq_get_health(kdq_handle h, double *out) {
    if (!h || !out) return fail(h, KDQ_EARG, "kdq_get_health: bad argument");
    if (!h->grp.empty()) return g_gather(h, out, 4, [](kdq_handle g, double *o) { return kdq_get_health(g, o); });
    const int n = h->nch;
    std::vector<double> err(n);
    std::vector<int> info(n), qinfo(n), pois(n);
    HIPCHK(hipMemcpyAsync(err.data(), h->d_wraperr, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(pois.data(), h->d_poison, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(info.data(), h->d_info, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(qinfo.data(), h->d_qinfo, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<double> zero(n, 0.0);
    HIPCHK(hipMemcpy(h->d_wraperr, zero.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    for (int c = 0; c < n; ++c) {
        out[4 * c] = err[c];
        out[4 * c + 1] = pois[c] ? 1.0 : 0.0;
        out[4 * c + 2] = info[c] ? 1.0 : 0.0;
        out[4 * c + 3] = qinfo[c] ? 1.0 : 0.0;
    }
    return KDQ_OK;
}

int kdq_check_G(kdq_handle h, double *out) {
    if (!h || !out) return fail(h, KDQ_EARG, "kdq_check_G: bad argument");
    if (!h->grp.empty()) return g_gather(h, out, 2, [](kdq_handle g, double *o) { return kdq_check_G(g, o); });
    if (h->scee) return fail(h, KDQ_EARG, "kdq_check_G: sweep handles only (not swap-replica handles)");
    const int n = h->nch;
    Opnd *G = h->t_G[h->cur];
    if (h->mixed)
        HIPCHK(launch_to_f64(h->stream, n, h->max_ns, h->d_chains, h->t_G32[h->cur32], h->t_Gm, DIM_NS, 1));
    else
        HIPCHK(launch_copy_mat(h->stream, n, h->max_ns, h->max_ns, h->d_chains, G, h->t_Gm, DIM_NS, DIM_NS, true));
    GemmArgs g = gargs(h, h->t_Gm, h->t_Gm, h->t_Tm, DIM_NS, DIM_NS, DIM_NS);   // G²
    if (int r = gemm(h, true, true, 0, 0, g)) return r;
    HIPCHK(hipMemsetAsync(h->d_lw, 0, sizeof(double) * n, h->stream));
    HIPCHK(launch_node_check(h->stream, n, h->max_ns, h->d_chains, h->t_Tm, h->t_Gm, DIM_NS, h->d_lw, nullptr));
    std::vector<double> e(n), tr(n);
    HIPCHK(hipMemcpyAsync(e.data(), h->d_lw, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(launch_trace(h->stream, n, h->d_chains, h->t_Gm, h->d_lw));
    HIPCHK(hipMemcpyAsync(tr.data(), h->d_lw, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < n; ++c) {
        const int ns = h->ch[c].ns;
        out[2 * c] = std::isfinite(tr[c]) ? e[c] : tr[c];
        out[2 * c + 1] = std::fabs(tr[c] - (ns - ns / 2));
    }
    return KDQ_OK;
}

int kdq_get_fields(kdq_handle h, int32_t c, uint8_t *s) {
    if (!h || !s || c < 0 || c >= h->nch) return fail(h, KDQ_EARG, "kdq_get_fields: bad argument");
    if (!h->grp.empty()) {
        kdq_handle sub;
        int l;
        g_chain(h, c, sub, l, "kdq_get_fields");
        if (int r = kdq_get_fields(sub, l, s)) return fail(h, r, sub->err);
        return KDQ_OK;
    }
    HIPCHK(hipMemcpyAsync(s, h->ch[c].s, (size_t)elems(h, c) * h->nt, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_initial_s(kdq_handle h) {
    if (!h) return fail(h, KDQ_EARG, "kdq_initial_s: null handle");
    if (!h->grp.empty()) return g_each(h, [](kdq_handle g) { return kdq_initial_s(g); });
    HIPCHK(launch_initial_s(h->stream, h->nch, h->max_elem, h->nt, h->d_chains));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

// A chain with replayed draws must hold the sweeps [first, first + count) of the call.
static int check_replay(kdq_handle h, uint32_t first, uint32_t count, const char *who) {
    for (int c = 0; c < (int)h->rp_buf.size(); ++c) {
        const ChainDev &d = h->cd_host[c];
        if (!d.rp_u || count == 0) continue;
        const uint32_t s = first - d.rp_sweep0;
        if (s >= d.rp_nsweeps || count > d.rp_nsweeps - s)
            return fail(h, KDQ_EARG, std::string(who) + ": the replayed draws of chain " + std::to_string(c) +
                                         " cover sweeps " + std::to_string(d.rp_sweep0) + ".." +
                                         std::to_string(d.rp_sweep0 + d.rp_nsweeps - 1) + ", the call needs " +
                                         std::to_string(first) + ".." + std::to_string(first + count - 1));
    }
    return KDQ_OK;
}

int kdq_phy_update(kdq_handle h, int32_t sweeps, int32_t record, double *rows) {
    if (!h || sweeps < 0) return fail(h, KDQ_EARG, "kdq_phy_update: bad argument");
    if (record && !rows) return fail(h, KDQ_EARG, "kdq_phy_update: rows must not be NULL when record != 0");
    if (!h->grp.empty()) return g_phy_update(h, sweeps, record, rows);
    if (int r = check_replay(h, h->sweep, (uint32_t)sweeps, "kdq_phy_update")) return r;
    return phy_update(h, sweeps, record, rows);
}

int kdq_set_replay(kdq_handle h, int32_t c, const uint8_t *choice, const double *u, int64_t nsweeps) {
    if (!h || c < 0 || c >= h->nch || nsweeps < 0 || (nsweeps > 0 && (!choice || !u)))
        return fail(h, KDQ_EARG, "kdq_set_replay: bad argument");
    if (!h->grp.empty()) {
        kdq_handle sub;
        int l;
        g_chain(h, c, sub, l, "kdq_set_replay");
        if (int r = kdq_set_replay(sub, l, choice, u, nsweeps)) return fail(h, r, sub->err);
        return KDQ_OK;
    }
    if (h->scee) return fail(h, KDQ_EARG, "kdq_set_replay: sweep handles only (not swap-replica handles)");
    if (nsweeps > 0xFFFFFFF) return fail(h, KDQ_EARG, "kdq_set_replay: too many sweeps");
    ChainDev &d = h->cd_host[c];
    const size_t count = (size_t)nsweeps * 2 * (size_t)h->nt * (size_t)d.n_elem;
    for (size_t q = 0; q < count; ++q)
        if (choice[q] > 2) return fail(h, KDQ_EARG, "kdq_set_replay: choice values must be 0, 1 or 2");
    if (h->rp_buf.empty()) h->rp_buf.assign(h->nch, nullptr);
    HIPCHK(hipStreamSynchronize(h->stream));   // no kernel may still read the old buffer
    if (h->rp_buf[c]) {
        HIPCHK(hipFree(h->rp_buf[c]));
        h->rp_buf[c] = nullptr;
    }
    d.rp_choice = nullptr;
    d.rp_u = nullptr;
    d.rp_sweep0 = d.rp_nsweeps = 0;
    if (count) {
        const size_t bytes_u = sizeof(double) * count, bytes = bytes_u + count;
        void *q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess) return fail(h, KDQ_EHIP, "kdq_set_replay: device allocation failed");
        h->rp_buf[c] = q;
        HIPCHK(hipMemcpy(q, u, bytes_u, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((char *)q + bytes_u, choice, count, hipMemcpyHostToDevice));
        d.rp_u = (const double *)q;
        d.rp_choice = (const uint8_t *)q + bytes_u;
        d.rp_sweep0 = h->sweep;
        d.rp_nsweeps = (uint32_t)nsweeps;
    }
    HIPCHK(hipMemcpy(h->d_chains + c, &d, sizeof(ChainDev), hipMemcpyHostToDevice));
    return KDQ_OK;
}

int kdq_get_G(kdq_handle h, int32_t c, double *G) {
    if (!h || !G || c < 0 || c >= h->nch) return fail(h, KDQ_EARG, "kdq_get_G: bad argument");
    if (!h->grp.empty()) {
        kdq_handle sub;
        int l;
        g_chain(h, c, sub, l, "kdq_get_G");
        if (int r = kdq_get_G(sub, l, G)) return fail(h, r, sub->err);
        return KDQ_OK;
    }
    const int n = h->ch[c].ns;
    const size_t nn = (size_t)n * n;
    std::vector<double> tmp(2 * nn);
    if (h->mixed) {   // the live G is the FP32 copy
        std::vector<float> f(2 * nn);
        HIPCHK(hipMemcpyAsync(f.data(), h->ch[c].G32[h->cur32], 2 * nn * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (size_t q = 0; q < 2 * nn; ++q) tmp[q] = f[q];
    } else {
        HIPCHK(hipMemcpyAsync(tmp.data(), h->ch[c].G[h->cur], 2 * nn * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    for (size_t q = 0; q < nn; ++q) {
        G[2 * q] = tmp[q];
        G[2 * q + 1] = tmp[nn + q];
    }
    return KDQ_OK;
}

int kdq_set_G(kdq_handle h, int32_t c, const double *G) {
    if (!h || !G || c < 0 || c >= h->nch) return fail(h, KDQ_EARG, "kdq_set_G: bad argument");
    if (!h->grp.empty()) {
        kdq_handle sub;
        int l;
        g_chain(h, c, sub, l, "kdq_set_G");
        if (int r = kdq_set_G(sub, l, G)) return fail(h, r, sub->err);
        return KDQ_OK;
    }
    const int n = h->ch[c].ns;
    const size_t nn = (size_t)n * n;
    std::vector<double> tmp(2 * nn);
    for (size_t q = 0; q < nn; ++q) {
        tmp[q] = G[2 * q];
        tmp[nn + q] = G[2 * q + 1];
    }
    HIPCHK(hipMemcpyAsync(h->ch[c].G[h->cur], tmp.data(), 2 * nn * 8, hipMemcpyHostToDevice, h->stream));
    if (int r = refresh_g32(h, false)) return r;
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_wrap(kdq_handle h, int32_t lt, int32_t forward) {
    if (h && !h->grp.empty()) return g_per_slice(h, "kdq_wrap");
    if (!h || lt < 1 || lt > h->nt) return fail(h, KDQ_EARG, "kdq_wrap: bad argument");
    if (int r = wrap_slice(h, h->t_G[h->cur], h->t_T, lt, forward != 0, KDQ_PROF_WRAP)) return r;
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_update_layer(kdq_handle h, int32_t lt, int32_t dir, uint32_t sweep) {
    if (h && !h->grp.empty()) return g_per_slice(h, "kdq_update_layer");
    if (!h || lt < 1 || lt > h->nt || (dir != 0 && dir != 1)) return fail(h, KDQ_EARG, "kdq_update_layer: bad argument");
    if (int r = check_replay(h, sweep, 1, "kdq_update_layer")) return r;
    if (h->kind == KDQ_MODEL_TV) {
        for (int k = 0; k < 3; ++k)   // forward pass visits families 3,2,1, backward 1,2,3
            if (int r = update_family_r2(h, lt, dir, sweep, dir == 0 ? 2 - k : k)) return r;
    } else if (int r = update_layer(h, lt, dir, sweep)) {
        return r;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_G4(kdq_handle h, int32_t c, int32_t tau1, int32_t tau2, double *G1, double *G2, double *G12,
           double *G21) {
    if (!h || h->scee || c < 0 || c >= h->nch || !G1 || !G2 || !G12 || !G21)
        return fail(h, KDQ_EARG, "kdq_G4: bad argument");
    if (!h->grp.empty()) {
        kdq_handle sub;
        int l;
        g_chain(h, c, sub, l, "kdq_G4");
        if (int r = kdq_G4(sub, l, tau1, tau2, G1, G2, G12, G21)) return fail(h, r, sub->err);
        return KDQ_OK;
    }
    if (h->kind != KDQ_MODEL_TU) return fail(h, KDQ_EARG, "kdq_G4: time-displaced G is for t-U chains");
    if (tau1 < 0 || tau1 > h->nt || tau2 < 0 || tau2 > h->nt) return fail(h, KDQ_EARG, "kdq_G4: slice out of range");
    const bool swap = tau1 < tau2;
    Opnd *res[4] = {};
    if (int r = g4_compute(h, swap ? tau2 : tau1, swap ? tau1 : tau2, res)) return r;
    if (int r = check_health(h)) return r;
    const int n = h->ch[c].ns;
    const size_t nn = (size_t)n * n;
    std::vector<double> tmp(2 * nn);
    auto fetch = [&](Opnd *tab, double *dst, bool minus_identity) -> int {
        Opnd o;
        HIPCHK(hipMemcpy(&o, tab + c, sizeof(Opnd), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(tmp.data(), o.re, 2 * nn * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < nn; ++q) {
            dst[2 * q] = tmp[q] - (minus_identity && q % (n + 1) == 0 ? 1.0 : 0.0);
            dst[2 * q + 1] = tmp[nn + q];
        }
        return KDQ_OK;
    };
    HIPCHK(hipStreamSynchronize(h->stream));
    if (tau1 == tau2) {   // G, G, -(I - G), G
        if (int r = fetch(res[0], G1, false)) return r;
        if (int r = fetch(res[0], G2, false)) return r;
        if (int r = fetch(res[0], G12, true)) return r;
        return fetch(res[0], G21, false);
    }
    if (int r = fetch(res[swap ? 1 : 0], G1, false)) return r;
    if (int r = fetch(res[swap ? 0 : 1], G2, false)) return r;
    if (int r = fetch(res[swap ? 3 : 2], G12, false)) return r;
    return fetch(res[swap ? 2 : 3], G21, false);
}

int kdq_init_chain(kdq_handle h) {
    if (!h) return fail(h, KDQ_EARG, "kdq_init_chain: null handle");
    if (!h->grp.empty()) return g_each(h, [](kdq_handle g) { return kdq_init_chain(g); });
    if (int r = init_chain(h)) return r;
    return check_health(h);
}

int kdq_measure(kdq_handle h, int32_t lt, double *out) {
    if (h && !h->grp.empty()) return g_per_slice(h, "kdq_measure");
    if (!h || !out || lt < 1 || lt > h->nt) return fail(h, KDQ_EARG, "kdq_measure: bad argument");
    HIPCHK(launch_zero_meas(h->stream, h->nch, h->d_chains));
    if (int r = measure(h, lt)) return r;
    return read_meas(h, out);
}

}  // extern "C"

// ---- measurement without device allocation -------------------------------------------------
// Working memory of kdq_measure_noalloc: a file-scope device array, one copy per GPU, held
// until the process ends. Host code never passes `scratch` itself to a kernel or a copy: the
// symbol's name is a host-side stand-in; the device address comes from hipGetSymbolAddress,
// looked up once per device and kept in g_scratch_ptr.
namespace {

constexpr size_t KDQ_NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[KDQ_NOALLOC_SCRATCH_BYTES];

constexpr int KDQ_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[KDQ_MAX_DEVICES] = {};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Device address of this device's scratch array (dev is the current device).
int scratch_base(kdq_handle h, int dev, unsigned char **out) {
    if (dev < 0 || dev >= KDQ_MAX_DEVICES) return fail(h, KDQ_EARG, "kdq_measure_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(h, KDQ_EHIP, std::string("kdq_measure_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return KDQ_OK;
}

// Scratch layout: [Opnd tables Gm[nch], Tm[nch]] [accumulators nch x KDQ_MEAS_SLOTS]
// [tile partial sums of the measurement] [per chain: Gm re, Gm im, Tm re, Tm im, each ns x ns
// doubles padded to 256 bytes].
struct NoallocLayout {
    size_t off_meas, off_part, off_mats, total;
};

NoallocLayout noalloc_layout(kdq_handle h) {
    NoallocLayout L{};
    L.off_meas = align256(2 * sizeof(Opnd) * (size_t)h->nch);
    L.off_part = L.off_meas + align256(sizeof(double) * KDQ_MEAS_SLOTS * (size_t)h->nch);
    L.off_mats = L.off_part + align256(sizeof(double) * meas_part_doubles(h->nch, h->max_ns));
    size_t t = L.off_mats;
    for (int c = 0; c < h->nch; ++c) t += 4 * align256(sizeof(double) * (size_t)h->ch[c].ns * h->ch[c].ns);
    L.total = t;
    return L;
}

// Fills the operand tables that point each chain's Gm / Tm into the scratch array (the same
// prefix sums as noalloc_layout, from the chains' ns) and clears the accumulators.
__global__ void k_noalloc_layout(const ChainDev *chains, int nch, unsigned char *base, size_t off_meas,
                                 size_t off_mats) {
    Opnd *tGm = (Opnd *)base, *tTm = tGm + nch;
    double *meas = (double *)(base + off_meas);
    for (int q = threadIdx.x; q < nch * KDQ_MEAS_SLOTS; q += blockDim.x) meas[q] = 0.0;
    for (int c = threadIdx.x; c < nch; c += blockDim.x) {
        size_t off = off_mats;
        for (int k = 0; k < c; ++k) off += 4 * (((size_t)chains[k].ns * chains[k].ns * sizeof(double) + 255) & ~size_t(255));
        const int n = chains[c].ns;
        const size_t plane = ((size_t)n * n * sizeof(double) + 255) & ~size_t(255);
        double *g = (double *)(base + off);
        tGm[c] = Opnd{g, (double *)(base + off + plane), n, 0};
        tTm[c] = Opnd{(double *)(base + off + 2 * plane), (double *)(base + off + 3 * plane), n, 0};
    }
}

}  // namespace

extern "C" {

int kdq_measure_noalloc(kdq_handle h, int32_t lt, double *out) {
    if (h && !h->grp.empty()) return g_per_slice(h, "kdq_measure_noalloc");
    if (!h || !out || lt < 1 || lt > h->nt) return fail(h, KDQ_EARG, "kdq_measure_noalloc: bad argument");
    if (h->scee) return fail(h, KDQ_EARG, "kdq_measure_noalloc: sweep handles only (not swap-replica handles)");
    int dev = -1, sdev = -1;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipStreamGetDevice(h->stream, &sdev));
    if (dev != sdev) return fail(h, KDQ_EARG, "kdq_measure_noalloc: the handle's device must be the current device");
    const NoallocLayout L = noalloc_layout(h);
    if (L.total > KDQ_NOALLOC_SCRATCH_BYTES)
        return fail(h, KDQ_EARG, "kdq_measure_noalloc: the batch needs " + std::to_string(L.total) +
                                     " bytes of working memory; the limit is " +
                                     std::to_string(KDQ_NOALLOC_SCRATCH_BYTES) + " (use kdq_measure)");
    unsigned char *base = nullptr;
    if (int r = scratch_base(h, dev, &base)) return r;
    hipLaunchKernelGGL(k_noalloc_layout, dim3(1), dim3(256), 0, h->stream, h->d_chains, h->nch, base, L.off_meas,
                       L.off_mats);
    HIPCHK(hipGetLastError());
    Opnd *tGm = (Opnd *)base, *tTm = tGm + h->nch;
    double *meas = (double *)(base + L.off_meas);
    const uint32_t mask = h->prof_mask;   // kernel timing would create events: off for this call
    h->prof_mask = 0;
    const int r = measure_into(h, lt, tGm, tTm, meas, (double *)(base + L.off_part));
    h->prof_mask = mask;
    if (r) return r;
    const int S = KDQ_MEAS_SLOTS, no = h->nobs;
    HIPCHK(hipMemcpyAsync(h->h_meas, meas, sizeof(double) * S * (size_t)h->nch, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < h->nch; ++c) {
        const double cnt = h->h_meas[S * c + S - 1];
        for (int k = 0; k < no; ++k) out[no * c + k] = cnt > 0 ? h->h_meas[S * c + k] / cnt : 0.0;
    }
    return KDQ_OK;
}

int kdq_profile(kdq_handle h, uint32_t mask) {
    if (!h) return fail(h, KDQ_EARG, "kdq_profile: null handle");
    if (!h->grp.empty()) return g_each(h, [mask](kdq_handle g) { return kdq_profile(g, mask); });
    harvest(h);
    h->prof_mask = mask;
    for (int f = 0; f < KDQ_PROF_FAMILIES; ++f) {
        h->prof_ms[f] = 0.0;
        h->prof_n[f] = 0;
    }
    return KDQ_OK;
}

int kdq_profile_read(kdq_handle h, double *ms, int64_t *launches) {
    if (!h || !ms || !launches) return fail(h, KDQ_EARG, "kdq_profile_read: bad argument");
    if (!h->grp.empty()) {   // summed over the groups
        for (int f = 0; f < KDQ_PROF_FAMILIES; ++f) {
            ms[f] = 0.0;
            launches[f] = 0;
        }
        return g_each(h, [&](kdq_handle g) {
            double m[KDQ_PROF_FAMILIES];
            int64_t n[KDQ_PROF_FAMILIES];
            if (int r = kdq_profile_read(g, m, n)) return r;
            for (int f = 0; f < KDQ_PROF_FAMILIES; ++f) {
                ms[f] += m[f];
                launches[f] += n[f];
            }
            return (int)KDQ_OK;
        });
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    harvest(h);
    for (int f = 0; f < KDQ_PROF_FAMILIES; ++f) {
        ms[f] = h->prof_ms[f];
        launches[f] = h->prof_n[f];
    }
    return KDQ_OK;
}

int kdq_set_option(kdq_handle h, int32_t key, int32_t value) {
    if (!h) return fail(h, KDQ_EARG, "kdq_set_option: null handle");
    if (!h->grp.empty()) {
        if (int r = g_each(h, [key, value](kdq_handle g) { return kdq_set_option(g, key, value); })) return r;
        if (key == KDQ_OPT_DELAY) h->delay = value;
        return KDQ_OK;
    }
    switch (key) {
        case KDQ_OPT_GRAPH:
            h->use_graph = value ? 1 : 0;
            break;
        case KDQ_OPT_DELAY: {
            const bool r2_scee = h->scee && h->kind != KDQ_MODEL_TU;
            const int kmax = h->scee ? (r2_scee ? 8 : 16) : 64;
            if ((h->kind != KDQ_MODEL_TU && !r2_scee) || value < 1 || value > kmax)
                return fail(h, KDQ_EARG, "kdq_set_option: delay is adjustable for t-U chains (1..64, 1..16 with "
                                         "replicas) and t-V / t-U-V replica chains (1..8 bonds or elements)");
            h->delay = value;
            if (h->scee) scee_set_kb(h, value);
            break;
        }
        case KDQ_OPT_GEMM_TILE:
            if (value != 0 && value != 32 && value != 64) return fail(h, KDQ_EARG, "kdq_set_option: GEMM tile is 0, 32 or 64");
            h->gemm_tile = value;
            break;
        case KDQ_OPT_DEFER:
            if (h->kind != KDQ_MODEL_TU) return fail(h, KDQ_EARG, "kdq_set_option: the deferred flush is for t-U chains");
            h->defer = value != 0;
            break;
        case KDQ_OPT_PIPE:
            h->pipe = value != 0;
            break;
        case KDQ_OPT_CIRC:
            h->circ = value != 0;
            break;
        case KDQ_OPT_MASIDE:
            h->maside = value != 0;
            break;
        case KDQ_OPT_GINV:
            h->ginv = value != 0;
            break;
        case KDQ_OPT_QINV:
            h->qinv = value != 0;
            break;
        case KDQ_OPT_PREWRAP:
            h->prewrap = value != 0;
            break;
        case KDQ_OPT_MCHAIN:
            h->mchain = value != 0;
            break;
        case KDQ_OPT_QSPLIT:
            h->qsplit = value != 0;
            break;
        case KDQ_OPT_ASIDE:
            h->aside = value != 0;
            break;
        case KDQ_OPT_FLUSH:
            if (value != 0 && value != 1) return fail(h, KDQ_EARG, "kdq_set_option: flush form is 0 (two launches) or 1 (one launch)");
            h->flush1 = value;
            break;
        default:
            return fail(h, KDQ_EARG, "kdq_set_option: unknown option");
    }
    drop_graph(h);
    return KDQ_OK;
}

int kdq_get_precision_error(kdq_handle h, double *out) {
    if (!h || !out) return fail(h, KDQ_EARG, "kdq_get_precision_error: bad argument");
    if (!h->grp.empty())
        return g_gather(h, out, 1, [](kdq_handle g, double *o) { return kdq_get_precision_error(g, o); });
    if (!h->mixed) {
        for (int c = 0; c < h->nch; ++c) out[c] = 0.0;
        return KDQ_OK;
    }
    HIPCHK(hipMemcpyAsync(out, h->d_maxerr, sizeof(double) * h->nch, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemsetAsync(h->d_maxerr, 0, sizeof(double) * h->nch, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_slmc_move(kdq_handle h, kdq_net net, int32_t nsteps, double *accepted) {
    if (h && !h->grp.empty()) return fail(h, KDQ_EARG, "kdq_slmc_move: the network takes one lattice and nt");
    if (!h || !net || !accepted || nsteps < 0 || h->kind != KDQ_MODEL_TU || h->scee)
        return fail(h, KDQ_EARG, "kdq_slmc_move: t-U batches with a network only");
    int nsn = 0, ntn = 0;
    slmc_geometry(net, &nsn, &ntn);
    for (int c = 0; c < h->nch; ++c)
        if (h->ch[c].ns != nsn || h->nt != ntn) return fail(h, KDQ_EARG, "kdq_slmc_move: network lattice differs from a chain's");
    const int nch = h->nch;
    const size_t ne = (size_t)nsn * h->nt;
    if (!h->d_sl_ptrs) {   // first move: scratch
        h->sl_trial.resize(nch);
        h->sl_backup.resize(nch);
        int bad = 0;
        for (int c = 0; c < nch; ++c) {
            void *q1 = nullptr, *q2 = nullptr;
            if (hipMalloc(&q1, ne) == hipSuccess) h->dev_misc.push_back(q1); else bad = 1;
            if (hipMalloc(&q2, ne) == hipSuccess) h->dev_misc.push_back(q2); else bad = 1;
            h->sl_trial[c] = (uint8_t *)q1;
            h->sl_backup[c] = (uint8_t *)q2;
        }
        void *q[4] = {};
        const size_t sz[4] = {sizeof(uint8_t *) * nch, sizeof(uint64_t) * nch, sizeof(uint32_t) * nch,
                              sizeof(double) * nch};
        for (int k = 0; k < 4; ++k) {
            if (hipMalloc(&q[k], sz[k]) == hipSuccess) h->dev_misc.push_back(q[k]); else bad = 1;
        }
        if (bad) return fail(h, KDQ_EHIP, "kdq_slmc_move: device allocation failed");
        h->d_sl_ptrs = (uint8_t **)q[0];
        h->d_sl_seed = (uint64_t *)q[1];
        h->d_sl_chain = (uint32_t *)q[2];
        h->d_sl_E = (double *)q[3];
        std::vector<uint64_t> sd(nch);
        std::vector<uint32_t> cid(nch);
        for (int c = 0; c < nch; ++c) {
            sd[c] = h->rd[c].seed;
            cid[c] = h->rd[c].chain;
        }
        HIPCHK(hipMemcpy(h->d_sl_ptrs, h->sl_trial.data(), sizeof(uint8_t *) * nch, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->d_sl_seed, sd.data(), sizeof(uint64_t) * nch, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->d_sl_chain, cid.data(), sizeof(uint32_t) * nch, hipMemcpyHostToDevice));
    }
    if (!h->sl_logw_valid) {
        h->sl_logw.assign(nch, 0.0);
        if (int r = log_weight(h, h->sl_logw.data())) return r;
        h->sl_logw_valid = true;
    }
    for (int c = 0; c < nch; ++c) {
        HIPCHK(hipMemcpyAsync(h->sl_trial[c], h->ch[c].s, ne, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->sl_backup[c], h->ch[c].s, ne, hipMemcpyDeviceToDevice, h->stream));
    }
    std::vector<double> E0(nch), E1(nch), lw1(nch);
    if (int r = slmc_inner(net, h->stream, nch, h->d_sl_ptrs, h->d_sl_seed, h->d_sl_chain, h->d_sl_E, nsteps,
                           h->sl_move, E0.data()))
        return fail(h, r, std::string("kdq_slmc_move: ") + kdq_net_last_error(net));
    HIPCHK(hipMemcpyAsync(E1.data(), h->d_sl_E, sizeof(double) * nch, hipMemcpyDeviceToHost, h->stream));
    for (int c = 0; c < nch; ++c)
        HIPCHK(hipMemcpyAsync(h->ch[c].s, h->sl_trial[c], ne, hipMemcpyDeviceToDevice, h->stream));
    if (int r = log_weight(h, lw1.data())) return r;   // synchronises
    // acceptance min(1, W(s')/W(s) · exp(E(s') - E(s))) with the exact FP64 weights
    for (int c = 0; c < nch; ++c) {
        uint32_t x[4];
        philox4x32_10(0u, (uint32_t)h->sl_move, 0x51C2u, (4u << 24) | (h->rd[c].chain & 0xFFFFFFu), (uint32_t)h->rd[c].seed,
                      (uint32_t)(h->rd[c].seed >> 32), x);
        const double u = ((double)(x[1] >> 5) * 67108864.0 + (double)(x[2] >> 6)) * (1.0 / 9007199254740992.0);
        const double la = lw1[c] - h->sl_logw[c] + (E1[c] - E0[c]);
        const bool acc = std::isfinite(la) && (la >= 0.0 || u < std::exp(la));
        accepted[c] = acc ? 1.0 : 0.0;
        if (acc) h->sl_logw[c] = lw1[c];
        else HIPCHK(hipMemcpyAsync(h->ch[c].s, h->sl_backup[c], ne, hipMemcpyDeviceToDevice, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->sl_move += 1;
    return check_health(h);
}

int kdq_log_weight(kdq_handle h, double *out) {
    if (!h || !out || h->kind != KDQ_MODEL_TU || h->scee) return fail(h, KDQ_EARG, "kdq_log_weight: t-U batches only");
    if (!h->grp.empty()) return g_gather(h, out, 1, [](kdq_handle g, double *o) { return kdq_log_weight(g, o); });
    if (int r = log_weight(h, out)) return r;
    return check_health(h);
}

int kdq_get_logdet(kdq_handle h, double *out) {
    if (!h || !out) return fail(h, KDQ_EARG, "kdq_get_logdet: bad argument");
    if (!h->grp.empty()) return g_gather(h, out, 1, [](kdq_handle g, double *o) { return kdq_get_logdet(g, o); });
    HIPCHK(hipMemcpyAsync(out, h->d_logdet, sizeof(double) * h->nch, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_get_counters(kdq_handle h, double *out) {
    if (!h || !out) return fail(h, KDQ_EARG, "kdq_get_counters: bad argument");
    if (!h->grp.empty()) return g_gather(h, out, 2, [](kdq_handle g, double *o) { return kdq_get_counters(g, o); });
    for (int c = 0; c < h->nch; ++c)
        HIPCHK(hipMemcpyAsync(out + 2 * c, h->ch[c].counters, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_get_near_ties(kdq_handle h, double *out) {
    if (!h || !out) return fail(h, KDQ_EARG, "kdq_get_near_ties: bad argument");
    if (!h->grp.empty()) return g_gather(h, out, 1, [](kdq_handle g, double *o) { return kdq_get_near_ties(g, o); });
    for (int c = 0; c < h->nch; ++c)
        HIPCHK(hipMemcpyAsync(out + c, h->ch[c].counters + 2, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

}  // extern "C"

// =================================================================== swap-replica Rényi-2
// ctrl_SCEEicr of the t-U model (src/tU/SCEE.jl:1-262). The handle holds replica r of chain
// c as entry c + r·nch (Philox stream 1 + r), so wraps, block products, chain
// orthonormalisation and get_G! run over all 2·nch entries in the same launches as the
// plain sweep. Per entry: G(τ) = Gt (the plain G buffers), G(Θ) = G0, G(τ,Θ) = Gt0,
// G(Θ,τ) = G0t (ping-pong pairs), and the block products BM / BM⁻¹ of every node block.
struct SceeState {
    int nch = 0, kb = 8;
    int nrep = 2;                  // replicas per chain: 2 (Rényi-2) or 1 (disorder operator)
    bool dop = false;
    int na_max = 0, nb_max = 0, nx_max = 0, m2_max = 0;   // m2: largest nA² + nB²
    Opnd *t_X[3][2] = {};          // G0, Gt0, G0t ping-pong tables over the 2·nch entries
    int cur[3] = {0, 0, 0};
    std::vector<Opnd *> t_BM, t_BMi;   // block idx (1-based) at idx - 1
    Opnd *t_M[2][2] = {}, *t_C[2] = {}, *t_Inv[2] = {}, *t_IX[2] = {};   // region tables (nch)
    SceeRegion *d_reg = nullptr;
    double *d_ld = nullptr;        // [2][nch] region log|det|
    double *d_ph = nullptr;        // [2][nch][2] region det phase (t-V keeps the signed det)
    int *d_info = nullptr, *d_piv = nullptr;
    double *h_osum = nullptr;      // pinned [nch][2]
    std::vector<SceeRegion> reg;   // host copy (device pointers)
    std::vector<void *> allocs;
};

namespace {

enum { SX_G0 = 0, SX_GT0 = 1, SX_G0T = 2 };

Opnd *xcur(kdq_handle h, int k) { return h->scee->t_X[k][h->scee->cur[k]]; }
Opnd *xspare(kdq_handle h, int k) { return h->scee->t_X[k][1 - h->scee->cur[k]]; }
void xflip(kdq_handle h, int k) { h->scee->cur[k] ^= 1; }

void scee_set_kb(kdq_handle h, int kb) { h->scee->kb = kb; }

void scee_free(kdq_handle h) {
    SceeState *S = h->scee;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (void *p : S->allocs) hipFree(p);
    if (S->h_osum) hipHostFree(S->h_osum);
    delete S;
    h->scee = nullptr;
}

// BM_F! / BMinv_F! (src/tU/GreenMatrix.jl:10-47) of block idx for every entry
int scee_blocks(kdq_handle h, int idx) {
    SceeState *S = h->scee;
    const int lo = (int)h->nodes[idx - 1] + 1, hi = (int)h->nodes[idx], nb = hi - lo + 1;
    Opnd *BM = S->t_BM[idx - 1], *BMi = S->t_BMi[idx - 1];
    if (h->kind == KDQ_MODEL_TV) {
        // src/tV/GreenMatrix.jl:12-67: per slice BM <- Π_{j=3..1}(UV_j D_j UV_j) eK BM and
        // BMinv <- BMinv eKinv Π_{j=3..1}(UV_j D_j⁻¹ UV_j)
        HIPCHK(launch_diag(h->stream, h->nch, h->max_ns, h->d_chains, BM, DIM_NS, 0.0, 1));
        HIPCHK(launch_diag(h->stream, h->nch, h->max_ns, h->d_chains, BMi, DIM_NS, 0.0, 1));
        for (int lt = lo; lt <= hi; ++lt) {
            GemmArgs g = gargs(h, h->t_eK, BM, h->t_T, DIM_NS, DIM_NS, DIM_NS);
            if (int r = gemm_m(h, true, false, false, true, 0, 0, g, KDQ_PROF_CHAIN)) return r;
            HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, h->t_T, BM, DIM_NS, DIM_NS, true));
            GemmArgs gi = gargs(h, BMi, h->t_eKinv, h->t_Tm, DIM_NS, DIM_NS, DIM_NS);
            if (int r = gemm_m(h, false, true, true, false, 0, 0, gi, KDQ_PROF_CHAIN)) return r;
            HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, h->t_Tm, BMi, DIM_NS, DIM_NS, true));
            for (int j = 2; j >= 0; --j) {
                if (int r = bondrot(h, BM, j, lt, +1, 1, DIM_NS, KDQ_PROF_CHAIN)) return r;
                if (int r = bondrot(h, BMi, j, lt, +1, 2, DIM_NS, KDQ_PROF_CHAIN)) return r;
            }
        }
        return KDQ_OK;
    }
    HIPCHK(launch_diag(h->stream, h->nch, h->max_ns, h->d_chains, h->t_Gm, DIM_NS, 0.0, 1));   // I
    Opnd *src = h->t_Gm, *srci = h->t_Gm;
    for (int s = 0; s < nb; ++s) {
        const int lt = lo + s;
        const bool last_even = ((nb - 1 - s) % 2) == 0;
        Opnd *out = last_even ? BM : h->t_T;
        GemmArgs g = gargs(h, h->t_eK, src, out, DIM_NS, DIM_NS, DIM_NS);   // D eK X
        g.rowS = hs_diag(h, lt, +1);
        if (int r = gemm_m(h, true, false, false, true, 0, 0, g, KDQ_PROF_CHAIN)) return r;
        Opnd *outi = last_even ? BMi : h->t_Tm;
        GemmArgs gi = gargs(h, srci, h->t_eKinv, outi, DIM_NS, DIM_NS, DIM_NS);   // X eK⁻¹ D⁻¹
        gi.colS = hs_diag(h, lt, -1);
        if (int r = gemm_m(h, false, true, true, false, 0, 0, gi, KDQ_PROF_CHAIN)) return r;
        src = out;
        srci = outi;
    }
    return KDQ_OK;
}

// BRMs[i] = QR(BMs[i-1] BRMs[i-1]) and BLMs[i] = RQ(BLMs[i+1] BMs[i]) (src/tU/SCEE.jl:60-86)
int scee_update_BR(kdq_handle h, int i) {
    GemmArgs g = gargs(h, h->scee->t_BM[i - 2], h->t_UR[i - 2], h->t_UR[i - 1], DIM_NS, DIM_NP, DIM_NS);
    if (int r = gemm(h, true, true, 0, 0, g, KDQ_PROF_CHAIN)) return r;
    return orthonormalise(h, h->t_UR[i - 1]);
}
int scee_update_BL(kdq_handle h, int i) {
    GemmArgs g = gargs(h, h->scee->t_BM[i - 1], h->t_ULH[i], h->t_ULH[i - 1], DIM_NS, DIM_NP, DIM_NS);
    if (int r = gemm(h, true, true, 2, 0, g, KDQ_PROF_CHAIN)) return r;   // BMᴴ U_Lᴴ
    return orthonormalise(h, h->t_ULH[i - 1]);
}

// G4! (src/public/Gupdate.jl:159-212) for every entry
int scee_G4(kdq_handle h, int idx, bool forward) {
    const int thidx = h->nn / 2 + 1;
    Opnd *Gt = h->t_G[h->cur];
    if (int r = get_G(h, idx, Gt)) return r;
    if (idx == thidx) {
        for (int k = 0; k < 3; ++k)
            HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, Gt, xcur(h, k), DIM_NS, DIM_NS, true));
        HIPCHK(launch_diag(h->stream, h->nch, h->max_ns, h->d_chains, xcur(h, forward ? SX_G0T : SX_GT0), DIM_NS, -1.0, 0));
        return KDQ_OK;
    }
    if (int r = get_G(h, thidx, xcur(h, SX_G0))) return r;
    HIPCHK(launch_diag(h->stream, h->nch, h->max_ns, h->d_chains, xcur(h, SX_GT0), DIM_NS, 0.0, 1));
    HIPCHK(launch_diag(h->stream, h->nch, h->max_ns, h->d_chains, xcur(h, SX_G0T), DIM_NS, 0.0, 1));
    // idx < Θidx: G0t <- BM_j G_j G0t, Gt0 <- Gt0 (I - G_j) BM_j⁻¹, then Gt0 = -Gt0;
    // idx > Θidx: the same with the roles of Gt0 and G0t exchanged.
    const int j0 = std::min(idx, thidx), j1 = std::max(idx, thidx) - 1;
    const int kp = idx < thidx ? SX_G0T : SX_GT0, km = idx < thidx ? SX_GT0 : SX_G0T;
    for (int j = j0; j <= j1; ++j) {
        Opnd *T2 = h->t_Tm;
        if (j == j0) T2 = idx < thidx ? Gt : xcur(h, SX_G0);
        else if (int r = get_G(h, j, h->t_Tm)) return r;
        GemmArgs g1 = gargs(h, T2, xcur(h, kp), h->t_T, DIM_NS, DIM_NS, DIM_NS);
        if (int r = gemm(h, true, true, 0, 0, g1, KDQ_PROF_GETG)) return r;
        GemmArgs g2 = gargs(h, h->scee->t_BM[j - 1], h->t_T, xspare(h, kp), DIM_NS, DIM_NS, DIM_NS);
        if (int r = gemm(h, true, true, 0, 0, g2, KDQ_PROF_GETG)) return r;
        xflip(h, kp);
        GemmArgs g3 = gargs(h, xcur(h, km), T2, h->t_Gm, DIM_NS, DIM_NS, DIM_NS);   // X - X G_j
        g3.mode = 2;
        g3.Cin = xcur(h, km);
        if (int r = gemm(h, true, true, 0, 0, g3, KDQ_PROF_GETG)) return r;
        GemmArgs g4 = gargs(h, h->t_Gm, h->scee->t_BMi[j - 1], xspare(h, km), DIM_NS, DIM_NS, DIM_NS);
        if (j == j1) g4.mode = 3;   // the final sign flip
        if (int r = gemm(h, true, true, 0, 0, g4, KDQ_PROF_GETG)) return r;
        xflip(h, km);
    }
    return KDQ_OK;
}

SceeArgs scee_args(kdq_handle h) {
    SceeState *S = h->scee;
    SceeArgs a{};
    a.chains = h->d_chains;
    a.Gt = h->t_G[h->cur];
    a.Gt0 = xcur(h, SX_GT0);
    a.G0t = xcur(h, SX_G0T);
    a.G0 = xcur(h, SX_G0);
    a.reg = S->d_reg;
    a.nch = S->nch;
    a.nx_max = S->nx_max;
    return a;
}

// GroverMatrix! of A and B, |det|² and the inverses (get_ABGM!, src/tU/SCEE.jl:264-276), or
// X = e^{iα}I + (1 - e^{iα})G0[X,X], |det X|² and X⁻¹ of the disorder operator (get_ABX!,
// src/tU/disorder_operate.jl:202-223)
int scee_grover(kdq_handle h) {
    SceeState *S = h->scee;
    if (S->nx_max == 0) return KDQ_OK;
    SceeArgs a = scee_args(h);
    HIPCHK(launch_scee_gather(h->stream, S->nch, S->nx_max, a));
    const int dims[2] = {DIM_NA, DIM_NB}, nmax[2] = {S->na_max, S->nb_max};
    for (int X = 0; X < 2 && !S->dop; ++X) {
        if (nmax[X] == 0) continue;
        GemmArgs g{};
        g.chains = h->d_chains;
        g.A = S->t_M[0][X]; g.B = S->t_M[1][X]; g.C = S->t_C[X];
        g.mdim = g.ndim = g.kdim = dims[X];
        HIPCHK(launch_gemm(h->stream, S->nch, nmax[X], nmax[X], nmax[X], true, true, 0, 0, g));
    }
    if (S->dop)
        HIPCHK(launch_dop_form(h->stream, S->nch, S->m2_max, S->d_reg));
    else
        HIPCHK(launch_scee_grover_form(h->stream, S->nch, S->m2_max, S->d_reg));
    for (int X = 0; X < 2; ++X) {
        if (nmax[X] == 0) continue;
        LuArgs la{h->d_chains, S->t_C[X], dims[X], S->nx_max, S->d_piv, S->d_ld + X * S->nch, S->d_info,
                  S->d_ph + 2 * X * S->nch};
        HIPCHK(launch_lu(h->stream, S->nch, nmax[X], la));
        LuSolveArgs sa{h->d_chains, S->t_C[X], S->t_IX[X], S->t_Inv[X], S->d_piv, dims[X], dims[X], S->nx_max, 0};
        HIPCHK(launch_lu_solve(h->stream, S->nch, nmax[X], nmax[X], sa));
    }
    if (h->kind == KDQ_MODEL_TV)   // src/tV/SCEE.jl:436: det g itself (real, signed)
        HIPCHK(launch_scee_detg_signed(h->stream, S->nch, S->d_reg, S->d_ld, S->d_ld + S->nch, S->d_ph,
                                       S->d_ph + 2 * S->nch));
    else
        HIPCHK(launch_scee_detg(h->stream, S->nch, S->d_reg, S->d_ld, S->d_ld + S->nch));
    return KDQ_OK;
}

int scee_get_ABGM(kdq_handle h, int idx, bool forward) {
    if (int r = scee_G4(h, idx, forward)) return r;
    return scee_grover(h);
}

// the four-matrix flush of one block (both replicas, every chain); kmax: the largest
// accepted set of the block (0: the t-U block size)
int scee_flush(kdq_handle h, int kmax = 0) {
    SceeState *S = h->scee;
    const int kb = kmax ? kmax : S->kb;
    Opnd *Gt = h->t_G[h->cur];
    FlushArgs f[4] = {
        {h->d_chains, Gt, h->t_G[1 - h->cur], kb, 0, nullptr, nullptr, h->t_FZ, h->t_FC},                    // Gt
        {h->d_chains, xcur(h, SX_G0T), xspare(h, SX_G0T), kb, 0, nullptr, Gt, h->t_FZ, h->t_FC},              // G0t
        {h->d_chains, xcur(h, SX_GT0), xspare(h, SX_GT0), kb, 1, Gt, nullptr, h->t_FZ, h->t_FC},              // Gt0
        {h->d_chains, xcur(h, SX_G0), xspare(h, SX_G0), kb, 1, xcur(h, SX_G0T), xcur(h, SX_GT0), h->t_FZ, h->t_FC},   // G0
    };
    ProfScope ps(h, KDQ_PROF_FLUSH);
    for (int k = 0; k < 4; ++k) {
        f[k].one_launch = h->flush1;
        HIPCHK(launch_flush(h->stream, h->nch, h->max_ns, f[k]));
    }
    h->cur ^= 1;
    for (int k = 0; k < 3; ++k) xflip(h, k);
    return KDQ_OK;
}

// UpdateSCEELayer! (src/tU/SCEE.jl:278-324) plus the per-slice Rényi accumulation
int scee_layer(kdq_handle h, int lt, int dir, uint32_t sweep) {
    SceeState *S = h->scee;
    for (int s0 = 0; s0 < h->max_ns; s0 += S->kb) {
        SceeArgs a = scee_args(h);
        a.lt = lt; a.dir = dir; a.s0 = s0; a.kb = S->kb; a.sweep = sweep;
        {
            ProfScope ps(h, KDQ_PROF_DECIDE);
            HIPCHK(launch_scee_decide(h->stream, a));
        }
        if (int r = scee_flush(h)) return r;
    }
    HIPCHK(launch_scee_accum(h->stream, S->nch, S->d_reg));
    return KDQ_OK;
}

// bonds (t-V) or elements (t-U-V) per delayed block of the replica decisions (0: element by
// element): 8, or 4, as the handle's delay allows and the block's LDS fits
int scee_r2_block(kdq_handle h) {
    const SceeState *S = h->scee;
    for (int kb : {8, 4}) {
        if (S->kb < kb) continue;
        if (h->kind == KDQ_MODEL_TV && scee_r2_block_lds_bytes(kb, S->nx_max) <= 150 * 1024) return kb;
        if (h->kind == KDQ_MODEL_TUV && scee_r2_block_uv_lds_bytes(kb, S->nx_max) <= 150 * 1024) return kb;
    }
    return 0;
}

// UpdateSCEELayer! of the rank-2 models: t-V bond family `family`, or every element (-1, t-U-V).
// Delayed blocks of bonds (t-V) or elements (t-U-V), each followed by the four-matrix flush;
// with a delay below 4 the element kernel updates the four matrices in place
int scee_layer_r2(kdq_handle h, int lt, int dir, uint32_t sweep, int family) {
    SceeState *S = h->scee;
    SceeR2Args a{};
    a.chains = h->d_chains;
    a.Gt = h->t_G[h->cur];
    a.Gt0 = xcur(h, SX_GT0);
    a.G0t = xcur(h, SX_G0T);
    a.G0 = xcur(h, SX_G0);
    a.reg = S->d_reg;
    a.nch = S->nch;
    a.lt = lt; a.dir = dir; a.family = family;
    a.kind = h->kind == KDQ_MODEL_TV ? 1 : 2;
    a.sweep = sweep;
    a.nx_max = S->nx_max;
    a.max_ns = h->max_ns;
    const int kb = scee_r2_block(h);
    if (kb) {
        const bool uv = family < 0;   // t-U-V: blocks of consecutive elements of the slice
        for (int s0 = 0; s0 < (uv ? h->max_elem : h->max_ns / 2); s0 += kb) {
            a.s0 = s0;
            a.Gt = h->t_G[h->cur];
            a.Gt0 = xcur(h, SX_GT0);
            a.G0t = xcur(h, SX_G0T);
            a.G0 = xcur(h, SX_G0);
            {
                ProfScope ps(h, KDQ_PROF_DECIDE);
                if (uv) HIPCHK(launch_scee_r2_block_uv(h->stream, a, kb));
                else HIPCHK(launch_scee_r2_block(h->stream, a, kb));
            }
            if (int r = scee_flush(h, 2 * kb)) return r;
        }
        return KDQ_OK;
    }
    ProfScope ps(h, KDQ_PROF_DECIDE);
    HIPCHK(launch_scee_r2(h->stream, a));
    return KDQ_OK;
}

// t-V: WrapK! (src/tV/SCEE.jl:509-517) of the four matrices: Gt <- eK Gt eKinv, Gt0 <- eK Gt0,
// G0t <- G0t eKinv (forward), eK and eKinv exchanged backward
int scee_wrapK_tv(kdq_handle h, bool forward) {
    if (int r = wrap_K(h, h->t_G[h->cur], h->t_T, forward, KDQ_PROF_WRAP)) return r;
    GemmArgs gl = gargs(h, forward ? h->t_eK : h->t_eKinv, xcur(h, SX_GT0), xspare(h, SX_GT0), DIM_NS, DIM_NS, DIM_NS);
    if (int r = gemm_m(h, true, false, false, true, 0, 0, gl, KDQ_PROF_WRAP)) return r;
    xflip(h, SX_GT0);
    GemmArgs gr = gargs(h, xcur(h, SX_G0T), forward ? h->t_eKinv : h->t_eK, xspare(h, SX_G0T), DIM_NS, DIM_NS, DIM_NS);
    if (int r = gemm_m(h, false, true, true, false, 0, 0, gr, KDQ_PROF_WRAP)) return r;
    xflip(h, SX_G0T);
    return KDQ_OK;
}

// t-V WrapV! of family j (src/tV/SCEE.jl:118-123, 205-210): "L" on Gt0, "B" on Gt, "R" on G0t
// with D (forward) or D⁻¹ (backward)
int scee_wrapV_tv(kdq_handle h, int lt, bool forward, int j) {
    const int sg = forward ? +1 : -1;
    if (int r = bondrot(h, xcur(h, SX_GT0), j, lt, sg, 1, DIM_NS, KDQ_PROF_WRAP)) return r;
    if (int r = bondrot(h, h->t_G[h->cur], j, lt, sg, 0, DIM_NS, KDQ_PROF_WRAP)) return r;
    return bondrot(h, xcur(h, SX_G0T), j, lt, sg, 2, DIM_NS, KDQ_PROF_WRAP);
}

// the "L", "B", "R" wraps of Gt0, Gt and G0t (WrapKV!, src/tU/GreenMatrix.jl:56-86)
int scee_wraps(kdq_handle h, int lt, bool forward) {
    GemmArgs gl, gr;
    if (forward) {   // Gt0 <- D eK Gt0, G0t <- G0t eK⁻¹ D⁻¹
        gl = gargs(h, h->t_eK, xcur(h, SX_GT0), xspare(h, SX_GT0), DIM_NS, DIM_NS, DIM_NS);
        gl.rowS = hs_diag(h, lt, +1);
        gr = gargs(h, xcur(h, SX_G0T), h->t_eKinv, xspare(h, SX_G0T), DIM_NS, DIM_NS, DIM_NS);
        gr.colS = hs_diag(h, lt, -1);
    } else {         // Gt0 <- eK⁻¹ D⁻¹ Gt0, G0t <- G0t D eK
        gl = gargs(h, h->t_eKinv, xcur(h, SX_GT0), xspare(h, SX_GT0), DIM_NS, DIM_NS, DIM_NS);
        gl.inS = hs_diag(h, lt, -1);
        gr = gargs(h, xcur(h, SX_G0T), h->t_eK, xspare(h, SX_G0T), DIM_NS, DIM_NS, DIM_NS);
        gr.inS = hs_diag(h, lt, +1);
    }
    if (int r = gemm_m(h, true, false, false, true, 0, 0, gl, KDQ_PROF_WRAP)) return r;
    xflip(h, SX_GT0);
    if (int r = wrap_B(h, h->t_G[h->cur], h->t_T, lt, forward)) return r;
    if (int r = gemm_m(h, false, true, true, false, 0, 0, gr, KDQ_PROF_WRAP)) return r;
    xflip(h, SX_G0T);
    return KDQ_OK;
}

// chain set-up of ctrl_SCEEicr (src/tU/SCEE.jl:49-86)
int scee_init(kdq_handle h) {
    const int NN = h->nn;
    for (int idx = 1; idx <= NN - 1; ++idx)
        if (int r = scee_blocks(h, idx)) return r;
    HIPCHK(load_P(h, h->t_UR[0]));
    HIPCHK(load_P(h, h->t_ULH[NN - 1]));
    for (int i = 1; i <= NN - 1; ++i) {
        if (int r = scee_update_BL(h, NN - i)) return r;
        if (int r = scee_update_BR(h, i + 1)) return r;
    }
    h->idx = 1;
    return scee_get_ABGM(h, 1, true);
}

int scee_sweeps(kdq_handle h, int sweeps, double *O) {
    SceeState *S = h->scee;
    const int NN = h->nn, nt = h->nt, thidx = NN / 2 + 1, nch = S->nch;
    if (int r = scee_init(h)) return r;
    for (int sw = 0; sw < sweeps; ++sw) {
        const uint32_t sweep = h->sweep;
        for (int c = 0; c < nch; ++c) HIPCHK(hipMemsetAsync(S->reg[c].osum, 0, 2 * sizeof(double), h->stream));
        for (int lt = 1; lt <= nt; ++lt) {
            if (h->kind == KDQ_MODEL_TV) {   // src/tV/SCEE.jl:103-138
                if (int r = scee_wrapK_tv(h, true)) return r;
                for (int j = 2; j >= 0; --j) {
                    if (int r = scee_wrapV_tv(h, lt, true, j)) return r;
                    if (int r = scee_layer_r2(h, lt, 0, sweep, j)) return r;
                }
                HIPCHK(launch_scee_accum(h->stream, S->nch, S->d_reg));
            } else if (h->kind == KDQ_MODEL_TUV) {   // src/tUV/SCEE.jl:90-122
                if (int r = scee_wraps(h, lt, true)) return r;
                if (int r = scee_layer_r2(h, lt, 0, sweep, -1)) return r;
                HIPCHK(launch_scee_accum(h->stream, S->nch, S->d_reg));
            } else {
                if (int r = scee_wraps(h, lt, true)) return r;
                if (int r = scee_layer(h, lt, 0, sweep)) return r;
            }
            if (is_node(h, lt)) {
                const int idx = ++h->idx;
                if (int r = scee_blocks(h, idx - 1)) return r;
                for (int i = idx; i <= std::max(thidx, idx); ++i)
                    if (int r = scee_update_BR(h, i)) return r;
                for (int i = idx - 1; i >= std::min(thidx, idx); --i)
                    if (int r = scee_update_BL(h, i)) return r;
                if (int r = scee_get_ABGM(h, idx, true)) return r;
            }
        }
        for (int lt = nt; lt >= 1; --lt) {
            // rank-2 models: the updates and the inverse wraps; before a node, where G4!
            // recomputes the four matrices, the wraps after the last update are skipped
            // (src/tV/SCEE.jl:177-214, src/tUV/SCEE.jl:156-183)
            if (h->kind == KDQ_MODEL_TV) {
                const bool node = is_node(h, lt - 1);
                for (int j = 0; j < 3; ++j) {
                    if (int r = scee_layer_r2(h, lt, 1, sweep, j)) return r;
                    if (!node || j < 2)
                        if (int r = scee_wrapV_tv(h, lt, false, j)) return r;
                }
                if (!node)
                    if (int r = scee_wrapK_tv(h, false)) return r;
                HIPCHK(launch_scee_accum(h->stream, S->nch, S->d_reg));
            } else if (h->kind == KDQ_MODEL_TUV) {
                if (int r = scee_layer_r2(h, lt, 1, sweep, -1)) return r;
                HIPCHK(launch_scee_accum(h->stream, S->nch, S->d_reg));
            } else if (int r = scee_layer(h, lt, 1, sweep)) {
                return r;
            }
            if (is_node(h, lt - 1)) {
                const int idx = --h->idx;
                if (int r = scee_blocks(h, idx)) return r;
                for (int i = idx; i >= std::min(thidx, idx); --i)
                    if (int r = scee_update_BL(h, i)) return r;
                for (int i = idx + 1; i <= std::max(thidx, idx); ++i)
                    if (int r = scee_update_BR(h, i)) return r;
                if (int r = scee_get_ABGM(h, idx, false)) return r;
            } else if (h->kind != KDQ_MODEL_TV) {
                if (int r = scee_wraps(h, lt, false)) return r;
            }
        }
        for (int c = 0; c < nch; ++c)
            HIPCHK(hipMemcpyAsync(S->h_osum + 2 * c, S->reg[c].osum, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (O)
            for (int c = 0; c < nch; ++c) O[(size_t)c * sweeps + sw] = S->h_osum[2 * c] / S->h_osum[2 * c + 1];
        h->sweep += 1;
    }
    return check_health(h);
}

template <class T>
int salloc(kdq_handle h, T **p, size_t count) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (count ? count : 1)) != hipSuccess) return 1;
    h->scee->allocs.push_back(q);
    hipMemset(q, 0, sizeof(T) * (count ? count : 1));
    *p = (T *)q;
    return 0;
}

int stable(kdq_handle h, Opnd **out, const std::vector<Opnd> &v) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(Opnd) * v.size()) != hipSuccess) return 1;
    h->scee->allocs.push_back(q);
    hipMemcpy(q, v.data(), sizeof(Opnd) * v.size(), hipMemcpyHostToDevice);
    *out = (Opnd *)q;
    return 0;
}

}  // namespace

// The swap-replica handle (nrep = 2) or the disorder-operator handle (nrep = 1, field
// potential α of X = e^{iα}I + (1 - e^{iα})G0).
static int scee_create(const kdq_model_desc *models, const kdq_run_desc *runs, const kdq_scee_desc *regions,
                int32_t nchains, int nrep, double alpha, kdq_handle *out) {
    if (!models || !runs || !regions || !out || nchains < 1)
        return fail(nullptr, KDQ_EARG, "kdq_scee_create: null argument or nchains < 1");
    *out = nullptr;
    int na_max = 0, nb_max = 0, m2_max = 0;
    for (int c = 0; c < nchains; ++c) {
        const kdq_scee_desc &g = regions[c];
        m2_max = std::max(m2_max, (int)(g.nA * g.nA + g.nB * g.nB));
        if (models[c].kind != models[0].kind)
            return fail(nullptr, KDQ_EARG, "kdq_scee_create: all chains of a batch must share the model kind");
        if (models[c].kind == KDQ_MODEL_TU && (runs[c].delay < 1 || runs[c].delay > 16))
            return fail(nullptr, KDQ_EARG, "kdq_scee_create: delay must be in 1..16");
        if (g.nA < 0 || g.nB < 0 || (g.nA && !g.indexA) || (g.nB && !g.indexB) || g.nlam < 1)
            return fail(nullptr, KDQ_EARG, "kdq_scee_create: invalid region description");
        for (int i = 0; i < g.nA; ++i)
            if (g.indexA[i] < 1 || g.indexA[i] > models[c].ns) return fail(nullptr, KDQ_EARG, "kdq_scee_create: region A index out of range");
        for (int i = 0; i < g.nB; ++i)
            if (g.indexB[i] < 1 || g.indexB[i] > models[c].ns) return fail(nullptr, KDQ_EARG, "kdq_scee_create: region B index out of range");
        na_max = std::max(na_max, (int)g.nA);
        nb_max = std::max(nb_max, (int)g.nB);
    }
    if (models[0].kind == KDQ_MODEL_TU) {
        const int ks = runs[0].delay <= 8 ? 8 : 16;
        if (scee_lds_bytes(ks, na_max + nb_max) > 163840)
            return fail(nullptr, KDQ_EARG, "kdq_scee_create: regions too large for the block size (reduce delay or |A|+|B|)");
    } else {
        int ns_max = 0;
        for (int c = 0; c < nchains; ++c) ns_max = std::max(ns_max, (int)models[c].ns);
        if (scee_r2_lds_bytes(na_max + nb_max, ns_max) > 163840)
            return fail(nullptr, KDQ_EARG, "kdq_scee_create: lattice and regions too large for the rank-2 decision kernel");
    }
    std::vector<kdq_model_desc> md(nrep * nchains);
    std::vector<kdq_run_desc> rd(nrep * nchains);
    std::vector<uint32_t> streams(nrep * nchains);
    for (int r = 0; r < nrep; ++r)
        for (int c = 0; c < nchains; ++c) {
            md[c + r * nchains] = models[c];
            rd[c + r * nchains] = runs[c];
            streams[c + r * nchains] = 1 + r;
        }
    kdq_handle h = nullptr;
    if (int r = create_impl(md.data(), rd.data(), nrep * nchains, streams.data(), &h)) return r;
    SceeState *S = new SceeState();
    h->scee = S;
    S->nch = nchains;
    S->nrep = nrep;
    S->dop = nrep == 1;
    S->kb = runs[0].delay;
    S->na_max = na_max;
    S->nb_max = nb_max;
    S->nx_max = na_max + nb_max;
    S->m2_max = m2_max;
    int bad = 0;
    // four-matrix buffers and block products per entry
    std::vector<Opnd> vx[3][2];
    std::vector<std::vector<Opnd>> vbm(h->nn - 1), vbmi(h->nn - 1);
    for (int e = 0; e < nrep * nchains; ++e) {
        const int n = h->ch[e].ns;
        const size_t nn = (size_t)n * n;
        for (int k = 0; k < 3; ++k)
            for (int b = 0; b < 2; ++b) {
                double *p = nullptr;
                bad |= salloc(h, &p, 2 * nn);
                vx[k][b].push_back(cplx(p, n, nn));
            }
        for (int i = 0; i < h->nn - 1; ++i) {
            double *p = nullptr, *q = nullptr;
            bad |= salloc(h, &p, 2 * nn);
            bad |= salloc(h, &q, 2 * nn);
            vbm[i].push_back(cplx(p, n, nn));
            vbmi[i].push_back(cplx(q, n, nn));
        }
    }
    for (int k = 0; k < 3; ++k)
        for (int b = 0; b < 2; ++b) bad |= stable(h, &S->t_X[k][b], vx[k][b]);
    S->t_BM.resize(h->nn - 1);
    S->t_BMi.resize(h->nn - 1);
    for (int i = 0; i < h->nn - 1; ++i) {
        bad |= stable(h, &S->t_BM[i], vbm[i]);
        bad |= stable(h, &S->t_BMi[i], vbmi[i]);
    }
    // regions per chain
    std::vector<Opnd> vm[2][2], vc[2], vinv[2], vix[2];
    S->reg.resize(nchains);
    std::vector<double> eye;
    for (int c = 0; c < nchains; ++c) {
        const kdq_scee_desc &g = regions[c];
        SceeRegion &R = S->reg[c];
        memset(&R, 0, sizeof(R));
        R.nA = g.nA; R.nB = g.nB; R.nlam = g.nlam; R.lambda = g.lambda;
        R.dop = nrep == 1;
        R.ea[0] = std::cos(alpha);
        R.ea[1] = std::sin(alpha);
        std::vector<int> idx;
        for (int i = 0; i < g.nA; ++i) idx.push_back((int)g.indexA[i] - 1);
        for (int i = 0; i < g.nB; ++i) idx.push_back((int)g.indexB[i] - 1);
        int *didx = nullptr;
        bad |= salloc(h, &didx, idx.size());
        if (!idx.empty()) hipMemcpy(didx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice);
        R.idx = didx;
        const int nx[2] = {g.nA, g.nB};
        for (int X = 0; X < 2; ++X) {
            const size_t m2 = (size_t)nx[X] * nx[X];
            for (int r = 0; r < 2; ++r) {
                bad |= salloc(h, &R.M[r][X], 2 * m2);
                vm[r][X].push_back(cplx(R.M[r][X], std::max(nx[X], 1), m2));
            }
            bad |= salloc(h, &R.gmInv[X], 2 * m2);
            bad |= salloc(h, &R.C[X], 2 * m2);
            double *I = nullptr;
            bad |= salloc(h, &I, 2 * m2);
            eye.assign(2 * m2, 0.0);
            for (int i = 0; i < nx[X]; ++i) eye[i + (size_t)i * nx[X]] = 1.0;
            if (m2) hipMemcpy(I, eye.data(), sizeof(double) * 2 * m2, hipMemcpyHostToDevice);
            vinv[X].push_back(cplx(R.gmInv[X], std::max(nx[X], 1), m2));
            vc[X].push_back(cplx(R.C[X], std::max(nx[X], 1), m2));
            vix[X].push_back(cplx(I, std::max(nx[X], 1), m2));
        }
        bad |= salloc(h, &R.detg, 2);
        bad |= salloc(h, &R.osum, 2);
        for (int r = 0; r < nrep; ++r) {
            h->cd_host[c + r * nchains].nra = g.nA;
            h->cd_host[c + r * nchains].nrb = g.nB;
        }
    }
    for (int X = 0; X < 2; ++X) {
        for (int r = 0; r < 2; ++r) bad |= stable(h, &S->t_M[r][X], vm[r][X]);
        bad |= stable(h, &S->t_C[X], vc[X]);
        bad |= stable(h, &S->t_Inv[X], vinv[X]);
        bad |= stable(h, &S->t_IX[X], vix[X]);
    }
    bad |= salloc(h, &S->d_ld, 2 * nchains);
    bad |= salloc(h, &S->d_ph, 4 * nchains);
    bad |= salloc(h, &S->d_info, nchains);
    bad |= salloc(h, &S->d_piv, (size_t)nchains * std::max(S->nx_max, 1));
    {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(SceeRegion) * nchains) == hipSuccess) {
            S->allocs.push_back(q);
            hipMemcpy(q, S->reg.data(), sizeof(SceeRegion) * nchains, hipMemcpyHostToDevice);
            S->d_reg = (SceeRegion *)q;
        } else {
            bad = 1;
        }
    }
    if (hipHostMalloc((void **)&S->h_osum, sizeof(double) * 2 * nchains, hipHostMallocDefault) != hipSuccess) bad = 1;
    hipMemcpy(h->d_chains, h->cd_host.data(), sizeof(ChainDev) * h->cd_host.size(), hipMemcpyHostToDevice);
    if (bad || hipDeviceSynchronize() != hipSuccess) {
        int r = fail(h, KDQ_EHIP, "kdq_scee_create: device allocation failed");
        kdq_destroy(h);
        return r;
    }
    *out = h;
    return KDQ_OK;
}

extern "C" {
int kdq_scee_create(const kdq_model_desc *models, const kdq_run_desc *runs, const kdq_scee_desc *regions,
                    int32_t nchains, kdq_handle *out) {
    return scee_create(models, runs, regions, nchains, 2, 0.0, out);
}

int kdq_dop_create(const kdq_model_desc *models, const kdq_run_desc *runs, const kdq_scee_desc *regions,
                   double alpha, int32_t nchains, kdq_handle *out) {
    if (!models || nchains < 1) return fail(nullptr, KDQ_EARG, "kdq_dop_create: null argument or nchains < 1");
    for (int c = 0; c < nchains; ++c)
        if (models[c].kind != KDQ_MODEL_TU) return fail(nullptr, KDQ_EARG, "kdq_dop_create: the disorder operator is for t-U chains");
    return scee_create(models, runs, regions, nchains, 1, alpha, out);
}

int kdq_scee_set_fields(kdq_handle h, int32_t c, int32_t replica, const uint8_t *s) {
    if (!h || !h->scee || replica < 0 || replica >= h->scee->nrep || c < 0 || c >= h->scee->nch)
        return fail(h, KDQ_EARG, "kdq_scee_set_fields: bad argument");
    return kdq_set_fields(h, c + replica * h->scee->nch, s);
}

int kdq_scee_get_fields(kdq_handle h, int32_t c, int32_t replica, uint8_t *s) {
    if (!h || !h->scee || replica < 0 || replica >= h->scee->nrep || c < 0 || c >= h->scee->nch)
        return fail(h, KDQ_EARG, "kdq_scee_get_fields: bad argument");
    return kdq_get_fields(h, c + replica * h->scee->nch, s);
}

int kdq_scee_init(kdq_handle h) {
    if (!h || !h->scee) return fail(h, KDQ_EARG, "kdq_scee_init: not a swap-replica handle");
    if (int r = scee_init(h)) return r;
    return check_health(h);
}

int kdq_scee_sweeps(kdq_handle h, int32_t sweeps, double *O) {
    if (!h || !h->scee || sweeps < 0) return fail(h, KDQ_EARG, "kdq_scee_sweeps: bad argument");
    return scee_sweeps(h, sweeps, O);
}

int kdq_scee_get_G4(kdq_handle h, int32_t c, int32_t replica, double *Gt, double *G0, double *Gt0, double *G0t) {
    if (!h || !h->scee || replica < 0 || replica >= h->scee->nrep || c < 0 || c >= h->scee->nch || !Gt || !G0 ||
        !Gt0 || !G0t)
        return fail(h, KDQ_EARG, "kdq_scee_get_G4: bad argument");
    const int e = c + replica * h->scee->nch, n = h->ch[e].ns;
    const size_t nn = (size_t)n * n;
    std::vector<Opnd> tab(h->nch);
    double *outs[4] = {Gt, G0, Gt0, G0t};
    const double *src[4];
    src[0] = h->ch[e].G[h->cur];
    for (int k = 0; k < 3; ++k) {
        HIPCHK(hipMemcpy(tab.data(), xcur(h, k), sizeof(Opnd) * tab.size(), hipMemcpyDeviceToHost));
        src[k + 1] = tab[e].re;
    }
    std::vector<double> tmp(2 * nn);
    const int order[4] = {0, 1, 2, 3};   // Gt, G0, Gt0, G0t
    for (int q = 0; q < 4; ++q) {
        HIPCHK(hipMemcpy(tmp.data(), src[order[q]], 2 * nn * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nn; ++i) {
            outs[q][2 * i] = tmp[i];
            outs[q][2 * i + 1] = tmp[nn + i];
        }
    }
    return KDQ_OK;
}

int kdq_scee_detg(kdq_handle h, double *out) {
    if (!h || !h->scee || !out) return fail(h, KDQ_EARG, "kdq_scee_detg: bad argument");
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < h->scee->nch; ++c) HIPCHK(hipMemcpy(out + 2 * c, h->scee->reg[c].detg, 2 * sizeof(double), hipMemcpyDeviceToHost));
    return KDQ_OK;
}
}  // extern "C"
