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

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
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
    double *G[2], *T, *Gm, *Tm, *P0, *P1, *M, *X;
    std::vector<double *> UR, ULH;
    uint8_t *s;
    int *acc_cnt, *acc_idx, *info;
    double *W, *counters, *meas, *logdet;
};

struct Table {  // one logical operand across chains
    Opnd *d = nullptr;
};

}  // namespace

struct kdq_handle_s {
    int nch = 0, nt = 0, nn = 0, max_ns = 0, max_np = 0, delay = 16;
    std::vector<int64_t> nodes;
    std::vector<ChainHost> ch;
    std::vector<kdq_model_desc> md;
    std::vector<kdq_run_desc> rd;
    hipStream_t stream = nullptr;
    ChainDev *d_chains = nullptr;
    // operand tables
    Opnd *t_eK, *t_eKinv, *t_hK, *t_hKinv, *t_Pt, *t_G[2], *t_T, *t_Gm, *t_Tm, *t_P0, *t_P1, *t_M, *t_X;
    std::vector<Opnd *> t_UR, t_ULH;
    std::vector<void *> dev_misc;
    double *d_logdet = nullptr;   // [nch] log|det(U_L U_R)| of the last get_G
    int *d_info = nullptr;        // [nch] sticky: nonzero after a singular overlap
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

    ~kdq_handle_s() {
        if (stream) hipStreamSynchronize(stream);
        for (auto &c : ch)
            for (void *p : c.allocs) hipFree(p);
        for (void *p : dev_misc) hipFree(p);
        if (h_meas) hipHostFree(h_meas);
        for (auto &p : ev_pending) {
            hipEventDestroy(p.second.first);
            hipEventDestroy(p.second.second);
        }
        for (hipEvent_t e : ev_pool) hipEventDestroy(e);
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
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
    HIPCHK(launch_gemm(h->stream, h->nch, maxdim(h, g.mdim), maxdim(h, g.ndim), ac, bc, opa, opb, g));
    return KDQ_OK;
}

DiagSpec phase(int lt, int sign) { return DiagSpec{1, lt, sign, 0}; }

// G <- D eK G eKinv D⁻¹ (forward) or eKinv D⁻¹ G D eK (backward), D = cis(α η_s) at slice lt.
int wrap_B(kdq_handle h, Opnd *Gt, Opnd *Tt, int lt, bool forward, int fam = KDQ_PROF_WRAP) {
    if (forward) {
        GemmArgs g1 = gargs(h, h->t_eK, Gt, Tt, DIM_NS, DIM_NS, DIM_NS);
        if (int r = gemm(h, false, true, 0, 0, g1, fam)) return r;
        GemmArgs g2 = gargs(h, Tt, h->t_eKinv, Gt, DIM_NS, DIM_NS, DIM_NS);
        g2.rowS = phase(lt, +1);
        g2.colS = phase(lt, -1);
        return gemm(h, true, false, 0, 0, g2, fam);
    }
    // D' = cis(-α η): T = eKinv · (D' G) · D'⁻¹, then G = T eK
    GemmArgs g1 = gargs(h, h->t_eKinv, Gt, Tt, DIM_NS, DIM_NS, DIM_NS);
    g1.inS = phase(lt, -1);
    g1.colS = phase(lt, +1);
    if (int r = gemm(h, false, true, 0, 0, g1, fam)) return r;
    GemmArgs g2 = gargs(h, Tt, h->t_eK, Gt, DIM_NS, DIM_NS, DIM_NS);
    return gemm(h, true, false, 0, 0, g2, fam);
}

// dst = QR_Q(B_block(idx) · src) for U_R: slices nodes[idx-1]+1 .. nodes[idx] (1-based idx)
int propagate_R(kdq_handle h, Opnd *src, Opnd *dst, int idx) {
    const int lo = (int)h->nodes[idx - 1] + 1, hi = (int)h->nodes[idx];
    Opnd *cur = src;
    for (int lt = lo; lt <= hi; ++lt) {
        Opnd *out = (lt == hi) ? dst : ((cur == h->t_P0) ? h->t_P1 : h->t_P0);
        GemmArgs g = gargs(h, h->t_eK, cur, out, DIM_NS, DIM_NP, DIM_NS);
        g.rowS = phase(lt, +1);
        if (int r = gemm(h, false, true, 0, 0, g, KDQ_PROF_CHAIN)) return r;
        cur = out;
    }
    if (lo > hi) HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, src, dst, DIM_NS, DIM_NP, true));
    QRArgs q{h->d_chains, dst, DIM_NS, DIM_NP};
    ProfScope ps(h, KDQ_PROF_QR);
    HIPCHK(launch_qr(h->stream, h->nch, q));
    return KDQ_OK;
}

// dst = QR_Q((U_L · B_block(idx))ᴴ) with U_L stored as U_Lᴴ: U_Lᴴ <- eKᵀ (conj(D_l) ⊙ U_Lᴴ)
// for l = nodes[idx] down to nodes[idx-1]+1.
int propagate_L(kdq_handle h, Opnd *src, Opnd *dst, int idx) {
    const int lo = (int)h->nodes[idx - 1] + 1, hi = (int)h->nodes[idx];
    Opnd *cur = src;
    for (int lt = hi; lt >= lo; --lt) {
        Opnd *out = (lt == lo) ? dst : ((cur == h->t_P0) ? h->t_P1 : h->t_P0);
        GemmArgs g = gargs(h, h->t_eK, cur, out, DIM_NS, DIM_NP, DIM_NS);
        g.inS = phase(lt, -1);
        if (int r = gemm(h, false, true, 1, 0, g, KDQ_PROF_CHAIN)) return r;
        cur = out;
    }
    if (lo > hi) HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, src, dst, DIM_NS, DIM_NP, true));
    QRArgs q{h->d_chains, dst, DIM_NS, DIM_NP};
    ProfScope ps(h, KDQ_PROF_QR);
    HIPCHK(launch_qr(h->stream, h->nch, q));
    return KDQ_OK;
}

// get_G!: G = I - U_R (U_L U_R)⁻¹ U_L into the current G buffer
int get_G(kdq_handle h, int idx, int *bad = nullptr) {
    Opnd *ULH = h->t_ULH[idx - 1], *UR = h->t_UR[idx - 1];
    GemmArgs g1 = gargs(h, ULH, UR, h->t_M, DIM_NP, DIM_NP, DIM_NS);
    if (int r = gemm(h, true, true, 2, 0, g1, KDQ_PROF_GETG)) return r;
    InvArgs iv{h->d_chains, h->t_M, DIM_NP, 0, h->d_logdet, h->d_info};
    {
        ProfScope ps(h, KDQ_PROF_INV);
        HIPCHK(launch_inv(h->stream, h->nch, iv));
    }
    GemmArgs g2 = gargs(h, h->t_M, ULH, h->t_X, DIM_NP, DIM_NS, DIM_NP);
    if (int r = gemm(h, true, true, 0, 2, g2, KDQ_PROF_GETG)) return r;
    GemmArgs g3 = gargs(h, UR, h->t_X, h->t_G[h->cur], DIM_NS, DIM_NS, DIM_NP);
    g3.mode = 1;
    (void)bad;
    return gemm(h, true, true, 0, 0, g3, KDQ_PROF_GETG);
}

int init_chain(kdq_handle h) {
    const int NN = h->nn;
    HIPCHK(launch_real_to_cplx(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, h->t_Pt, h->t_UR[0], DIM_NS, DIM_NP));
    HIPCHK(launch_real_to_cplx(h->stream, h->nch, h->max_ns, h->max_np, h->d_chains, h->t_Pt, h->t_ULH[NN - 1], DIM_NS, DIM_NP));
    for (int idx = NN - 1; idx >= 1; --idx)
        if (int r = propagate_L(h, h->t_ULH[idx], h->t_ULH[idx - 1], idx)) return r;
    h->idx = 1;
    return get_G(h, 1);
}

int update_layer(kdq_handle h, int lt, int dir, uint32_t sweep) {
    const int kb = h->delay;
    for (int s0 = 0; s0 < h->max_ns; s0 += kb) {
        DecideArgs a{};
        a.chains = h->d_chains;
        a.G = h->t_G[h->cur];
        a.lt = lt; a.dir = dir; a.s0 = s0; a.kb = kb; a.sweep = sweep;
        {
            ProfScope ps(h, KDQ_PROF_DECIDE);
            HIPCHK(launch_decide_tu(h->stream, h->nch, a));
        }
        FlushArgs f{h->d_chains, h->t_G[h->cur], h->t_G[1 - h->cur]};
        {
            ProfScope ps(h, KDQ_PROF_FLUSH);
            HIPCHK(launch_flush(h->stream, h->nch, h->max_ns, f));
        }
        h->cur ^= 1;
    }
    return KDQ_OK;
}

// phy_measure: wrap a copy of G to τ = Nt/2, HalfeK sandwich, observables (accumulated).
int measure(kdq_handle h, int lt) {
    HIPCHK(launch_copy_mat(h->stream, h->nch, h->max_ns, h->max_ns, h->d_chains, h->t_G[h->cur], h->t_Gm, DIM_NS, DIM_NS, true));
    const int nt = h->nt;
    if (lt > nt / 2.0) {
        for (int t = lt; t >= nt / 2 + 1; --t)
            if (int r = wrap_B(h, h->t_Gm, h->t_Tm, t, false, KDQ_PROF_MEAS_WRAP)) return r;
    } else {
        for (int t = lt + 1; t <= nt / 2; ++t)
            if (int r = wrap_B(h, h->t_Gm, h->t_Tm, t, true, KDQ_PROF_MEAS_WRAP)) return r;
    }
    GemmArgs g1 = gargs(h, h->t_hK, h->t_Gm, h->t_Tm, DIM_NS, DIM_NS, DIM_NS);
    if (int r = gemm(h, false, true, 0, 0, g1, KDQ_PROF_MEAS_WRAP)) return r;
    GemmArgs g2 = gargs(h, h->t_Tm, h->t_hKinv, h->t_Gm, DIM_NS, DIM_NS, DIM_NS);
    if (int r = gemm(h, true, false, 0, 0, g2, KDQ_PROF_MEAS_WRAP)) return r;
    MeasArgs m{h->d_chains, h->t_Gm};
    ProfScope ps(h, KDQ_PROF_MEASURE);
    HIPCHK(launch_measure_tu(h->stream, h->nch, m));
    return KDQ_OK;
}

bool is_node(kdq_handle h, int t) {
    for (int64_t v : h->nodes)
        if (v == t) return true;
    return false;
}

int read_meas(kdq_handle h, double *dst /* nch x 6 */) {
    for (int c = 0; c < h->nch; ++c)
        HIPCHK(hipMemcpyAsync(h->h_meas + 8 * c, h->ch[c].meas, 8 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < h->nch; ++c) {
        const double cnt = h->h_meas[8 * c + 6];
        for (int k = 0; k < 6; ++k) dst[6 * c + k] = cnt > 0 ? h->h_meas[8 * c + k] / cnt : 0.0;
    }
    return KDQ_OK;
}

// Reads the sticky singular-overlap flags; KDQ_ENUM names the first failing chain.
int check_health(kdq_handle h) {
    std::vector<int> info(h->nch);
    HIPCHK(hipMemcpyAsync(info.data(), h->d_info, sizeof(int) * h->nch, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < h->nch; ++c)
        if (info[c]) return fail(h, KDQ_ENUM, "singular overlap U_L U_R in chain " + std::to_string(c));
    return KDQ_OK;
}

int phy_update(kdq_handle h, int sweeps, int record, double *rows) {
    const int NN = h->nn, nt = h->nt, thidx = NN / 2 + 1;
    if (int r = init_chain(h)) return r;
    std::vector<double> tmp(6 * h->nch);
    for (int sw = 0; sw < sweeps; ++sw) {
        const uint32_t sweep = h->sweep;
        if (record) HIPCHK(launch_zero_meas(h->stream, h->nch, h->d_chains));
        for (int lt = 1; lt <= nt; ++lt) {
            if (int r = wrap_B(h, h->t_G[h->cur], h->t_T, lt, true)) return r;
            if (int r = update_layer(h, lt, 0, sweep)) return r;
            if (record && thidx - h->idx >= 0 && thidx - h->idx <= 1)
                if (int r = measure(h, lt)) return r;
            if (is_node(h, lt)) {
                h->idx += 1;
                if (int r = propagate_R(h, h->t_UR[h->idx - 2], h->t_UR[h->idx - 1], h->idx - 1)) return r;
                if (int r = get_G(h, h->idx)) return r;
            }
        }
        for (int lt = nt; lt >= 1; --lt) {
            if (int r = update_layer(h, lt, 1, sweep)) return r;
            if (record && h->idx - thidx >= 0 && h->idx - thidx <= 1)
                if (int r = measure(h, lt)) return r;
            if (int r = wrap_B(h, h->t_G[h->cur], h->t_T, lt, false)) return r;
            if (is_node(h, lt - 1)) {
                h->idx -= 1;
                if (int r = propagate_L(h, h->t_ULH[h->idx], h->t_ULH[h->idx - 1], h->idx)) return r;
                if (int r = get_G(h, h->idx)) return r;
            }
        }
        if (record && rows) {
            if (int r = read_meas(h, tmp.data())) return r;
            for (int c = 0; c < h->nch; ++c)
                for (int k = 0; k < 6; ++k) rows[((size_t)c * sweeps + sw) * 6 + k] = tmp[6 * c + k];
        }
        h->sweep += 1;
    }
    return check_health(h);
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char *kdq_last_error(kdq_handle h) {
    if (h) return h->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err.c_str();
}

int kdq_create(const kdq_model_desc *models, const kdq_run_desc *runs, int32_t nchains, kdq_handle *out) {
    if (!models || !runs || !out || nchains < 1) return fail(nullptr, KDQ_EARG, "kdq_create: null argument or nchains < 1");
    *out = nullptr;
    for (int c = 0; c < nchains; ++c) {
        const kdq_model_desc &m = models[c];
        if (m.kind != KDQ_MODEL_TU) return fail(nullptr, KDQ_EARG, "kdq_create: only the t-U model is available in this build");
        if (m.ns < 2 || m.ns % 2 || m.nt < 1 || m.n_nodes < 2 || !m.nodes || !m.alpha || !m.eta || !m.gamma ||
            !m.K || !m.eK || !m.eKinv || !m.halfeK || !m.halfeKinv || !m.Pt)
            return fail(nullptr, KDQ_EARG, "kdq_create: invalid model description for chain " + std::to_string(c));
        if (m.ns / 2 > 256) return fail(nullptr, KDQ_EARG, "kdq_create: ns/2 above 256 is not supported");
        if (m.nt != models[0].nt || m.n_nodes != models[0].n_nodes)
            return fail(nullptr, KDQ_EARG, "kdq_create: all chains of a batch must share nt and nodes");
        for (int k = 0; k < m.n_nodes; ++k)
            if (m.nodes[k] != models[0].nodes[k]) return fail(nullptr, KDQ_EARG, "kdq_create: nodes differ between chains");
        if (m.nodes[0] != 0 || m.nodes[m.n_nodes - 1] != m.nt) return fail(nullptr, KDQ_EARG, "kdq_create: nodes must start at 0 and end at nt");
        if (runs[c].delay < 1 || runs[c].delay > 64) return fail(nullptr, KDQ_EARG, "kdq_create: delay must be in 1..64");
        if (runs[c].delay != runs[0].delay) return fail(nullptr, KDQ_EARG, "kdq_create: delay must be equal across chains");
        if (runs[c].precision != KDQ_PREC_FP64) return fail(nullptr, KDQ_EARG, "kdq_create: only FP64 precision is available in this build");
        if (runs[c].chain >= (1u << 24)) return fail(nullptr, KDQ_EARG, "kdq_create: chain id must be < 2^24");
    }
    kdq_handle h = new kdq_handle_s();
    h->nch = nchains;
    h->nt = models[0].nt;
    h->nn = models[0].n_nodes;
    h->nodes.assign(models[0].nodes, models[0].nodes + h->nn);
    h->delay = runs[0].delay;
    h->md.assign(models, models + nchains);
    h->rd.assign(runs, runs + nchains);
    h->sweep = runs[0].sweep0;
    h->ch.resize(nchains);
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        int r = fail(h, KDQ_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        delete h;
        return r;
    }
    std::vector<ChainDev> cd(nchains);
    std::vector<Opnd> v_eK, v_eKinv, v_hK, v_hKinv, v_Pt, v_G0, v_G1, v_T, v_Gm, v_Tm, v_P0, v_P1, v_M, v_X;
    std::vector<std::vector<Opnd>> v_UR(h->nn), v_ULH(h->nn);
    for (int c = 0; c < nchains; ++c) {
        const kdq_model_desc &m = models[c];
        ChainHost &C = h->ch[c];
        const int n = m.ns, np = m.ns / 2;
        C.ns = n; C.np = np;
        h->max_ns = std::max(h->max_ns, n);
        h->max_np = std::max(h->max_np, np);
        const size_t nn = (size_t)n * n, nP = (size_t)n * np, pp = (size_t)np * np;
        bool ok = true;
        ok &= dalloc(C, &C.eK, nn) == hipSuccess;
        ok &= dalloc(C, &C.eKinv, nn) == hipSuccess;
        ok &= dalloc(C, &C.hK, nn) == hipSuccess;
        ok &= dalloc(C, &C.hKinv, nn) == hipSuccess;
        ok &= dalloc(C, &C.K, nn) == hipSuccess;
        ok &= dalloc(C, &C.Pt, nP) == hipSuccess;
        ok &= dalloc(C, &C.alpha, (size_t)m.nt) == hipSuccess;
        ok &= dalloc(C, &C.G[0], 2 * nn) == hipSuccess;
        ok &= dalloc(C, &C.G[1], 2 * nn) == hipSuccess;
        ok &= dalloc(C, &C.T, 2 * nn) == hipSuccess;
        ok &= dalloc(C, &C.Gm, 2 * nn) == hipSuccess;
        ok &= dalloc(C, &C.Tm, 2 * nn) == hipSuccess;
        ok &= dalloc(C, &C.P0, 2 * nP) == hipSuccess;
        ok &= dalloc(C, &C.P1, 2 * nP) == hipSuccess;
        ok &= dalloc(C, &C.M, 2 * pp) == hipSuccess;
        ok &= dalloc(C, &C.X, 2 * nP) == hipSuccess;
        C.UR.resize(h->nn);
        C.ULH.resize(h->nn);
        for (int k = 0; k < h->nn; ++k) {
            ok &= dalloc(C, &C.UR[k], 2 * nP) == hipSuccess;
            ok &= dalloc(C, &C.ULH[k], 2 * nP) == hipSuccess;
        }
        ok &= dalloc(C, &C.s, (size_t)n * m.nt) == hipSuccess;
        ok &= dalloc(C, &C.acc_cnt, 1) == hipSuccess;
        ok &= dalloc(C, &C.acc_idx, 64) == hipSuccess;
        ok &= dalloc(C, &C.info, 1) == hipSuccess;
        ok &= dalloc(C, &C.W, 2 * 64 * 64) == hipSuccess;
        ok &= dalloc(C, &C.counters, 2) == hipSuccess;
        ok &= dalloc(C, &C.meas, 8) == hipSuccess;
        ok &= dalloc(C, &C.logdet, 1) == hipSuccess;
        if (!ok) {
            int r = fail(h, KDQ_EHIP, "kdq_create: device allocation failed");
            delete h;
            return r;
        }
        hipMemcpy(C.eK, m.eK, nn * 8, hipMemcpyHostToDevice);
        hipMemcpy(C.eKinv, m.eKinv, nn * 8, hipMemcpyHostToDevice);
        hipMemcpy(C.hK, m.halfeK, nn * 8, hipMemcpyHostToDevice);
        hipMemcpy(C.hKinv, m.halfeKinv, nn * 8, hipMemcpyHostToDevice);
        hipMemcpy(C.K, m.K, nn * 8, hipMemcpyHostToDevice);
        hipMemcpy(C.Pt, m.Pt, nP * 8, hipMemcpyHostToDevice);
        hipMemcpy(C.alpha, m.alpha, (size_t)m.nt * 8, hipMemcpyHostToDevice);

        ChainDev &d = cd[c];
        memset(&d, 0, sizeof(d));
        d.ns = n; d.np = np; d.nt = m.nt;
        d.alpha = C.alpha;
        for (int k = 0; k < 4; ++k) { d.eta[k] = m.eta[k]; d.gamma[k] = m.gamma[k]; }
        d.s = C.s;
        d.ht = m.ht;
        d.lattice = m.lattice; d.L1 = m.site[0]; d.L2 = m.site[1];
        d.K = C.K;
        d.seed = runs[c].seed;
        d.chain_id = runs[c].chain;
        d.acc_cnt = C.acc_cnt; d.acc_idx = C.acc_idx; d.W = C.W;
        d.counters = C.counters; d.meas = C.meas;

        v_eK.push_back(real(C.eK, n)); v_eKinv.push_back(real(C.eKinv, n));
        v_hK.push_back(real(C.hK, n)); v_hKinv.push_back(real(C.hKinv, n));
        v_Pt.push_back(real(C.Pt, n));
        v_G0.push_back(cplx(C.G[0], n, nn)); v_G1.push_back(cplx(C.G[1], n, nn));
        v_T.push_back(cplx(C.T, n, nn)); v_Gm.push_back(cplx(C.Gm, n, nn)); v_Tm.push_back(cplx(C.Tm, n, nn));
        v_P0.push_back(cplx(C.P0, n, nP)); v_P1.push_back(cplx(C.P1, n, nP));
        v_M.push_back(cplx(C.M, np, pp)); v_X.push_back(cplx(C.X, np, nP));
        for (int k = 0; k < h->nn; ++k) {
            v_UR[k].push_back(cplx(C.UR[k], n, nP));
            v_ULH[k].push_back(cplx(C.ULH[k], n, nP));
        }
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
    }
    int r = 0;
    r |= make_table(h, &h->t_eK, v_eK); r |= make_table(h, &h->t_eKinv, v_eKinv);
    r |= make_table(h, &h->t_hK, v_hK); r |= make_table(h, &h->t_hKinv, v_hKinv);
    r |= make_table(h, &h->t_Pt, v_Pt);
    r |= make_table(h, &h->t_G[0], v_G0); r |= make_table(h, &h->t_G[1], v_G1);
    r |= make_table(h, &h->t_T, v_T); r |= make_table(h, &h->t_Gm, v_Gm); r |= make_table(h, &h->t_Tm, v_Tm);
    r |= make_table(h, &h->t_P0, v_P0); r |= make_table(h, &h->t_P1, v_P1);
    r |= make_table(h, &h->t_M, v_M); r |= make_table(h, &h->t_X, v_X);
    h->t_UR.resize(h->nn);
    h->t_ULH.resize(h->nn);
    for (int k = 0; k < h->nn; ++k) {
        r |= make_table(h, &h->t_UR[k], v_UR[k]);
        r |= make_table(h, &h->t_ULH[k], v_ULH[k]);
    }
    {
        void *q1 = nullptr, *q2 = nullptr;
        if (hipMalloc(&q1, sizeof(double) * nchains) == hipSuccess) h->dev_misc.push_back(q1); else r = 1;
        if (hipMalloc(&q2, sizeof(int) * nchains) == hipSuccess) h->dev_misc.push_back(q2); else r = 1;
        h->d_logdet = (double *)q1;
        h->d_info = (int *)q2;
        if (!r) hipMemset(q2, 0, sizeof(int) * nchains);
    }
    if (r || hipHostMalloc((void **)&h->h_meas, sizeof(double) * 8 * nchains, hipHostMallocDefault) != hipSuccess) {
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

int kdq_destroy(kdq_handle h) {
    delete h;
    return KDQ_OK;
}

int kdq_num_chains(kdq_handle h) { return h ? h->nch : 0; }

int kdq_set_fields(kdq_handle h, int32_t c, const uint8_t *s) {
    if (!h || !s || c < 0 || c >= h->nch) return fail(h, KDQ_EARG, "kdq_set_fields: bad argument");
    const size_t n = (size_t)h->ch[c].ns * h->nt;
    for (size_t i = 0; i < n; ++i)
        if (s[i] < 1 || s[i] > 4) return fail(h, KDQ_EARG, "kdq_set_fields: field values must be 1..4");
    HIPCHK(hipMemcpyAsync(h->ch[c].s, s, n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_get_fields(kdq_handle h, int32_t c, uint8_t *s) {
    if (!h || !s || c < 0 || c >= h->nch) return fail(h, KDQ_EARG, "kdq_get_fields: bad argument");
    HIPCHK(hipMemcpyAsync(s, h->ch[c].s, (size_t)h->ch[c].ns * h->nt, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_initial_s(kdq_handle h) {
    if (!h) return fail(h, KDQ_EARG, "kdq_initial_s: null handle");
    HIPCHK(launch_initial_s(h->stream, h->nch, h->max_ns, h->nt, h->d_chains));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_phy_update(kdq_handle h, int32_t sweeps, int32_t record, double *rows) {
    if (!h || sweeps < 0) return fail(h, KDQ_EARG, "kdq_phy_update: bad argument");
    if (record && !rows) return fail(h, KDQ_EARG, "kdq_phy_update: rows must not be NULL when record != 0");
    return phy_update(h, sweeps, record, rows);
}

int kdq_get_G(kdq_handle h, int32_t c, double *G) {
    if (!h || !G || c < 0 || c >= h->nch) return fail(h, KDQ_EARG, "kdq_get_G: bad argument");
    const int n = h->ch[c].ns;
    const size_t nn = (size_t)n * n;
    std::vector<double> tmp(2 * nn);
    HIPCHK(hipMemcpyAsync(tmp.data(), h->ch[c].G[h->cur], 2 * nn * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t q = 0; q < nn; ++q) {
        G[2 * q] = tmp[q];
        G[2 * q + 1] = tmp[nn + q];
    }
    return KDQ_OK;
}

int kdq_set_G(kdq_handle h, int32_t c, const double *G) {
    if (!h || !G || c < 0 || c >= h->nch) return fail(h, KDQ_EARG, "kdq_set_G: bad argument");
    const int n = h->ch[c].ns;
    const size_t nn = (size_t)n * n;
    std::vector<double> tmp(2 * nn);
    for (size_t q = 0; q < nn; ++q) {
        tmp[q] = G[2 * q];
        tmp[nn + q] = G[2 * q + 1];
    }
    HIPCHK(hipMemcpyAsync(h->ch[c].G[h->cur], tmp.data(), 2 * nn * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_wrap(kdq_handle h, int32_t lt, int32_t forward) {
    if (!h || lt < 1 || lt > h->nt) return fail(h, KDQ_EARG, "kdq_wrap: bad argument");
    if (int r = wrap_B(h, h->t_G[h->cur], h->t_T, lt, forward != 0)) return r;
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_update_layer(kdq_handle h, int32_t lt, int32_t dir, uint32_t sweep) {
    if (!h || lt < 1 || lt > h->nt || (dir != 0 && dir != 1)) return fail(h, KDQ_EARG, "kdq_update_layer: bad argument");
    if (int r = update_layer(h, lt, dir, sweep)) return r;
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

int kdq_init_chain(kdq_handle h) {
    if (!h) return fail(h, KDQ_EARG, "kdq_init_chain: null handle");
    if (int r = init_chain(h)) return r;
    return check_health(h);
}

int kdq_measure(kdq_handle h, int32_t lt, double *out) {
    if (!h || !out || lt < 1 || lt > h->nt) return fail(h, KDQ_EARG, "kdq_measure: bad argument");
    HIPCHK(launch_zero_meas(h->stream, h->nch, h->d_chains));
    if (int r = measure(h, lt)) return r;
    return read_meas(h, out);
}

int kdq_profile(kdq_handle h, uint32_t mask) {
    if (!h) return fail(h, KDQ_EARG, "kdq_profile: null handle");
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
    HIPCHK(hipStreamSynchronize(h->stream));
    harvest(h);
    for (int f = 0; f < KDQ_PROF_FAMILIES; ++f) {
        ms[f] = h->prof_ms[f];
        launches[f] = h->prof_n[f];
    }
    return KDQ_OK;
}

int kdq_get_counters(kdq_handle h, double *out) {
    if (!h || !out) return fail(h, KDQ_EARG, "kdq_get_counters: bad argument");
    for (int c = 0; c < h->nch; ++c)
        HIPCHK(hipMemcpyAsync(out + 2 * c, h->ch[c].counters, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return KDQ_OK;
}

}  // extern "C"
