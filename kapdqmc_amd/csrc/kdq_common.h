// Shared device/host definitions of the MI355X projector-DQMC library.
//
// Complex matrices are stored planar (a real plane and an imaginary plane, each
// column-major with leading dimension ld) because CDNA4 matrix cores have no complex
// type: a real x complex product is two real GEMMs sharing one operand and a
// complex x complex product is four. Every kernel is batched over independent chains
// (blockIdx.z = chain) and reads per-chain sizes and pointers from a ChainDev table in
// device memory, so chains of different lattice size share one launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define KDQ_DEV __device__ __forceinline__
#define KDQ_HD __host__ __device__ __forceinline__

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// One operand of a batched GEMM: pointers of chain c in a device table.
struct Opnd {
    double *re;
    double *im;  // nullptr for real operands
    int ld;
    int pad;
};
// FP32 operand (mixed-precision G and kinetic factors); same layout rules as Opnd.
struct OpndF {
    float *re;
    float *im;
    int ld;
    int pad;
};
template <class T> struct OpndOf;
template <> struct OpndOf<double> { typedef Opnd type; };
template <> struct OpndOf<float> { typedef OpndF type; };

// Per-chain constants and state pointers (device table, one entry per chain).
struct ChainDev {
    int ns, np, nt, pad0;
    const double *alpha;   // [nt]
    double eta[4];
    double gamma[4];
    uint8_t *s;            // ns x nt column-major, values 1..4
    double ht;
    int lattice, L1, L2, pad1;
    const double *K;       // ns x ns
    uint64_t seed;
    uint32_t chain_id, stream;   // Philox stream: 0 plain sweep, 1 + replica in SCEE
    // local-update scratch
    int *acc_cnt;          // [2]: accepted count of a block, slot by block parity (t-U sweep)
    int *acc_idx;          // [2][64]: accepted sites of a block, by slot
    int *pw_fld;           // [64]: field values of a prewrapped block 0 (old | new << 8 | accepted << 16), written back by its flush
    double *W;             // [2 slots][2 planes (re, im)] of 64 x 64, ld 64
    double *counters;      // [3]: proposals, accepted, near ties (|u - p| < KDQ_NEAR_TIE)
    double *meas;          // [KDQ_MEAS_SLOTS]: accumulated observables, count in the last slot
    // model kind and field layout
    int kind;              // 0 t-U, 1 t-V, 2 t-U-V
    int n_elem;            // field elements per slice (t-U: ns, t-V: 3 ns/2, t-U-V: ns z)
    int is_complex;        // t-U-V: HS phases exp(i·) (U > 0) or exp(·) (U < 0)
    int nb;                // t-V: bonds per family (ns/2)
    const int *ex, *ey;    // [n_elem] 0-based sites of element e (t-V, t-U-V)
    const int *inc;        // t-U-V: [ns * 6] incidence, entry = elem * 2 + (0: coefficient 1, 1: coefficient a)
    double a, U, V;        // t-U-V constants (a = V/g), couplings used by the observables
    int nra, nrb;          // SCEE region sizes (0 outside SCEE)
    // deferred flush of a t-U slice: rk[b] = rank of the factors U, V before block b (rk[0]
    // is never read: 0), *rank = rank after the latest block
    int *rk;
    int *rank;
    // replayed draws (kdq_set_replay): nullptr = the Philox stream. Entry
    // ((s * 2 + dir) * nt + lt - 1) * n_elem + e is the proposal of element e at slice lt of
    // pass dir of sweep rp_sweep0 + s.
    const uint8_t *rp_choice;
    const double *rp_u;
    uint32_t rp_sweep0, rp_nsweeps;
    const double *Kim;     // imaginary part of K (complex hopping), nullptr when K is real
    // block-circulant kinetic factors: with ns = 2 L² sites in L x L periodic cells (site
    // 2·(x + L·y) + b), eK and eK⁻¹ act as 2 x 2 matrices per cell momentum k. circ holds, for
    // eK then eK⁻¹, Λ(k) = (1/L²) Σ_d E(d) e^{+i k·d} ([k][a][b] complex, k = kx + L·ky) with
    // E_ab(d) = the entry of row a (cell 0), column 2d + b. cL = L, or 0 (dense products only).
    const double *circ;
    int cL, pad2;
};

#define KDQ_MEAS_SLOTS 16
// A decision whose uniform draw u and acceptance probability p differ by less than this is a
// near tie: rounding differences between two correct implementations can flip it.
#define KDQ_NEAR_TIE 1e-12

// DIM_ACC: the accepted count of the current update block (device counter), the rank of a
// delayed-update flush.
// DIM_RANK: the rank of the deferred factors after the latest block.
enum { DIM_NS = 0, DIM_NP = 1, DIM_NA = 2, DIM_NB = 3, DIM_ACC = 4, DIM_RANK = 5 };

KDQ_DEV int dim_of(const ChainDev &c, int code) {
    switch (code) {
        case DIM_NS: return c.ns;
        case DIM_NP: return c.np;
        case DIM_NA: return c.nra;
        case DIM_NB: return c.nrb;
        case DIM_ACC: return *c.acc_cnt;
        default: return *c.rank;
    }
}

// Diagonal factor applied to GEMM rows, columns or the contraction index.
// kind 0: none; kind 1: t-U phase cis(sign * alpha[lt-1] * eta[s[i, lt-1] - 1]);
// kind 2: t-U-V site factor of the directed-bond field, x = sign * Σ_b coef_b η[s_b],
//         exp(i x) (complex chains) or exp(x) (real chains).
struct DiagSpec {
    int kind;
    int lt;    // 1-based slice
    int sign;  // +1 or -1
    int conj;  // 1: complex conjugate of the value (D^H for the left chain)
};

KDQ_DEV void diag_value(const ChainDev &c, const DiagSpec &d, int i, double &re, double &im) {
    if (d.kind == 1) {
        int sv = c.s[(size_t)i + (size_t)(d.lt - 1) * c.n_elem];
        double x = d.sign * c.alpha[d.lt - 1] * c.eta[sv - 1];
        sincos(x, &im, &re);
    } else if (d.kind == 2) {
        const uint8_t *sl = c.s + (size_t)(d.lt - 1) * c.n_elem;
        double x1 = 0.0, x2 = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int e = c.inc[i * 6 + k];
            const double v = c.eta[sl[e >> 1] - 1];
            if (e & 1) x2 += v;
            else x1 += v;
        }
        const double x = d.sign * (x1 + c.a * x2);
        if (c.is_complex) {
            sincos(x, &im, &re);
        } else {
            re = exp(x);
            im = 0.0;
        }
    } else {
        re = 1.0;
        im = 0.0;
    }
    if (d.conj) im = -im;
}

template <class T>
KDQ_HD void cmulT(T ar, T ai, T br, T bi, T &cr, T &ci) {
    T r = ar * br - ai * bi;
    T i = ar * bi + ai * br;
    cr = r;
    ci = i;
}

KDQ_HD void cmul(double ar, double ai, double br, double bi, double &cr, double &ci) {
    double r = ar * br - ai * bi;
    double i = ar * bi + ai * br;
    cr = r;
    ci = i;
}

// ---------------------------------------------------------------- Philox4x32-10
// Identical stream layout to the CPU restatement: key = {seed_lo, seed_hi},
// ctr = {elem, slice_tag, sweep, (stream << 24) | chain}.
KDQ_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                          uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Proposal draw: choice in {0,1,2} (index into the other three field values, increasing
// order) and a uniform double in [0,1).
KDQ_HD void proposal_draw(uint64_t seed, uint32_t chain, uint32_t stream, uint32_t sweep,
                          uint32_t slice_tag, uint32_t elem, int &choice, double &u) {
    uint32_t x[4];
    philox4x32_10(elem, slice_tag, sweep, (stream << 24) | (chain & 0xFFFFFFu), (uint32_t)seed,
                  (uint32_t)(seed >> 32), x);
    choice = (int)(((uint64_t)x[0] * 3u) >> 32);
    u = ((double)(x[1] >> 5) * 67108864.0 + (double)(x[2] >> 6)) * (1.0 / 9007199254740992.0);
}

// Draw of a local-update proposal of the plain sweep (Philox stream 0), or the caller's replayed
// draw when the chain has a replay buffer. A sweep outside the buffer reads as a rejected
// proposal (u = +inf); the host refuses calls that would reach one.
KDQ_HD void sweep_draw(const ChainDev &ch, uint32_t sweep, int lt, int dir, uint32_t elem, int &choice, double &u) {
    if (ch.rp_u) {
        const uint32_t s = sweep - ch.rp_sweep0;
        if (s < ch.rp_nsweeps) {
            const size_t q = (((size_t)s * 2 + (size_t)dir) * (size_t)ch.nt + (size_t)(lt - 1)) * (size_t)ch.n_elem + elem;
            choice = ch.rp_choice[q] % 3;
            u = ch.rp_u[q];
        } else {
            choice = 0;
            u = __builtin_huge_val();
        }
        return;
    }
    proposal_draw(ch.seed, ch.chain_id, 0, sweep, (uint32_t)lt | ((uint32_t)dir << 31), elem, choice, u);
}

KDQ_HD int init_field_value(uint64_t seed, uint32_t chain, uint32_t stream, uint32_t slice,
                            uint32_t elem) {
    uint32_t x[4];
    philox4x32_10(elem, slice, 0xFFFFFFFFu, (stream << 24) | (chain & 0xFFFFFFu), (uint32_t)seed,
                  (uint32_t)(seed >> 32), x);
    return 1 + (int)(((uint64_t)x[0] * 4u) >> 32);
}

// ------------------------------------------------------------------ reductions
KDQ_DEV double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

KDQ_DEV double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
