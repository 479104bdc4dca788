// Self-learning Monte Carlo for the t-U model: a translation-equivariant convolutional
// effective model of the HS field, trained on (field, log W) pairs the sweep produces, and
// the global moves it proposes (no counterpart in the reference).
//
// Field tensor: site i of the honeycomb lattice is unit cell i/2 = x + L1 y, sublattice i%2;
// channel 4·sublattice + (s - 1) is the one-hot value of s at (slice t, y, x). Layout
// [b][channel][t][y][x], x fastest. Network (BF16 weights, FP32 biases, activations and
// accumulation):
//   h1 = tanh(conv(X, W1) + b1),  h2 = tanh(conv(h1, W2) + b2),  o = conv(h2, W3),
//   E(s) = Σ_q o_q,
// 3 x 3 x 3 kernels, periodic in x and y (lattice translations), zero padding in t (the
// projection has open ends in imaginary time). Training minimises the batch variance of
// E + log W (only ratios enter the moves). All reductions run in a fixed order, so E is a
// deterministic function of the field, which the cumulative global move requires.
// The optimiser keeps FP32 master parameters; the kernels read a BF16 copy of the weights
// (round to nearest even), refreshed whenever the parameters change, so E is the network of
// the rounded weights and the weight gradient is taken at them.
#include "../../include/kapdqmc.h"
#include "kdq_kernels.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int CIN = 8;

struct Geo {
    int T, Y, X, P;   // slices, y cells, x cells, positions
};

KDQ_DEV int tap_pos(const Geo &g, int q, int k, bool &ok) {
    const int x = q % g.X, y = (q / g.X) % g.Y, t = q / (g.X * g.Y);
    const int dt = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
    const int tt = t + dt;
    ok = tt >= 0 && tt < g.T;
    const int yy = (y + dy + g.Y) % g.Y, xx = (x + dx + g.X) % g.X;
    return (tt * g.Y + yy) * g.X + xx;
}

// one-hot embedding; fields [b][n * nt], column-major per configuration (s[i + t n])
__global__ void k_embed(const uint8_t *fields, float *X0, Geo g, int n, int B) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= B * n * g.T) return;
    const int b = e / (n * g.T), r = e % (n * g.T), t = r / n, i = r % n;
    const int v = fields[(size_t)b * n * g.T + r] - 1, sub = i & 1, cell = i >> 1;
    const int q = (t * g.Y + cell / g.X) * g.X + cell % g.X;
    for (int k = 0; k < 4; ++k) X0[((size_t)b * CIN + sub * 4 + k) * g.P + q] = (k == v) ? 1.f : 0.f;
}

// ---- per-position evaluation, shared by the full forward pass and the local step so that
// both produce bit-identical activations (fixed tap order, explicit fma, same tanhf)
struct NetW {
    const uint16_t *W1, *W2, *W3;   // BF16 bit patterns
    const float *b1, *b2;
    int C;
};

KDQ_DEV float bf16f(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }

// FP32 -> BF16, round to nearest even (finite inputs)
KDQ_DEV uint16_t f2bf16(float x) {
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// value (1..4) of the field at position p, sublattice sub, of configuration F (ns x nt,
// column-major: s[i + t ns])
KDQ_DEV int field_at(const Geo &g, const uint8_t *F, int n, int p, int sub) {
    const int xy = g.X * g.Y;
    return F[2 * (p % xy) + sub + (size_t)(p / xy) * n];
}

template <class FV>
KDQ_DEV float h1_eval(const Geo &g, const NetW &w, int co, int q, FV fv) {
    float s = w.b1[co];
    for (int k = 0; k < 27; ++k) {
        bool ok;
        const int p = tap_pos(g, q, k, ok);
        if (!ok) continue;
        s += bf16f(w.W1[((size_t)co * CIN + fv(0, p) - 1) * 27 + k]);
        s += bf16f(w.W1[((size_t)co * CIN + 4 + fv(1, p) - 1) * 27 + k]);
    }
    return tanhf(s);
}

// Σ_ci Σ_k W[co][ci][k] in(ci, q + off_k), started from s0: taps resolved once (tag(p)),
// inputs read as iv(ci, tag); the fma sequence is the same whatever the accessors
template <class TAG, class IV>
KDQ_DEV float conv_eval(const Geo &g, const uint16_t *W, int cin, int co, int q, float s0, TAG tag, IV iv) {
    int src[27];
    bool ok[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const int p = tap_pos(g, q, k, ok[k]);
        src[k] = tag(p);
    }
    float s = s0;
    for (int ci = 0; ci < cin; ++ci) {
        const uint16_t *wk = W + ((size_t)co * cin + ci) * 27;
#pragma unroll
        for (int k = 0; k < 27; ++k)
            if (ok[k]) s = fmaf(bf16f(wk[k]), iv(ci, src[k]), s);
    }
    return s;
}

// h1[b][co][q] from the fields of configuration b (device pointer table)
__global__ void k_layer1(uint8_t *const *fp, NetW w, float *h1, Geo g, int n, int B) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)B * w.C * g.P) return;
    const int q = e % g.P, co = (e / g.P) % w.C, b = e / ((size_t)g.P * w.C);
    const uint8_t *F = fp[b];
    h1[e] = h1_eval(g, w, co, q, [&](int sub, int p) { return field_at(g, F, n, p, sub); });
}

// out[b][co][q] = act(bias[co] + conv(in)[co][q])
__global__ void k_conv(const float *in, const uint16_t *W, const float *bias, float *out, Geo g, int B, int cin,
                       int cout, int act) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)B * cout * g.P) return;
    const int q = e % g.P, co = (e / g.P) % cout, b = e / ((size_t)g.P * cout);
    const float *src = in + (size_t)b * cin * g.P;
    const float s = conv_eval(g, W, cin, co, q, bias ? bias[co] : 0.f, [](int p) { return p; },
                              [&](int ci, int p) { return src[(size_t)ci * g.P + p]; });
    out[e] = act ? tanhf(s) : s;
}

// per-block partial sums of o[b][q] in a fixed tree order; E[b] = Σ partials in order
constexpr int RT = 256;
__global__ void k_partial(const float *o, double *part, int P, int nblk) {
    const int b = blockIdx.y, blk = blockIdx.x;
    __shared__ double sh[RT];
    double s = 0.0;
    for (int q = blk * RT + threadIdx.x; q < P; q += nblk * RT) s += o[(size_t)b * P + q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = RT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)b * nblk + blk] = sh[0];
}
__global__ void k_final(const double *part, double *E, int nblk, int B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double s = 0.0;
    for (int k = 0; k < nblk; ++k) s += part[(size_t)b * nblk + k];
    E[b] = s;
}

// din[b][ci][q] = (Σ_co Σ_k w[co][ci][k] dout[b][co][q - off_k]) · (1 - h[b][ci][q]²) (h: the
// tanh output that `in` was; null: no derivative)
__global__ void k_conv_bwd_data(const float *dout, const uint16_t *w, const float *h, float *din, Geo g, int B,
                                int cin, int cout) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)B * cin * g.P) return;
    const int q = e % g.P, ci = (e / g.P) % cin, b = e / ((size_t)g.P * cin);
    int pos[27];
    bool ok[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) pos[k] = tap_pos(g, q, 26 - k, ok[k]);   // q - off_k
    float s = 0.f;
    for (int co = 0; co < cout; ++co) {
        const float *src = dout + ((size_t)b * cout + co) * g.P;
        const uint16_t *wk = w + ((size_t)co * cin + ci) * 27;
#pragma unroll
        for (int k = 0; k < 27; ++k)
            if (ok[k]) s += bf16f(wk[k]) * src[pos[k]];
    }
    if (h) {
        const float hv = h[e];
        s *= 1.f - hv * hv;
    }
    din[e] = s;
}

// dw[co][ci][k] = Σ_{b,q} dout[b][co][q] in[b][ci][q + off_k]; db[co] = Σ_{b,q} dout[b][co][q]
// (block (co, ci·27 + k); the bias rows use ci·27 + k = cin·27)
__global__ void k_conv_bwd_w(const float *dout, const float *in, float *dw, float *db, Geo g, int B, int cin,
                             int cout) {
    const int co = blockIdx.y, idx = blockIdx.x;
    const bool bias = idx == cin * 27;
    const int ci = bias ? 0 : idx / 27, k = bias ? 0 : idx % 27;
    __shared__ double sh[RT];
    double s = 0.0;
    for (int b = 0; b < B; ++b) {
        const float *d = dout + ((size_t)b * cout + co) * g.P;
        const float *src = in + ((size_t)b * cin + ci) * g.P;
        for (int q = threadIdx.x; q < g.P; q += RT) {
            if (bias) {
                s += d[q];
            } else {
                bool ok;
                const int p = tap_pos(g, q, k, ok);
                if (ok) s += (double)d[q] * src[p];
            }
        }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = RT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (bias) {
            if (db) db[co] = (float)sh[0];
        } else {
            dw[((size_t)co * cin + ci) * 27 + k] = (float)sh[0];
        }
    }
}

__global__ void k_fill(float *x, size_t n, const double *gb, int P) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) x[e] = (float)gb[e / P];
}

__global__ void k_fill_i(int *x, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) x[e] = 0;
}

// Adam on the FP32 master parameters; the BF16 copy follows
__global__ void k_adam(float *p, uint16_t *wb, const float *g, float *m, float *v, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gv = g[e];
    const float mv = b1 * m[e] + (1.f - b1) * gv, vv = b2 * v[e] + (1.f - b2) * gv * gv;
    m[e] = mv;
    v[e] = vv;
    const float pv = p[e] - lr * (mv / c1) / (sqrtf(vv / c2) + eps);
    p[e] = pv;
    wb[e] = f2bf16(pv);
}

__global__ void k_to_bf16(const float *p, uint16_t *wb, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) wb[e] = f2bf16(p[e]);
}

// ---- inner chain of a global move: local Metropolis steps on exp(-E)
//
// One step picks a slice class r in [0, 7) and proposes one single-field change (uniform
// site, uniform new value among the other three) in every slice t = r + 7 j. A change at
// position p alters h1 within Chebyshev distance 1 of p, h2 within 2 and o within 3, so its
// energy change depends on the fields within 6 slices: changes 7 slices apart are
// independent and run in parallel (one workgroup each), and each step is a mixture over r of
// products of commuting reversible kernels, hence reversible for exp(-E). The energy change is
// Σ over the changed outputs of (o_new - o_old) in double, with o recomputed by the same code
// as the full pass, so stored activations stay bit-identical to a fresh forward pass.

// distinct positions within Chebyshev radius r of (t, y, x): open in t, periodic in y and x
struct Box {
    int t0, nt, y0, ny, x0, nx;
};
KDQ_DEV Box box_of(const Geo &g, int t, int y, int x, int r) {
    Box b;
    b.t0 = max(0, t - r);
    b.nt = min(g.T - 1, t + r) - b.t0 + 1;
    if (2 * r + 1 >= g.Y) { b.y0 = 0; b.ny = g.Y; } else { b.y0 = (y - r + g.Y) % g.Y; b.ny = 2 * r + 1; }
    if (2 * r + 1 >= g.X) { b.x0 = 0; b.nx = g.X; } else { b.x0 = (x - r + g.X) % g.X; b.nx = 2 * r + 1; }
    return b;
}
KDQ_DEV int box_pos(const Geo &g, const Box &b, int j) {
    const int jx = j % b.nx, jy = (j / b.nx) % b.ny, jt = j / (b.nx * b.ny);
    return ((b.t0 + jt) * g.Y + (b.y0 + jy) % g.Y) * g.X + (b.x0 + jx) % g.X;
}
KDQ_DEV int box_index(const Geo &g, const Box &b, int q) {
    const int x = q % g.X, y = (q / g.X) % g.Y, t = q / (g.X * g.Y);
    const int jt = t - b.t0;
    if (jt < 0 || jt >= b.nt) return -1;
    const int jy = (y - b.y0 + g.Y) % g.Y;
    if (jy >= b.ny) return -1;
    const int jx = (x - b.x0 + g.X) % g.X;
    if (jx >= b.nx) return -1;
    return (jt * b.ny + jy) * b.nx + jx;
}

constexpr int SLICE_SEP = 7;
constexpr int STEP_THREADS = 256;

struct StepArgs {
    uint8_t **fields;        // [nch] device field arrays (changed in place)
    const uint64_t *seed;    // [nch]
    const uint32_t *chain;   // [nch]
    float *h1, *h2, *o;      // stored activations [b][C][P], [b][C][P], [b][P]
    int *nacc;               // [nch] accepted single-field changes
    NetW w;
    Geo g;
    int n, move, step;
};

// LDS: new h1 (27 C), new h2 (125 C), new o (343) floats, then STEP_THREADS doubles
inline size_t step_lds_bytes(int C) {
    return ((size_t)(27 + 125) * C + 343 + 1) / 2 * 2 * sizeof(float) + STEP_THREADS * sizeof(double);
}

__global__ __launch_bounds__(STEP_THREADS) void k_slmc_step(StepArgs a) {
    extern __shared__ double lds_d[];
    const int c = blockIdx.y, j = blockIdx.x, tid = threadIdx.x;
    const Geo g = a.g;
    const int C = a.w.C, n = a.n;
    const uint64_t sd = a.seed[c];
    const uint32_t cw = (3u << 24) | (a.chain[c] & 0xFFFFFFu);
    uint32_t x[4];
    philox4x32_10(0xFFFFFFFFu, (uint32_t)a.step, (uint32_t)a.move, cw, (uint32_t)sd, (uint32_t)(sd >> 32), x);
    const int r = (int)(((uint64_t)x[0] * SLICE_SEP) >> 32);
    const int t = r + SLICE_SEP * j;
    if (t >= g.T) return;   // block-uniform
    philox4x32_10((uint32_t)t, (uint32_t)a.step, (uint32_t)a.move, cw, (uint32_t)sd, (uint32_t)(sd >> 32), x);
    const int site = (int)(((uint64_t)x[0] * (uint32_t)n) >> 32);
    const int ch3 = (int)(((uint64_t)x[1] * 3u) >> 32);
    const double u = ((double)(x[2] >> 5) * 67108864.0 + (double)(x[3] >> 6)) * (1.0 / 9007199254740992.0);
    uint8_t *F = a.fields[c];
    const int old = F[site + (size_t)t * n];
    const int nv = ch3 + 1 + ((ch3 + 1) >= old ? 1 : 0);
    const int sub = site & 1, cell = site >> 1, cy = cell / g.X, cx = cell % g.X;
    const int q0 = (t * g.Y + cy) * g.X + cx;
    const Box B1 = box_of(g, t, cy, cx, 1), B2 = box_of(g, t, cy, cx, 2), B3 = box_of(g, t, cy, cx, 3);
    const int n1 = B1.nt * B1.ny * B1.nx, n2 = B2.nt * B2.ny * B2.nx, n3 = B3.nt * B3.ny * B3.nx;
    float *nh1 = (float *)(lds_d + STEP_THREADS);
    float *nh2 = nh1 + 27 * C;
    float *no = nh2 + 125 * C;
    double *red = lds_d;
    const size_t cP = (size_t)C * g.P;
    const float *H1 = a.h1 + c * cP, *H2 = a.h2 + c * cP, *O = a.o + (size_t)c * g.P;
    // new h1 around the change; LDS is channel-major (nh1[co n1 + j]) and lanes run over
    // positions, so weight reads are uniform across a wave. A tap's tag is -1 - j for
    // position j of the inner box (new value in LDS) or the global position p >= 0.
    for (int e = tid; e < n1 * C; e += STEP_THREADS) {
        const int co = e / n1, jp = e % n1;
        const int q = box_pos(g, B1, jp);
        nh1[e] = h1_eval(g, a.w, co, q, [&](int sb, int p) {
            return (p == q0 && sb == sub) ? nv : field_at(g, F, n, p, sb);
        });
    }
    __syncthreads();
    for (int e = tid; e < n2 * C; e += STEP_THREADS) {
        const int co = e / n2, jp = e % n2;
        const int q = box_pos(g, B2, jp);
        const float s = conv_eval(
            g, a.w.W2, C, co, q, a.w.b2[co],
            [&](int p) {
                const int jj = box_index(g, B1, p);
                return jj >= 0 ? -1 - jj : p;
            },
            [&](int ci, int tg) { return tg < 0 ? nh1[ci * n1 - 1 - tg] : H1[(size_t)ci * g.P + tg]; });
        nh2[e] = tanhf(s);
    }
    __syncthreads();
    double d = 0.0;
    for (int e = tid; e < n3; e += STEP_THREADS) {
        const int q = box_pos(g, B3, e);
        const float on = conv_eval(
            g, a.w.W3, C, 0, q, 0.f,
            [&](int p) {
                const int jj = box_index(g, B2, p);
                return jj >= 0 ? -1 - jj : p;
            },
            [&](int ci, int tg) { return tg < 0 ? nh2[ci * n2 - 1 - tg] : H2[(size_t)ci * g.P + tg]; });
        no[e] = on;
        d += (double)on - (double)O[q];
    }
    red[tid] = d;
    __syncthreads();
    for (int w = STEP_THREADS / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    const double dE = red[0];
    if (!(u < exp(-dE))) return;   // block-uniform
    float *H1w = a.h1 + c * cP, *H2w = a.h2 + c * cP, *Ow = a.o + (size_t)c * g.P;
    for (int e = tid; e < n1 * C; e += STEP_THREADS) H1w[(size_t)(e / n1) * g.P + box_pos(g, B1, e % n1)] = nh1[e];
    for (int e = tid; e < n2 * C; e += STEP_THREADS) H2w[(size_t)(e / n2) * g.P + box_pos(g, B2, e % n2)] = nh2[e];
    for (int e = tid; e < n3; e += STEP_THREADS) Ow[box_pos(g, B3, e)] = no[e];
    if (tid == 0) {
        F[site + (size_t)t * n] = (uint8_t)nv;
        a.nacc[c] += 1;
    }
}

// ---- training on BF16 matrix cores (C = 16 or 32 channels): layer 2 forward and backward-data,
// weight and bias gradients of layers 1 and 2
//
// The 3 x 3 x 3 convolution as an implicit GEMM: out[m][q] = Σ_(c, k) A[m][(c, k)] in[c][tap(q, k)],
// the reduction running over (input channel, tap) pairs, on MFMA f32_16x16x32_bf16 (lane l holds A[l%16][8(l/16) + v],
// B[8(l/16) + v][l%16] and D[4(l/16) + v][l%16], checked on the device). The weights are BF16
// already; each FP32 activation or gradient x is split into x_hi = bf16(x) and x_lo =
// bf16(x − x_hi), and A·x = A·x_hi + A·x_lo (products of two split operands add the
// hi·lo and lo·hi terms), so the result carries FP32-level accuracy (≈ 2^-16 relative per
// term) with FP32 accumulation. Training only: the energies of the inner chain and the global
// moves come from the per-position VALU code, which reproduces them bit for bit.
typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

KDQ_DEV void split_bf16(float x, bf16x8 &hi, bf16x8 &lo, int v) {
    const uint16_t h = f2bf16(x);
    hi[v] = (short)h;
    lo[v] = (short)f2bf16(x - bf16f(h));
}

// position (x, y, t) of q and its tap k: the same result as tap_pos, with the coordinates
// decomposed once by the caller and the periodic wrap done by compare (|dx|, |dy| <= 1)
struct Coord {
    int x, y, t;
};
KDQ_DEV Coord coord_of(const Geo &g, int q) { return Coord{q % g.X, (q / g.X) % g.Y, q / (g.X * g.Y)}; }
KDQ_DEV int tap_at(const Geo &g, Coord c, int k, bool &ok) {
    const int tt = c.t + k / 9 - 1;
    int yy = c.y + (k / 3) % 3 - 1, xx = c.x + k % 3 - 1;
    yy += yy < 0 ? g.Y : (yy >= g.Y ? -g.Y : 0);
    xx += xx < 0 ? g.X : (xx >= g.X ? -g.X : 0);
    ok = tt >= 0 && tt < g.T;
    return (tt * g.Y + yy) * g.X + xx;
}

constexpr int MF_NQ = 64;   // positions per workgroup (4 waves x 16)

// The reduction index is tap-major, kk = k·C + c, so each lane's run of 8 consecutive kk shares
// one tap and reads 8 channels at one position.
// MODE 0: forward, A[co][k·C + ci] = W[co][ci][k], out = tanh(bias + acc).
// MODE 1: backward data, A[ci][k·C + co] = W[co][ci][k] (taps mirrored), out = acc · (1 − h²).
template <int C, int MODE>
__global__ __launch_bounds__(256) void k_conv_mfma(const float *in, const uint16_t *W, const float *bias,
                                                   const float *h, float *out, Geo g, int B) {
    constexpr int K = C * 27, KP = (K + 31) / 32 * 32, MT = C / 16;
    __shared__ uint16_t As[C][KP + 8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = blockIdx.y, q0 = blockIdx.x * MF_NQ;
    for (int e = tid; e < C * KP; e += 256) {
        const int m = e / KP, kk = e % KP;
        uint16_t v = 0;
        if (kk < K) {
            const int k = kk / C, c2 = kk % C;
            v = MODE == 0 ? W[((size_t)m * C + c2) * 27 + k] : W[((size_t)c2 * C + m) * 27 + k];
        }
        As[m][kk] = v;
    }
    __syncthreads();
    const int j = lane & 15, grp = lane >> 4;
    const int q = q0 + wave * 16 + j;
    const bool qok = q < g.P;
    const Coord cq = coord_of(g, qok ? q : 0);
    const float *src = in + (size_t)b * C * g.P;
    f32x4 acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < KP; k0 += 32) {
        const int kk0 = k0 + 8 * grp, k = kk0 / C, c0 = kk0 % C;
        bool ok = false;
        int p = 0;
        if (qok && kk0 < K) p = tap_at(g, cq, MODE == 0 ? k : 26 - k, ok);
        bf16x8 bh, bl;
#pragma unroll
        for (int v = 0; v < 8; ++v) split_bf16(ok ? src[(size_t)(c0 + v) * g.P + p] : 0.f, bh, bl, v);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            bf16x8 a;
#pragma unroll
            for (int v = 0; v < 8; ++v) a[v] = (short)As[mt * 16 + j][kk0 + v];
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh, acc[mt], 0, 0, 0);
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl, acc[mt], 0, 0, 0);
        }
    }
    if (!qok) return;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = mt * 16 + 4 * grp + v;
            const size_t o = ((size_t)b * C + m) * g.P + q;
            if (MODE == 0) {
                out[o] = tanhf(bias[m] + acc[mt][v]);
            } else {
                const float hv = h[o];
                out[o] = acc[mt][v] * (1.f - hv * hv);
            }
        }
}

// Weight and bias gradient partials of a layer with CO outputs and CI inputs:
// part[chunk][co][n] = Σ_{positions of the chunk} dout[co][q] x_n(q), with x_n(q) = in[ci][tap(q, k)]
// for n = ci·27 + k < CI·27 and x_n = 1 for the bias column n = CI·27; M = CO, N = CI·27 + 1 in
// blocks of 64 columns, K = positions (flattened over the configurations).
constexpr int WG_CHUNK = 2048, WG_SUB = 256;
template <int CO, int CI>
__global__ __launch_bounds__(256) void k_wgrad_mfma(const float *dout, const float *in, float *part, Geo g, int B) {
    constexpr int N = CI * 27 + 1, MT = CO / 16;
    // dout of WG_SUB positions, split once and shared by the four waves
    __shared__ uint16_t Ah[CO][WG_SUB + 8], Al[CO][WG_SUB + 8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, j = lane & 15, grp = lane >> 4;
    const int n = blockIdx.x * 64 + wave * 16 + j;   // this lane's B column
    const bool nok = n < N, isb = n == CI * 27;
    const int ci = nok && !isb ? n / 27 : 0, k = nok && !isb ? n % 27 : 0;
    const size_t total = (size_t)B * g.P;
    f32x4 acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int s0 = 0; s0 < WG_CHUNK; s0 += WG_SUB) {
        const size_t sb = (size_t)blockIdx.y * WG_CHUNK + s0;
        {
            const size_t qg = sb + tid;   // one position per thread (WG_SUB = blockDim)
            const bool in_range = qg < total;
            const int bb = in_range ? (int)(qg / g.P) : 0, qq = in_range ? (int)(qg % g.P) : 0;
            __syncthreads();   // the previous sub-chunk's reads are done
            for (int co = 0; co < CO; ++co) {
                const float x = in_range ? dout[((size_t)bb * CO + co) * g.P + qq] : 0.f;
                const uint16_t h = f2bf16(x);
                Ah[co][tid] = h;
                Al[co][tid] = f2bf16(x - bf16f(h));
            }
            __syncthreads();
        }
        for (int k0 = 0; k0 < WG_SUB; k0 += 32) {
            bf16x8 bh, bl;
            // the lane's 8 consecutive positions: decompose the first, then step x with carries
            const size_t q0g = sb + k0 + 8 * grp;
            int b0 = (int)(q0g / g.P), r0 = (int)(q0g % g.P);
            Coord c = coord_of(g, r0);
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                float x = 0.f;
                if (q0g + v < total && nok) {
                    if (isb) {
                        x = 1.f;
                    } else {
                        bool ok;
                        const int p = tap_at(g, c, k, ok);
                        if (ok) x = in[((size_t)b0 * CI + ci) * g.P + p];
                    }
                }
                split_bf16(x, bh, bl, v);
                if (++r0 == g.P) {
                    r0 = 0;
                    ++b0;
                    c = Coord{0, 0, 0};
                } else if (++c.x == g.X) {
                    c.x = 0;
                    if (++c.y == g.Y) {
                        c.y = 0;
                        ++c.t;
                    }
                }
            }
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                bf16x8 ah, al;
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    ah[v] = (short)Ah[mt * 16 + j][k0 + 8 * grp + v];
                    al[v] = (short)Al[mt * 16 + j][k0 + 8 * grp + v];
                }
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[mt], 0, 0, 0);
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[mt], 0, 0, 0);
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[mt], 0, 0, 0);
            }
        }
    }
    if (!nok) return;
    float *pc = part + (size_t)blockIdx.y * CO * N;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) pc[(size_t)(mt * 16 + 4 * grp + v) * N + n] = acc[mt][v];
}

// Σ over the chunks in double, fixed order: dw[co][n] (n < ncol − 1) and db[co] (n = ncol − 1)
__global__ void k_wgrad_reduce(const float *part, int nchunk, int cout, int ncol, float *dw, float *db) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cout * ncol) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * cout * ncol + e];
    const int co = e / ncol, n = e % ncol;
    if (n < ncol - 1)
        dw[(size_t)co * (ncol - 1) + n] = (float)s;
    else
        db[co] = (float)s;
}

inline int nb(size_t n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

// =================================================================== network object
struct kdq_net_s {
    Geo g{};
    int n = 0, C = 0, maxB = 0, nblk = 0;
    int np = 0;                  // parameter count
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    uint16_t *d_wb = nullptr;        // [np] BF16 copy of d_p (the kernels read its weight entries)
    float *d_X0 = nullptr, *d_h1 = nullptr, *d_h2 = nullptr, *d_o = nullptr;
    float *d_d2 = nullptr, *d_d1 = nullptr, *d_do = nullptr;
    double *d_part = nullptr, *d_E = nullptr, *d_gb = nullptr;
    float *d_wpart = nullptr;        // weight-gradient partials of the matrix-core layer 2 (C = 16, 32)
    int nwchunk = 0;
    uint8_t *d_fields = nullptr;     // [maxB][ns nt] fields uploaded by the host entry points
    uint8_t **d_fptr = nullptr;      // [maxB] device pointers into d_fields
    uint64_t *d_seed = nullptr;      // [maxB] inner-chain seeds of kdq_net_sample
    uint32_t *d_chain = nullptr;     // [maxB]
    int *d_nacc = nullptr;           // [maxB]
    int step = 0, sample_calls = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<void *> allocs;
    // parameter offsets
    size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0;
    ~kdq_net_s() {
        if (stream) hipStreamSynchronize(stream);
        for (void *p : allocs) hipFree(p);
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {

std::string g_net_err;

int nfail(kdq_net n, const std::string &m) {
    if (n) n->err = m;
    g_net_err = m;
    return KDQ_EARG;
}

#define NCHK(expr)                                                                \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            (void)hipGetLastError();                                              \
            g_net_err = std::string(#expr) + ": " + hipGetErrorString(e_);        \
            if (n) n->err = g_net_err;                                            \
            return KDQ_EHIP;                                                      \
        }                                                                         \
    } while (0)

template <class T>
int nalloc(kdq_net n, T **p, size_t count) {
    void *q = nullptr;
    NCHK(hipMalloc(&q, sizeof(T) * (count ? count : 1)));
    n->allocs.push_back(q);
    NCHK(hipMemset(q, 0, sizeof(T) * (count ? count : 1)));
    *p = (T *)q;
    return KDQ_OK;
}

NetW net_weights(kdq_net n) {
    const float *p = n->d_p;
    const uint16_t *wb = n->d_wb;
    return NetW{wb + n->oW1, wb + n->oW2, wb + n->oW3, p + n->ob1, p + n->ob2, n->C};
}

// refresh the BF16 weight copy after the master parameters changed
int net_sync_bf16(kdq_net n) {
    hipLaunchKernelGGL(k_to_bf16, dim3(nb(n->np, 256)), dim3(256), 0, n->stream, n->d_p, n->d_wb, n->np);
    NCHK(hipGetLastError());
    NCHK(hipStreamSynchronize(n->stream));
    return KDQ_OK;
}

// E[b] = Σ_q o[b][q] in a fixed order
int net_energy_sum(kdq_net n, int B, hipStream_t st) {
    hipLaunchKernelGGL(k_partial, dim3(n->nblk, B), dim3(RT), 0, st, n->d_o, n->d_part, n->g.P, n->nblk);
    hipLaunchKernelGGL(k_final, dim3(nb(B, 64)), dim3(64), 0, st, n->d_part, n->d_E, n->nblk, B);
    NCHK(hipGetLastError());
    return KDQ_OK;
}

// forward pass of the B configurations whose fields the device table fp points to ->
// h1, h2, o and E
// (matrix: layer 2 on the BF16 matrix cores, C = 16 or 32; training only)
int net_forward(kdq_net n, uint8_t *const *fp, int B, hipStream_t st, bool matrix = false) {
    const Geo g = n->g;
    const int C = n->C;
    const NetW w = net_weights(n);
    hipLaunchKernelGGL(k_layer1, dim3(nb((size_t)B * C * g.P, 256)), dim3(256), 0, st, fp, w, n->d_h1, g, n->n, B);
    if (matrix && C == 32)
        hipLaunchKernelGGL((k_conv_mfma<32, 0>), dim3(nb(g.P, MF_NQ), B), dim3(256), 0, st, n->d_h1, w.W2, w.b2,
                           (const float *)nullptr, n->d_h2, g, B);
    else if (matrix && C == 16)
        hipLaunchKernelGGL((k_conv_mfma<16, 0>), dim3(nb(g.P, MF_NQ), B), dim3(256), 0, st, n->d_h1, w.W2, w.b2,
                           (const float *)nullptr, n->d_h2, g, B);
    else
        hipLaunchKernelGGL(k_conv, dim3(nb((size_t)B * C * g.P, 256)), dim3(256), 0, st, n->d_h1, w.W2, w.b2, n->d_h2, g,
                           B, C, C, 1);
    hipLaunchKernelGGL(k_conv, dim3(nb((size_t)B * g.P, 256)), dim3(256), 0, st, n->d_h2, w.W3, (const float *)nullptr,
                       n->d_o, g, B, C, 1, 0);
    NCHK(hipGetLastError());
    return net_energy_sum(n, B, st);
}

int net_upload(kdq_net n, const uint8_t *host_fields, int B) {
    const size_t per = (size_t)n->n * n->g.T;
    for (size_t e = 0; e < per * B; ++e)   // the kernels index weights by the field value
        if (host_fields[e] < 1 || host_fields[e] > 4) return nfail(n, "field values must be 1..4");
    NCHK(hipMemcpyAsync(n->d_fields, host_fields, per * B, hipMemcpyHostToDevice, n->stream));
    return KDQ_OK;
}

// one-hot input tensor of the uploaded fields (training only: the weight gradient of layer 1)
int net_embed(kdq_net n, int B) {
    const size_t per = (size_t)n->n * n->g.T;
    hipLaunchKernelGGL(k_embed, dim3(nb(per * B, 256)), dim3(256), 0, n->stream, n->d_fields, n->d_X0, n->g, n->n, B);
    NCHK(hipGetLastError());
    return KDQ_OK;
}

// nsteps inner-chain steps on the configurations of fp (activations of a forward pass of
// the same fields in d_h1/d_h2/d_o), then E of the final fields
int net_steps(kdq_net n, uint8_t **fp, const uint64_t *d_seed, const uint32_t *d_chain, int B, int nsteps, int move,
              hipStream_t st) {
    StepArgs a{};
    a.fields = fp;
    a.seed = d_seed;
    a.chain = d_chain;
    a.h1 = n->d_h1;
    a.h2 = n->d_h2;
    a.o = n->d_o;
    a.nacc = n->d_nacc;
    a.w = net_weights(n);
    a.g = n->g;
    a.n = n->n;
    a.move = move;
    const dim3 grid(nb(n->g.T, SLICE_SEP), B);
    const size_t lds = step_lds_bytes(n->C);
    hipLaunchKernelGGL(k_fill_i, dim3(nb(B, 64)), dim3(64), 0, st, n->d_nacc, B);
    for (int k = 0; k < nsteps; ++k) {
        a.step = k;
        hipLaunchKernelGGL(k_slmc_step, grid, dim3(STEP_THREADS), lds, st, a);
    }
    NCHK(hipGetLastError());
    return net_energy_sum(n, B, st);
}

// gradient of loss = Var_b(E_b + logw_b) into d_g; returns the loss
int net_grad(kdq_net n, const uint8_t *fields, const double *logw, int B, double *loss) {
    if (int r = net_upload(n, fields, B)) return r;
    if (int r = net_embed(n, B)) return r;
    const bool matrix = n->d_wpart != nullptr;
    if (int r = net_forward(n, n->d_fptr, B, n->stream, matrix)) return r;
    std::vector<double> E(B), gb(B);
    NCHK(hipMemcpyAsync(E.data(), n->d_E, sizeof(double) * B, hipMemcpyDeviceToHost, n->stream));
    NCHK(hipStreamSynchronize(n->stream));
    double mean = 0.0;
    for (int b = 0; b < B; ++b) mean += E[b] + logw[b];
    mean /= B;
    double L = 0.0;
    for (int b = 0; b < B; ++b) {
        const double d = E[b] + logw[b] - mean;
        L += d * d;
        gb[b] = 2.0 * d / B;   // dL/dE_b (the mean's own derivative sums to zero)
    }
    *loss = L / B;
    const Geo g = n->g;
    const int C = n->C;
    float *gr = n->d_g;
    NCHK(hipMemcpyAsync(n->d_gb, gb.data(), sizeof(double) * B, hipMemcpyHostToDevice, n->stream));
    const size_t no = (size_t)B * g.P;
    hipLaunchKernelGGL(k_fill, dim3(nb(no, 256)), dim3(256), 0, n->stream, n->d_do, no, n->d_gb, g.P);
    // layer 3 (no bias, linear): dW3, d(h2) through tanh
    hipLaunchKernelGGL(k_conv_bwd_w, dim3(C * 27, 1), dim3(RT), 0, n->stream, n->d_do, n->d_h2, gr + n->oW3,
                       (float *)nullptr, g, B, C, 1);
    hipLaunchKernelGGL(k_conv_bwd_data, dim3(nb((size_t)B * C * g.P, 256)), dim3(256), 0, n->stream, n->d_do,
                       n->d_wb + n->oW3, n->d_h2, n->d_d2, g, B, C, 1);
    // layer 2
    if (matrix) {
        const int N = C * 27 + 1;
        const dim3 gw(nb(N, 64), n->nwchunk), gd(nb(g.P, MF_NQ), B);
        if (C == 32) {
            hipLaunchKernelGGL((k_wgrad_mfma<32, 32>), gw, dim3(256), 0, n->stream, n->d_d2, n->d_h1, n->d_wpart, g, B);
            hipLaunchKernelGGL((k_conv_mfma<32, 1>), gd, dim3(256), 0, n->stream, n->d_d2, n->d_wb + n->oW2,
                               (const float *)nullptr, n->d_h1, n->d_d1, g, B);
        } else {
            hipLaunchKernelGGL((k_wgrad_mfma<16, 16>), gw, dim3(256), 0, n->stream, n->d_d2, n->d_h1, n->d_wpart, g, B);
            hipLaunchKernelGGL((k_conv_mfma<16, 1>), gd, dim3(256), 0, n->stream, n->d_d2, n->d_wb + n->oW2,
                               (const float *)nullptr, n->d_h1, n->d_d1, g, B);
        }
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(nb((size_t)C * N, 256)), dim3(256), 0, n->stream, n->d_wpart,
                           n->nwchunk, C, N, gr + n->oW2, gr + n->ob2);
    } else {
        hipLaunchKernelGGL(k_conv_bwd_w, dim3(C * 27 + 1, C), dim3(RT), 0, n->stream, n->d_d2, n->d_h1, gr + n->oW2,
                           gr + n->ob2, g, B, C, C);
        hipLaunchKernelGGL(k_conv_bwd_data, dim3(nb((size_t)B * C * g.P, 256)), dim3(256), 0, n->stream, n->d_d2,
                           n->d_wb + n->oW2, n->d_h1, n->d_d1, g, B, C, C);
    }
    // layer 1
    if (matrix) {
        const int N = CIN * 27 + 1;
        const dim3 gw(nb(N, 64), n->nwchunk);
        if (C == 32)
            hipLaunchKernelGGL((k_wgrad_mfma<32, CIN>), gw, dim3(256), 0, n->stream, n->d_d1, n->d_X0, n->d_wpart, g, B);
        else
            hipLaunchKernelGGL((k_wgrad_mfma<16, CIN>), gw, dim3(256), 0, n->stream, n->d_d1, n->d_X0, n->d_wpart, g, B);
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(nb((size_t)C * N, 256)), dim3(256), 0, n->stream, n->d_wpart,
                           n->nwchunk, C, N, gr + n->oW1, gr + n->ob1);
    } else {
        hipLaunchKernelGGL(k_conv_bwd_w, dim3(CIN * 27 + 1, C), dim3(RT), 0, n->stream, n->d_d1, n->d_X0, gr + n->oW1,
                           gr + n->ob1, g, B, CIN, C);
    }
    NCHK(hipGetLastError());
    return KDQ_OK;
}

}  // namespace

extern "C" {
const char *kdq_net_last_error(kdq_net n) { return n ? n->err.c_str() : g_net_err.c_str(); }

int kdq_net_create(int32_t L1, int32_t L2, int32_t nt, int32_t channels, uint64_t seed, int32_t max_batch,
                   kdq_net *out) {
    if (!out || L1 < 1 || L2 < 1 || nt < 1 || channels < 1 || channels > 64 || max_batch < 1)
        return nfail(nullptr, "kdq_net_create: bad argument");
    *out = nullptr;
    kdq_net n = new kdq_net_s();
    n->g = Geo{nt, L2, L1, nt * L1 * L2};
    n->n = 2 * L1 * L2;
    n->C = channels;
    n->maxB = max_batch;
    n->nblk = std::min(64, nb(n->g.P, RT));
    const int C = channels;
    n->oW1 = 0;
    n->ob1 = n->oW1 + (size_t)C * CIN * 27;
    n->oW2 = n->ob1 + C;
    n->ob2 = n->oW2 + (size_t)C * C * 27;
    n->oW3 = n->ob2 + C;
    n->np = (int)(n->oW3 + (size_t)C * 27);
    const size_t P = n->g.P, B = max_batch;
    int r = 0;
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) r = KDQ_EHIP;
    r |= nalloc(n, &n->d_p, n->np); r |= nalloc(n, &n->d_g, n->np); r |= nalloc(n, &n->d_wb, n->np);
    r |= nalloc(n, &n->d_m, n->np); r |= nalloc(n, &n->d_v, n->np);
    r |= nalloc(n, &n->d_X0, B * CIN * P); r |= nalloc(n, &n->d_h1, B * C * P); r |= nalloc(n, &n->d_h2, B * C * P);
    r |= nalloc(n, &n->d_o, B * P); r |= nalloc(n, &n->d_d2, B * C * P); r |= nalloc(n, &n->d_d1, B * C * P);
    r |= nalloc(n, &n->d_do, B * P);
    r |= nalloc(n, &n->d_part, B * n->nblk); r |= nalloc(n, &n->d_E, B); r |= nalloc(n, &n->d_gb, B);
    r |= nalloc(n, &n->d_fields, B * (size_t)n->n * nt);
    r |= nalloc(n, &n->d_fptr, B); r |= nalloc(n, &n->d_seed, B); r |= nalloc(n, &n->d_chain, B);
    r |= nalloc(n, &n->d_nacc, B);
    if (C == 16 || C == 32) {
        n->nwchunk = nb(B * P, WG_CHUNK);
        r |= nalloc(n, &n->d_wpart, (size_t)n->nwchunk * C * (C * 27 + 1));
    }
    if (r) {
        nfail(n, "kdq_net_create: device allocation failed");
        delete n;
        return KDQ_EHIP;
    }
    // Glorot-uniform initial weights from a host generator seeded by `seed`; zero biases
    std::vector<float> p(n->np, 0.f);
    uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](size_t off, size_t cnt, int fan_in, int fan_out) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (size_t q = 0; q < cnt; ++q) p[off + q] = (float)((2.0 * uni() - 1.0) * a);
    };
    fill(n->oW1, (size_t)C * CIN * 27, CIN * 27, C * 27);
    fill(n->oW2, (size_t)C * C * 27, C * 27, C * 27);
    fill(n->oW3, (size_t)C * 27, C * 27, 27);
    std::vector<uint8_t *> fpt(B);
    std::vector<uint32_t> cid(B);
    for (size_t b = 0; b < B; ++b) {
        fpt[b] = n->d_fields + b * (size_t)n->n * nt;
        cid[b] = (uint32_t)b;
    }
    if (hipMemcpy(n->d_p, p.data(), sizeof(float) * n->np, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->d_fptr, fpt.data(), sizeof(uint8_t *) * B, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->d_chain, cid.data(), sizeof(uint32_t) * B, hipMemcpyHostToDevice) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_slmc_step, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)step_lds_bytes(C)) != hipSuccess ||
        net_sync_bf16(n) != KDQ_OK) {
        delete n;
        nfail(nullptr, "kdq_net_create: parameter upload failed");
        return KDQ_EHIP;
    }
    *out = n;
    return KDQ_OK;
}

int kdq_net_destroy(kdq_net n) {
    delete n;
    return KDQ_OK;
}

int kdq_net_num_params(kdq_net n) { return n ? n->np : 0; }

int kdq_net_get_params(kdq_net n, float *out) {
    if (!n || !out) return nfail(n, "kdq_net_get_params: bad argument");
    NCHK(hipMemcpy(out, n->d_p, sizeof(float) * n->np, hipMemcpyDeviceToHost));
    return KDQ_OK;
}

int kdq_net_set_params(kdq_net n, const float *in) {
    if (!n || !in) return nfail(n, "kdq_net_set_params: bad argument");
    NCHK(hipMemcpy(n->d_p, in, sizeof(float) * n->np, hipMemcpyHostToDevice));
    return net_sync_bf16(n);
}

int kdq_net_energy(kdq_net n, const uint8_t *fields, int32_t batch, double *E) {
    if (!n || !fields || !E || batch < 1 || batch > n->maxB) return nfail(n, "kdq_net_energy: bad argument");
    if (int r = net_upload(n, fields, batch)) return r;
    if (int r = net_forward(n, n->d_fptr, batch, n->stream)) return r;
    NCHK(hipMemcpyAsync(E, n->d_E, sizeof(double) * batch, hipMemcpyDeviceToHost, n->stream));
    NCHK(hipStreamSynchronize(n->stream));
    return KDQ_OK;
}

int kdq_net_sample(kdq_net n, uint8_t *fields, int32_t batch, int32_t nsteps, uint64_t seed, double *E,
                   int32_t *accepted) {
    if (!n || !fields || batch < 1 || batch > n->maxB || nsteps < 0) return nfail(n, "kdq_net_sample: bad argument");
    if (int r = net_upload(n, fields, batch)) return r;
    std::vector<uint64_t> sd(batch, seed);
    NCHK(hipMemcpyAsync(n->d_seed, sd.data(), sizeof(uint64_t) * batch, hipMemcpyHostToDevice, n->stream));
    if (int r = net_forward(n, n->d_fptr, batch, n->stream)) return r;
    if (int r = net_steps(n, n->d_fptr, n->d_seed, n->d_chain, batch, nsteps, n->sample_calls, n->stream)) return r;
    n->sample_calls += 1;
    NCHK(hipMemcpyAsync(fields, n->d_fields, (size_t)n->n * n->g.T * batch, hipMemcpyDeviceToHost, n->stream));
    if (E) NCHK(hipMemcpyAsync(E, n->d_E, sizeof(double) * batch, hipMemcpyDeviceToHost, n->stream));
    if (accepted) NCHK(hipMemcpyAsync(accepted, n->d_nacc, sizeof(int) * batch, hipMemcpyDeviceToHost, n->stream));
    NCHK(hipStreamSynchronize(n->stream));
    return KDQ_OK;
}

int kdq_net_loss_grad(kdq_net n, const uint8_t *fields, const double *logw, int32_t batch, double *loss,
                      float *grad) {
    if (!n || !fields || !logw || !loss || batch < 2 || batch > n->maxB) return nfail(n, "kdq_net_loss_grad: bad argument");
    if (int r = net_grad(n, fields, logw, batch, loss)) return r;
    if (grad) NCHK(hipMemcpyAsync(grad, n->d_g, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->stream));
    NCHK(hipStreamSynchronize(n->stream));
    return KDQ_OK;
}

int kdq_net_train_step(kdq_net n, const uint8_t *fields, const double *logw, int32_t batch, float lr, double *loss) {
    if (!n || !fields || !logw || !loss || batch < 2 || batch > n->maxB || !(lr > 0.f))
        return nfail(n, "kdq_net_train_step: bad argument");
    if (int r = net_grad(n, fields, logw, batch, loss)) return r;
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
    hipLaunchKernelGGL(k_adam, dim3(nb(n->np, 256)), dim3(256), 0, n->stream, n->d_p, n->d_wb, n->d_g, n->d_m, n->d_v,
                       n->np, lr, b1, b2, 1e-8f, c1, c2);
    NCHK(hipGetLastError());
    NCHK(hipStreamSynchronize(n->stream));
    return KDQ_OK;
}
}  // extern "C"

// ---- the inner chain of a global move (called by the engine on its stream): forward pass of
// the chains' trial fields (E0 to the host), nsteps local steps, final energies in d_E1
int slmc_inner(kdq_net n, hipStream_t st, int nch, uint8_t **d_field_ptrs, const uint64_t *d_seed,
               const uint32_t *d_chain, double *d_E1, int nsteps, int move, double *E0_host) {
    if (nch > n->maxB) return nfail(n, "slmc move: more chains than the network's batch");
    if (int r = net_forward(n, d_field_ptrs, nch, st)) return r;
    NCHK(hipMemcpyAsync(E0_host, n->d_E, sizeof(double) * nch, hipMemcpyDeviceToHost, st));
    if (int r = net_steps(n, d_field_ptrs, d_seed, d_chain, nch, nsteps, move, st)) return r;
    NCHK(hipMemcpyAsync(d_E1, n->d_E, sizeof(double) * nch, hipMemcpyDeviceToDevice, st));
    return KDQ_OK;
}

int slmc_geometry(kdq_net n, int *ns, int *nt) {
    *ns = n->n;
    *nt = n->g.T;
    return KDQ_OK;
}
