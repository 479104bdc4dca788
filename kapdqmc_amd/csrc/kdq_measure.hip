// Equal-time observables of phy_measure (src/tU/phy_update.jl:154-254, src/tV/phy_update.jl:
// 207-314, src/tUV/phy_update.jl:205-355) on G0, the G wrapped to τ = Nt/2 and sandwiched with
// HalfeK (done with the GEMM kernels).
//
// The structure factors are reorganised by what they measure. With the staggered density
// m(c) = n(c, A) − n(c, B) of unit cell c (n = 1 − G_ii for spin up, conj(G_ii) for the
// particle-hole transformed spin down of t-U; n↓ = n↑ for t-U-V with U < 0), the reference's
// sums over shifts r and cells c are
//   Σ_r tmp1(r) w(r) = Σ_{c,c'} w(c' − c) [m↑(c) m↑(c') + X(c, c')],
//   Σ_r tmp2(r) w(r) = Σ_{c,c'} w(c' − c) m↑(c) m↓(c'),
// w = 1 or cos(q·r), X(c, c') = Σ_ab ε_a ε_b (δ − G)[ca, c'b] G[c'b, ca] (ε_A = 1, ε_B = −1),
// where only the first expression of each `tmp += … / +adjoint(…)` line pair accumulates (the
// second line is a separate statement whose value is discarded). cos(q·(c' − c)) =
// Re e^{iθ(c')} e^{−iθ(c)} with θ(c) = 2π x/L1 + 2π y/L2 separates, so the m terms are products
// of O(N) sums, and the pair term X is an elementwise sum over G. t-V's R0/R1 are the same with
// the four sublattice pairs kept apart and ⟨n_i n_j⟩ = n_i n_j + (δ − G_ji) G_ij.
//
// Two launches: k_meas_tiles reduces the O(N²) elementwise sums (K∘G and the pair terms) over
// 32 x 32 tiles of G, many workgroups per chain, each writing its partial sums; k_meas_finish
// (one workgroup per chain) adds the partials in a fixed order, forms the O(N) density sums and
// the bond sums, and accumulates the row. No atomics: results are bitwise reproducible.
#include "kdq_kernels.h"

namespace {

constexpr int TS = 32;     // tile edge
constexpr int NP = 9;      // partial sums per tile
constexpr int NT = 256;

// cos θ_i, sin θ_i of the unit cell of site i (site i is in cell i / 2 = x + L1 y)
KDQ_DEV void cell_phase(const ChainDev &ch, int i, double &cs, double &sn) {
    const int cell = i >> 1, x = cell % ch.L1, y = cell / ch.L1;
    sincos(2.0 * M_PI * x / ch.L1 + 2.0 * M_PI * y / ch.L2, &sn, &cs);
}

// Partials of tile (I, J): t-U / t-U-V: [Re Σ K∘G, Re Σ X, Re Σ cos X]; t-V: [Σ K∘G,
// Σ Y by sublattice pair (AA, BB, BA, AB), the same weighted by cos(θ_j − θ_i)].
__global__ __launch_bounds__(NT) void k_meas_tiles(MeasArgs a) {
    const int c = blockIdx.z;
    const ChainDev &ch = a.chains[c];
    const int n = ch.ns;
    const int i0 = blockIdx.y * TS, j0 = blockIdx.x * TS;
    double *out = a.part + (((size_t)c * a.T + blockIdx.y) * a.T + blockIdx.x) * NP;
    const int tid = threadIdx.x;
    if (i0 >= n || j0 >= n) {   // tiles beyond a smaller chain's G
        if (tid < NP) out[tid] = 0.0;
        return;
    }
    const Opnd G = a.G[c];
    const bool cx = G.im != nullptr && ch.kind != 1;
    __shared__ double Ar[TS][TS + 1], Ai[TS][TS + 1];   // G[i0 + r, j0 + q]
    __shared__ double Br[TS][TS + 1], Bi[TS][TS + 1];   // G[j0 + r, i0 + q]
    __shared__ double ci[TS], si[TS], cj[TS], sj[TS];
    for (int e = tid; e < TS * TS; e += NT) {
        const int r = e % TS, q = e / TS;
        const bool oa = i0 + r < n && j0 + q < n, ob = j0 + r < n && i0 + q < n;
        const size_t pa = (size_t)(i0 + r) + (size_t)(j0 + q) * G.ld, pb = (size_t)(j0 + r) + (size_t)(i0 + q) * G.ld;
        Ar[r][q] = oa ? G.re[pa] : 0.0;
        Br[r][q] = ob ? G.re[pb] : 0.0;
        Ai[r][q] = (oa && cx) ? G.im[pa] : 0.0;
        Bi[r][q] = (ob && cx) ? G.im[pb] : 0.0;
    }
    if (tid < TS && i0 + tid < n) cell_phase(ch, i0 + tid, ci[tid], si[tid]);
    if (tid >= 64 && tid < 64 + TS && j0 + tid - 64 < n) cell_phase(ch, j0 + tid - 64, cj[tid - 64], sj[tid - 64]);
    __syncthreads();
    const bool honey = ch.lattice <= 1;
    double v[NP] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = tid; e < TS * TS; e += NT) {
        const int r = e % TS, q = e / TS, i = i0 + r, j = j0 + q;
        if (i >= n || j >= n) continue;
        const size_t pk = (size_t)i + (size_t)j * n;
        const double gr = Ar[r][q], gi = Ai[r][q];          // G_ij
        const double tr = Br[q][r], ti = Bi[q][r];          // G_ji
        v[0] += ch.K[pk] * gr;
        if (ch.Kim) v[0] -= ch.Kim[pk] * gi;
        if (!honey) continue;
        const double w = cj[q] * ci[r] + sj[q] * si[r];      // cos(θ_j − θ_i)
        const double d = (i == j) ? 1.0 : 0.0;
        if (ch.kind == 1) {   // t-V: Y = (δ − G_ji) G_ij, bin by (i, j) sublattices
            const double y = (d - tr) * gr;
            const int k = (i & 1) == (j & 1) ? (i & 1) : ((i & 1) ? 2 : 3);
            v[1 + k] += y;
            v[5 + k] += w * y;
        } else {              // X = ε_i ε_j (δ − G_ij) G_ji, real part
            const double x = ((d - gr) * tr + gi * ti) * (((i ^ j) & 1) ? -1.0 : 1.0);
            v[1] += x;
            v[2] += w * x;
        }
    }
    __shared__ double red[NP][NT / 64];
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) red[k][wave] = s;
    }
    __syncthreads();
    if (tid < NP) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; ++w) s += red[tid][w];
        out[tid] = s;
    }
}

struct Cx {
    double r, i;
};
KDQ_DEV Cx cm(Cx a, Cx b) { return Cx{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }

// Per-chain finish: adds the tile partials, forms the density sums (plain and Fourier at ±q)
// and the bond sums, and accumulates the observables of the row.
__global__ __launch_bounds__(NT) void k_meas_finish(MeasArgs a) {
    const int c = blockIdx.x, tid = threadIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = ch.ns;
    const Opnd G = a.G[c];
    const bool cx = G.im != nullptr && ch.kind != 1;
    const bool cplx_spin = ch.kind == 0 || (ch.kind == 2 && ch.is_complex);   // n↓ = conj(G_ii)
    const double *part = a.part + (size_t)c * a.T * a.T * NP;
    constexpr int NV = 25;
    // v[0..8]: tile partials; t-U / t-U-V: v[9..12] Σ m↑, Σ m↓ (re, im), v[13..16] Σ m↑ e^{∓iθ},
    // v[17..20] Σ m↓ e^{∓iθ}, v[21] Eu; t-V: v[9..13] sublattice A: Σ n, Σ n e^{−iθ} (re, im),
    // Σ n e^{+iθ} (re, im), v[14..18] the same for B; v[24] bond sum (Ev)
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    for (int t = tid; t < a.T * a.T; t += NT)
#pragma unroll
        for (int k = 0; k < NP; ++k) v[k] += part[(size_t)t * NP + k];
    for (int i = tid; i < n; i += NT) {
        const double gr = G.re[(size_t)i + (size_t)i * G.ld], gi = cx ? G.im[(size_t)i + (size_t)i * G.ld] : 0.0;
        double cs, sn;
        cell_phase(ch, i, cs, sn);
        if (ch.kind == 1) {
            const double nn = 1.0 - gr;
            const int b = (i & 1) * 5;
            v[9 + b] += nn;
            v[10 + b] += nn * cs;
            v[11 + b] -= nn * sn;
            v[12 + b] += nn * cs;
            v[13 + b] += nn * sn;
        } else {
            v[21] += cplx_spin ? (1.0 - gr) * gr - gi * gi : (1.0 - gr) * (1.0 - gr);   // Eu
            const double e = (i & 1) ? -1.0 : 1.0;
            const Cx up{e * (1.0 - gr), -e * gi};
            const Cx dn = cplx_spin ? Cx{e * gr, -e * gi} : up;
            const Cx em{cs, -sn}, ep{cs, sn};
            const Cx a1 = cm(up, em), a2 = cm(up, ep), b1 = cm(dn, em), b2 = cm(dn, ep);
            v[9] += up.r;  v[10] += up.i;  v[11] += dn.r;  v[12] += dn.i;
            v[13] += a1.r; v[14] += a1.i;  v[15] += a2.r;  v[16] += a2.i;
            v[17] += b1.r; v[18] += b1.i;  v[19] += b2.r;  v[20] += b2.i;
        }
    }
    if (ch.kind != 0) {   // bond sums: t-V Ev, t-U-V Ev
#define GR(p, q) G.re[(size_t)(p) + (size_t)(q) * G.ld]
#define GI(p, q) (cx ? G.im[(size_t)(p) + (size_t)(q) * G.ld] : 0.0)
        for (int k = tid; k < ch.n_elem; k += NT) {
            const int x = ch.ex[k], y = ch.ey[k];
            if (ch.kind == 1) {
                v[24] += (1 - GR(x, x)) * (1 - GR(y, y)) - GR(x, y) * GR(y, x);
            } else if (ch.is_complex) {
                double t0r, t0i, t1r, t1i;
                // (1-gxx)(1-gyy) - gxy gyx  +  conj(gxx gyy - gxy gyx)  +  (1-gyy) conj(gxx) + (1-gxx) conj(gyy)
                cmul(1.0 - GR(x, x), -GI(x, x), 1.0 - GR(y, y), -GI(y, y), t0r, t0i);
                cmul(GR(x, y), GI(x, y), GR(y, x), GI(y, x), t1r, t1i);
                v[24] += t0r - t1r;
                cmul(GR(x, x), GI(x, x), GR(y, y), GI(y, y), t0r, t0i);
                v[24] += t0r - t1r;
                cmul(1.0 - GR(y, y), -GI(y, y), GR(x, x), -GI(x, x), t0r, t0i);
                v[24] += t0r;
                cmul(1.0 - GR(x, x), -GI(x, x), GR(y, y), -GI(y, y), t0r, t0i);
                v[24] += t0r;
            } else {
                v[24] += 2.0 * ((1.0 - GR(x, x)) * (1.0 - GR(y, y)) - GR(x, y) * GR(y, x));
                v[24] += 2.0 * (1.0 - GR(y, y)) * (1.0 - GR(x, x));
            }
        }
#undef GR
#undef GI
    }
    __shared__ double red[NV][NT / 64];
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) red[k][wave] = s;
    }
    __syncthreads();
    if (tid != 0) return;
    double S[NV];
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; ++w) s += red[k][w];
        S[k] = s;
    }
    const double n2 = (double)n * (double)n;
    double *acc = a.meas ? a.meas + (size_t)c * KDQ_MEAS_SLOTS : ch.meas;
    const bool honey = ch.lattice <= 1;
    if (ch.kind == 1) {   // t-V: [Ek Ev R0[4] R1[4]], sublattice pairs (cell c, cell c + r): AA, BB, BA, AB
        acc[0] += ch.ht * S[0];
        acc[1] += S[24];
        const double f = 4.0 / n2;
        const int pa[4] = {0, 1, 1, 0}, pb[4] = {0, 1, 0, 1};
        for (int k = 0; k < 4; ++k) {
            const double *A = S + 9 + 5 * pa[k], *B = S + 9 + 5 * pb[k];
            const Cx am{A[1], A[2]}, ap{A[3], A[4]}, bm{B[1], B[2]}, bp{B[3], B[4]};
            const double r0 = A[0] * B[0] + S[1 + k];
            const double r1 = 0.5 * (cm(am, bp).r + cm(ap, bm).r) + S[5 + k];
            acc[2 + k] += honey ? f * r0 : 0.0;
            acc[6 + k] += honey ? f * r1 : 0.0;
        }
        acc[KDQ_MEAS_SLOTS - 1] += 1.0;
        return;
    }
    // t-U, t-U-V: Σ_r tmp1, tmp2 plain and cos-weighted
    const Cx M{S[9], S[10]}, D{S[11], S[12]}, Mm{S[13], S[14]}, Mp{S[15], S[16]}, Dm{S[17], S[18]}, Dp{S[19], S[20]};
    const double f = (ch.kind == 2 && !ch.is_complex) ? 2.0 : 1.0;
    const double t1_0 = f * (cm(M, M).r + S[1]), t1_1 = f * (0.5 * (cm(Mm, Mp).r + cm(Mp, Mm).r) + S[2]);
    const double t2_0 = f * cm(M, D).r, t2_1 = f * 0.5 * (cm(Mm, Dp).r + cm(Mp, Dm).r);
    const double cdw0 = honey ? (t1_0 + t2_0) / n2 : 0.0, cdw1 = honey ? (t1_1 + t2_1) / n2 : 0.0;
    const double sdw0 = honey ? (t1_0 - t2_0) / n2 : 0.0, sdw1 = honey ? (t1_1 - t2_1) / n2 : 0.0;
    if (ch.kind == 0) {   // [Ek Eu CDW0 CDW1 SDW0 SDW1]
        acc[0] += -2.0 * ch.ht * S[0];
        acc[1] += S[21];
        acc[2] += cdw0; acc[3] += cdw1; acc[4] += sdw0; acc[5] += sdw1;
    } else {              // [Ek Eu Ev CDW0 CDW1 SDW0 SDW1]
        acc[0] += 2.0 * ch.ht * S[0];
        acc[1] += ch.U * S[21];
        acc[2] += ch.V * S[24];
        acc[3] += cdw0; acc[4] += cdw1; acc[5] += sdw0; acc[6] += sdw1;
    }
    acc[KDQ_MEAS_SLOTS - 1] += 1.0;
}

}  // namespace

int meas_tiles(int max_ns) { return (max_ns + TS - 1) / TS; }
size_t meas_part_doubles(int nchains, int max_ns) {
    const int T = meas_tiles(max_ns);
    return (size_t)nchains * T * T * NP;
}

hipError_t launch_measure(hipStream_t st, int nchains, const MeasArgs &m) {
    hipLaunchKernelGGL(k_meas_tiles, dim3(m.T, m.T, nchains), dim3(NT), 0, st, m);
    if (hipError_t e = hipGetLastError()) return e;
    hipLaunchKernelGGL(k_meas_finish, dim3(nchains), dim3(NT), 0, st, m);
    return hipGetLastError();
}
