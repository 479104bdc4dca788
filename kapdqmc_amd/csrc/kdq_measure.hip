// Equal-time observables of phy_measure (src/tU/phy_update.jl:154-254), t-U model.
//
// Input is G0 after the wrap to τ = Nt/2 and the HalfeK sandwich (done with the GEMM
// kernels). One workgroup per chain reduces, in a fixed order,
//   Ek = -2 Ht Re Σ_ij K_ij G0_ij,   Eu = Re Σ_i (1 - G0_ii) conj(G0_ii),
// and the honeycomb CDW/SDW sums. The reference accumulates only the first expression of
// each `tmp += (...)` / `+adjoint(...)` line pair (the second line is a separate
// statement whose value is discarded); the same terms are accumulated here. Since the
// structure-factor sums are linear in the per-pair terms, every (r, cell) pair adds its
// term weighted by cos(q·r) directly; only real parts survive (real(CDW0)/Ns^2 etc.).
#include "kdq_kernels.h"

namespace {

constexpr int NT = 256;

__global__ __launch_bounds__(NT) void k_measure_tu(MeasArgs a) {
    const int c = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = ch.ns;
    const Opnd G = a.G[c];
    const size_t ld = G.ld;
    const int tid = threadIdx.x;
    double ek = 0.0, eu = 0.0, c0 = 0.0, c1 = 0.0, s0 = 0.0, s1 = 0.0;
    for (int q = tid; q < n * n; q += NT) {
        const int i = q % n, j = q / n;
        ek += ch.K[q] * G.re[i + j * ld];
        if (ch.Kim) ek -= ch.Kim[q] * G.im[i + j * ld];   // Re Σ K∘G with complex hopping
    }
    for (int i = tid; i < n; i += NT) {
        const double gr = G.re[i + i * ld], gi = G.im[i + i * ld];
        eu += (1.0 - gr) * gr - gi * gi;
    }
    if (ch.lattice <= 1) {  // honeycomb
        const int L1 = ch.L1, L2 = ch.L2, nc = L1 * L2;
        const double tp1 = 2.0 * M_PI / L1, tp2 = 2.0 * M_PI / L2;
        for (int p = tid; p < nc * nc; p += NT) {
            const int cell = p % nc, rr = p / nc;
            const int ix = cell % L1, iy = cell / L1;          // 0-based
            const int rx = rr % L1 + 1, ry = rr / L1 + 1;       // 1-based shift
            const int a0 = 2 * (ix + iy * L1);
            const int b0 = 2 * ((ix + rx) % L1 + ((iy + ry) % L2) * L1);
            const double delta = (a0 == b0) ? 1.0 : 0.0;
#define GR(x, y) G.re[(size_t)(x) + (size_t)(y) * ld]
#define GI(x, y) G.im[(size_t)(x) + (size_t)(y) * ld]
            double t1r = 0.0, t1i = 0.0, t2r = 0.0, t2i = 0.0, xr, xi;
            // tmp1 += (1 - g_aa)(1 - g_bb) + (δ - g_ab) g_ba
            cmul(1.0 - GR(a0, a0), -GI(a0, a0), 1.0 - GR(b0, b0), -GI(b0, b0), xr, xi);
            t1r += xr; t1i += xi;
            cmul(delta - GR(a0, b0), -GI(a0, b0), GR(b0, a0), GI(b0, a0), xr, xi);
            t1r += xr; t1i += xi;
            // tmp1 += (1 - g_a'a')(1 - g_b'b') + (δ - g_a'b') g_b'a'
            cmul(1.0 - GR(a0 + 1, a0 + 1), -GI(a0 + 1, a0 + 1), 1.0 - GR(b0 + 1, b0 + 1), -GI(b0 + 1, b0 + 1), xr, xi);
            t1r += xr; t1i += xi;
            cmul(delta - GR(a0 + 1, b0 + 1), -GI(a0 + 1, b0 + 1), GR(b0 + 1, a0 + 1), GI(b0 + 1, a0 + 1), xr, xi);
            t1r += xr; t1i += xi;
            // tmp1 -= (1 - g_a'a')(1 - g_bb) - g_a'b g_ba'
            cmul(1.0 - GR(a0 + 1, a0 + 1), -GI(a0 + 1, a0 + 1), 1.0 - GR(b0, b0), -GI(b0, b0), xr, xi);
            t1r -= xr; t1i -= xi;
            cmul(GR(a0 + 1, b0), GI(a0 + 1, b0), GR(b0, a0 + 1), GI(b0, a0 + 1), xr, xi);
            t1r += xr; t1i += xi;
            // tmp1 -= (1 - g_aa)(1 - g_b'b') - g_ab' g_b'a
            cmul(1.0 - GR(a0, a0), -GI(a0, a0), 1.0 - GR(b0 + 1, b0 + 1), -GI(b0 + 1, b0 + 1), xr, xi);
            t1r -= xr; t1i -= xi;
            cmul(GR(a0, b0 + 1), GI(a0, b0 + 1), GR(b0 + 1, a0), GI(b0 + 1, a0), xr, xi);
            t1r += xr; t1i += xi;
            // tmp2 += (1 - g_aa) conj(g_bb) ; += (1 - g_a'a') conj(g_b'b')
            cmul(1.0 - GR(a0, a0), -GI(a0, a0), GR(b0, b0), -GI(b0, b0), xr, xi);
            t2r += xr; t2i += xi;
            cmul(1.0 - GR(a0 + 1, a0 + 1), -GI(a0 + 1, a0 + 1), GR(b0 + 1, b0 + 1), -GI(b0 + 1, b0 + 1), xr, xi);
            t2r += xr; t2i += xi;
            // tmp2 -= (1 - g_aa) conj(g_b'b') ; -= (1 - g_a'a') conj(g_bb)
            cmul(1.0 - GR(a0, a0), -GI(a0, a0), GR(b0 + 1, b0 + 1), -GI(b0 + 1, b0 + 1), xr, xi);
            t2r -= xr; t2i -= xi;
            cmul(1.0 - GR(a0 + 1, a0 + 1), -GI(a0 + 1, a0 + 1), GR(b0, b0), -GI(b0, b0), xr, xi);
            t2r -= xr; t2i -= xi;
#undef GR
#undef GI
            const double w = cos(tp1 * rx + tp2 * ry);
            c0 += t1r + t2r;
            c1 += w * (t1r + t2r);
            s0 += t1r - t2r;
            s1 += w * (t1r - t2r);
            (void)t1i;
            (void)t2i;
        }
    }
    __shared__ double red[6][NT / 64];
    double v[6] = {ek, eu, c0, c1, s0, s1};
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) red[k][wave] = s;
    }
    __syncthreads();
    if (tid == 0) {
        double out[6];
        for (int k = 0; k < 6; ++k) {
            double s = 0.0;
            for (int w = 0; w < NT / 64; ++w) s += red[k][w];
            out[k] = s;
        }
        const double n2 = (double)n * (double)n;
        double *acc = a.meas ? a.meas + (size_t)c * KDQ_MEAS_SLOTS : ch.meas;
        acc[0] += -2.0 * ch.ht * out[0];
        acc[1] += out[1];
        acc[2] += out[2] / n2;
        acc[3] += out[3] / n2;
        acc[4] += out[4] / n2;
        acc[5] += out[5] / n2;
        acc[KDQ_MEAS_SLOTS - 1] += 1.0;
    }
}

}  // namespace

hipError_t launch_measure_tu(hipStream_t st, int nchains, const MeasArgs &m) {
    hipLaunchKernelGGL(k_measure_tu, dim3(nchains), dim3(NT), 0, st, m);
    return hipGetLastError();
}
