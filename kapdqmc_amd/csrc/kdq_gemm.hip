// Batched planar-complex GEMM on CDNA4 FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// Tile 64 x 64 x 16, 256 threads = 4 wave64s in a 2 x 2 arrangement, each wave owning a
// 32 x 32 block of C as 2 x 2 MFMA 16x16 blocks. Operands are staged through LDS as
// k-major tiles. The MFMA is fed with op(B) as its first operand and op(A) as its second,
// so each lane's four accumulator registers hold four *columns* of one row and every store
// instruction writes 16 consecutive rows: coalesced stores into the column-major output.
//
// Diagonal factors of the projector-DQMC algebra are fused: the HS phases of a time slice
// (cis(±α η_s)) scale rows/columns in the epilogue or the contraction index while the
// complex operand is staged, so WrapKV! (src/tU/GreenMatrix.jl:56-86) is two launches.
#include "kdq_kernels.h"

namespace {

constexpr int BM = 64, BN = 64, BK = 16, LDP = 64 + 2;

template <int AC, int BC, int OPA, int OPB>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    constexpr bool OC = AC || BC;
    const int c = blockIdx.z;
    const ChainDev &ch = g.chains[c];
    const int M = dim_of(ch, g.mdim), N = dim_of(ch, g.ndim), K = dim_of(ch, g.kdim);
    const int i0 = blockIdx.y * BM, j0 = blockIdx.x * BN;
    if (i0 >= M || j0 >= N) return;
    const Opnd A = g.A[c], B = g.B[c];

    __shared__ double As[AC ? 2 : 1][BK][LDP];
    __shared__ double Bs[BC ? 2 : 1][BK][LDP];
    __shared__ double Sk[2][BK];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    v4d accR[2][2], accI[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accR[a][b] = v4d{0, 0, 0, 0};
            accI[a][b] = v4d{0, 0, 0, 0};
        }
    const bool scale_k = g.inS.kind != 0;

    for (int k0 = 0; k0 < K; k0 += BK) {
        if (scale_k) {
            if (tid < BK) {
                double sr = 1.0, si = 0.0;
                if (k0 + tid < K) diag_value(ch, g.inS, k0 + tid, sr, si);
                Sk[0][tid] = sr;
                Sk[1][tid] = si;
            }
            __syncthreads();
        }
        // ---- stage op(A)[i0:i0+64, k0:k0+16] as As[k][i]
#pragma unroll
        for (int q = 0; q < (BM * BK) / 256; ++q) {
            const int e = tid + q * 256;
            int ii, kk;
            if (OPA == 0) { ii = e & 63; kk = e >> 6; }
            else { kk = e & 15; ii = e >> 4; }
            const int gi = i0 + ii, gk = k0 + kk;
            double vr = 0.0, vi = 0.0;
            if (gi < M && gk < K) {
                const size_t off = OPA == 0 ? (size_t)gi + (size_t)gk * A.ld : (size_t)gk + (size_t)gi * A.ld;
                vr = A.re[off];
                if (AC) {
                    vi = A.im[off];
                    if (OPA == 2) vi = -vi;
                }
            }
            if (AC && scale_k) cmul(vr, vi, Sk[0][kk], Sk[1][kk], vr, vi);
            As[0][kk][ii] = vr;
            if (AC) As[1][kk][ii] = vi;
        }
        // ---- stage op(B)[k0:k0+16, j0:j0+64] as Bs[k][j]
#pragma unroll
        for (int q = 0; q < (BN * BK) / 256; ++q) {
            const int e = tid + q * 256;
            int jj, kk;
            if (OPB == 0) { kk = e & 15; jj = e >> 4; }
            else { jj = e & 63; kk = e >> 6; }
            const int gj = j0 + jj, gk = k0 + kk;
            double vr = 0.0, vi = 0.0;
            if (gj < N && gk < K) {
                const size_t off = OPB == 0 ? (size_t)gk + (size_t)gj * B.ld : (size_t)gj + (size_t)gk * B.ld;
                vr = B.re[off];
                if (BC) {
                    vi = B.im[off];
                    if (OPB == 2) vi = -vi;
                }
            }
            if (BC && !AC && scale_k) cmul(vr, vi, Sk[0][kk], Sk[1][kk], vr, vi);
            Bs[0][kk][jj] = vr;
            if (BC) Bs[1][kk][jj] = vi;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4), cc = lane & 15;
            double ar[2], ai[2], br[2], bi[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                ar[t] = As[0][kr][wm * 32 + t * 16 + cc];
                ai[t] = AC ? As[1][kr][wm * 32 + t * 16 + cc] : 0.0;
                br[t] = Bs[0][kr][wn * 32 + t * 16 + cc];
                bi[t] = BC ? Bs[1][kr][wn * 32 + t * 16 + cc] : 0.0;
            }
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) {
                    accR[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[ni], ar[mi], accR[mi][ni], 0, 0, 0);
                    if (AC && BC) {
                        accR[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bi[ni], ai[mi], accR[mi][ni], 0, 0, 0);
                        accI[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(bi[ni], ar[mi], accI[mi][ni], 0, 0, 0);
                        accI[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[ni], ai[mi], accI[mi][ni], 0, 0, 0);
                    } else if (BC) {
                        accI[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(bi[ni], ar[mi], accI[mi][ni], 0, 0, 0);
                    } else if (AC) {
                        accI[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[ni], ai[mi], accI[mi][ni], 0, 0, 0);
                    }
                }
        }
        __syncthreads();
    }

    // ---- epilogue. For v_mfma_f64_16x16x4f64 lane l, register v holds D[(l/16) + 4v][l%16]
    // of the 16x16 block (checked on gfx950); D = op(B)ᵀ op(A)ᵀ here, so the lane writes
    // C[i][j] with i = l%16 and j = l/16 + 4v: 16 consecutive rows per store.
    const Opnd Cm = g.C[c];
    Opnd Ci;
    if (g.mode == 2) Ci = g.Cin[c];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
        const int i = i0 + wm * 32 + mi * 16 + (lane & 15);
        if (i >= M) continue;
        double rsr = 1.0, rsi = 0.0;
        if (g.rowS.kind) diag_value(ch, g.rowS, i, rsr, rsi);
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + wn * 32 + ni * 16 + (lane >> 4) + 4 * v;
                if (j >= N) continue;
                double cr = accR[mi][ni][v], ci = OC ? accI[mi][ni][v] : 0.0;
                if (g.rowS.kind) cmul(cr, ci, rsr, rsi, cr, ci);
                if (g.colS.kind) {
                    double csr, csi;
                    diag_value(ch, g.colS, j, csr, csi);
                    cmul(cr, ci, csr, csi, cr, ci);
                }
                const size_t off = (size_t)i + (size_t)j * Cm.ld;
                if (g.mode == 1) {
                    cr = (i == j ? 1.0 : 0.0) - cr;
                    ci = -ci;
                } else if (g.mode == 2) {
                    const size_t oi = (size_t)i + (size_t)j * Ci.ld;
                    cr = Ci.re[oi] - cr;
                    if (OC) ci = Ci.im[oi] - ci;
                }
                Cm.re[off] = cr;
                if (OC) Cm.im[off] = ci;
            }
    }
}

template <int AC, int BC, int OPA, int OPB>
hipError_t run(hipStream_t st, int nchains, int max_m, int max_n, const GemmArgs &g) {
    dim3 grid((max_n + BN - 1) / BN, (max_m + BM - 1) / BM, nchains);
    hipLaunchKernelGGL((k_gemm<AC, BC, OPA, OPB>), grid, dim3(256), 0, st, g);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gemm(hipStream_t st, int nchains, int max_m, int max_n, bool a_cplx, bool b_cplx,
                       int opa, int opb, const GemmArgs &g) {
    const int code = (a_cplx ? 1000 : 0) + (b_cplx ? 100 : 0) + opa * 10 + opb;
    switch (code) {
        case 100: return run<0, 1, 0, 0>(st, nchains, max_m, max_n, g);     // eK · G
        case 110: return run<0, 1, 1, 0>(st, nchains, max_m, max_n, g);     // eKᵀ · X
        case 1000: return run<1, 0, 0, 0>(st, nchains, max_m, max_n, g);    // T · eKinv
        case 1100: return run<1, 1, 0, 0>(st, nchains, max_m, max_n, g);    // U_R · X
        case 1120: return run<1, 1, 2, 0>(st, nchains, max_m, max_n, g);    // U_Lᴴᴴ · U_R
        case 1102: return run<1, 1, 0, 2>(st, nchains, max_m, max_n, g);    // M⁻¹ · U_Lᴴᴴ
        case 0: return run<0, 0, 0, 0>(st, nchains, max_m, max_n, g);
        case 10: return run<0, 0, 1, 0>(st, nchains, max_m, max_n, g);
        default: return hipErrorInvalidValue;
    }
}
