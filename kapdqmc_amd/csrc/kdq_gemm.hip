// Batched planar-complex GEMM on CDNA4 FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// Tile BM x BN x 16 (BM = BN = 64 or 32), 256 threads = 4 wave64s in a 2 x 2 arrangement,
// each wave owning a (BM/2) x (BN/2) block of C as MFMA 16x16 blocks. Operands are staged
// through double-buffered LDS as k-major tiles; the next k-tile is fetched into registers
// while the current one is multiplied. The MFMA is fed with op(B) as its first operand and
// op(A) as its second, so each lane's four accumulator registers hold four *columns* of one
// row and every store instruction writes 16 consecutive rows of the column-major output.
//
// Diagonal factors of the projector-DQMC algebra are fused: the HS phases of a time slice
// (cis(±α η_s)) scale rows/columns in the epilogue or the contraction index while the
// complex operand is staged, so WrapKV! (src/tU/GreenMatrix.jl:56-86) is two launches.
// The factors are evaluated once per workgroup into LDS (rows, columns, whole K range).
#include "kdq_kernels.h"

namespace {

constexpr int BK = 16;
constexpr int KMAX = 512;   // longest contraction with a diagonal factor (N <= 512)

KDQ_DEV v4d mma(double b, double a, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, c, 0, 0, 0); }
KDQ_DEV v4f mma(float b, float a, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(b, a, c, 0, 0, 0); }
template <class T> struct Acc;
template <> struct Acc<double> { typedef v4d type; };
template <> struct Acc<float> { typedef v4f type; };
// column (within the 16x16 block) held in accumulator register v of lane l: the FP64 MFMA
// spreads a row over l/16 + 4v, the FP32 one over 4(l/16) + v (both checked on gfx950)
template <class T> KDQ_DEV int acc_col(int lane, int v);
template <> KDQ_DEV int acc_col<double>(int lane, int v) { return (lane >> 4) + 4 * v; }
template <> KDQ_DEV int acc_col<float>(int lane, int v) { return 4 * (lane >> 4) + v; }

// PF: mode 2 with its Cin entries loaded before the k loop (all at once, overlapping the first
// k-tile) instead of one by one between the epilogue's stores.
// Workgroups are dealt to the 8 XCDs round-robin by linear id. The launch is one-dimensional
// and linear id L is given tile g of the chain-major (chain, row tile, column tile) order, with
// each XCD taking one contiguous run of g: every XCD works through its own chains, whose
// operands are then fetched into that XCD's L2 only (a bijection for any tile count).
constexpr int NXCD = 8;
KDQ_DEV int xcd_tile(int L, int total) {
    const int x = L % NXCD, k = L / NXCD, q = total / NXCD, r = total % NXCD;
    return x * q + min(x, r) + k;
}

template <class T, int AC, int BC, int OPA, int OPB, int BM, int BN, bool PF>
__global__ __launch_bounds__(256) void k_gemm(GemmArgsT<T> g, int tx, int ty) {
    typedef typename Acc<T>::type VA;
    constexpr bool OC = AC || BC;
    constexpr int WMB = BM / 32, WNB = BN / 32;   // 16x16 MFMA blocks per wave (each dim)
    // LDS row strides: FP64 odd, so that the transposed staging stores (16 lanes of 8 bytes
    // down one column) hit distinct banks; FP32 2 mod 4 (32 lanes over two columns)
    constexpr int PAD = sizeof(T) == 8 ? 1 : 2;
    constexpr int LPA = BM + PAD, LPB = BN + PAD;
    constexpr int EA = (BM * BK) / 256, EB = (BN * BK) / 256;   // staged elements per thread
    const int tile = xcd_tile(blockIdx.x, gridDim.x), c = tile / (tx * ty);
    const ChainDev &ch = g.chains[c];
    const int M = dim_of(ch, g.mdim), N = dim_of(ch, g.ndim), K = dim_of(ch, g.kdim);
    const int i0 = (tile % (tx * ty)) / tx * BM, j0 = tile % tx * BN;
    if (i0 >= M || j0 >= N) return;
    const auto A = g.A[c], B = g.B[c];

    __shared__ T As[2][AC ? 2 : 1][BK][LPA];
    __shared__ T Bs[2][BC ? 2 : 1][BK][LPB];
    extern __shared__ __align__(16) char gemm_dyn[];   // k-index factors (only when scaled)
    T *const Sk0 = reinterpret_cast<T *>(gemm_dyn), *const Sk1 = Sk0 + K;
    __shared__ T Sr[2][BM], Sc[2][BN];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const bool scale_k = g.inS.kind != 0;
    // diagonal factors, once per workgroup
    if (scale_k)
        for (int k = tid; k < K; k += 256) {
            double dr, di;
            diag_value(ch, g.inS, k, dr, di);
            Sk0[k] = (T)dr;
            Sk1[k] = (T)di;
        }
    if (g.rowS.kind && tid < BM) {
        double dr = 1.0, di = 0.0;
        if (i0 + tid < M) diag_value(ch, g.rowS, i0 + tid, dr, di);
        Sr[0][tid] = (T)dr;
        Sr[1][tid] = (T)di;
    }
    if (g.colS.kind && tid >= 128 && tid < 128 + BN) {
        const int t = tid - 128;
        double dr = 1.0, di = 0.0;
        if (j0 + t < N) diag_value(ch, g.colS, j0 + t, dr, di);
        Sc[0][t] = (T)dr;
        Sc[1][t] = (T)di;
    }

    VA accR[WMB][WNB], accI[WMB][WNB];
#pragma unroll
    for (int a = 0; a < WMB; ++a)
#pragma unroll
        for (int b = 0; b < WNB; ++b) {
            accR[a][b] = VA{0, 0, 0, 0};
            accI[a][b] = VA{0, 0, 0, 0};
        }

    // register staging of one k-tile: element e = tid + 256 q of op(A)[BM x BK], op(B)[BK x BN]
    T ra[EA], rai[EA], rb[EB], rbi[EB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < EA; ++q) {
            const int e = tid + q * 256;
            int ii, kk;
            if (OPA == 0) { ii = e % BM; kk = e / BM; }
            else { kk = e % BK; ii = e / BK; }
            const int gi = i0 + ii, gk = k0 + kk;
            T vr = 0, vi = 0;
            if (gi < M && gk < K) {
                const size_t off = OPA == 0 ? (size_t)gi + (size_t)gk * A.ld : (size_t)gk + (size_t)gi * A.ld;
                vr = A.re[off];
                if (AC) {
                    vi = A.im[off];
                    if (OPA == 2) vi = -vi;
                }
            }
            ra[q] = vr;
            rai[q] = vi;
        }
#pragma unroll
        for (int q = 0; q < EB; ++q) {
            const int e = tid + q * 256;
            int jj, kk;
            if (OPB == 0) { kk = e % BK; jj = e / BK; }
            else { jj = e % BN; kk = e / BN; }
            const int gj = j0 + jj, gk = k0 + kk;
            T vr = 0, vi = 0;
            if (gj < N && gk < K) {
                const size_t off = OPB == 0 ? (size_t)gk + (size_t)gj * B.ld : (size_t)gj + (size_t)gk * B.ld;
                vr = B.re[off];
                if (BC) {
                    vi = B.im[off];
                    if (OPB == 2) vi = -vi;
                }
            }
            rb[q] = vr;
            rbi[q] = vi;
        }
    };
    auto stash = [&](int buf, int k0) {
#pragma unroll
        for (int q = 0; q < EA; ++q) {
            const int e = tid + q * 256;
            int ii, kk;
            if (OPA == 0) { ii = e % BM; kk = e / BM; }
            else { kk = e % BK; ii = e / BK; }
            T vr = ra[q], vi = rai[q];
            if (AC && scale_k && k0 + kk < K) cmulT(vr, vi, Sk0[k0 + kk], Sk1[k0 + kk], vr, vi);
            As[buf][0][kk][ii] = vr;
            if (AC) As[buf][1][kk][ii] = vi;
        }
#pragma unroll
        for (int q = 0; q < EB; ++q) {
            const int e = tid + q * 256;
            int jj, kk;
            if (OPB == 0) { kk = e % BK; jj = e / BK; }
            else { jj = e % BN; kk = e / BN; }
            T vr = rb[q], vi = rbi[q];
            if (BC && !AC && scale_k && k0 + kk < K) cmulT(vr, vi, Sk0[k0 + kk], Sk1[k0 + kk], vr, vi);
            Bs[buf][0][kk][jj] = vr;
            if (BC) Bs[buf][1][kk][jj] = vi;
        }
    };

    T pcr[PF ? WMB : 1][PF ? WNB : 1][4], pci[PF ? WMB : 1][PF ? WNB : 1][4];
    if constexpr (PF) {
        const auto Ci = g.Cin[c];
#pragma unroll
        for (int mi = 0; mi < WMB; ++mi)
#pragma unroll
            for (int ni = 0; ni < WNB; ++ni)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = i0 + wm * (BM / 2) + mi * 16 + (lane & 15);
                    const int j = j0 + wn * (BN / 2) + ni * 16 + acc_col<T>(lane, v);
                    const bool ok = i < M && j < N;
                    const size_t oi = (size_t)i + (size_t)j * Ci.ld;
                    pcr[mi][ni][v] = ok ? Ci.re[oi] : T(0);
                    pci[mi][ni][v] = (ok && OC) ? Ci.im[oi] : T(0);
                }
    }
    fetch(0);
    __syncthreads();   // diagonal factors visible
    stash(0, 0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += BK) {
        const bool more = k0 + BK < K;
        if (more) fetch(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4), cc = lane & 15;
            T ar[WMB], ai[WMB], br[WNB], bi[WNB];
#pragma unroll
            for (int t = 0; t < WMB; ++t) {
                ar[t] = As[buf][0][kr][wm * (BM / 2) + t * 16 + cc];
                ai[t] = AC ? As[buf][AC ? 1 : 0][kr][wm * (BM / 2) + t * 16 + cc] : T(0);
            }
#pragma unroll
            for (int t = 0; t < WNB; ++t) {
                br[t] = Bs[buf][0][kr][wn * (BN / 2) + t * 16 + cc];
                bi[t] = BC ? Bs[buf][BC ? 1 : 0][kr][wn * (BN / 2) + t * 16 + cc] : T(0);
            }
#pragma unroll
            for (int mi = 0; mi < WMB; ++mi)
#pragma unroll
                for (int ni = 0; ni < WNB; ++ni) {
                    accR[mi][ni] = mma(br[ni], ar[mi], accR[mi][ni]);
                    if (AC && BC) {
                        accR[mi][ni] = mma(-bi[ni], ai[mi], accR[mi][ni]);
                        accI[mi][ni] = mma(bi[ni], ar[mi], accI[mi][ni]);
                        accI[mi][ni] = mma(br[ni], ai[mi], accI[mi][ni]);
                    } else if (BC) {
                        accI[mi][ni] = mma(bi[ni], ar[mi], accI[mi][ni]);
                    } else if (AC) {
                        accI[mi][ni] = mma(br[ni], ai[mi], accI[mi][ni]);
                    }
                }
        }
        if (more) stash(buf ^ 1, k0 + BK);
        __syncthreads();
        buf ^= 1;
    }

    // ---- epilogue. Lane l, register v holds D[acc_col(l, v)][l%16] of the 16x16 block;
    // D = op(B)ᵀ op(A)ᵀ here, so the lane writes C[i][j] with i = l%16 and j = acc_col(l, v):
    // 16 consecutive rows per store.
    const auto Cm = g.C[c];
    typename OpndOf<T>::type Ci{};
    if (g.mode == 2) Ci = g.Cin[c];
#pragma unroll
    for (int mi = 0; mi < WMB; ++mi) {
        const int il = wm * (BM / 2) + mi * 16 + (lane & 15), i = i0 + il;
        if (i >= M) continue;
        T rsr = 1, rsi = 0;
        if (g.rowS.kind) {
            rsr = Sr[0][il];
            rsi = Sr[1][il];
        }
#pragma unroll
        for (int ni = 0; ni < WNB; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int jl = wn * (BN / 2) + ni * 16 + acc_col<T>(lane, v), j = j0 + jl;
                if (j >= N) continue;
                T cr = accR[mi][ni][v], ci = OC ? accI[mi][ni][v] : T(0);
                if (g.rowS.kind) cmulT(cr, ci, rsr, rsi, cr, ci);
                if (g.colS.kind) cmulT(cr, ci, Sc[0][jl], Sc[1][jl], cr, ci);
                const size_t off = (size_t)i + (size_t)j * Cm.ld;
                if (g.mode == 1) {
                    cr = (i == j ? T(1) : T(0)) - cr;
                    ci = -ci;
                } else if (g.mode == 2) {
                    if constexpr (PF) {
                        cr = pcr[mi][ni][v] - cr;
                        if (OC) ci = pci[mi][ni][v] - ci;
                    } else {
                        const size_t oi = (size_t)i + (size_t)j * Ci.ld;
                        cr = Ci.re[oi] - cr;
                        if (OC) ci = Ci.im[oi] - ci;
                    }
                } else if (g.mode == 3) {
                    cr = -cr;
                    ci = -ci;
                }
                Cm.re[off] = cr;
                if (OC) Cm.im[off] = ci;
            }
    }
}

template <class T, int AC, int BC, int OPA, int OPB, bool PF = false>
hipError_t run(hipStream_t st, int nchains, int max_m, int max_n, int max_k, const GemmArgsT<T> &g, int tile) {
    if (g.inS.kind && max_k > KMAX) return hipErrorInvalidValue;
    const size_t dyn = g.inS.kind ? 2 * sizeof(T) * (size_t)max_k : 0;
    // 32 x 32 tiles (more workgroups in flight hide the operand latency better: 64 batched
    // N = 288 wraps run in ~90 µs against ~180 µs with 64 x 64) unless the 32 x 32 grid is
    // very large, or as forced by `tile`
    const long t32 = (long)((max_n + 31) / 32) * ((max_m + 31) / 32) * nchains;
    if (tile == 64 || (tile == 0 && t32 > 32768)) {
        const int tx = (max_n + 63) / 64, ty = (max_m + 63) / 64;
        hipLaunchKernelGGL((k_gemm<T, AC, BC, OPA, OPB, 64, 64, PF>), dim3(tx * ty * nchains), dim3(256), dyn, st, g, tx, ty);
    } else {
        const int tx = (max_n + 31) / 32, ty = (max_m + 31) / 32;
        hipLaunchKernelGGL((k_gemm<T, AC, BC, OPA, OPB, 32, 32, PF>), dim3(tx * ty * nchains), dim3(256), dyn, st, g, tx, ty);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gemm(hipStream_t st, int nchains, int max_m, int max_n, int max_k, bool a_cplx, bool b_cplx,
                       int opa, int opb, const GemmArgs &g, int tile) {
    const int code = (a_cplx ? 1000 : 0) + (b_cplx ? 100 : 0) + opa * 10 + opb;
    switch (code) {
        case 100: return run<double, 0, 1, 0, 0>(st, nchains, max_m, max_n, max_k, g, tile);     // eK · G
        case 110: return run<double, 0, 1, 1, 0>(st, nchains, max_m, max_n, max_k, g, tile);     // eKᵀ · X
        case 1000: return run<double, 1, 0, 0, 0>(st, nchains, max_m, max_n, max_k, g, tile);    // T · eKinv
        case 1100: return run<double, 1, 1, 0, 0>(st, nchains, max_m, max_n, max_k, g, tile);    // U_R · X
        case 1120: return run<double, 1, 1, 2, 0>(st, nchains, max_m, max_n, max_k, g, tile);    // U_Lᴴᴴ · U_R, Uᴴ U
        case 1102: return run<double, 1, 1, 0, 2>(st, nchains, max_m, max_n, max_k, g, tile);
        case 1101: return run<double, 1, 1, 0, 1, true>(st, nchains, max_m, max_n, max_k, g, tile);   // G0 - U Vᵀ (mode 2)
        case 0: return run<double, 0, 0, 0, 0>(st, nchains, max_m, max_n, max_k, g, tile);
        case 10: return run<double, 0, 0, 1, 0>(st, nchains, max_m, max_n, max_k, g, tile);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gemm_f32(hipStream_t st, int nchains, int max_m, int max_n, int max_k, bool a_cplx, bool b_cplx,
                           int opa, int opb, const GemmArgsF &g, int tile) {
    const int code = (a_cplx ? 1000 : 0) + (b_cplx ? 100 : 0) + opa * 10 + opb;
    switch (code) {
        case 100: return run<float, 0, 1, 0, 0>(st, nchains, max_m, max_n, max_k, g, tile);    // eK · G
        case 1000: return run<float, 1, 0, 0, 0>(st, nchains, max_m, max_n, max_k, g, tile);   // T · eKinv
        case 1101: return run<float, 1, 1, 0, 1, true>(st, nchains, max_m, max_n, max_k, g, tile);   // G0 - U Vᵀ (mode 2)
        default: return hipErrorInvalidValue;
    }
}
