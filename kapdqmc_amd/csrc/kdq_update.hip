// Local HS-field update of the t-U model as delayed (block) Sherman-Morrison updates.
//
// Reference: UpdatePhyLayer! (src/tU/phy_update.jl:256-266) visits the sites of slice lt
// in order; each proposal needs the current G_ii (get_r!, src/tU/GreenMatrix.jl:49-54)
// and each accepted flip applies the rank-1 update of Gupdate! (src/public/Gupdate.jl:5-11)
//     G <- G - G[:,i] r (e_i - G[i,:]),   r = Δ / (1 + Δ (1 - G_ii)).
// Here the sites are processed in blocks S of k consecutive sites:
//   k_decide_tu (one workgroup per chain) loads G[S,S] into registers, runs the k sequential
//   proposals on that k x k working submatrix (rank-1 updates restricted to S x S) and
//   builds the k_a x k_a factor W with
//       G_new = G - G[:,A] W (E_A - G[A,:]),   W = Γ⁻¹,  Γ = diag(1/Δ_a) + I - G[A,A]
//   (A = accepted sites in acceptance order), which equals the product of the sequential
//   rank-1 updates. W grows by bordering: accepting site m with r = 1/(Γ's Schur complement)
//       W' = [[W + x r y, x r], [r y, r]],   x = W G[A,m],  y = G[m,A] W.
//   Both vectors follow from the current working matrix Gc without any product:
//       Gc[A,m] = (I + G[A,A] W) G[A,m] = (D + I) x   ->  x_t = Δ_t/(1+Δ_t) · Gc[a_t, m]
//       Gc[m,A] = G[m,A] (I - W (I - G[A,A])) = y D  ->  y_u = Δ_u · Gc[m, a_u]
//   (D = diag(1/Δ_a); m is not in A because every site is visited once per slice).
//   k_flush applies the rank-|A| correction to the whole N x N matrix (all workgroups).
#include "kdq_kernels.h"

namespace {

constexpr int KB = 64;   // maximum block size
constexpr int LP = KB + 1;

// Layout sized to the block: KBT in {8, 16, 32, 64} (smallest >= k), NTH = min(1024, KBT²)
// threads, thread (r = tid % KBT, cg = tid / KBT) owns row r, columns cg·CPT .. cg·CPT+CPT-1
// (CPT = KBT²/NTH) of both the working submatrix and W. The accept path is bound by VALU
// issue, so no wave works on padding: at k = 32 each of 1024 threads owns one entry.
// Everything a proposal needs except the current diagonal (Δ, γ'/γ, Δ/(1+Δ), the uniform,
// the new value) is computed for all k sites in parallel before the sequential loop.
template <int KBT>
__global__ __launch_bounds__(KBT >= 32 ? 1024 : KBT * KBT) void k_decide_tu(DecideArgs a) {
    constexpr int NTH = KBT >= 32 ? 1024 : KBT * KBT;
    constexpr int CPT = KBT * KBT / NTH;
    const int c = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = ch.ns;
    const int s0 = a.s0;
    const int kb = min(a.kb, n - s0);
    const int tid = threadIdx.x;
    if (kb <= 0) {
        if (tid == 0) *ch.acc_cnt = 0;
        return;
    }
    __shared__ double colr[KBT], coli[KBT], rowr[KBT], rowi[KBT];   // published column / row
    __shared__ double dgr[KBT], dgi[KBT];                           // current diagonal
    __shared__ double dR[KBT], dI[KBT], pg[KBT], us[KBT];           // per site: Δ, γ'/γ, u
    __shared__ double qxr[KBT], qxi[KBT];                           // per site: Δ/(1+Δ)
    __shared__ int accl[KBT], sv[KBT];

    const Opnd G = a.G[c];
    const int lt = a.lt;
    const int r = tid % KBT, c0 = (tid / KBT) * CPT;
    double wr[CPT], wi[CPT], Wr[CPT], Wi[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int cc = c0 + q;
        const bool ok = r < kb && cc < kb;
        if (a.Gf) {   // FP32 G: the decisions still run in FP64
            const OpndF Gf = a.Gf[c];
            const size_t off = (size_t)(s0 + r) + (size_t)(s0 + cc) * Gf.ld;
            wr[q] = ok ? (double)Gf.re[off] : 0.0;
            wi[q] = ok ? (double)Gf.im[off] : 0.0;
        } else {
            const size_t off = (size_t)(s0 + r) + (size_t)(s0 + cc) * G.ld;
            wr[q] = ok ? G.re[off] : 0.0;
            wi[q] = ok ? G.im[off] : 0.0;
        }
        Wr[q] = 0.0;
        Wi[q] = 0.0;
    }
    const uint32_t tag = (uint32_t)lt | ((uint32_t)a.dir << 31);
    if (tid < KBT) {
        accl[tid] = 0;   // entries beyond the accepted count stay valid indices
        if (tid < kb) {
            int chv;
            double u;
            const uint32_t sweep = a.sweep_dev ? *a.sweep_dev : a.sweep;
            proposal_draw(ch.seed, ch.chain_id, 0, sweep, tag, (uint32_t)(s0 + tid), chv, u);
            const int si = ch.s[(size_t)(s0 + tid) + (size_t)(lt - 1) * ch.n_elem];
            const int sx = chv + 1 + ((chv + 1) >= si ? 1 : 0);
            const double x = ch.alpha[lt - 1] * (ch.eta[sx - 1] - ch.eta[si - 1]);
            double sn, cs;
            sincos(x, &sn, &cs);
            dR[tid] = cs - 1.0;
            dI[tid] = sn;
            pg[tid] = ch.gamma[sx - 1] / ch.gamma[si - 1];
            // Δ/(1+Δ) = 1 - 1/(1+Δ), with 1+Δ = cis(x)
            qxr[tid] = 1.0 - cs;
            qxi[tid] = sn;
            us[tid] = u;
            sv[tid] = si | (sx << 8);
        }
    }
#pragma unroll
    for (int q = 0; q < CPT; ++q)
        if (r < kb && r == c0 + q) {
            dgr[r] = wr[q];
            dgi[r] = wi[q];
        }
    __syncthreads();

    int na = 0;
    for (int m = 0; m < kb; ++m) {
        const double Dr = dR[m], Di = dI[m];
        // p = 1 + Δ (1 - G_mm)
        const double omr = 1.0 - dgr[m], omi = -dgi[m];
        double pr, pi;
        cmul(Dr, Di, omr, omi, pr, pi);
        pr += 1.0;
        const double p2 = pr * pr + pi * pi;
        const double prob = p2 * pg[m];
        if (!(us[m] < prob)) continue;
        // r = Δ / p
        const double ip2 = 1.0 / p2;
        const double rr = (Dr * pr + Di * pi) * ip2, ri = (Di * pr - Dr * pi) * ip2;
        // publish column m and row m of the working submatrix (state before this flip)
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            if (r < kb && m == c0 + q) {
                colr[r] = wr[q];
                coli[r] = wi[q];
            }
            if (r == m) {
                rowr[c0 + q] = wr[q];
                rowi[c0 + q] = wi[q];
            }
        }
        if (tid == 0) {
            accl[na] = m;
            sv[m] |= 1 << 16;   // accepted: the new value is written back
        }
        __syncthreads();
        if (r < kb) {
            // LDS operands first: x_r = Δ/(1+Δ) Gc[a_r, m], y_u = Δ_u Gc[m, a_u]
            const double clr = colr[r], cli = coli[r];
            const int ar_ = accl[r];
            const double cxr = colr[ar_], cxi = coli[ar_], qr = qxr[ar_], qi = qxi[ar_];
            double rwr[CPT], rwi[CPT], ayr[CPT], ayi[CPT];
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                rwr[q] = rowr[c0 + q];
                rwi[q] = rowi[c0 + q];
                const int au = accl[c0 + q];
                cmul(dR[au], dI[au], rowr[au], rowi[au], ayr[q], ayi[q]);
            }
            // rank-1 update of the working submatrix
            double fr, fi;
            cmul(clr, cli, rr, ri, fr, fi);
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const double vr = (c0 + q == m ? 1.0 : 0.0) - rwr[q], vi = -rwi[q];
                double tr, ti;
                cmul(fr, fi, vr, vi, tr, ti);
                wr[q] -= tr;
                wi[q] -= ti;
            }
            // bordering of W: rows t < na get W + x r y and x r in the new column, row na
            // gets r y and r
            const bool old_row = r < na, new_row = r == na;
            double xr, xi, xrr, xri;
            cmul(qr, qi, cxr, cxi, xr, xi);
            cmul(xr, xi, rr, ri, xrr, xri);
            const double lr = old_row ? xrr : rr, li = old_row ? xri : ri;   // left factor
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int u = c0 + q;
                const double yr = (u < na) ? ayr[q] : 1.0, yi = (u < na) ? ayi[q] : 0.0;
                double tr, ti;
                cmul(lr, li, yr, yi, tr, ti);
                const bool keep = old_row && u < na;
                const bool set = (old_row && u == na) || (new_row && u <= na);
                Wr[q] = keep ? Wr[q] + tr : (set ? tr : Wr[q]);
                Wi[q] = keep ? Wi[q] + ti : (set ? ti : Wi[q]);
            }
#pragma unroll
            for (int q = 0; q < CPT; ++q)
                if (r == c0 + q) {
                    dgr[r] = wr[q];
                    dgi[r] = wi[q];
                }
        }
        __syncthreads();
        ++na;
    }
    // outputs
    if (tid == 0) {
        *ch.acc_cnt = na;
        ch.counters[0] += kb;
        ch.counters[1] += na;
    }
    if (tid < na) ch.acc_idx[tid] = s0 + accl[tid];
    if (r < na) {
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int u = c0 + q;
            if (u < na) {
                ch.W[r + u * KB] = Wr[q];
                ch.W[KB * KB + r + u * KB] = Wi[q];
            }
        }
    }
    if (tid < kb) {
        const int v = sv[tid];
        ch.s[(size_t)(s0 + tid) + (size_t)(lt - 1) * ch.n_elem] = (uint8_t)((v >> 16) ? (v >> 8) & 0xff : v & 0xff);
    }
}

// Xout[I,J] = Xin[I,J] - sgn · Cin[I,A] · (W · Y[:,J]),  Y[t, j] = edge · δ(A_t, j) - sgn · Rin[A_t, j]
// (mode 0: sgn = edge = 1; mode 1: sgn = -1, edge = 0, so the product is added).
// Flush, first launch: per chain and chunk of PC columns, Z[:, J] = W · Y[:, J] with
// Y = edge·E_A - sgn·Rin[A, :], and rows J of the gathered columns Cs = sgn·Cin[:, A]. The
// accepted sites of a block are consecutive indices, so the row gather Rin[A, j] reads one or
// two cache lines per column.
constexpr int PC = 64;
template <class T>
__global__ __launch_bounds__(256) void k_flush_prep(FlushArgsT<T> f) {
    const int c = blockIdx.y;
    const ChainDev &ch = f.chains[c];
    const int n = ch.ns, na = *ch.acc_cnt;
    const int j0 = blockIdx.x * PC;
    if (na == 0 || j0 >= n) return;   // rank 0: the GEMM copies Xin
    const auto Gi = f.Gin[c];
    const auto Ci = f.Cin ? f.Cin[c] : Gi, Ri = f.Rin ? f.Rin[c] : Gi;
    const auto Z = f.Zs[c], Cg = f.Cs[c];
    const T sgn = f.mode == 0 ? T(1) : T(-1);
    const T edge = f.mode == 0 ? T(1) : T(0);
    const int K = f.kmax;
    extern __shared__ __align__(8) unsigned char lds_raw[];
    T *Wr = reinterpret_cast<T *>(lds_raw), *Wi = Wr + K * K;   // W [t][u], stride K
    T *Yr = Wi + K * K, *Yi = Yr + K * PC;                       // Y [u][jj]
    __shared__ int aidx[KB];
    const int tid = threadIdx.x;
    if (tid < na) aidx[tid] = ch.acc_idx[tid];
    for (int e = tid; e < na * na; e += 256) {
        const int t = e % na, u = e / na;
        Wr[t * K + u] = (T)ch.W[t + u * KB];
        Wi[t * K + u] = (T)ch.W[KB * KB + t + u * KB];
    }
    __syncthreads();
    for (int e = tid; e < na * PC; e += 256) {
        const int jj = e % PC, u = e / PC, j = j0 + jj;
        T yr = 0, yi = 0;
        if (j < n) {
            const size_t off = (size_t)aidx[u] + (size_t)j * Ri.ld;
            yr = (aidx[u] == j ? edge : T(0)) - sgn * Ri.re[off];
            yi = -sgn * Ri.im[off];
            const size_t oc = (size_t)j + (size_t)aidx[u] * Ci.ld, og = (size_t)j + (size_t)u * Cg.ld;
            Cg.re[og] = sgn * Ci.re[oc];
            Cg.im[og] = sgn * Ci.im[oc];
        }
        Yr[u * PC + jj] = yr;
        Yi[u * PC + jj] = yi;
    }
    __syncthreads();
    for (int e = tid; e < na * PC; e += 256) {
        const int jj = e % PC, t = e / PC, j = j0 + jj;
        if (j >= n) continue;
        T sr = 0, si = 0;
        for (int u = 0; u < na; ++u) {
            const T ar = Wr[t * K + u], ai = Wi[t * K + u];
            const T br = Yr[u * PC + jj], bi = Yi[u * PC + jj];
            sr += ar * br - ai * bi;
            si += ar * bi + ai * br;
        }
        const size_t oz = (size_t)t + (size_t)j * Z.ld;
        Z.re[oz] = sr;
        Z.im[oz] = si;
    }
}

// Flush, second launch: Xout = Xin - Cs · Zs of rank |A|. Tile 32 rows x 64 columns; the
// thread's Xin entries (row ii, columns 8 jg .. 8 jg + 7) are loaded first, together with the
// first chunk of Cs[I, t] and Zs[t, J] (16 ranks per chunk, staged in LDS); the rank loop then
// reads Cs rows (consecutive across the wave) and Zs columns (one address per half-wave) from LDS.
constexpr int AM = 32, AN = 64, AQ = 8, AK = 16;
template <class T>
__global__ __launch_bounds__(256) void k_flush_apply(FlushArgsT<T> f) {
    const int c = blockIdx.z;
    const ChainDev &ch = f.chains[c];
    const int n = ch.ns;
    const int i0 = blockIdx.y * AM, j0 = blockIdx.x * AN;
    if (i0 >= n || j0 >= n) return;
    const int tid = threadIdx.x, ii = tid & (AM - 1), jg = tid >> 5;
    const int i = i0 + ii;
    const auto Gi = f.Gin[c], Go = f.Gout[c], Z = f.Zs[c], Cg = f.Cs[c];
    __shared__ T Csr[AK][AM], Csi[AK][AM];
    __shared__ T Zsr[AK][AN + 1], Zsi[AK][AN + 1];
    T xr[AQ], xi[AQ], ar[AQ], ai[AQ];
#pragma unroll
    for (int q = 0; q < AQ; ++q) {
        const int j = j0 + jg * AQ + q;
        const bool ok = i < n && j < n;
        const size_t oi = (size_t)i + (size_t)j * Gi.ld;
        xr[q] = ok ? Gi.re[oi] : T(0);
        xi[q] = ok ? Gi.im[oi] : T(0);
        ar[q] = T(0);
        ai[q] = T(0);
    }
    const int na = *ch.acc_cnt;
    for (int t0 = 0; t0 < na; t0 += AK) {
        const int tk = min(AK, na - t0);
        // stage: Cs[I, t0 + t] (2 x 16 x 32) and Zs[t0 + t, J] (2 x 16 x 64)
        for (int e = tid; e < AK * AM; e += 256) {
            const int r = e & (AM - 1), t = e / AM;
            T vr = 0, vi = 0;
            if (t < tk && i0 + r < n) {
                const size_t oc = (size_t)(i0 + r) + (size_t)(t0 + t) * Cg.ld;
                vr = Cg.re[oc];
                vi = Cg.im[oc];
            }
            Csr[t][r] = vr;
            Csi[t][r] = vi;
        }
        for (int e = tid; e < AK * AN; e += 256) {
            const int t = e % AK, jj = e / AK;
            T vr = 0, vi = 0;
            if (t < tk && j0 + jj < n) {
                const size_t oz = (size_t)(t0 + t) + (size_t)(j0 + jj) * Z.ld;
                vr = Z.re[oz];
                vi = Z.im[oz];
            }
            Zsr[t][jj] = vr;
            Zsi[t][jj] = vi;
        }
        __syncthreads();
        for (int t = 0; t < tk; ++t) {
            const T cr = Csr[t][ii], ci = Csi[t][ii];
#pragma unroll
            for (int q = 0; q < AQ; ++q) {
                const T zr = Zsr[t][jg * AQ + q], zi = Zsi[t][jg * AQ + q];
                ar[q] = fma(cr, zr, fma(-ci, zi, ar[q]));
                ai[q] = fma(cr, zi, fma(ci, zr, ai[q]));
            }
        }
        __syncthreads();
    }
    if (i >= n) return;
#pragma unroll
    for (int q = 0; q < AQ; ++q) {
        const int j = j0 + jg * AQ + q;
        if (j < n) {
            const size_t oo = (size_t)i + (size_t)j * Go.ld;
            Go.re[oo] = xr[q] - ar[q];
            Go.im[oo] = xi[q] - ai[q];
        }
    }
}

}  // namespace

hipError_t launch_decide_tu(hipStream_t st, int nchains, const DecideArgs &a) {
    if (a.kb < 1 || a.kb > KB) return hipErrorInvalidValue;
    if (a.kb <= 8) hipLaunchKernelGGL(k_decide_tu<8>, dim3(nchains), dim3(64), 0, st, a);
    else if (a.kb <= 16) hipLaunchKernelGGL(k_decide_tu<16>, dim3(nchains), dim3(256), 0, st, a);
    else if (a.kb <= 32) hipLaunchKernelGGL(k_decide_tu<32>, dim3(nchains), dim3(1024), 0, st, a);
    else hipLaunchKernelGGL(k_decide_tu<64>, dim3(nchains), dim3(1024), 0, st, a);
    return hipGetLastError();
}

template <class T>
hipError_t flush_t(hipStream_t st, int nchains, int max_ns, const FlushArgsT<T> &f) {
    if (f.kmax < 1 || f.kmax > KB || !f.Zs || !f.Cs) return hipErrorInvalidValue;
    const size_t K = (size_t)f.kmax;
    const size_t bytes = (2 * K * K + 2 * K * PC) * sizeof(T);
    if (hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_flush_prep<T>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes))
        return e;
    hipLaunchKernelGGL(k_flush_prep<T>, dim3((max_ns + PC - 1) / PC, nchains), dim3(256), bytes, st, f);
    if (hipError_t e = hipGetLastError()) return e;
    hipLaunchKernelGGL(k_flush_apply<T>, dim3((max_ns + AN - 1) / AN, (max_ns + AM - 1) / AM, nchains), dim3(256), 0,
                       st, f);
    return hipGetLastError();
}

hipError_t launch_flush(hipStream_t st, int nchains, int max_ns, const FlushArgs &f) {
    return flush_t<double>(st, nchains, max_ns, f);
}

hipError_t launch_flush_f32(hipStream_t st, int nchains, int max_ns, const FlushArgsF &f) {
    return flush_t<float>(st, nchains, max_ns, f);
}
