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

// Layout sized to the block: KBT in {8, 16, 32, 64} (smallest >= k), NTH threads (KBT² up to
// 1024; 256 at KBT = 32), thread (r = tid % KBT, cg = tid / KBT) owns row r, columns
// cg·CPT .. cg·CPT+CPT-1 (CPT = KBT²/NTH) of both the working submatrix and W. No wave works on
// padding. At k = 32, 256 threads (4 entries each) decide as fast as 1024 and leave most of
// their compute unit to the flush kernels running beside them in the pipelined sweep (345 vs
// 348 ms per sweep of the 64-chain N = 288 batch, same box; 512 threads: no different).
// Everything a proposal needs except the current diagonal (Δ, γ'/γ, Δ/(1+Δ), the uniform,
// the new value) is computed for all k sites in parallel before the sequential loop.
template <int KBT, int NTH = (KBT >= 32 ? 1024 : KBT * KBT)>
__global__ __launch_bounds__(NTH) void k_decide_tu(DecideArgs a) {
    constexpr int CPT = KBT * KBT / NTH;
    const int c = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = ch.ns;
    const int s0 = a.s0;
    const int kb = min(a.kb, n - s0);
    const int tid = threadIdx.x;
    if (kb <= 0) {
        if (tid == 0) {
            ch.acc_cnt[a.slot] = 0;
            if (a.deferred) {
                const int rk = a.blk > 0 ? ch.rk[a.blk] : 0;
                ch.rk[a.blk + 1] = rk;
                *ch.rank = rk;
            }
        }
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
    // the proposals (draws, phases, weight ratios) first: they depend on nothing below, so their
    // latency overlaps the loads of the working submatrix and the pending-block correction
    if (tid < KBT) {
        accl[tid] = 0;   // entries beyond the accepted count stay valid indices
        if (tid < kb) {
            int chv;
            double u;
            const uint32_t sweep = a.sweep_dev ? *a.sweep_dev : a.sweep;
            sweep_draw(ch, sweep, lt, a.dir, (uint32_t)(s0 + tid), chv, u);
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
    if (a.prewrap) {
        // the block of the wrapped G: line ⊙ (T[S,:] E[:,S]) ⊙ out, the product summed over the
        // chunk partials of k_pw_part in chunk order
        const double *P = a.pw_part + (size_t)c * a.pw_nkc * 2048;
        double lr = 1.0, li = 0.0;
        if (r < kb) diag_value(ch, a.pw_line, s0 + r, lr, li);
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int cc = c0 + q;
            double sr = 0.0, si = 0.0;
            if (r < 32 && cc < 32)
                for (int kc = 0; kc < a.pw_nkc; ++kc) {
                    sr += P[(size_t)kc * 2048 + r * 32 + cc];
                    si += P[(size_t)kc * 2048 + 1024 + r * 32 + cc];
                }
            double orr = 1.0, oi = 0.0;
            if (cc < kb) diag_value(ch, a.pw_out, s0 + cc, orr, oi);
            double tr_, ti_;
            cmul(lr, li, sr, si, tr_, ti_);
            double vr, vi;
            cmul(tr_, ti_, orr, oi, vr, vi);
            const bool ok = r < kb && cc < kb;
            wr[q] = ok ? vr : 0.0;
            wi[q] = ok ? vi : 0.0;
            Wr[q] = 0.0;
            Wi[q] = 0.0;
        }
    } else
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
    // a rank-rk product not yet applied to G (deferred flush, or the previous block's flush in
    // the pipelined sweep): the block works on G[S,S] - U[S,:] V[:,S]
    if ((a.deferred && a.blk > 0) || a.pre) {
        constexpr int DCH = KBT >= 32 ? 32 : 16;
        __shared__ double Usr[KBT][DCH + 1], Usi[KBT][DCH + 1], Vsr[DCH][KBT + 1], Vsi[DCH][KBT + 1];
        __shared__ double Upr[KBT][DCH + 1], Upi[KBT][DCH + 1];   // pipelined: G[S, A] chunk 0, prefetched
        const int ps = 1 - a.slot;
        const int rk = a.pre ? ch.acc_cnt[ps] : ch.rk[a.blk];
        const int *Ap = ch.acc_idx + ps * KB;
        const double *Wp = ch.W + (size_t)ps * 2 * KB * KB;
        auto g_at = [&](int i, int j, double &vr, double &vi) {   // the current G, either precision
            if (a.Gf) {
                const OpndF Gf = a.Gf[c];
                const size_t o = (size_t)i + (size_t)j * Gf.ld;
                vr = Gf.re[o];
                vi = Gf.im[o];
            } else {
                const size_t o = (size_t)i + (size_t)j * G.ld;
                vr = G.re[o];
                vi = G.im[o];
            }
        };
        // pipelined sweep, phase 1: V = -W Y (rk x kb) with Y = G[A, S], in registers of the
        // entry layout (thread (r, c0 + q) holds V[r][c0 + q]), over chunks of DCH of the
        // contraction index staged in LDS (W columns in Usr, Y rows in Vsr)
        double pvr[CPT], pvi[CPT];
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            pvr[q] = 0.0;
            pvi[q] = 0.0;
        }
        if (a.pre) {
            for (int u0 = 0; u0 < rk; u0 += DCH) {
                const int uc = min(DCH, rk - u0);
                for (int e = tid; e < KBT * DCH; e += NTH) {
                    const int t = e % KBT, u = e / KBT;
                    const bool ok = t < rk && u < uc;
                    Usr[t][u] = ok ? Wp[t + (u0 + u) * KB] : 0.0;
                    Usi[t][u] = ok ? Wp[KB * KB + t + (u0 + u) * KB] : 0.0;
                }
                for (int e = tid; e < DCH * KBT; e += NTH) {
                    const int j = e % KBT, u = e / KBT;
                    double yr = 0.0, yi = 0.0;
                    if (j < kb && u < uc) g_at(Ap[u0 + u], s0 + j, yr, yi);
                    Vsr[u][j] = yr;
                    Vsi[u][j] = yi;
                }
                if (u0 == 0)   // the first chunk of U = G[S, A] for the second phase, in the same round trip
                    for (int e = tid; e < KBT * DCH; e += NTH) {
                        const int i = e % KBT, t = e / KBT;
                        double ur = 0.0, ui = 0.0;
                        if (i < kb && t < min(DCH, rk)) g_at(s0 + i, Ap[t], ur, ui);
                        Upr[i][t] = ur;
                        Upi[i][t] = ui;
                    }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < CPT; ++q)
                    for (int u = 0; u < uc; ++u) {
                        const double wr_ = Usr[r][u], wi_ = Usi[r][u], yr = Vsr[u][c0 + q], yi = Vsi[u][c0 + q];
                        pvr[q] -= wr_ * yr - wi_ * yi;
                        pvi[q] -= wr_ * yi + wi_ * yr;
                    }
                __syncthreads();
            }
        }
        // subtract U[S, :rk] V[:rk, S] over chunks of DCH ranks staged in LDS. Deferred flush:
        // the slice's factors (V transposed, V[t, j] at j + t·ld). Pipelined sweep: the previous
        // block's update G - G[:,A] W (E_A - G[A,:]) with its accepted sites A and W (slot
        // 1 - slot), so U = G[S, A] and V = -W G[A, S] from phase 1 (E_A vanishes on S)
        for (int t0 = 0; t0 < rk; t0 += DCH) {
            const int tc = min(DCH, rk - t0);
            const bool pref = a.pre && t0 == 0;   // staged with the first phase
            double (*UR)[DCH + 1] = pref ? Upr : Usr, (*UI)[DCH + 1] = pref ? Upi : Usi;
            if (!pref)
            for (int e = tid; e < KBT * DCH; e += NTH) {
                const int i = e % KBT, t = e / KBT;
                double ur = 0.0, ui = 0.0;
                if (i < kb && t < tc) {
                    if (a.pre) {
                        g_at(s0 + i, Ap[t0 + t], ur, ui);
                    } else if (a.Uf) {
                        const OpndF U = a.Uf[c];
                        const size_t o = (size_t)(s0 + i) + (size_t)(t0 + t) * U.ld;
                        ur = U.re[o];
                        ui = U.im[o];
                    } else {
                        const Opnd U = a.U[c];
                        const size_t o = (size_t)(s0 + i) + (size_t)(t0 + t) * U.ld;
                        ur = U.re[o];
                        ui = U.im[o];
                    }
                }
                Usr[i][t] = ur;
                Usi[i][t] = ui;
            }
            if (a.pre) {   // rows t0 .. t0 + tc of V from the threads that hold them
                if (r >= t0 && r < t0 + DCH)
#pragma unroll
                    for (int q = 0; q < CPT; ++q) {
                        const bool ok = r < t0 + tc && c0 + q < kb;
                        Vsr[r - t0][c0 + q] = ok ? pvr[q] : 0.0;
                        Vsi[r - t0][c0 + q] = ok ? pvi[q] : 0.0;
                    }
            } else {
                for (int e = tid; e < DCH * KBT; e += NTH) {
                    const int j = e % KBT, t = e / KBT;
                    double vr = 0.0, vi = 0.0;
                    if (j < kb && t < tc) {
                        if (a.Vf) {
                            const OpndF V = a.Vf[c];
                            const size_t o = (size_t)(s0 + j) + (size_t)(t0 + t) * V.ld;
                            vr = V.re[o];
                            vi = V.im[o];
                        } else {
                            const Opnd V = a.V[c];
                            const size_t o = (size_t)(s0 + j) + (size_t)(t0 + t) * V.ld;
                            vr = V.re[o];
                            vi = V.im[o];
                        }
                    }
                    Vsr[t][j] = vr;
                    Vsi[t][j] = vi;
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int cc = c0 + q;
                double sr = 0.0, si = 0.0;
                for (int t = 0; t < tc; ++t) {
                    sr += UR[r][t] * Vsr[t][cc] - UI[r][t] * Vsi[t][cc];
                    si += UR[r][t] * Vsi[t][cc] + UI[r][t] * Vsr[t][cc];
                }
                wr[q] -= sr;
                wi[q] -= si;
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int q = 0; q < CPT; ++q)
        if (r < kb && r == c0 + q) {
            dgr[r] = wr[q];
            dgi[r] = wi[q];
        }
    __syncthreads();

    int na = 0, nties = 0;
    for (int m = 0; m < kb; ++m) {
        const double Dr = dR[m], Di = dI[m];
        // p = 1 + Δ (1 - G_mm)
        const double omr = 1.0 - dgr[m], omi = -dgi[m];
        double pr, pi;
        cmul(Dr, Di, omr, omi, pr, pi);
        pr += 1.0;
        const double p2 = pr * pr + pi * pi;
        const double prob = p2 * pg[m];
        nties += fabs(us[m] - prob) < KDQ_NEAR_TIE;
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
        ch.acc_cnt[a.slot] = na;
        ch.counters[0] += kb;
        ch.counters[1] += na;
        ch.counters[2] += nties;
        if (a.deferred) {
            const int rk = (a.blk > 0 ? ch.rk[a.blk] : 0) + na;
            ch.rk[a.blk + 1] = rk;
            *ch.rank = rk;
        }
    }
    if (tid < na) ch.acc_idx[a.slot * KB + tid] = s0 + accl[tid];
    if (r < na) {
        double *Wo = ch.W + (size_t)a.slot * 2 * KB * KB;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int u = c0 + q;
            if (u < na) {
                Wo[r + u * KB] = Wr[q];
                Wo[KB * KB + r + u * KB] = Wi[q];
            }
        }
    }
    if (a.prewrap) {   // the wrap's second product may still read this slice's fields
        if (tid < KBT) ch.pw_fld[tid] = tid < kb ? sv[tid] : 0;
    } else if (tid < kb) {
        const int v = sv[tid];
        ch.s[(size_t)(s0 + tid) + (size_t)(lt - 1) * ch.n_elem] = (uint8_t)((v >> 16) ? (v >> 8) & 0xff : v & 0xff);
    }
}

// prewrap, first part: per chain (blockIdx.y) and chunk K = [32 kc, 32 kc + 32) of the contraction
// index (blockIdx.x), P[kc] = T[S, K] E[K, S] for the block's sites S = s0 .. s0 + 31 (rows
// and columns beyond the block are 0), both operands staged in LDS with one global round trip;
// thread (r = tid % 32, c0 = 4 (tid / 32)) sums four entries over k in order.
__global__ __launch_bounds__(256) void k_pw_part(DecideArgs a) {
    const int c = blockIdx.y, kc = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = ch.ns, s0 = a.s0, kb = min(a.kb, n - s0), k0 = 32 * kc;
    const int kn = min(32, n - k0);
    double *P = a.pw_part + ((size_t)c * a.pw_nkc + kc) * 2048;
    const int tid = threadIdx.x, r = tid % 32, c0 = (tid / 32) * 4;
    if (kn <= 0 || kb <= 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            P[r * 32 + c0 + q] = 0.0;
            P[1024 + r * 32 + c0 + q] = 0.0;
        }
        return;
    }
    const Opnd T = a.T[c], E = a.E[c];
    const bool ecx = E.im != nullptr;
    __shared__ double Tr[32][33], Ti[32][33], Er[32][33], Ei[32][33];
    for (int e = tid; e < 32 * 32; e += 256) {
        const int i = e % 32, k = e / 32;   // consecutive threads: consecutive rows of a column of T
        const bool ok = i < kb && k < kn;
        const size_t o = (size_t)(s0 + i) + (size_t)(k0 + k) * T.ld;
        Tr[i][k] = ok ? T.re[o] : 0.0;
        Ti[i][k] = ok ? T.im[o] : 0.0;
        const int kk = e % 32, j = e / 32;  // consecutive entries of a column of E
        const bool ok2 = j < kb && kk < kn;
        const size_t o2 = (size_t)(k0 + kk) + (size_t)(s0 + j) * E.ld;
        Er[kk][j] = ok2 ? E.re[o2] : 0.0;
        Ei[kk][j] = ok2 && ecx ? E.im[o2] : 0.0;
    }
    __syncthreads();
    double sr[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < kn; ++k) {
        const double tr_ = Tr[r][k], ti_ = Ti[r][k];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double er = Er[k][c0 + q], ei = Ei[k][c0 + q];
            sr[q] = fma(tr_, er, fma(-ti_, ei, sr[q]));
            si[q] = fma(tr_, ei, fma(ti_, er, si[q]));
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        P[r * 32 + c0 + q] = sr[q];
        P[1024 + r * 32 + c0 + q] = si[q];
    }
}

// Xout[I,J] = Xin[I,J] - sgn · Cin[I,A] · (W · Y[:,J]),  Y[t, j] = edge · δ(A_t, j) - sgn · Rin[A_t, j]
// (mode 0: sgn = edge = 1; mode 1: sgn = -1, edge = 0, so the product is added).
// Flush, first launch: per chain and chunk of PC columns, Z[:, J] = W · Y[:, J] with
// Y = edge·E_A - sgn·Rin[A, :], and rows J of the gathered columns Cs = sgn·Cin[:, A]. The
// accepted sites of a block are consecutive indices, so the row gather Rin[A, j] reads one or
// two cache lines per column. The launch is latency-bound (three dependent global reads, then
// the product), so chunks are narrow: 18 workgroups per chain at N = 288 (64-column chunks
// measured 431, 32: 412, 16: 407, 8: 413 ms per sweep of the 64-chain N = 288 batch).
constexpr int PC = 16;
template <class T>
__global__ __launch_bounds__(256) void k_flush_prep(FlushArgsT<T> f) {
    const int c = blockIdx.y;
    const ChainDev &ch = f.chains[c];
    const int n = ch.ns, na = ch.acc_cnt[f.slot];
    const int j0 = blockIdx.x * PC;
    if (f.commit_lt > 0 && blockIdx.x == 0 && threadIdx.x < KB) {   // accepted fields of a prewrapped block 0
        const int v = ch.pw_fld[threadIdx.x], s = f.commit_s0 + threadIdx.x;
        if ((v >> 16) && s < n) ch.s[(size_t)s + (size_t)(f.commit_lt - 1) * ch.n_elem] = (uint8_t)((v >> 8) & 0xff);
    }
    if (na == 0 || j0 >= n) return;   // rank 0: the GEMM copies Xin
    const auto Gi = f.Gin[c];
    const auto Ri = f.Rin ? f.Rin[c] : Gi;
    const auto Z = f.Zs[c];
    const T sgn = f.mode == 0 ? T(1) : T(-1);
    const T edge = f.mode == 0 ? T(1) : T(0);
    const int K = f.kmax;
    extern __shared__ __align__(8) unsigned char lds_raw[];
    T *Wr = reinterpret_cast<T *>(lds_raw), *Wi = Wr + K * K;   // W [t][u], stride K
    T *Yr = Wi + K * K, *Yi = Yr + K * PC;                       // Y [u][jj]
    __shared__ int aidx[KB];
    const int tid = threadIdx.x;
    const double *Wg = ch.W + (size_t)f.slot * 2 * KB * KB;   // the block's W and accepted sites
    if (tid < na) aidx[tid] = ch.acc_idx[f.slot * KB + tid];
    for (int e = tid; e < na * na; e += 256) {
        const int t = e % na, u = e / na;
        Wr[t * K + u] = (T)Wg[t + u * KB];
        Wi[t * K + u] = (T)Wg[KB * KB + t + u * KB];
    }
    __syncthreads();
    for (int e = tid; e < na * PC; e += 256) {
        const int jj = e % PC, u = e / PC, j = j0 + jj;
        T yr = 0, yi = 0;
        if (j < n) {
            const size_t off = (size_t)aidx[u] + (size_t)j * Ri.ld;
            yr = (aidx[u] == j ? edge : T(0)) - sgn * Ri.re[off];
            yi = -sgn * Ri.im[off];
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
            sr = fma(ar, br, fma(-ai, bi, sr));   // explicit order: the one-launch form repeats it
            si = fma(ar, bi, fma(ai, br, si));
        }
        const size_t oz = (size_t)t + (size_t)j * Z.ld;
        Z.re[oz] = sr;
        Z.im[oz] = si;
    }
}

// Flush, second launch: Xout = Xin - Cs · Zs of rank |A|. Tile 32 rows x 64 columns; thread
// (ri = tid % 16, cg = tid / 16) owns rows ri and ri + 16 and columns cg + 16 b (b < 4), so one
// load instruction covers 16 consecutive rows of four columns. Its Xin entries are loaded
// first; Cs[I, t] and Zs[t, J] (16 ranks per chunk) are staged in LDS as interleaved complex
// pairs, so each rank step reads 2 + 4 complex values (16-byte LDS reads) for 8 complex MACs.
// Workgroups are dealt to the 8 XCDs round-robin by linear id; with xcd_map set (batches of at
// least 8 chains) every tile of chain c runs on XCD c % 8, so each chain's Cs / Zs operands are
// fetched into one L2 only.
constexpr int AM = 32, AN = 64, AK = 16;
constexpr int NXCD = 8;
template <class T> struct Cplx2;
template <> struct Cplx2<double> { typedef double2 type; };
template <> struct Cplx2<float> { typedef float2 type; };
template <class T>
__global__ __launch_bounds__(256) void k_flush_apply(FlushArgsT<T> f, int nch, int tx, int ty, int xcd_map) {
    typedef typename Cplx2<T>::type C2;
    int c, bx, by;
    if (xcd_map) {
        const int L = blockIdx.x, x = L % NXCD, k = L / NXCD, tiles = tx * ty;
        c = x + NXCD * (k / tiles);
        bx = (k % tiles) % tx;
        by = (k % tiles) / tx;
        if (c >= nch) return;
    } else {
        c = blockIdx.z;
        bx = blockIdx.x;
        by = blockIdx.y;
    }
    const ChainDev &ch = f.chains[c];
    const int n = ch.ns;
    const int i0 = by * AM, j0 = bx * AN;
    if (i0 >= n || j0 >= n) return;
    const int tid = threadIdx.x, ri = tid & 15, cg = tid >> 4;
    const auto Gi = f.Gin[c], Go = f.Gout[c], Z = f.Zs[c];
    const auto Ci = f.Cin ? f.Cin[c] : Gi;
    const T sgn = f.mode == 0 ? T(1) : T(-1);
    const int *aidx = ch.acc_idx + f.slot * KB;   // entries at or beyond the count hold valid indices
    __shared__ C2 Cs2[AK][AM];
    __shared__ C2 Zs2[AK][AN + 1];   // padded: the staging stores run down a column (t)
    constexpr int RA = AM / 16, RB = AN / 16;
    T xr[RA][RB], xi[RA][RB], ar[RA][RB], ai[RA][RB];
#pragma unroll
    for (int a = 0; a < RA; ++a)
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int i = i0 + ri + 16 * a, j = j0 + cg + 16 * b;
            const bool ok = i < n && j < n;
            const size_t oi = (size_t)i + (size_t)j * Gi.ld;
            xr[a][b] = ok ? Gi.re[oi] : T(0);
            xi[a][b] = ok ? Gi.im[oi] : T(0);
            ar[a][b] = T(0);
            ai[a][b] = T(0);
        }
    const int na = ch.acc_cnt[f.slot];
    for (int t0 = 0; t0 < na || t0 == 0; t0 += AK) {
        // the loads do not wait for the count: ranks t0 .. t0 + 15 are inside the 64-rank
        // scratch whatever it is, and entries at or beyond the count are zeroed by a select
        C2 cst[(AK * AM) / 256], zst[(AK * AN) / 256];
#pragma unroll
        for (int q = 0; q < (AK * AM) / 256; ++q) {   // Cs = sgn · Cin[:, A], gathered here
            const int e = tid + 256 * q, r = e & (AM - 1), t = e / AM;
            const int ir = min(i0 + r, n - 1);
            const size_t oc = (size_t)ir + (size_t)aidx[t0 + t] * Ci.ld;
            cst[q].x = sgn * Ci.re[oc];
            cst[q].y = sgn * Ci.im[oc];
        }
#pragma unroll
        for (int q = 0; q < (AK * AN) / 256; ++q) {
            const int e = tid + 256 * q, t = e % AK, jj = e / AK;
            const int jc = min(j0 + jj, n - 1);
            const size_t oz = (size_t)(t0 + t) + (size_t)jc * Z.ld;
            zst[q].x = Z.re[oz];
            zst[q].y = Z.im[oz];
        }
        const int tk = min(AK, na - t0);
#pragma unroll
        for (int q = 0; q < (AK * AM) / 256; ++q) {
            const int e = tid + 256 * q, r = e & (AM - 1), t = e / AM;
            const bool ok = t < tk && i0 + r < n;
            C2 v;
            v.x = ok ? cst[q].x : T(0);
            v.y = ok ? cst[q].y : T(0);
            Cs2[t][r] = v;
        }
#pragma unroll
        for (int q = 0; q < (AK * AN) / 256; ++q) {
            const int e = tid + 256 * q, t = e % AK, jj = e / AK;
            const bool ok = t < tk && j0 + jj < n;
            C2 v;
            v.x = ok ? zst[q].x : T(0);
            v.y = ok ? zst[q].y : T(0);
            Zs2[t][jj] = v;
        }
        __syncthreads();
        for (int t = 0; t < tk; ++t) {
            C2 cv[RA], zv[RB];
#pragma unroll
            for (int a = 0; a < RA; ++a) cv[a] = Cs2[t][ri + 16 * a];
#pragma unroll
            for (int b = 0; b < RB; ++b) zv[b] = Zs2[t][cg + 16 * b];
#pragma unroll
            for (int a = 0; a < RA; ++a)
#pragma unroll
                for (int b = 0; b < RB; ++b) {
                    ar[a][b] = fma(cv[a].x, zv[b].x, fma(-cv[a].y, zv[b].y, ar[a][b]));
                    ai[a][b] = fma(cv[a].x, zv[b].y, fma(cv[a].y, zv[b].x, ai[a][b]));
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < RA; ++a)
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int i = i0 + ri + 16 * a, j = j0 + cg + 16 * b;
            if (i < n && j < n) {
                const size_t oo = (size_t)i + (size_t)j * Go.ld;
                Go.re[oo] = xr[a][b] - ar[a][b];
                Go.im[oo] = xi[a][b] - ai[a][b];
            }
        }
}

// Flush in one launch: workgroup (chain c, strip sx) owns columns J = [32 sx, 32 sx + 32).
//  1. Y[u, J] = edge·δ(A_u, J) - sgn·Rin[A_u, J] (the strip's entries of the accepted rows)
//     into LDS, then Z[:, J] = sgn · W · Y[:, J] (thread: one column, ranks t = t_g + 8q),
//     with W staged in LDS 16 columns at a time.
//  2. The strip's rows in chunks of SRM = 96: thread (ri = tid % 32, cg = tid / 32) owns rows
//     ri + 32a (a < 3) and columns cg + 8b (b < 4); the chunk's entries of the gathered
//     columns Cin[I, A] pass through LDS 16 ranks at a time, each stage loaded into registers
//     while the previous one is used, and Xout = Xin - Cin[I,A]·Z.
// The sign sits on Z here and on the gathered columns in the two-launch form: multiplying by
// ±1 is exact, and both forms run the same fma sequences, so they give the same bits. No
// scratch operands and no second launch.
constexpr int SW = 32, SRA = 3, SRM = 32 * SRA, SCP = AK * SRM / 256;   // SCP: staged C entries per thread
template <class T>
size_t strip_lds_bytes(int K) {
    typedef typename Cplx2<T>::type C2;
    const size_t z = (size_t)K * (SW + 1), y = (size_t)K * (SW + 1), u = y > (size_t)AK * SRM ? y : (size_t)AK * SRM;
    return (z + u) * sizeof(C2);
}
template <class T>
__global__ __launch_bounds__(256, 3) void k_flush_strip(FlushArgsT<T> f, int nch, int strips, int rparts, int rpw,
                                                        int xcd_map) {
    typedef typename Cplx2<T>::type C2;
    int c, tile;
    if (xcd_map) {   // every tile of chain c on XCD c % 8 (one L2 holds the chain's operands)
        const int L = blockIdx.x, x = L % NXCD, k = L / NXCD;
        c = x + NXCD * (k / (strips * rparts));
        tile = k % (strips * rparts);
        if (c >= nch) return;
    } else {
        c = blockIdx.y;
        tile = blockIdx.x;
    }
    const int sx = tile % strips, rp = tile / strips;
    const ChainDev &ch = f.chains[c];
    const int n = ch.ns, j0 = sx * SW, rbeg = rp * rpw, rend = min(n, rbeg + rpw);
    if (j0 >= n || rbeg >= n) return;   // uniform per workgroup
    const auto Gi = f.Gin[c], Go = f.Gout[c];
    const auto Ci = f.Cin ? f.Cin[c] : Gi, Ri = f.Rin ? f.Rin[c] : Gi;
    const int tid = threadIdx.x, ri = tid & 31, cg = tid >> 5;
    T xr[SRA][4], xi[SRA][4];
    auto load_x = [&](int r0) {
#pragma unroll
        for (int a = 0; a < SRA; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int i = r0 + ri + 32 * a, j = j0 + cg + 8 * b;
                const bool ok = i < rend && j < n;
                const size_t oi = (size_t)i + (size_t)j * Gi.ld;
                xr[a][b] = ok ? Gi.re[oi] : T(0);
                xi[a][b] = ok ? Gi.im[oi] : T(0);
            }
    };
    const int K = f.kmax;
    const int na = min(ch.acc_cnt[f.slot], K);   // the decisions accept at most kmax sites
    const T sgn = f.mode == 0 ? T(1) : T(-1);
    const T edge = f.mode == 0 ? T(1) : T(0);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C2 *Zs = reinterpret_cast<C2 *>(lds_raw);   // [t][jj], stride SW + 1 (W chunks [t][u] before)
    C2 *Us = Zs + (size_t)K * (SW + 1);           // Y [u][jj] (stride SW + 1), then C stages [t][r] (stride SRM)
    __shared__ int aidx[KB];
    if (na > 0) {
        if (tid < na) aidx[tid] = ch.acc_idx[f.slot * KB + tid];
        __syncthreads();
        for (int e = tid; e < na * SW; e += 256) {
            const int u = e % na, jj = e / na, j = j0 + jj;   // u fastest: the accepted rows are neighbours
            C2 y;
            y.x = T(0);
            y.y = T(0);
            if (j < n) {
                const size_t off = (size_t)aidx[u] + (size_t)j * Ri.ld;
                y.x = (aidx[u] == j ? edge : T(0)) - sgn * Ri.re[off];
                y.y = -sgn * Ri.im[off];
            }
            Us[u * (SW + 1) + jj] = y;
        }
        const double *Wg = ch.W + (size_t)f.slot * 2 * KB * KB;
        const int jj = tid % SW, tg = tid / SW;
        T sr[KB / 8], si[KB / 8];
#pragma unroll
        for (int q = 0; q < KB / 8; ++q) {
            sr[q] = T(0);
            si[q] = T(0);
        }
        C2 *Ws = Zs;   // W[t, u0 + uu] at Ws[t * 17 + uu] (padded: the stores run down a column)
        for (int u0 = 0; u0 < na; u0 += 16) {
            const int uc = min(16, na - u0);
            __syncthreads();   // Y stored (first chunk); the previous chunk's reads done
            for (int e = tid; e < na * 16; e += 256) {
                const int t = e % na, uu = e / na;
                C2 w;
                w.x = T(0);
                w.y = T(0);
                if (uu < uc) {
                    w.x = (T)Wg[t + (u0 + uu) * KB];
                    w.y = (T)Wg[KB * KB + t + (u0 + uu) * KB];
                }
                Ws[t * 17 + uu] = w;
            }
            __syncthreads();
            for (int uu = 0; uu < uc; ++uu) {
                const C2 y = Us[(u0 + uu) * (SW + 1) + jj];
#pragma unroll
                for (int q = 0; q < KB / 8; ++q) {
                    const int t = tg + 8 * q;
                    if (t < na) {
                        const C2 w = Ws[t * 17 + uu];
                        sr[q] = fma(w.x, y.x, fma(-w.y, y.y, sr[q]));
                        si[q] = fma(w.x, y.y, fma(w.y, y.x, si[q]));
                    }
                }
            }
        }
        __syncthreads();   // W chunks and Y are dead
#pragma unroll
        for (int q = 0; q < KB / 8; ++q) {
            const int t = tg + 8 * q;
            if (t < na) {
                C2 z;
                z.x = sgn * sr[q];
                z.y = sgn * si[q];
                Zs[t * (SW + 1) + jj] = z;
            }
        }
    }
    // stages (row chunk r0, ranks t0 .. t0 + 15) in order; stage s+1 is loaded into registers
    // while stage s is used
    const int nrk = (na + AK - 1) / AK;
    C2 cst[SCP];
    auto load_c = [&](int r0, int t0) {
        const int tk = min(AK, na - t0);
#pragma unroll
        for (int q = 0; q < SCP; ++q) {
            const int e = tid + 256 * q, r = e % SRM, t = e / SRM;
            cst[q].x = T(0);
            cst[q].y = T(0);
            if (t < tk && r0 + r < n) {
                const size_t oc = (size_t)(r0 + r) + (size_t)aidx[t0 + t] * Ci.ld;
                cst[q].x = Ci.re[oc];
                cst[q].y = Ci.im[oc];
            }
        }
    };
    if (nrk > 0) load_c(rbeg, 0);
    for (int r0 = rbeg; r0 < rend; r0 += SRM) {
        T ar[SRA][4], ai[SRA][4];
#pragma unroll
        for (int a = 0; a < SRA; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                ar[a][b] = T(0);
                ai[a][b] = T(0);
            }
        for (int k = 0; k < nrk; ++k) {
            const int t0 = k * AK, tk = min(AK, na - t0);
            __syncthreads();   // Z stored / the previous stage's reads done
#pragma unroll
            for (int q = 0; q < SCP; ++q) Us[tid + 256 * q] = cst[q];
            __syncthreads();
            if (k + 1 < nrk) load_c(r0, t0 + AK);
            else if (r0 + SRM < rend) load_c(r0 + SRM, 0);
            for (int t = 0; t < tk; ++t) {
                C2 cv[SRA], zv[4];
#pragma unroll
                for (int a = 0; a < SRA; ++a) cv[a] = Us[t * SRM + ri + 32 * a];
#pragma unroll
                for (int b = 0; b < 4; ++b) zv[b] = Zs[(t0 + t) * (SW + 1) + cg + 8 * b];
#pragma unroll
                for (int a = 0; a < SRA; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        ar[a][b] = fma(cv[a].x, zv[b].x, fma(-cv[a].y, zv[b].y, ar[a][b]));
                        ai[a][b] = fma(cv[a].x, zv[b].y, fma(cv[a].y, zv[b].x, ai[a][b]));
                    }
            }
        }
        load_x(r0);   // after the products: the registers go to the accumulators and stages
#pragma unroll
        for (int a = 0; a < SRA; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int i = r0 + ri + 32 * a, j = j0 + cg + 8 * b;
                if (i < rend && j < n) {
                    const size_t oo = (size_t)i + (size_t)j * Go.ld;
                    Go.re[oo] = xr[a][b] - ar[a][b];
                    Go.im[oo] = xi[a][b] - ai[a][b];
                }
            }
    }
}

// Deferred flush, per block: chunk of XR rows (for X) and XR columns (for T = G[A, J]) of one
// chain per workgroup; the rank sum over the factors runs in chunks of XCH staged in LDS. After
// it, T and W share the staging memory and Y = W (E_A - T) is formed for the same columns.
constexpr int XR = 32, XCH = 16;
template <class T, int KBT>
size_t defer_lds_elems() {
    const size_t stage = 2 * ((size_t)XR * (XCH + 1) + XCH * (KBT + 1) + KBT * (XCH + 1) + XCH * (XR + 1));
    const size_t post = 2 * ((size_t)KBT * (XR + 1) + KBT * (KBT + 1));
    return stage > post ? stage : post;
}
template <class T, int KBT>
__global__ __launch_bounds__(256) void k_defer_xy(DeferArgsT<T> a) {
    constexpr int NQ = XR * KBT / 256;   // X (and T) entries per thread
    const int c = blockIdx.y;
    const ChainDev &ch = a.chains[c];
    const int n = ch.ns, i0 = blockIdx.x * XR;
    if (i0 >= n) return;
    const int na = *ch.acc_cnt;
    if (na == 0) return;
    const int rk = a.blk > 0 ? ch.rk[a.blk] : 0;
    const auto G = a.G0[c], U = a.U[c], V = a.V[c];
    extern __shared__ __align__(8) unsigned char lds_raw[];
    T *sm = reinterpret_cast<T *>(lds_raw);
    T *Uir = sm, *Uii = Uir + XR * (XCH + 1);                   // U[I, t]      [XR][XCH+1]
    T *VAr = Uii + XR * (XCH + 1), *VAi = VAr + XCH * (KBT + 1);  // V[t, A]      [XCH][KBT+1]
    T *UAr = VAi + XCH * (KBT + 1), *UAi = UAr + KBT * (XCH + 1); // U[A, t]      [KBT][XCH+1]
    T *Vjr = UAi + KBT * (XCH + 1), *Vji = Vjr + XCH * (XR + 1);  // V[t, J]      [XCH][XR+1]
    T *Tsr = sm, *Tsi = Tsr + KBT * (XR + 1);                     // T[A, J]      [KBT][XR+1]
    T *Wsr = Tsi + KBT * (XR + 1), *Wsi = Wsr + KBT * (KBT + 1);  // W            [KBT][KBT+1]
    __shared__ int aidx[KBT];
    const int tid = threadIdx.x;
    if (tid < na) aidx[tid] = ch.acc_idx[tid];
    __syncthreads();
    T xr[NQ], xi[NQ], tr[NQ], ti[NQ];   // X[i0 + ii][aa] and T[aa][i0 + ii] of entry e = tid + 256 q
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = tid + 256 * q, ii = e % XR, aa = e / XR, i = i0 + ii;
        const bool ok = aa < na && i < n;
        const int ac = ok ? aidx[aa] : 0;
        xr[q] = ok ? G.re[(size_t)i + (size_t)ac * G.ld] : T(0);
        xi[q] = ok ? G.im[(size_t)i + (size_t)ac * G.ld] : T(0);
        tr[q] = ok ? G.re[(size_t)ac + (size_t)i * G.ld] : T(0);
        ti[q] = ok ? G.im[(size_t)ac + (size_t)i * G.ld] : T(0);
    }
    for (int t0 = 0; t0 < rk; t0 += XCH) {
        const int tc = min(XCH, rk - t0);
        for (int e = tid; e < XR * XCH; e += 256) {
            const int ii = e % XR, t = e / XR;
            const bool ok = i0 + ii < n && t < tc;
            const size_t o = (size_t)(i0 + ii) + (size_t)(t0 + t) * U.ld;
            Uir[ii * (XCH + 1) + t] = ok ? U.re[o] : T(0);
            Uii[ii * (XCH + 1) + t] = ok ? U.im[o] : T(0);
        }
        for (int e = tid; e < XCH * KBT; e += 256) {
            const int aa = e % KBT, t = e / KBT;
            const bool ok = aa < na && t < tc;
            const size_t o = (size_t)(ok ? aidx[aa] : 0) + (size_t)(t0 + t) * V.ld;
            VAr[t * (KBT + 1) + aa] = ok ? V.re[o] : T(0);
            VAi[t * (KBT + 1) + aa] = ok ? V.im[o] : T(0);
        }
        for (int e = tid; e < KBT * XCH; e += 256) {
            const int aa = e % KBT, t = e / KBT;
            const bool ok = aa < na && t < tc;
            const size_t o = (size_t)(ok ? aidx[aa] : 0) + (size_t)(t0 + t) * U.ld;
            UAr[aa * (XCH + 1) + t] = ok ? U.re[o] : T(0);
            UAi[aa * (XCH + 1) + t] = ok ? U.im[o] : T(0);
        }
        for (int e = tid; e < XCH * XR; e += 256) {
            const int jj = e % XR, t = e / XR;
            const bool ok = i0 + jj < n && t < tc;
            const size_t o = (size_t)(i0 + jj) + (size_t)(t0 + t) * V.ld;
            Vjr[t * (XR + 1) + jj] = ok ? V.re[o] : T(0);
            Vji[t * (XR + 1) + jj] = ok ? V.im[o] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = tid + 256 * q, ii = e % XR, aa = e / XR;
            T sxr = 0, sxi = 0, str = 0, sti = 0;
            for (int t = 0; t < tc; ++t) {
                const T ur = Uir[ii * (XCH + 1) + t], ui = Uii[ii * (XCH + 1) + t];
                const T vr = VAr[t * (KBT + 1) + aa], vi = VAi[t * (KBT + 1) + aa];
                sxr += ur * vr - ui * vi;
                sxi += ur * vi + ui * vr;
                const T pr = UAr[aa * (XCH + 1) + t], pi = UAi[aa * (XCH + 1) + t];
                const T wr = Vjr[t * (XR + 1) + ii], wi = Vji[t * (XR + 1) + ii];
                str += pr * wr - pi * wi;
                sti += pr * wi + pi * wr;
            }
            xr[q] -= sxr;
            xi[q] -= sxi;
            tr[q] -= str;
            ti[q] -= sti;
        }
        __syncthreads();
    }
    // X -> columns rk + aa of U; T and W -> LDS
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = tid + 256 * q, ii = e % XR, aa = e / XR, i = i0 + ii;
        if (aa < na && i < n) {
            const size_t o = (size_t)i + (size_t)(rk + aa) * U.ld;
            U.re[o] = xr[q];
            U.im[o] = xi[q];
        }
        Tsr[aa * (XR + 1) + ii] = tr[q];
        Tsi[aa * (XR + 1) + ii] = ti[q];
    }
    for (int e = tid; e < na * na; e += 256) {
        const int u = e % na, v = e / na;
        Wsr[u * (KBT + 1) + v] = (T)ch.W[u + v * KB];
        Wsi[u * (KBT + 1) + v] = (T)ch.W[KB * KB + u + v * KB];
    }
    __syncthreads();
    // Y[aa, j] = Σ_b W[aa][b] (δ(A_b, j) - T[b, j]) -> rows rk + aa of V
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = tid + 256 * q, jj = e % XR, aa = e / XR, j = i0 + jj;
        if (aa >= na || j >= n) continue;
        T sr = 0, si = 0;
        for (int b = 0; b < na; ++b) {
            const T wr = Wsr[aa * (KBT + 1) + b], wi = Wsi[aa * (KBT + 1) + b];
            const T er = (aidx[b] == j ? T(1) : T(0)) - Tsr[b * (XR + 1) + jj], ei = -Tsi[b * (XR + 1) + jj];
            sr += wr * er - wi * ei;
            si += wr * ei + wi * er;
        }
        const size_t o = (size_t)j + (size_t)(rk + aa) * V.ld;
        V.re[o] = sr;
        V.im[o] = si;
    }
}

template <class T, int KBT>
hipError_t defer_launch(hipStream_t st, int nchains, int max_ns, const DeferArgsT<T> &a) {
    const size_t bytes = defer_lds_elems<T, KBT>() * sizeof(T);
    if (hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_defer_xy<T, KBT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes))
        return e;
    hipLaunchKernelGGL((k_defer_xy<T, KBT>), dim3((max_ns + XR - 1) / XR, nchains), dim3(256), bytes, st, a);
    return hipGetLastError();
}

template <class T>
hipError_t defer_t(hipStream_t st, int nchains, int max_ns, const DeferArgsT<T> &a) {
    if (a.kmax < 1 || a.kmax > KB) return hipErrorInvalidValue;
    if (a.kmax <= 8) return defer_launch<T, 8>(st, nchains, max_ns, a);
    if (a.kmax <= 16) return defer_launch<T, 16>(st, nchains, max_ns, a);
    if (a.kmax <= 32) return defer_launch<T, 32>(st, nchains, max_ns, a);
    return defer_launch<T, 64>(st, nchains, max_ns, a);
}

}  // namespace

hipError_t launch_decide_tu(hipStream_t st, int nchains, const DecideArgs &a) {
    if (a.kb < 1 || a.kb > KB || a.slot < 0 || a.slot > 1) return hipErrorInvalidValue;
    if (a.kb <= 8) hipLaunchKernelGGL(k_decide_tu<8>, dim3(nchains), dim3(64), 0, st, a);
    else if (a.kb <= 16) hipLaunchKernelGGL(k_decide_tu<16>, dim3(nchains), dim3(256), 0, st, a);
    else if (a.kb <= 32) hipLaunchKernelGGL((k_decide_tu<32, 256>), dim3(nchains), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_decide_tu<64>, dim3(nchains), dim3(1024), 0, st, a);
    return hipGetLastError();
}

template <class T>
hipError_t flush_t(hipStream_t st, int nchains, int max_ns, const FlushArgsT<T> &f) {
    if (f.kmax < 1 || f.kmax > KB || f.slot < 0 || f.slot > 1) return hipErrorInvalidValue;
    if (f.one_launch) {
        const size_t bytes = strip_lds_bytes<T>(f.kmax);
        if (hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_flush_strip<T>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes))
            return e;
        // whole strips: splitting a strip's rows over several workgroups measured slower
        // (48.2 against 47.4 µs per block, 64 chains at N = 288)
        const int strips = (max_ns + SW - 1) / SW, rpw = ((max_ns + SRM - 1) / SRM) * SRM, rparts = 1;
        if (nchains >= NXCD) {
            const int groups = (nchains + NXCD - 1) / NXCD;
            hipLaunchKernelGGL(k_flush_strip<T>, dim3(NXCD * groups * strips * rparts), dim3(256), bytes, st, f, nchains,
                               strips, rparts, rpw, 1);
        } else {
            hipLaunchKernelGGL(k_flush_strip<T>, dim3(strips * rparts, nchains), dim3(256), bytes, st, f, nchains, strips,
                               rparts, rpw, 0);
        }
        return hipGetLastError();
    }
    if (!f.Zs || !f.Cs) return hipErrorInvalidValue;
    const size_t K = (size_t)f.kmax;
    const size_t bytes = (2 * K * K + 2 * K * PC) * sizeof(T);
    if (hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_flush_prep<T>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes))
        return e;
    hipLaunchKernelGGL(k_flush_prep<T>, dim3((max_ns + PC - 1) / PC, nchains), dim3(256), bytes, st, f);
    if (hipError_t e = hipGetLastError()) return e;
    const int tx = (max_ns + AN - 1) / AN, ty = (max_ns + AM - 1) / AM;
    if (nchains >= NXCD) {
        const int groups = (nchains + NXCD - 1) / NXCD;
        hipLaunchKernelGGL(k_flush_apply<T>, dim3(NXCD * groups * tx * ty), dim3(256), 0, st, f, nchains, tx, ty, 1);
    } else {
        hipLaunchKernelGGL(k_flush_apply<T>, dim3(tx, ty, nchains), dim3(256), 0, st, f, nchains, tx, ty, 0);
    }
    return hipGetLastError();
}

hipError_t launch_prewrap_part(hipStream_t st, int nchains, const DecideArgs &a) {
    if (!a.pw_part || a.pw_nkc < 1 || a.kb > 32) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pw_part, dim3(a.pw_nkc, nchains), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_flush(hipStream_t st, int nchains, int max_ns, const FlushArgs &f) {
    return flush_t<double>(st, nchains, max_ns, f);
}

hipError_t launch_flush_f32(hipStream_t st, int nchains, int max_ns, const FlushArgsF &f) {
    return flush_t<float>(st, nchains, max_ns, f);
}

hipError_t launch_defer_xy(hipStream_t st, int nchains, int max_ns, const DeferArgsT<double> &a) {
    return defer_t<double>(st, nchains, max_ns, a);
}

hipError_t launch_defer_xy(hipStream_t st, int nchains, int max_ns, const DeferArgsT<float> &a) {
    return defer_t<float>(st, nchains, max_ns, a);
}
