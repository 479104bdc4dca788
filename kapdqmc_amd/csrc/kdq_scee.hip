// Swap-replica Rényi-2 (SCEE) decisions of the t-U model.
//
// Reference: UpdateSCEELayer! (src/tU/SCEE.jl:278-324) visits the sites of slice lt in
// order and, per site, replica 1 then replica 2. A proposal of replica r is weighed by
// |1 + Δ(1 - Gt_ii)|² γ'/γ · |Tau_A|^{2λ} |Tau_B|^{2(1-λ)} with the Grover factors of
// get_abTau1!/get_abTau2! (src/public/Gupdate.jl:91-132), which read the *other* replica's
// G(Θ) on the region. An accepted flip updates detg, the inverse Grover matrices
// (GMupdate!, Gupdate.jl:45-51) and the four matrices of the replica (G4update!,
// Gupdate.jl:23-37).
//
// The four matrices are the blocks of Big = [[Gt, Gt0], [G0t, G0]], and G4update! is the
// rank-1 Gupdate! of Big with the unit vector in the first block, so a block of accepted
// sites A of one replica gives, with the same W = (diag(1/Δ_a) + I - Gt[A,A])⁻¹ as the plain
// sweep, Gt -= Gt[:,A] W (E_A - Gt[A,:]), G0t -= G0t[:,A] W (E_A - Gt[A,:]),
// Gt0 += Gt[:,A] W Gt0[A,:], G0 += G0t[:,A] W Gt0[A,:] (k_flush modes 0 and 1).
// k_scee_decide therefore works on a block of KS sites of both replicas: Gt[S,S],
// Gt0[S,X], G0t[X,S] (X = A ∪ B) in LDS, the region blocks M[r][X] of G0 and the inverse
// Grover matrices in global memory, all kept current with the sequential rank-1 updates.
//
// The disorder operator (ctrl_SCDOPicr, src/tU/disorder_operate.jl:1-247) runs the same
// kernel on one replica: UpdateSCDOPLayer! (:225-247) weighs a proposal with |Tau_A|^{2λ}
// |Tau_B|^{2(1-λ)} of get_abTau! for DOPBuffer_ (src/public/Gupdate.jl:70-82),
// a = (1 - e^{iα}) G0t[X, m] r, b = Gt0[m, X] X⁻¹, Tau = 1 + b a, and GMupdate! (:52-58)
// updates X⁻¹ the same way as the inverse Grover matrix; no G0 region blocks are needed.
#include "kdq_kernels.h"

namespace {

constexpr int ST = 512;   // threads of k_scee_decide

KDQ_DEV void block_sum4(double v[4], double *red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
        for (int k = 0; k < 4; ++k) red[wave * 4 + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double s = 0.0;
        for (int w = 0; w < ST / 64; ++w) s += red[w * 4 + k];
        v[k] = s;
    }
    __syncthreads();
}

template <int KS>
__global__ __launch_bounds__(ST) void k_scee_decide(SceeArgs a) {
    const int c = blockIdx.x, nch = a.nch, tid = threadIdx.x;
    const ChainDev &ch0 = a.chains[c];
    const SceeRegion &R = a.reg[c];
    const int n = ch0.ns, s0 = a.s0, lt = a.lt;
    const int kb = min(a.kb, n - s0);
    const int NR = R.dop ? 1 : 2;   // replicas
    if (kb <= 0) {
        if (tid < NR) *a.chains[c + tid * nch].acc_cnt = 0;
        return;
    }
    const int nA = R.nA, nB = R.nB, nX = nA + nB, NX = a.nx_max;
    extern __shared__ double lds[];
    double *gtr = lds, *gti = gtr + 2 * KS * KS;               // Gt[S,S]   [r][i][j]
    double *t0r = gti + 2 * KS * KS, *t0i = t0r + 2 * KS * NX;  // Gt0[S,X]  [r][i][x]
    double *o0r = t0i + 2 * KS * NX, *o0i = o0r + 2 * KS * NX;  // G0t[X,S]  [r][x][i]
    double *var_ = o0i + 2 * KS * NX, *vai = var_ + NX;         // a
    double *vbr = vai + NX, *vbi = vbr + NX;                    // b
    double *vzr = vbi + NX, *vzi = vzr + NX;                    // zN, then gmInv a
    double *cxr = vzi + NX, *cxi = cxr + NX;                    // G0t[X, m]
    double *rxr = cxi + NX, *rxi = rxr + NX;                    // Gt0[m, X]
    double *csr = rxi + NX, *csi = csr + KS;                    // Gt[S, m]
    double *rsr = csi + KS, *rsi = rsr + KS;                    // Gt[m, S]
    double *Wr = rsi + KS, *Wi = Wr + 2 * KS * KS;              // W [r][t][u]
    double *pd = Wi + 2 * KS * KS;                              // per site [6][2][KS]
    double *red = pd + 12 * KS;                                 // ST/64 x 4
    __shared__ int sv[2][KS], accl[2][KS];
#define GT(r, i, j) ((r) * KS * KS + (i) * KS + (j))
#define T0(r, i, x) ((r) * KS * NX + (i) * NX + (x))
#define O0(r, x, i) ((r) * NX * KS + (x) * KS + (i))
#define PD(k, r, q) pd[((k) * 2 + (r)) * KS + (q)]

    const uint32_t tag = (uint32_t)lt | ((uint32_t)a.dir << 31);
    if (tid < NR * KS) {
        const int r = tid / KS, q = tid % KS;
        accl[r][q] = 0;
        if (q < kb) {
            const ChainDev &chr = a.chains[c + r * nch];
            int chv;
            double u;
            proposal_draw(chr.seed, chr.chain_id, chr.stream, a.sweep, tag, (uint32_t)(s0 + q), chv, u);
            const int si = chr.s[(size_t)(s0 + q) + (size_t)(lt - 1) * chr.n_elem];
            const int sx = chv + 1 + ((chv + 1) >= si ? 1 : 0);
            const double x = chr.alpha[lt - 1] * (chr.eta[sx - 1] - chr.eta[si - 1]);
            double sn, cs;
            sincos(x, &sn, &cs);
            PD(0, r, q) = cs - 1.0;   // Δ
            PD(1, r, q) = sn;
            PD(2, r, q) = chr.gamma[sx - 1] / chr.gamma[si - 1];
            PD(3, r, q) = 1.0 - cs;   // Δ/(1+Δ), 1+Δ = cis(x)
            PD(4, r, q) = sn;
            PD(5, r, q) = u;
            sv[r][q] = si | (sx << 8);
        }
    }
    for (int e = tid; e < 2 * KS * KS; e += ST) {
        const int r = e / (KS * KS), i = (e / KS) % KS, j = e % KS;
        double vr = 0.0, vi = 0.0;
        if (i < kb && j < kb && r < NR) {
            const Opnd G = a.Gt[c + r * nch];
            const size_t o = (size_t)(s0 + i) + (size_t)(s0 + j) * G.ld;
            vr = G.re[o];
            vi = G.im[o];
        }
        gtr[e] = vr;
        gti[e] = vi;
        Wr[e] = 0.0;
        Wi[e] = 0.0;
    }
    for (int e = tid; e < 2 * KS * nX; e += ST) {
        const int r = e / (KS * nX), rem = e % (KS * nX), x = rem / KS, i = rem % KS;
        double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
        if (i < kb && r < NR) {
            const Opnd T = a.Gt0[c + r * nch], O = a.G0t[c + r * nch];
            const int sx = R.idx[x];
            const size_t ot = (size_t)(s0 + i) + (size_t)sx * T.ld, oo = (size_t)sx + (size_t)(s0 + i) * O.ld;
            ar = T.re[ot]; ai = T.im[ot];
            br = O.re[oo]; bi = O.im[oo];
        }
        t0r[T0(r, i, x)] = ar; t0i[T0(r, i, x)] = ai;
        o0r[O0(r, x, i)] = br; o0i[O0(r, x, i)] = bi;
    }
    // region blocks of G0 of both replicas (the Grover factors read the other one's)
    const int mA = nA * nA, mAB = mA + nB * nB;
    for (int e = tid; e < (R.dop ? 0 : 2 * mAB); e += ST) {
        const int r = e / mAB, f = e % mAB;
        const int X = f < mA ? 0 : 1, nr = X ? nB : nA, base = X ? nA : 0, g = X ? f - mA : f;
        const int p = g % nr, q = g / nr;
        const Opnd G0 = a.G0[c + r * nch];
        const size_t o = (size_t)R.idx[base + p] + (size_t)R.idx[base + q] * G0.ld;
        double *M = R.M[r][X];
        M[p + (size_t)q * nr] = G0.re[o];
        M[(size_t)nr * nr + p + (size_t)q * nr] = G0.im[o];
    }
    __syncthreads();

    int na[2] = {0, 0}, nties[2] = {0, 0};
    const double car = 1.0 - R.ea[0], cai = -R.ea[1];   // 1 - e^{iα} (disorder operator)
    for (int m = 0; m < kb; ++m) {
        for (int r = 0; r < NR; ++r) {
            const int o = 1 - r;
            const double Dr = PD(0, r, m), Di = PD(1, r, m);
            const double omr = 1.0 - gtr[GT(r, m, m)], omi = -gti[GT(r, m, m)];
            double pr, pi;
            cmul(Dr, Di, omr, omi, pr, pi);
            pr += 1.0;
            const double p2 = pr * pr + pi * pi, ip2 = 1.0 / p2;
            double prob = p2 * PD(2, r, m);
            const double rr = (Dr * pr + Di * pi) * ip2, ri = (Di * pr - Dr * pi) * ip2;
            // Grover factors, phase 1: zN (replica 1) or a (replica 2)
            for (int q = tid; q < nX; q += ST) {
                const int X = q < nA ? 0 : 1, base = X ? nA : 0, nr = X ? nB : nA, ql = q - base;
                const double *Mo = R.M[o][X], *Moi = Mo + (size_t)nr * nr;
                double sr = 0.0, si = 0.0;
                if (R.dop) {
                    // zN = Gt0[m, X], a_q = (1 - e^{iα}) (G0t[X_q, m] r)
                    vzr[q] = t0r[T0(r, m, q)]; vzi[q] = t0i[T0(r, m, q)];
                    double tr_, ti_;
                    cmul(o0r[O0(r, q, m)], o0i[O0(r, q, m)], rr, ri, tr_, ti_);
                    cmul(car, cai, tr_, ti_, var_[q], vai[q]);
                } else if (r == 0) {
                    // zN_q = Σ_k Gt0[m, X_k] (2 Mo[k, q] - δ)
                    for (int k = 0; k < nr; ++k) {
                        const double mr = 2.0 * Mo[k + (size_t)ql * nr] - (k == ql ? 1.0 : 0.0);
                        const double mi = 2.0 * Moi[k + (size_t)ql * nr];
                        const double xr = t0r[T0(r, m, base + k)], xi = t0i[T0(r, m, base + k)];
                        sr += xr * mr - xi * mi;
                        si += xr * mi + xi * mr;
                    }
                    vzr[q] = sr; vzi[q] = si;
                    cmul(o0r[O0(r, q, m)], o0i[O0(r, q, m)], rr, ri, var_[q], vai[q]);
                } else {
                    // a_q = Σ_k (2 Mo[q, k] - δ) G0t[X_k, m]
                    for (int k = 0; k < nr; ++k) {
                        const double mr = 2.0 * Mo[ql + (size_t)k * nr] - (k == ql ? 1.0 : 0.0);
                        const double mi = 2.0 * Moi[ql + (size_t)k * nr];
                        const double xr = o0r[O0(r, base + k, m)], xi = o0i[O0(r, base + k, m)];
                        sr += mr * xr - mi * xi;
                        si += mr * xi + mi * xr;
                    }
                    var_[q] = sr; vai[q] = si;
                    cmul(rr, ri, t0r[T0(r, m, q)], t0i[T0(r, m, q)], vzr[q], vzi[q]);
                }
            }
            __syncthreads();
            // phase 2: b = zN gmInv, Tau = 1 + b a per region
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            for (int q = tid; q < nX; q += ST) {
                const int X = q < nA ? 0 : 1, base = X ? nA : 0, nr = X ? nB : nA, ql = q - base;
                const double *Gv = R.gmInv[X], *Gvi = Gv + (size_t)nr * nr;
                double br = 0.0, bi = 0.0;
                for (int k = 0; k < nr; ++k) {
                    const double gr = Gv[k + (size_t)ql * nr], gi = Gvi[k + (size_t)ql * nr];
                    const double zr = vzr[base + k], zi = vzi[base + k];
                    br += zr * gr - zi * gi;
                    bi += zr * gi + zi * gr;
                }
                vbr[q] = br; vbi[q] = bi;
                double tr, ti;
                cmul(br, bi, var_[q], vai[q], tr, ti);
                acc[2 * X] += tr;
                acc[2 * X + 1] += ti;
            }
            block_sum4(acc, red);
            const double tAr = 1.0 + acc[0], tAi = acc[1], tBr = 1.0 + acc[2], tBi = acc[3];
            const double dA = tAr * tAr + tAi * tAi, dB = tBr * tBr + tBi * tBi;
            prob *= pow(dA, R.lambda) * pow(dB, 1.0 - R.lambda);
            nties[r] += fabs(PD(5, r, m) - prob) < KDQ_NEAR_TIE;
            if (!(PD(5, r, m) < prob)) continue;
            // ---- accepted: gmInv a, and publish the replica's rows / columns at m
            for (int q = tid; q < nX; q += ST) {
                const int X = q < nA ? 0 : 1, base = X ? nA : 0, nr = X ? nB : nA, ql = q - base;
                const double *Gv = R.gmInv[X], *Gvi = Gv + (size_t)nr * nr;
                double sr = 0.0, si = 0.0;
                for (int k = 0; k < nr; ++k) {
                    const double gr = Gv[ql + (size_t)k * nr], gi = Gvi[ql + (size_t)k * nr];
                    const double xr = var_[base + k], xi = vai[base + k];
                    sr += gr * xr - gi * xi;
                    si += gr * xi + gi * xr;
                }
                vzr[q] = sr; vzi[q] = si;
                cxr[q] = o0r[O0(r, q, m)]; cxi[q] = o0i[O0(r, q, m)];
                rxr[q] = t0r[T0(r, m, q)]; rxi[q] = t0i[T0(r, m, q)];
            }
            if (tid < KS) {
                csr[tid] = gtr[GT(r, tid, m)]; csi[tid] = gti[GT(r, tid, m)];
                rsr[tid] = gtr[GT(r, m, tid)]; rsi[tid] = gti[GT(r, m, tid)];
            }
            if (tid == 0) {
                R.detg[0] *= dA;
                R.detg[1] *= dB;
            }
            __syncthreads();
            // inverse Grover matrices and the replica's region blocks of G0
            const double itA2 = 1.0 / dA, itB2 = 1.0 / dB;
            for (int e = tid; e < mAB; e += ST) {
                const int X = e < mA ? 0 : 1, nr = X ? nB : nA, base = X ? nA : 0, g = X ? e - mA : e;
                const int p = g % nr, q = g / nr;
                // 1/Tau
                const double tr_ = X ? tBr : tAr, ti_ = X ? tBi : tAi, it2 = X ? itB2 : itA2;
                double bt_r, bt_i;   // b_q / Tau
                cmul(vbr[base + q], vbi[base + q], tr_ * it2, -ti_ * it2, bt_r, bt_i);
                double ur, ui;
                cmul(vzr[base + p], vzi[base + p], bt_r, bt_i, ur, ui);
                double *Gv = R.gmInv[X];
                Gv[p + (size_t)q * nr] -= ur;
                Gv[(size_t)nr * nr + p + (size_t)q * nr] -= ui;
                if (R.dop) continue;   // no region blocks of G0 in the disorder operator
                double cr_, ci_, vr_, vi_;
                cmul(cxr[base + p], cxi[base + p], rr, ri, cr_, ci_);
                cmul(cr_, ci_, rxr[base + q], rxi[base + q], vr_, vi_);
                double *M = R.M[r][X];
                M[p + (size_t)q * nr] += vr_;
                M[(size_t)nr * nr + p + (size_t)q * nr] += vi_;
            }
            // working blocks of the replica
            for (int e = tid; e < KS * KS; e += ST) {
                const int i = e % KS, j = e / KS;
                double fr, fi, tr, ti;
                cmul(csr[i], csi[i], rr, ri, fr, fi);
                cmul(fr, fi, (j == m ? 1.0 : 0.0) - rsr[j], -rsi[j], tr, ti);
                gtr[GT(r, i, j)] -= tr;
                gti[GT(r, i, j)] -= ti;
            }
            for (int e = tid; e < nX * KS; e += ST) {
                const int x = e % nX, j = e / nX;
                double fr, fi, tr, ti;
                cmul(cxr[x], cxi[x], rr, ri, fr, fi);
                cmul(fr, fi, (j == m ? 1.0 : 0.0) - rsr[j], -rsi[j], tr, ti);
                o0r[O0(r, x, j)] -= tr;
                o0i[O0(r, x, j)] -= ti;
            }
            for (int e = tid; e < KS * nX; e += ST) {
                const int i = e % KS, x = e / KS;
                double fr, fi, tr, ti;
                cmul(csr[i], csi[i], rr, ri, fr, fi);
                cmul(fr, fi, rxr[x], rxi[x], tr, ti);
                t0r[T0(r, i, x)] += tr;
                t0i[T0(r, i, x)] += ti;
            }
            // bordering of the replica's W (same factor for all four matrices)
            if (tid < KS * KS) {
                const int t = tid % KS, u = tid / KS, nr_ = na[r];
                const int at = accl[r][t], au = accl[r][u];
                double xr = 0.0, xi = 0.0;   // x_t = Δ_t/(1+Δ_t) Gt[a_t, m]
                if (t < nr_) cmul(PD(3, r, at), PD(4, r, at), csr[at], csi[at], xr, xi);
                double yr = 1.0, yi = 0.0;   // y_u = Δ_u Gt[m, a_u]
                if (u < nr_) cmul(PD(0, r, au), PD(1, r, au), rsr[au], rsi[au], yr, yi);
                const bool old_row = t < nr_, new_row = t == nr_;
                double lr = rr, li = ri;
                if (old_row) cmul(xr, xi, rr, ri, lr, li);
                double vr_, vi_;
                cmul(lr, li, yr, yi, vr_, vi_);
                const int w = r * KS * KS + t * KS + u;
                if (old_row && u < nr_) {
                    Wr[w] += vr_;
                    Wi[w] += vi_;
                } else if ((old_row && u == nr_) || (new_row && u <= nr_)) {
                    Wr[w] = vr_;
                    Wi[w] = vi_;
                }
            }
            if (tid == 0) {
                accl[r][na[r]] = m;
                sv[r][m] |= 1 << 16;
            }
            __syncthreads();
            ++na[r];
        }
    }
    for (int r = 0; r < NR; ++r) {
        const ChainDev &chr = a.chains[c + r * nch];
        if (tid == 0) {
            *chr.acc_cnt = na[r];
            chr.counters[0] += kb;
            chr.counters[1] += na[r];
            chr.counters[2] += nties[r];
        }
        if (tid < na[r]) chr.acc_idx[tid] = s0 + accl[r][tid];
        for (int e = tid; e < na[r] * na[r]; e += ST) {
            const int t = e % na[r], u = e / na[r];
            chr.W[t + u * 64] = Wr[r * KS * KS + t * KS + u];
            chr.W[64 * 64 + t + u * 64] = Wi[r * KS * KS + t * KS + u];
        }
        if (tid < kb) {
            const int v = sv[r][tid];
            chr.s[(size_t)(s0 + tid) + (size_t)(lt - 1) * chr.n_elem] =
                (uint8_t)((v >> 16) ? (v >> 8) & 0xff : v & 0xff);
        }
    }
#undef GT
#undef T0
#undef O0
#undef PD
}

// M[r][X] = G0_r[X, X] for both replicas (one with the disorder operator) and regions
// (grid: elements x chains)
__global__ void k_scee_gather(SceeArgs a) {
    const int c = blockIdx.y;
    const SceeRegion &R = a.reg[c];
    const int nA = R.nA, nB = R.nB, mA = nA * nA, mAB = mA + nB * nB;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (R.dop ? 1 : 2) * mAB) return;
    const int r = e / mAB, f = e % mAB;
    const int X = f < mA ? 0 : 1, nr = X ? nB : nA, base = X ? nA : 0, g = X ? f - mA : f;
    const int p = g % nr, q = g / nr;
    const Opnd G0 = a.G0[c + r * a.nch];
    const size_t o = (size_t)R.idx[base + p] + (size_t)R.idx[base + q] * G0.ld;
    double *M = R.M[r][X];
    M[p + (size_t)q * nr] = G0.re[o];
    M[(size_t)nr * nr + p + (size_t)q * nr] = G0.im[o];
}

// GroverMatrix! (src/public/GF.jl:101-109): C = 2 C - M1 - M2 + I with C = M1 M2 on entry
__global__ void k_scee_grover_form(const SceeRegion *reg) {
    const int c = blockIdx.y;
    const SceeRegion &R = reg[c];
    const int nA = R.nA, nB = R.nB, mA = nA * nA, mAB = mA + nB * nB;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mAB) return;
    const int X = e < mA ? 0 : 1, nr = X ? nB : nA, g = X ? e - mA : e;
    const int p = g % nr, q = g / nr;
    const size_t o = p + (size_t)q * nr, pl = (size_t)nr * nr;
    double *C = R.C[X];
    const double *M1 = R.M[0][X], *M2 = R.M[1][X];
    C[o] = 2.0 * C[o] - M1[o] - M2[o] + (p == q ? 1.0 : 0.0);
    C[pl + o] = 2.0 * C[pl + o] - M1[pl + o] - M2[pl + o];
}

// get_ABX! (src/tU/disorder_operate.jl:202-223): C = e^{iα} I + (1 - e^{iα}) G0[X, X]
__global__ void k_dop_form(const SceeRegion *reg) {
    const int c = blockIdx.y;
    const SceeRegion &R = reg[c];
    const int nA = R.nA, nB = R.nB, mA = nA * nA, mAB = mA + nB * nB;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mAB) return;
    const int X = e < mA ? 0 : 1, nr = X ? nB : nA, g = X ? e - mA : e;
    const int p = g % nr, q = g / nr;
    const size_t o = p + (size_t)q * nr, pl = (size_t)nr * nr;
    const double *M = R.M[0][X];
    double vr, vi;
    cmul(1.0 - R.ea[0], -R.ea[1], M[o], M[pl + o], vr, vi);
    R.C[X][o] = vr + (p == q ? R.ea[0] : 0.0);
    R.C[X][pl + o] = vi + (p == q ? R.ea[1] : 0.0);
}

__global__ void k_scee_detg(const SceeRegion *reg, int nch, const double *la, const double *lb) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    const SceeRegion &R = reg[c];
    R.detg[0] = R.nA ? exp(2.0 * la[c]) : 1.0;
    R.detg[1] = R.nB ? exp(2.0 * lb[c]) : 1.0;
}

__global__ void k_scee_accum(const SceeRegion *reg, int nch) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    const SceeRegion &R = reg[c];
    R.osum[0] += pow(R.detg[0] / R.detg[1], 1.0 / R.nlam);
    R.osum[1] += 1.0;
}

__global__ void k_diag(const ChainDev *chains, const Opnd *X, int ndim, double v, int set) {
    const int c = blockIdx.z;
    const int n = dim_of(chains[c], ndim);
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    const Opnd A = X[c];
    const size_t o = (size_t)i + (size_t)j * A.ld;
    if (set) {
        A.re[o] = i == j ? 1.0 : 0.0;
        A.im[o] = 0.0;
    } else if (i == j) {
        A.re[o] += v;
    }
}

}  // namespace

size_t scee_lds_bytes(int ks, int nx) {
    return (size_t)(4 * ks * ks + 8 * ks * nx + 10 * nx + 4 * ks + 4 * ks * ks + 12 * ks + 4 * (ST / 64)) * sizeof(double);
}

hipError_t launch_scee_decide(hipStream_t st, const SceeArgs &a) {
    const int ks = a.kb <= 8 ? 8 : 16;
    if (a.kb < 1 || a.kb > 16) return hipErrorInvalidValue;
    const size_t b = scee_lds_bytes(ks, a.nx_max);
    if (ks == 8) {
        if (hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_scee_decide<8>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)b))
            return e;
        hipLaunchKernelGGL(k_scee_decide<8>, dim3(a.nch), dim3(ST), b, st, a);
    } else {
        if (hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_scee_decide<16>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)b))
            return e;
        hipLaunchKernelGGL(k_scee_decide<16>, dim3(a.nch), dim3(ST), b, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_scee_gather(hipStream_t st, int nch, int nx_max, const SceeArgs &a) {
    const int tot = 2 * nx_max * nx_max;
    hipLaunchKernelGGL(k_scee_gather, dim3((tot + 255) / 256, nch), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_scee_grover_form(hipStream_t st, int nch, int total, const SceeRegion *reg) {
    const int tot = total;   // largest nA² + nB² over the chains
    hipLaunchKernelGGL(k_scee_grover_form, dim3((tot + 255) / 256, nch), dim3(256), 0, st, reg);
    return hipGetLastError();
}

hipError_t launch_dop_form(hipStream_t st, int nch, int total, const SceeRegion *reg) {
    hipLaunchKernelGGL(k_dop_form, dim3((total + 255) / 256, nch), dim3(256), 0, st, reg);
    return hipGetLastError();
}

hipError_t launch_scee_detg(hipStream_t st, int nch, const SceeRegion *reg, const double *la, const double *lb) {
    hipLaunchKernelGGL(k_scee_detg, dim3((nch + 63) / 64), dim3(64), 0, st, reg, nch, la, lb);
    return hipGetLastError();
}

namespace {

// ------------------------------------------------------------ rank-2 models (t-V, t-U-V)
// UpdateSCEELayer! of t-V (src/tV/SCEE.jl:447-494, the bonds of one family) and t-U-V
// (src/tUV/SCEE.jl:355-402, every directed bond): one workgroup per chain walks the elements
// in order, replica 1 then replica 2 per element. The proposal weight is get_r!'s det r (t-V,
// src/tV/GreenMatrix.jl:76-89) or |det r|² (t-U-V, src/tUV/GreenMatrix.jl:73-89), times γ'/γ
// and the Grover factors det Tau (t-V) or |det Tau|² (t-U-V) of get_abTau1!/2!
// (src/public/Gupdate.jl:91-132) to the powers λ and 1 - λ. An accepted change is applied at
// once: det g and GMupdate! (Gupdate.jl:45-51) of both regions, G4update! (Gupdate.jl:23-37)
// on the replica's four matrices. The other replica's G0 on a region is read from its matrix.
constexpr int RT = 512;
// threads of the delayed-block rank-2 decision kernel: one wave per SIMD leaves 512 registers
// per lane, which its 2x2 algebra needs without spilling (the element kernel keeps RT: its
// full-matrix updates need the threads more than it loses to spills)
constexpr int BT = 256;

KDQ_DEV void c22_inv(const double ar[2][2], const double ai[2][2], double br[2][2], double bi[2][2], double &dr,
                     double &di) {
    double t1r, t1i, t2r, t2i;
    cmul(ar[0][0], ai[0][0], ar[1][1], ai[1][1], t1r, t1i);
    cmul(ar[0][1], ai[0][1], ar[1][0], ai[1][0], t2r, t2i);
    dr = t1r - t2r;
    di = t1i - t2i;
    const double n2 = dr * dr + di * di, qr = dr / n2, qi = -di / n2;   // 1 / det
    cmul(ar[1][1], ai[1][1], qr, qi, br[0][0], bi[0][0]);
    cmul(-ar[0][1], -ai[0][1], qr, qi, br[0][1], bi[0][1]);
    cmul(-ar[1][0], -ai[1][0], qr, qi, br[1][0], bi[1][0]);
    cmul(ar[0][0], ai[0][0], qr, qi, br[1][1], bi[1][1]);
}

// sum of 8 doubles over the block (NT threads)
template <int NT = RT>
KDQ_DEV void block_sum8(double v[8], double *red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
        for (int k = 0; k < 8; ++k) red[wave * 8 + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; ++w) s += red[w * 8 + k];
        v[k] = s;
    }
}

__global__ __launch_bounds__(RT) void k_scee_r2(SceeR2Args a) {
    const int c = blockIdx.x, nch = a.nch, tid = threadIdx.x;
    const SceeRegion &R = a.reg[c];
    const int nA = R.nA, nB = R.nB, nX = nA + nB, NXM = max(a.nx_max, 1), NM = a.max_ns;
    const int n = a.chains[c].ns, lt = a.lt;
    extern __shared__ double lds[];
    // region work vectors over the concatenated index q in [0, nX): region A then B
    double *zr = lds, *zi = zr + 2 * NXM;            // zN [t][q]
    double *ar_ = zi + 2 * NXM, *ai_ = ar_ + 2 * NXM;  // a  [u][q]
    double *br_ = ai_ + 2 * NXM, *bi_ = br_ + 2 * NXM; // b  [t][q]
    double *nr_ = bi_ + 2 * NXM, *ni_ = nr_ + 2 * NXM; // gmInv a [u][q]
    double *wr_ = ni_ + 2 * NXM, *wi_ = wr_ + 2 * NXM; // Tau⁻¹ b [t][q]
    double *c0r = wi_ + 2 * NXM, *c0i = c0r + 2 * NM;  // G0t[:, sub] [u][i]
    double *c1r = c0i + 2 * NM, *c1i = c1r + 2 * NM;   // Gt[:, sub]  [u][i]
    double *z1r = c1i + 2 * NM, *z1i = z1r + 2 * NM;   // r Gt0[sub, :]          [t][j]
    double *z2r = z1i + 2 * NM, *z2i = z2r + 2 * NM;   // r (E_sub - Gt[sub, :]) [t][j]
    double *red = z2i + 2 * NM;                        // RT/64 x 16
    __shared__ double rmr[2][2], rmi[2][2], tinv[2][2][2][2], dreg[2], prob_s;   // tinv [X][re/im][t][u]
    __shared__ int acc_s, sx_s;
    const uint32_t tag = (uint32_t)lt | ((uint32_t)a.dir << 31);
    const uint32_t sweep = a.sweep_dev ? *a.sweep_dev : a.sweep;
    const int e0 = a.family >= 0 ? a.family * a.chains[c].nb : 0;
    const int ne = a.family >= 0 ? a.chains[c].nb : a.chains[c].n_elem;
    for (int el = e0; el < e0 + ne; ++el) {
        const int sub0 = a.chains[c].ex[el], sub1 = a.chains[c].ey[el];
        auto sub = [&](int k) { return k ? sub1 : sub0; };
        for (int r = 0; r < 2; ++r) {
            const int e = c + r * nch, o = c + (1 - r) * nch;
            const ChainDev &ch = a.chains[e];
            const Opnd Gt = a.Gt[e], Gt0 = a.Gt0[e], G0t = a.G0t[e], G0 = a.G0[e], G0o = a.G0[o];
            const size_t sidx = (size_t)el + (size_t)(lt - 1) * ch.n_elem;
            if (tid == 0) {
                int chv;
                double u;
                proposal_draw(ch.seed, ch.chain_id, ch.stream, sweep, tag, (uint32_t)el, chv, u);
                const int si = ch.s[sidx];
                const int sx = chv + 1 + ((chv + 1) >= si ? 1 : 0);
                double Dr[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, Di[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
                if (a.kind == 1) {
                    const double h = 0.7071067811865476, uv[2][2] = {{-h, -h}, {-h, h}};
                    const double ds = ch.alpha[lt - 1] * (ch.eta[sx - 1] - ch.eta[si - 1]);
                    const double t2[2] = {exp(ds) - 1.0, exp(-ds) - 1.0};
#pragma unroll
                    for (int p = 0; p < 2; ++p)
#pragma unroll
                        for (int q = 0; q < 2; ++q) Dr[p][q] = uv[p][0] * t2[0] * uv[0][q] + uv[p][1] * t2[1] * uv[1][q];
                } else {
                    const double ds = ch.eta[sx - 1] - ch.eta[si - 1], v[2] = {ds, ch.a * ds};
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        if (ch.is_complex) {
                            double sn, cs;
                            sincos(v[q], &sn, &cs);
                            Dr[q][q] = cs - 1.0;
                            Di[q][q] = sn;
                        } else {
                            Dr[q][q] = exp(v[q]) - 1.0;
                        }
                    }
                }
                // rr = I + Δ (I - Gt[sub, sub])
                double rr[2][2], ri[2][2];
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        double sr = Dr[p][q] + (p == q ? 1.0 : 0.0), sim = Di[p][q];
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            const size_t og = (size_t)sub(k) + (size_t)sub(q) * Gt.ld;
                            double tr, ti;
                            cmul(Dr[p][k], Di[p][k], Gt.re[og], Gt.im[og], tr, ti);
                            sr -= tr;
                            sim -= ti;
                        }
                        rr[p][q] = sr;
                        ri[p][q] = sim;
                    }
                double ivr[2][2], ivi[2][2], dr, di;
                c22_inv(rr, ri, ivr, ivi, dr, di);
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        double sr = 0.0, sim = 0.0;
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            double tr, ti;
                            cmul(ivr[p][k], ivi[p][k], Dr[k][q], Di[k][q], tr, ti);
                            sr += tr;
                            sim += ti;
                        }
                        rmr[p][q] = sr;
                        rmi[p][q] = sim;
                    }
                const double pw = a.kind == 1 ? dr : dr * dr + di * di;
                prob_s = pw * (ch.gamma[sx - 1] / ch.gamma[si - 1]);
                sx_s = sx | (si << 8);
                acc_s = 0;
                // u is kept in the prob comparison below
                dreg[0] = u;
            }
            __syncthreads();
            const double u = dreg[0];
            // ---- Grover factors of both regions
            // which = 1: zN[t][q] = Σ_k Gt0[sub_t, X_k] (2 G0o[X_k, X_q] - δ); a[u][q] = Σ_t G0t[X_q, sub_t] r[t][u]
            // which = 2: a[u][q] = Σ_k (2 G0o[X_q, X_k] - δ) G0t[X_k, sub_u]; zN[t][q] = Σ_t' r[t][t'] Gt0[sub_t', X_q]
            for (int it = tid; it < 2 * nX; it += RT) {
                const int t = it / nX, q = it % nX;
                const int k0 = q < nA ? 0 : nA, k1 = q < nA ? nA : nX;
                const int xq = R.idx[q];
                double sr = 0.0, sim = 0.0, pr = 0.0, pi = 0.0;
                if (r == 0) {
                    for (int k = k0; k < k1; ++k) {
                        const int xk = R.idx[k];
                        const size_t og = (size_t)sub(t) + (size_t)xk * Gt0.ld, oo = (size_t)xk + (size_t)xq * G0o.ld;
                        double mr = 2.0 * G0o.re[oo] - (k == q ? 1.0 : 0.0), mi = 2.0 * G0o.im[oo], tr, ti;
                        cmul(Gt0.re[og], Gt0.im[og], mr, mi, tr, ti);
                        sr += tr;
                        sim += ti;
                    }
                    // a[u = t][q]
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const size_t oc = (size_t)xq + (size_t)sub(k) * G0t.ld;
                        double tr, ti;
                        cmul(G0t.re[oc], G0t.im[oc], rmr[k][t], rmi[k][t], tr, ti);
                        pr += tr;
                        pi += ti;
                    }
                } else {
                    for (int k = k0; k < k1; ++k) {
                        const int xk = R.idx[k];
                        const size_t oo = (size_t)xq + (size_t)xk * G0o.ld, oc = (size_t)xk + (size_t)sub(t) * G0t.ld;
                        double mr = 2.0 * G0o.re[oo] - (k == q ? 1.0 : 0.0), mi = 2.0 * G0o.im[oo], tr, ti;
                        cmul(mr, mi, G0t.re[oc], G0t.im[oc], tr, ti);
                        pr += tr;
                        pi += ti;
                    }
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const size_t og = (size_t)sub(k) + (size_t)xq * Gt0.ld;
                        double tr, ti;
                        cmul(rmr[t][k], rmi[t][k], Gt0.re[og], Gt0.im[og], tr, ti);
                        sr += tr;
                        sim += ti;
                    }
                }
                zr[t * NXM + q] = sr;
                zi[t * NXM + q] = sim;
                ar_[t * NXM + q] = pr;
                ai_[t * NXM + q] = pi;
            }
            __syncthreads();
            // b[t][q] = Σ_k zN[t][k] gmInv_X[k][q] (k over q's region)
            for (int it = tid; it < 2 * nX; it += RT) {
                const int t = it / nX, q = it % nX;
                const int X = q < nA ? 0 : 1, off = X ? nA : 0, m = X ? nB : nA;
                const double *gr = R.gmInv[X], *gi = gr + (size_t)m * m;
                double sr = 0.0, sim = 0.0;
                for (int k = 0; k < m; ++k) {
                    const size_t og = (size_t)k + (size_t)(q - off) * m;
                    double tr, ti;
                    cmul(zr[t * NXM + off + k], zi[t * NXM + off + k], gr[og], gi[og], tr, ti);
                    sr += tr;
                    sim += ti;
                }
                br_[t * NXM + q] = sr;
                bi_[t * NXM + q] = sim;
            }
            __syncthreads();
            // Tau_X[t][u] = δ + Σ_{q in X} b[t][q] a[u][q]: 2 regions x 4 complex sums
            double v[8] = {0, 0, 0, 0, 0, 0, 0, 0}, w8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int q = tid; q < nX; q += RT) {
                const bool inA = q < nA;
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int uu = 0; uu < 2; ++uu) {
                        double tr, ti;
                        cmul(br_[t * NXM + q], bi_[t * NXM + q], ar_[uu * NXM + q], ai_[uu * NXM + q], tr, ti);
                        const int k = 2 * (t * 2 + uu);
                        v[k] += inA ? tr : 0.0;
                        v[k + 1] += inA ? ti : 0.0;
                        w8[k] += inA ? 0.0 : tr;
                        w8[k + 1] += inA ? 0.0 : ti;
                    }
            }
            block_sum8(v, red);
            block_sum8(w8, red + 8 * (RT / 64));
            if (tid == 0) {
                const int m[2] = {nA, nB};
                double dX[2];
#pragma unroll
                for (int X = 0; X < 2; ++X) {
                    double tr[2][2], ti[2][2], ir[2][2], ii[2][2], dr = 1.0, di = 0.0;
#pragma unroll
                    for (int t = 0; t < 2; ++t)
#pragma unroll
                        for (int uu = 0; uu < 2; ++uu) {
                            const int k = 2 * (t * 2 + uu);
                            tr[t][uu] = (X ? w8[k] : v[k]) + (t == uu ? 1.0 : 0.0);
                            ti[t][uu] = X ? w8[k + 1] : v[k + 1];
                        }
                    if (m[X]) c22_inv(tr, ti, ir, ii, dr, di);
                    else {
                        ir[0][0] = ir[1][1] = 1.0;
                        ir[0][1] = ir[1][0] = 0.0;
                        ii[0][0] = ii[0][1] = ii[1][0] = ii[1][1] = 0.0;
                    }
#pragma unroll
                    for (int t = 0; t < 2; ++t)
#pragma unroll
                        for (int uu = 0; uu < 2; ++uu) {
                            tinv[X][0][t][uu] = ir[t][uu];
                            tinv[X][1][t][uu] = ii[t][uu];
                        }
                    dX[X] = m[X] ? (a.kind == 1 ? dr : dr * dr + di * di) : 1.0;
                }
                const double p = prob_s * pow(dX[0], R.lambda) * pow(dX[1], 1.0 - R.lambda);
                const int ok = u < p;
                if (fabs(u - p) < KDQ_NEAR_TIE) ch.counters[2] += 1.0;
                acc_s = ok;
                dreg[0] = dX[0];
                dreg[1] = dX[1];
                if (ok) {
                    R.detg[0] *= dX[0];
                    R.detg[1] *= dX[1];
                }
            }
            __syncthreads();
            if (!acc_s) continue;   // block-uniform
            // ---- GMupdate!: gmInv_X -= (gmInv a)(Tau⁻¹ b)
            for (int it = tid; it < 2 * nX; it += RT) {
                const int uu = it / nX, q = it % nX;
                const int X = q < nA ? 0 : 1, off = X ? nA : 0, m = X ? nB : nA;
                const double *gr = R.gmInv[X], *gi = gr + (size_t)m * m;
                double sr = 0.0, sim = 0.0;
                for (int k = 0; k < m; ++k) {
                    const size_t og = (size_t)(q - off) + (size_t)k * m;
                    double tr, ti;
                    cmul(gr[og], gi[og], ar_[uu * NXM + off + k], ai_[uu * NXM + off + k], tr, ti);
                    sr += tr;
                    sim += ti;
                }
                nr_[uu * NXM + q] = sr;
                ni_[uu * NXM + q] = sim;
                // Tau⁻¹ b [t = uu][q]
                double wr = 0.0, wi = 0.0;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    double tr, ti;
                    cmul(tinv[X][0][uu][k], tinv[X][1][uu][k], br_[k * NXM + q], bi_[k * NXM + q], tr, ti);
                    wr += tr;
                    wi += ti;
                }
                wr_[uu * NXM + q] = wr;
                wi_[uu * NXM + q] = wi;
            }
            // ---- G4update! operands from the current matrices
            for (int it = tid; it < 2 * n; it += RT) {
                const int uu = it / n, i = it % n;
                const size_t o0 = (size_t)i + (size_t)sub(uu) * G0t.ld, o1 = (size_t)i + (size_t)sub(uu) * Gt.ld;
                c0r[uu * NM + i] = G0t.re[o0];
                c0i[uu * NM + i] = G0t.im[o0];
                c1r[uu * NM + i] = Gt.re[o1];
                c1i[uu * NM + i] = Gt.im[o1];
                // column i of r Gt0[sub, :] and of r (E_sub - Gt[sub, :]) for row t = uu
                double s1r = 0.0, s1i = 0.0, s2r = 0.0, s2i = 0.0;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const size_t og = (size_t)sub(k) + (size_t)i * Gt0.ld, ot = (size_t)sub(k) + (size_t)i * Gt.ld;
                    double tr, ti;
                    cmul(rmr[uu][k], rmi[uu][k], Gt0.re[og], Gt0.im[og], tr, ti);
                    s1r += tr;
                    s1i += ti;
                    const double er = (i == sub(k) ? 1.0 : 0.0) - Gt.re[ot], ei = -Gt.im[ot];
                    cmul(rmr[uu][k], rmi[uu][k], er, ei, tr, ti);
                    s2r += tr;
                    s2i += ti;
                }
                z1r[uu * NM + i] = s1r;
                z1i[uu * NM + i] = s1i;
                z2r[uu * NM + i] = s2r;
                z2i[uu * NM + i] = s2i;
            }
            __syncthreads();
            for (int it = tid; it < nX * nX; it += RT) {   // both regions' nX_A² + nX_B² entries
                const int q = it % nX, cc = it / nX;
                const bool inA = q < nA && cc < nA, inB = q >= nA && cc >= nA;
                if (!inA && !inB) continue;
                const int X = inA ? 0 : 1, off = X ? nA : 0, m = X ? nB : nA;
                double *gr = R.gmInv[X], *gi = gr + (size_t)m * m;
                const size_t og = (size_t)(q - off) + (size_t)(cc - off) * m;
                double sr = 0.0, sim = 0.0;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    double tr, ti;
                    cmul(nr_[k * NXM + q], ni_[k * NXM + q], wr_[k * NXM + cc], wi_[k * NXM + cc], tr, ti);
                    sr += tr;
                    sim += ti;
                }
                gr[og] -= sr;
                gi[og] -= sim;
            }
            // G0 += G0t[:, sub] z1; Gt0 += Gt[:, sub] z1; G0t -= G0t[:, sub] z2; Gt -= Gt[:, sub] z2
            for (int it = tid; it < n * n; it += RT) {
                const int i = it % n, j = it / n;
                double p0r = 0.0, p0i = 0.0, p1r = 0.0, p1i = 0.0, q0r = 0.0, q0i = 0.0, q1r = 0.0, q1i = 0.0;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    double tr, ti;
                    cmul(c0r[k * NM + i], c0i[k * NM + i], z1r[k * NM + j], z1i[k * NM + j], tr, ti);
                    p0r += tr; p0i += ti;
                    cmul(c1r[k * NM + i], c1i[k * NM + i], z1r[k * NM + j], z1i[k * NM + j], tr, ti);
                    p1r += tr; p1i += ti;
                    cmul(c0r[k * NM + i], c0i[k * NM + i], z2r[k * NM + j], z2i[k * NM + j], tr, ti);
                    q0r += tr; q0i += ti;
                    cmul(c1r[k * NM + i], c1i[k * NM + i], z2r[k * NM + j], z2i[k * NM + j], tr, ti);
                    q1r += tr; q1i += ti;
                }
                const size_t oa = (size_t)i + (size_t)j * G0.ld, ob = (size_t)i + (size_t)j * Gt0.ld;
                const size_t oc = (size_t)i + (size_t)j * G0t.ld, od = (size_t)i + (size_t)j * Gt.ld;
                G0.re[oa] += p0r; G0.im[oa] += p0i;
                Gt0.re[ob] += p1r; Gt0.im[ob] += p1i;
                G0t.re[oc] -= q0r; G0t.im[oc] -= q0i;
                Gt.re[od] -= q1r; Gt.im[od] -= q1i;
            }
            if (tid == 0) {
                ch.s[sidx] = (uint8_t)(sx_s & 0xff);
                ch.counters[1] += 1.0;
            }
            __syncthreads();
        }
        if (tid == 0) {
            a.chains[c].counters[0] += 1.0;
            a.chains[c + nch].counters[0] += 1.0;
        }
    }
}


// GMupdate! of both regions and the replica's region blocks of G0 after an accepted rank-2
// change: gmInv_X -= (gmInv a)(Tau⁻¹ b), M_r += G0t[X, s] w Gt0[s, X]. Read-modify-writes of
// global memory: each thread loads UB entries' old values before it stores any, so the round
// trips overlap (one after another they serialise on the possible aliasing of the stores).
constexpr int UB = 8;
KDQ_DEV void region_update(const SceeRegion &R, int r, int nA, int nB, int NX, const double *nr_,
                           const double *ni_, const double *wr_, const double *wi_, const double *cxr,
                           const double *cxi, const double *rxr, const double *rxi, const double (*rmr)[2],
                           const double (*rmi)[2]) {
    const int mA = nA * nA, mAB = mA + nB * nB;
    for (int it0 = threadIdx.x; it0 < mAB; it0 += BT * UB) {
        double gvr[UB], gvi[UB], mvr[UB], mvi[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int it = it0 + u * BT;
            if (it >= mAB) continue;
            const int X = it < mA ? 0 : 1, nr = X ? nB : nA, g = X ? it - mA : it;
            const size_t o = (size_t)(g % nr) + (size_t)(g / nr) * nr, pl = (size_t)nr * nr;
            gvr[u] = R.gmInv[X][o];
            gvi[u] = R.gmInv[X][pl + o];
            mvr[u] = R.M[r][X][o];
            mvi[u] = R.M[r][X][pl + o];
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int it = it0 + u * BT;
            if (it >= mAB) continue;
            const int X = it < mA ? 0 : 1, nr = X ? nB : nA, base = X ? nA : 0, g = X ? it - mA : it;
            const int p = g % nr, q = g / nr;
            double sr = 0.0, sim = 0.0, mr = 0.0, mi = 0.0;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double tr, ti;
                cmul(nr_[k * NX + base + p], ni_[k * NX + base + p], wr_[k * NX + base + q], wi_[k * NX + base + q], tr,
                     ti);
                sr += tr;
                sim += ti;
                // (w Gt0[s, X_q])_k
                double zr_ = 0.0, zi_ = 0.0;
#pragma unroll
                for (int k2 = 0; k2 < 2; ++k2) {
                    double ur, ui;
                    cmul(rmr[k][k2], rmi[k][k2], rxr[k2 * NX + base + q], rxi[k2 * NX + base + q], ur, ui);
                    zr_ += ur;
                    zi_ += ui;
                }
                cmul(cxr[k * NX + base + p], cxi[k * NX + base + p], zr_, zi_, tr, ti);
                mr += tr;
                mi += ti;
            }
            const size_t o = (size_t)p + (size_t)q * nr, pl = (size_t)nr * nr;
            R.gmInv[X][o] = gvr[u] - sr;
            R.gmInv[X][pl + o] = gvi[u] - sim;
            R.M[r][X][o] = mvr[u] + mr;
            R.M[r][X][pl + o] = mvi[u] + mi;
        }
    }
}

// ---- delayed blocks of the t-V rank-2 decisions (UpdateSCEELayer!, src/tV/SCEE.jl:447-494)
// A block holds KB consecutive bonds of one family. The bonds of a family are disjoint, so the
// block's 2·KB sites S are distinct. Gt[S,S], Gt0[S,X] and G0t[X,S] of both replicas live in
// LDS; they, the region blocks M[r][X] = G0_r[X,X] and gmInv are kept current through the
// sequential rank-2 G4update! / GMupdate! (src/public/Gupdate.jl:23-51), and the decisions
// are those of k_scee_r2, element by element, replica 1 then 2. The four full matrices are
// left to the four-matrix flush: the sequential updates of a block equal one Woodbury update
// with the accepted sites A (in acceptance order) and W = (I + Δ_A (I - Gt[A,A]))⁻¹ Δ_A,
// Δ_A block-diagonal. W grows by bordering with 2x2 blocks, from the working matrix Gc before
// each accepted bond s with w = r⁻¹Δ: x = Q_A Gc[A,s] with Q_a = (I + Δ_a)⁻¹ Δ_a,
// y = Gc[s,A] Δ_A, W' = [[W + x w y, x w], [w y, w]].
template <int KB>
__global__ __launch_bounds__(BT) void k_scee_r2_block(SceeR2Args a) {
    constexpr int KS = 2 * KB;
    const int c = blockIdx.x, nch = a.nch, tid = threadIdx.x;
    const SceeRegion &R = a.reg[c];
    const ChainDev &ch0 = a.chains[c];
    const int nA = R.nA, nB = R.nB, nX = nA + nB, NX = max(a.nx_max, 1);
    const int lt = a.lt, mA = nA * nA, mAB = mA + nB * nB;
    const int e0 = a.family * ch0.nb + a.s0;
    const int kb = min(KB, ch0.nb - a.s0);
    if (kb <= 0) {
        if (tid < 2) *a.chains[c + tid * nch].acc_cnt = 0;
        return;
    }
    extern __shared__ double lds[];
    double *gtr = lds, *gti = gtr + 2 * KS * KS;                 // Gt[S,S]  [r][i][j]
    double *t0r = gti + 2 * KS * KS, *t0i = t0r + 2 * KS * NX;  // Gt0[S,X] [r][i][x]
    double *o0r = t0i + 2 * KS * NX, *o0i = o0r + 2 * KS * NX;  // G0t[X,S] [r][x][i]
    double *zr = o0i + 2 * KS * NX, *zi = zr + 2 * NX;          // zN [t][q]
    double *ar_ = zi + 2 * NX, *ai_ = ar_ + 2 * NX;             // a [u][q]
    double *br_ = ai_ + 2 * NX, *bi_ = br_ + 2 * NX;            // b [t][q]
    double *nr_ = bi_ + 2 * NX, *ni_ = nr_ + 2 * NX;            // gmInv a [u][q]
    double *wr_ = ni_ + 2 * NX, *wi_ = wr_ + 2 * NX;            // Tau⁻¹ b [t][q]
    double *cxr = wi_ + 2 * NX, *cxi = cxr + 2 * NX;            // G0t[X, s_t] [t][x]
    double *rxr = cxi + 2 * NX, *rxi = rxr + 2 * NX;            // Gt0[s_t, X] [t][x]
    double *csr = rxi + 2 * NX, *csi = csr + 2 * KS;            // Gt[S, s_t]  [t][i]
    double *rsr = csi + 2 * KS, *rsi = rsr + 2 * KS;            // Gt[s_t, S]  [t][j]
    double *Wr = rsi + 2 * KS, *Wi = Wr + 2 * KS * KS;          // W [r][t][u]
    double *red = Wi + 2 * KS * KS;                             // BT/64 x 16
    __shared__ double dlr[2][KB][2][2], dli[2][KB][2][2];       // Δ of each bond's proposal
    __shared__ double qlr[2][KB][2][2], qli[2][KB][2][2];       // (I + Δ)⁻¹ Δ
    __shared__ double pgam[2][KB], pu[2][KB];                   // γ'/γ, uniform draw
    __shared__ int sv[2][KB], site[KS], accl[2][KS], accb[2][KB];
    __shared__ double rmr[2][2], rmi[2][2], tinv[2][2][2][2], prob_s;
    __shared__ int acc_s;
#define GT(r, i, j) ((r) * KS * KS + (i) * KS + (j))
#define T0(r, i, x) ((r) * KS * NX + (i) * NX + (x))
#define O0(r, x, i) ((r) * NX * KS + (x) * KS + (i))
    const uint32_t tag = (uint32_t)lt | ((uint32_t)a.dir << 31);
    const uint32_t sweep = a.sweep_dev ? *a.sweep_dev : a.sweep;
    if (tid < 2 * KB) {
        const int r = tid / KB, q = tid % KB;
        if (r == 0) {
            site[2 * q] = q < kb ? ch0.ex[e0 + q] : 0;
            site[2 * q + 1] = q < kb ? ch0.ey[e0 + q] : 0;
        }
        if (q < kb) {
            const ChainDev &chr = a.chains[c + r * nch];
            const int el = e0 + q;
            int chv;
            double u;
            proposal_draw(chr.seed, chr.chain_id, chr.stream, sweep, tag, (uint32_t)el, chv, u);
            const int si = chr.s[(size_t)el + (size_t)(lt - 1) * chr.n_elem];
            const int sx = chv + 1 + ((chv + 1) >= si ? 1 : 0);
            // Δ = uv diag(e^{ds} - 1, e^{-ds} - 1) uv (src/tV/GreenMatrix.jl:76-89), and
            // (I + Δ)⁻¹ Δ = uv diag(1 - e^{-ds}, 1 - e^{ds}) uv since uv² = I
            const double h = 0.7071067811865476, uv[2][2] = {{-h, -h}, {-h, h}};
            const double ds = chr.alpha[lt - 1] * (chr.eta[sx - 1] - chr.eta[si - 1]);
            const double t2[2] = {exp(ds) - 1.0, exp(-ds) - 1.0};
            const double q2[2] = {1.0 - exp(-ds), 1.0 - exp(ds)};
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    dlr[r][q][p][s] = uv[p][0] * t2[0] * uv[0][s] + uv[p][1] * t2[1] * uv[1][s];
                    dli[r][q][p][s] = 0.0;
                    qlr[r][q][p][s] = uv[p][0] * q2[0] * uv[0][s] + uv[p][1] * q2[1] * uv[1][s];
                    qli[r][q][p][s] = 0.0;
                }
            pgam[r][q] = chr.gamma[sx - 1] / chr.gamma[si - 1];
            pu[r][q] = u;
            sv[r][q] = si | (sx << 8);
        }
    }
    __syncthreads();
    // working blocks of both replicas and the region blocks of G0 (M[r][X] = G0_r[X,X])
    for (int e = tid; e < 2 * KS * KS; e += BT) {
        const int r = e / (KS * KS), i = (e / KS) % KS, j = e % KS;
        double vr = 0.0, vi = 0.0;
        if (i < 2 * kb && j < 2 * kb) {
            const Opnd G = a.Gt[c + r * nch];
            const size_t o = (size_t)site[i] + (size_t)site[j] * G.ld;
            vr = G.re[o];
            vi = G.im[o];
        }
        gtr[e] = vr;
        gti[e] = vi;
        Wr[e] = 0.0;
        Wi[e] = 0.0;
    }
    for (int e = tid; e < 2 * KS * nX; e += BT) {
        const int r = e / (KS * nX), rem = e % (KS * nX), x = rem / KS, i = rem % KS;
        double tr = 0.0, ti = 0.0, orr = 0.0, oi = 0.0;
        if (i < 2 * kb) {
            const Opnd T = a.Gt0[c + r * nch], O = a.G0t[c + r * nch];
            const int sx = R.idx[x];
            const size_t ot = (size_t)site[i] + (size_t)sx * T.ld, oo = (size_t)sx + (size_t)site[i] * O.ld;
            tr = T.re[ot]; ti = T.im[ot];
            orr = O.re[oo]; oi = O.im[oo];
        }
        t0r[T0(r, i, x)] = tr; t0i[T0(r, i, x)] = ti;
        o0r[O0(r, x, i)] = orr; o0i[O0(r, x, i)] = oi;
    }
    for (int e = tid; e < 2 * mAB; e += BT) {
        const int r = e / mAB, f = e % mAB;
        const int X = f < mA ? 0 : 1, nr = X ? nB : nA, base = X ? nA : 0, g = X ? f - mA : f;
        const int p = g % nr, q = g / nr;
        const Opnd G0 = a.G0[c + r * nch];
        const size_t o = (size_t)R.idx[base + p] + (size_t)R.idx[base + q] * G0.ld;
        double *M = R.M[r][X];
        M[p + (size_t)q * nr] = G0.re[o];
        M[(size_t)nr * nr + p + (size_t)q * nr] = G0.im[o];
    }
    __syncthreads();

    int na[2] = {0, 0}, nties[2] = {0, 0};
    for (int qb = 0; qb < kb; ++qb) {
        const int s0l = 2 * qb;   // local indices of the bond's sites: s0l, s0l + 1
        for (int r = 0; r < 2; ++r) {
            const int o = 1 - r;
            if (tid == 0) {
                // rr = I + Δ (I - Gt[sub, sub]); w = rr⁻¹ Δ
                double rr[2][2], ri[2][2];
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        double sr = dlr[r][qb][p][q] + (p == q ? 1.0 : 0.0), sim = dli[r][qb][p][q];
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            double tr, ti;
                            cmul(dlr[r][qb][p][k], dli[r][qb][p][k], gtr[GT(r, s0l + k, s0l + q)],
                                 gti[GT(r, s0l + k, s0l + q)], tr, ti);
                            sr -= tr;
                            sim -= ti;
                        }
                        rr[p][q] = sr;
                        ri[p][q] = sim;
                    }
                double ivr[2][2], ivi[2][2], dr, di;
                c22_inv(rr, ri, ivr, ivi, dr, di);
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        double sr = 0.0, sim = 0.0;
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            double tr, ti;
                            cmul(ivr[p][k], ivi[p][k], dlr[r][qb][k][q], dli[r][qb][k][q], tr, ti);
                            sr += tr;
                            sim += ti;
                        }
                        rmr[p][q] = sr;
                        rmi[p][q] = sim;
                    }
                const double pw = a.kind == 1 ? dr : dr * dr + di * di;
                prob_s = pw * pgam[r][qb];
                acc_s = 0;
            }
            __syncthreads();
            // ---- Grover factors of both regions (as k_scee_r2, from the working blocks)
            for (int it = tid; it < 2 * nX; it += BT) {
                const int t = it / nX, q = it % nX;
                const int X = q < nA ? 0 : 1, k0 = X ? nA : 0, k1 = X ? nX : nA, nr = X ? nB : nA;
                const double *Mo = R.M[o][X], *Moi = Mo + (size_t)nr * nr;
                const int ql = q - k0;
                double sr = 0.0, sim = 0.0, pr = 0.0, pi = 0.0;
                if (r == 0) {
                    // zN[t][q] = Σ_k Gt0[s_t, X_k] (2 G0o[X_k, X_q] - δ); a[u = t][q] = Σ_k G0t[X_q, s_k] w[k][t]
#pragma unroll 8
                    for (int k = k0; k < k1; ++k) {
                        const int kl = k - k0;
                        const double mr = 2.0 * Mo[kl + (size_t)ql * nr] - (k == q ? 1.0 : 0.0);
                        const double mi = 2.0 * Moi[kl + (size_t)ql * nr];
                        double tr, ti;
                        cmul(t0r[T0(r, s0l + t, k)], t0i[T0(r, s0l + t, k)], mr, mi, tr, ti);
                        sr += tr;
                        sim += ti;
                    }
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(o0r[O0(r, q, s0l + k)], o0i[O0(r, q, s0l + k)], rmr[k][t], rmi[k][t], tr, ti);
                        pr += tr;
                        pi += ti;
                    }
                } else {
                    // a[u = t][q] = Σ_k (2 G0o[X_q, X_k] - δ) G0t[X_k, s_t]; zN[t][q] = Σ_k w[t][k] Gt0[s_k, X_q]
#pragma unroll 8
                    for (int k = k0; k < k1; ++k) {
                        const int kl = k - k0;
                        const double mr = 2.0 * Mo[ql + (size_t)kl * nr] - (k == q ? 1.0 : 0.0);
                        const double mi = 2.0 * Moi[ql + (size_t)kl * nr];
                        double tr, ti;
                        cmul(mr, mi, o0r[O0(r, k, s0l + t)], o0i[O0(r, k, s0l + t)], tr, ti);
                        pr += tr;
                        pi += ti;
                    }
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(rmr[t][k], rmi[t][k], t0r[T0(r, s0l + k, q)], t0i[T0(r, s0l + k, q)], tr, ti);
                        sr += tr;
                        sim += ti;
                    }
                }
                zr[t * NX + q] = sr;
                zi[t * NX + q] = sim;
                ar_[t * NX + q] = pr;
                ai_[t * NX + q] = pi;
            }
            __syncthreads();
            // b[t][q] = Σ_k zN[t][k] gmInv_X[k][q] (k over q's region)
            for (int it = tid; it < 2 * nX; it += BT) {
                const int t = it / nX, q = it % nX;
                const int X = q < nA ? 0 : 1, off = X ? nA : 0, m = X ? nB : nA;
                const double *gr = R.gmInv[X], *gi = gr + (size_t)m * m;
                double sr = 0.0, sim = 0.0;
#pragma unroll 8
                for (int k = 0; k < m; ++k) {
                    const size_t og = (size_t)k + (size_t)(q - off) * m;
                    double tr, ti;
                    cmul(zr[t * NX + off + k], zi[t * NX + off + k], gr[og], gi[og], tr, ti);
                    sr += tr;
                    sim += ti;
                }
                br_[t * NX + q] = sr;
                bi_[t * NX + q] = sim;
            }
            __syncthreads();
            // Tau_X[t][u] = δ + Σ_{q in X} b[t][q] a[u][q]
            double v[8] = {0, 0, 0, 0, 0, 0, 0, 0}, w8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int q = tid; q < nX; q += BT) {
                const bool inA = q < nA;
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int uu = 0; uu < 2; ++uu) {
                        double tr, ti;
                        cmul(br_[t * NX + q], bi_[t * NX + q], ar_[uu * NX + q], ai_[uu * NX + q], tr, ti);
                        const int k = 2 * (t * 2 + uu);
                        v[k] += inA ? tr : 0.0;
                        v[k + 1] += inA ? ti : 0.0;
                        w8[k] += inA ? 0.0 : tr;
                        w8[k + 1] += inA ? 0.0 : ti;
                    }
            }
            block_sum8<BT>(v, red);
            block_sum8<BT>(w8, red + 8 * (BT / 64));
            if (tid == 0) {
                const int m[2] = {nA, nB};
                double dX[2];
#pragma unroll
                for (int X = 0; X < 2; ++X) {
                    double tr[2][2], ti[2][2], ir[2][2], ii[2][2], dr = 1.0, di = 0.0;
#pragma unroll
                    for (int t = 0; t < 2; ++t)
#pragma unroll
                        for (int uu = 0; uu < 2; ++uu) {
                            const int k = 2 * (t * 2 + uu);
                            tr[t][uu] = (X ? w8[k] : v[k]) + (t == uu ? 1.0 : 0.0);
                            ti[t][uu] = X ? w8[k + 1] : v[k + 1];
                        }
                    if (m[X]) c22_inv(tr, ti, ir, ii, dr, di);
                    else {
                        ir[0][0] = ir[1][1] = 1.0;
                        ir[0][1] = ir[1][0] = 0.0;
                        ii[0][0] = ii[0][1] = ii[1][0] = ii[1][1] = 0.0;
                    }
#pragma unroll
                    for (int t = 0; t < 2; ++t)
#pragma unroll
                        for (int uu = 0; uu < 2; ++uu) {
                            tinv[X][0][t][uu] = ir[t][uu];
                            tinv[X][1][t][uu] = ii[t][uu];
                        }
                    dX[X] = m[X] ? (a.kind == 1 ? dr : dr * dr + di * di) : 1.0;
                }
                const double u = pu[r][qb];
                const double p = prob_s * pow(dX[0], R.lambda) * pow(dX[1], 1.0 - R.lambda);
                const int ok = u < p;
                if (fabs(u - p) < KDQ_NEAR_TIE) ++nties[r];
                acc_s = ok;
                if (ok) {
                    R.detg[0] *= dX[0];
                    R.detg[1] *= dX[1];
                }
            }
            __syncthreads();
            if (!acc_s) continue;   // block-uniform
            // ---- accepted: GMupdate! operands, and the working rows / columns at s (before the update)
            for (int it = tid; it < 2 * nX; it += BT) {
                const int uu = it / nX, q = it % nX;
                const int X = q < nA ? 0 : 1, off = X ? nA : 0, m = X ? nB : nA;
                const double *gr = R.gmInv[X], *gi = gr + (size_t)m * m;
                double sr = 0.0, sim = 0.0;
#pragma unroll 8
                for (int k = 0; k < m; ++k) {
                    const size_t og = (size_t)(q - off) + (size_t)k * m;
                    double tr, ti;
                    cmul(gr[og], gi[og], ar_[uu * NX + off + k], ai_[uu * NX + off + k], tr, ti);
                    sr += tr;
                    sim += ti;
                }
                nr_[uu * NX + q] = sr;
                ni_[uu * NX + q] = sim;
                double wr = 0.0, wi = 0.0;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    double tr, ti;
                    cmul(tinv[X][0][uu][k], tinv[X][1][uu][k], br_[k * NX + q], bi_[k * NX + q], tr, ti);
                    wr += tr;
                    wi += ti;
                }
                wr_[uu * NX + q] = wr;
                wi_[uu * NX + q] = wi;
                cxr[uu * NX + q] = o0r[O0(r, q, s0l + uu)];
                cxi[uu * NX + q] = o0i[O0(r, q, s0l + uu)];
                rxr[uu * NX + q] = t0r[T0(r, s0l + uu, q)];
                rxi[uu * NX + q] = t0i[T0(r, s0l + uu, q)];
            }
            if (tid < 2 * KS) {
                const int t = tid / KS, i = tid % KS;
                csr[t * KS + i] = gtr[GT(r, i, s0l + t)];
                csi[t * KS + i] = gti[GT(r, i, s0l + t)];
                rsr[t * KS + i] = gtr[GT(r, s0l + t, i)];
                rsi[t * KS + i] = gti[GT(r, s0l + t, i)];
            }
            __syncthreads();
            // inverse Grover matrices: gmInv_X -= (gmInv a)(Tau⁻¹ b); the replica's region blocks
            // of G0: M_r += G0t[X, s] w Gt0[s, X]
            region_update(R, r, nA, nB, NX, nr_, ni_, wr_, wi_, cxr, cxi, rxr, rxi, rmr, rmi);
            // working blocks: with f = w (E_s - Gt[s, S]) (2 x KS) and g = w Gt0[s, X] (2 x nX),
            // Gt[S,S] -= Gt[S,s] f, G0t[X,S] -= G0t[X,s] f, Gt0[S,X] += Gt[S,s] g
            for (int e = tid; e < KS * KS; e += BT) {
                const int i = e / KS, j = e % KS;
                double sr = 0.0, sim = 0.0;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    double fr = 0.0, fi = 0.0;
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(rmr[t][k], rmi[t][k], (j == s0l + k ? 1.0 : 0.0) - rsr[k * KS + j], -rsi[k * KS + j],
                             tr, ti);
                        fr += tr;
                        fi += ti;
                    }
                    double tr, ti;
                    cmul(csr[t * KS + i], csi[t * KS + i], fr, fi, tr, ti);
                    sr += tr;
                    sim += ti;
                }
                gtr[GT(r, i, j)] -= sr;
                gti[GT(r, i, j)] -= sim;
            }
            for (int e = tid; e < nX * KS; e += BT) {
                const int x = e / KS, j = e % KS;
                double sr = 0.0, sim = 0.0;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    double fr = 0.0, fi = 0.0;
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(rmr[t][k], rmi[t][k], (j == s0l + k ? 1.0 : 0.0) - rsr[k * KS + j], -rsi[k * KS + j],
                             tr, ti);
                        fr += tr;
                        fi += ti;
                    }
                    double tr, ti;
                    cmul(cxr[t * NX + x], cxi[t * NX + x], fr, fi, tr, ti);
                    sr += tr;
                    sim += ti;
                }
                o0r[O0(r, x, j)] -= sr;
                o0i[O0(r, x, j)] -= sim;
            }
            for (int e = tid; e < KS * nX; e += BT) {
                const int i = e / nX, x = e % nX;
                double sr = 0.0, sim = 0.0;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    double gr = 0.0, gi = 0.0;
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(rmr[t][k], rmi[t][k], rxr[k * NX + x], rxi[k * NX + x], tr, ti);
                        gr += tr;
                        gi += ti;
                    }
                    double tr, ti;
                    cmul(csr[t * KS + i], csi[t * KS + i], gr, gi, tr, ti);
                    sr += tr;
                    sim += ti;
                }
                t0r[T0(r, i, x)] += sr;
                t0i[T0(r, i, x)] += sim;
            }
            // bordering of W with the 2x2 block of this bond (rows / columns nA_ and nA_ + 1)
            const int n0 = na[r];
            for (int e = tid; e < (n0 + 2) * (n0 + 2); e += BT) {
                const int t = e % (n0 + 2), u = e / (n0 + 2);
                // left factor L[t][k] (k = 0, 1): x w for old rows, w for new rows
                // right factor Rr[k][u]: y for old columns, δ for new columns
                double lr[2], li[2], rr_[2], ri_[2];
                if (t < n0) {
                    // x[t][k2] = Σ_{t'} Q[t][t'] Gc[A_t', s_k2] over t's 2x2 block
                    const int pb = t >> 1, tl = t & 1, qq = accb[r][pb];
                    double xr[2], xi[2];
#pragma unroll
                    for (int k2 = 0; k2 < 2; ++k2) {
                        double sr = 0.0, sim = 0.0;
#pragma unroll
                        for (int tp = 0; tp < 2; ++tp) {
                            double tr, ti;
                            const int ia = accl[r][2 * pb + tp];
                            cmul(qlr[r][qq][tl][tp], qli[r][qq][tl][tp], csr[k2 * KS + ia], csi[k2 * KS + ia], tr, ti);
                            sr += tr;
                            sim += ti;
                        }
                        xr[k2] = sr;
                        xi[k2] = sim;
                    }
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double sr = 0.0, sim = 0.0;
#pragma unroll
                        for (int k2 = 0; k2 < 2; ++k2) {
                            double tr, ti;
                            cmul(xr[k2], xi[k2], rmr[k2][k], rmi[k2][k], tr, ti);
                            sr += tr;
                            sim += ti;
                        }
                        lr[k] = sr;
                        li[k] = sim;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        lr[k] = rmr[t - n0][k];
                        li[k] = rmi[t - n0][k];
                    }
                }
                if (u < n0) {
                    // y[k][u] = Σ_{u'} Gc[s_k, A_u'] Δ[u'][u] over u's 2x2 block
                    const int pb = u >> 1, ul = u & 1, qq = accb[r][pb];
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double sr = 0.0, sim = 0.0;
#pragma unroll
                        for (int up = 0; up < 2; ++up) {
                            double tr, ti;
                            const int ia = accl[r][2 * pb + up];
                            cmul(rsr[k * KS + ia], rsi[k * KS + ia], dlr[r][qq][up][ul], dli[r][qq][up][ul], tr, ti);
                            sr += tr;
                            sim += ti;
                        }
                        rr_[k] = sr;
                        ri_[k] = sim;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        rr_[k] = (k == u - n0) ? 1.0 : 0.0;
                        ri_[k] = 0.0;
                    }
                }
                double vr = 0.0, vi = 0.0;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    double tr, ti;
                    cmul(lr[k], li[k], rr_[k], ri_[k], tr, ti);
                    vr += tr;
                    vi += ti;
                }
                const int w = r * KS * KS + t * KS + u;
                if (t < n0 && u < n0) {
                    Wr[w] += vr;
                    Wi[w] += vi;
                } else {
                    Wr[w] = vr;
                    Wi[w] = vi;
                }
            }
            if (tid == 0) {
                accl[r][n0] = s0l;
                accl[r][n0 + 1] = s0l + 1;
                accb[r][n0 >> 1] = qb;
                sv[r][qb] |= 1 << 16;
            }
            __syncthreads();
            na[r] += 2;
        }
    }
    for (int r = 0; r < 2; ++r) {
        const ChainDev &chr = a.chains[c + r * nch];
        if (tid == 0) {
            *chr.acc_cnt = na[r];
            chr.counters[0] += kb;
            chr.counters[1] += na[r] / 2;
            chr.counters[2] += nties[r];
        }
        if (tid < na[r]) chr.acc_idx[tid] = site[accl[r][tid]];
        for (int e = tid; e < na[r] * na[r]; e += BT) {
            const int t = e % na[r], u = e / na[r];
            chr.W[t + u * 64] = Wr[r * KS * KS + t * KS + u];
            chr.W[64 * 64 + t + u * 64] = Wi[r * KS * KS + t * KS + u];
        }
        if (tid < kb) {
            const int v = sv[r][tid];
            chr.s[(size_t)(e0 + tid) + (size_t)(lt - 1) * chr.n_elem] =
                (uint8_t)((v >> 16) ? (v >> 8) & 0xff : v & 0xff);
        }
    }
#undef GT
#undef T0
#undef O0
}

// ---- delayed blocks of the t-U-V rank-2 decisions (UpdateSCEELayer!, src/tUV/SCEE.jl:355-402)
// A block holds KB consecutive elements (directed bonds) of a slice. Consecutive elements share
// sites, so the block's sites S are the distinct sites of its elements (at most 2·KB), and one
// site may be changed more than once. The decisions are those of k_scee_r2 element by element,
// replica 1 then 2, from the working blocks Gt[S,S], Gt0[S,X], G0t[X,S] (LDS), the region
// blocks M[r][X] and gmInv, all kept current with the sequential rank-2 updates. The four full
// matrices are left to the four-matrix flush, whose form G ← G0 − G0[:,A] W (E_A − G0[A,:])
// (Big-matrix rows and columns as in k_scee_r2_block) holds for any sequence of rank-2
// updates, repeated sites included: with A the accepted elements' site pairs in acceptance
// order (repeats kept), G_(t-1)[:,a] = G0[:,(A,a)] [−x; I] and E_a − G_(t-1)[a,:] =
// [y, I] (E_(A,a) − G0[(A,a),:]), where x = W (E_A[:,a] − G0[A,a]) and y = G0[a,A] W are read
// from the block-start copy G0[S,S], so W' = [[W − x w y, −x w], [w y, w]] with w = r⁻¹Δ.
template <int KB>
__global__ __launch_bounds__(BT) void k_scee_r2_block_uv(SceeR2Args a) {
    constexpr int KS = 2 * KB;
    const int c = blockIdx.x, nch = a.nch, tid = threadIdx.x;
    const SceeRegion &R = a.reg[c];
    const ChainDev &ch0 = a.chains[c];
    const int nA = R.nA, nB = R.nB, nX = nA + nB, NX = max(a.nx_max, 1);
    const int lt = a.lt, mA = nA * nA, mAB = mA + nB * nB;
    const int e0 = a.s0;
    const int kb = min(KB, ch0.n_elem - a.s0);
    if (kb <= 0) {
        if (tid < 2) *a.chains[c + tid * nch].acc_cnt = 0;
        return;
    }
    extern __shared__ double lds[];
    double *gtr = lds, *gti = gtr + 2 * KS * KS;                 // Gt[S,S] current [r][i][j]
    double *g0r = gti + 2 * KS * KS, *g0i = g0r + 2 * KS * KS;  // Gt[S,S] at the block start
    double *t0r = g0i + 2 * KS * KS, *t0i = t0r + 2 * KS * NX;  // Gt0[S,X] [r][i][x]
    double *o0r = t0i + 2 * KS * NX, *o0i = o0r + 2 * KS * NX;  // G0t[X,S] [r][x][i]
    double *zr = o0i + 2 * KS * NX, *zi = zr + 2 * NX;          // zN [t][q]
    double *ar_ = zi + 2 * NX, *ai_ = ar_ + 2 * NX;             // a [u][q]
    double *br_ = ai_ + 2 * NX, *bi_ = br_ + 2 * NX;            // b [t][q]
    double *nr_ = bi_ + 2 * NX, *ni_ = nr_ + 2 * NX;            // gmInv a [u][q]
    double *wr_ = ni_ + 2 * NX, *wi_ = wr_ + 2 * NX;            // Tau⁻¹ b [t][q]
    double *cxr = wi_ + 2 * NX, *cxi = cxr + 2 * NX;            // G0t[X, s_t] [t][x]
    double *rxr = cxi + 2 * NX, *rxi = rxr + 2 * NX;            // Gt0[s_t, X] [t][x]
    double *csr = rxi + 2 * NX, *csi = csr + 2 * KS;            // Gt[S, s_t]  [t][i]
    double *rsr = csi + 2 * KS, *rsi = rsr + 2 * KS;            // Gt[s_t, S]  [t][j]
    double *xr = rsi + 2 * KS, *xi = xr + 2 * KS;               // x [t][k] (W rows x 2)
    double *yr = xi + 2 * KS, *yi = yr + 2 * KS;                // y [k][u] (2 x W columns)
    double *Wr = yi + 2 * KS, *Wi = Wr + 2 * KS * KS;           // W [r][t][u]
    double *red = Wi + 2 * KS * KS;                             // BT/64 x 16
    __shared__ double dgr[2][KB][2], dgi[2][KB][2];             // diagonal Δ of each element's proposal
    __shared__ double pgam[2][KB], pu[2][KB];                   // γ'/γ, uniform draw
    __shared__ int sv[2][KB], site[KS], loc[KB][2], accl[2][KS];
    __shared__ int nsl_s;
    __shared__ double rmr[2][2], rmi[2][2], tinv[2][2][2][2], prob_s;
    __shared__ int acc_s;
#define GT(r, i, j) ((r) * KS * KS + (i) * KS + (j))
#define T0(r, i, x) ((r) * KS * NX + (i) * NX + (x))
#define O0(r, x, i) ((r) * NX * KS + (x) * KS + (i))
    const uint32_t tag = (uint32_t)lt | ((uint32_t)a.dir << 31);
    const uint32_t sweep = a.sweep_dev ? *a.sweep_dev : a.sweep;
    if (tid == 0) {   // distinct sites of the block, in first-appearance order
        int ns = 0;
        for (int q = 0; q < kb; ++q)
            for (int k = 0; k < 2; ++k) {
                const int sg = k ? ch0.ey[e0 + q] : ch0.ex[e0 + q];
                int j = 0;
                while (j < ns && site[j] != sg) ++j;
                if (j == ns) site[ns++] = sg;
                loc[q][k] = j;
            }
        for (int j = ns; j < KS; ++j) site[j] = 0;
        nsl_s = ns;
    }
    if (tid < 2 * KB) {
        const int r = tid / KB, q = tid % KB;
        if (q < kb) {
            const ChainDev &chr = a.chains[c + r * nch];
            const int el = e0 + q;
            int chv;
            double u;
            proposal_draw(chr.seed, chr.chain_id, chr.stream, sweep, tag, (uint32_t)el, chv, u);
            const int si = chr.s[(size_t)el + (size_t)(lt - 1) * chr.n_elem];
            const int sx = chv + 1 + ((chv + 1) >= si ? 1 : 0);
            // Δ = diag(e^{ds} − 1, e^{a ds} − 1), phases e^{i v} for the complex case
            const double ds = chr.eta[sx - 1] - chr.eta[si - 1], v[2] = {ds, chr.a * ds};
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                if (chr.is_complex) {
                    double sn, cs;
                    sincos(v[k], &sn, &cs);
                    dgr[r][q][k] = cs - 1.0;
                    dgi[r][q][k] = sn;
                } else {
                    dgr[r][q][k] = exp(v[k]) - 1.0;
                    dgi[r][q][k] = 0.0;
                }
            }
            pgam[r][q] = chr.gamma[sx - 1] / chr.gamma[si - 1];
            pu[r][q] = u;
            sv[r][q] = si | (sx << 8);
        }
    }
    __syncthreads();
    const int nsl = nsl_s;
    for (int e = tid; e < 2 * KS * KS; e += BT) {
        const int r = e / (KS * KS), i = (e / KS) % KS, j = e % KS;
        double vr = 0.0, vi = 0.0;
        if (i < nsl && j < nsl) {
            const Opnd G = a.Gt[c + r * nch];
            const size_t o = (size_t)site[i] + (size_t)site[j] * G.ld;
            vr = G.re[o];
            vi = G.im[o];
        }
        gtr[e] = vr;
        gti[e] = vi;
        g0r[e] = vr;
        g0i[e] = vi;
        Wr[e] = 0.0;
        Wi[e] = 0.0;
    }
    for (int e = tid; e < 2 * KS * nX; e += BT) {
        const int r = e / (KS * nX), rem = e % (KS * nX), x = rem / KS, i = rem % KS;
        double tr = 0.0, ti = 0.0, orr = 0.0, oi = 0.0;
        if (i < nsl) {
            const Opnd T = a.Gt0[c + r * nch], O = a.G0t[c + r * nch];
            const int sx = R.idx[x];
            const size_t ot = (size_t)site[i] + (size_t)sx * T.ld, oo = (size_t)sx + (size_t)site[i] * O.ld;
            tr = T.re[ot]; ti = T.im[ot];
            orr = O.re[oo]; oi = O.im[oo];
        }
        t0r[T0(r, i, x)] = tr; t0i[T0(r, i, x)] = ti;
        o0r[O0(r, x, i)] = orr; o0i[O0(r, x, i)] = oi;
    }
    for (int e = tid; e < 2 * mAB; e += BT) {
        const int r = e / mAB, f = e % mAB;
        const int X = f < mA ? 0 : 1, nr = X ? nB : nA, base = X ? nA : 0, g = X ? f - mA : f;
        const int p = g % nr, q = g / nr;
        const Opnd G0 = a.G0[c + r * nch];
        const size_t o = (size_t)R.idx[base + p] + (size_t)R.idx[base + q] * G0.ld;
        double *M = R.M[r][X];
        M[p + (size_t)q * nr] = G0.re[o];
        M[(size_t)nr * nr + p + (size_t)q * nr] = G0.im[o];
    }
    __syncthreads();

    int na[2] = {0, 0}, nties[2] = {0, 0};
    for (int qb = 0; qb < kb; ++qb) {
        const int l0 = loc[qb][0], l1 = loc[qb][1];
        auto ls = [&](int k) { return k ? l1 : l0; };
        for (int r = 0; r < 2; ++r) {
            const int o = 1 - r;
            if (tid == 0) {
                // rr = I + Δ (I - Gt[sub, sub]) = (I + Δ) - Δ Gt[sub, sub] (Δ diagonal; the
                // element kernel's order of operations); w = rr⁻¹ Δ
                double rr[2][2], ri[2][2];
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const double sr = p == q ? dgr[r][qb][p] + 1.0 : 0.0, sim = p == q ? dgi[r][qb][p] : 0.0;
                        double tr, ti;
                        cmul(dgr[r][qb][p], dgi[r][qb][p], gtr[GT(r, ls(p), ls(q))], gti[GT(r, ls(p), ls(q))], tr, ti);
                        rr[p][q] = sr - tr;
                        ri[p][q] = sim - ti;
                    }
                double ivr[2][2], ivi[2][2], dr, di;
                c22_inv(rr, ri, ivr, ivi, dr, di);
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q) cmul(ivr[p][q], ivi[p][q], dgr[r][qb][q], dgi[r][qb][q], rmr[p][q], rmi[p][q]);
                const double pw = a.kind == 1 ? dr : dr * dr + di * di;
                prob_s = pw * pgam[r][qb];
                acc_s = 0;
            }
            __syncthreads();
            // ---- Grover factors of both regions (as k_scee_r2, from the working blocks)
            for (int it = tid; it < 2 * nX; it += BT) {
                const int t = it / nX, q = it % nX;
                const int X = q < nA ? 0 : 1, k0 = X ? nA : 0, k1 = X ? nX : nA, nr = X ? nB : nA;
                const double *Mo = R.M[o][X], *Moi = Mo + (size_t)nr * nr;
                const int ql = q - k0;
                double sr = 0.0, sim = 0.0, pr = 0.0, pi = 0.0;
                if (r == 0) {
#pragma unroll 8
                    for (int k = k0; k < k1; ++k) {
                        const int kl = k - k0;
                        const double mr = 2.0 * Mo[kl + (size_t)ql * nr] - (k == q ? 1.0 : 0.0);
                        const double mi = 2.0 * Moi[kl + (size_t)ql * nr];
                        double tr, ti;
                        cmul(t0r[T0(r, ls(t), k)], t0i[T0(r, ls(t), k)], mr, mi, tr, ti);
                        sr += tr;
                        sim += ti;
                    }
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(o0r[O0(r, q, ls(k))], o0i[O0(r, q, ls(k))], rmr[k][t], rmi[k][t], tr, ti);
                        pr += tr;
                        pi += ti;
                    }
                } else {
#pragma unroll 8
                    for (int k = k0; k < k1; ++k) {
                        const int kl = k - k0;
                        const double mr = 2.0 * Mo[ql + (size_t)kl * nr] - (k == q ? 1.0 : 0.0);
                        const double mi = 2.0 * Moi[ql + (size_t)kl * nr];
                        double tr, ti;
                        cmul(mr, mi, o0r[O0(r, k, ls(t))], o0i[O0(r, k, ls(t))], tr, ti);
                        pr += tr;
                        pi += ti;
                    }
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(rmr[t][k], rmi[t][k], t0r[T0(r, ls(k), q)], t0i[T0(r, ls(k), q)], tr, ti);
                        sr += tr;
                        sim += ti;
                    }
                }
                zr[t * NX + q] = sr;
                zi[t * NX + q] = sim;
                ar_[t * NX + q] = pr;
                ai_[t * NX + q] = pi;
            }
            __syncthreads();
            for (int it = tid; it < 2 * nX; it += BT) {
                const int t = it / nX, q = it % nX;
                const int X = q < nA ? 0 : 1, off = X ? nA : 0, m = X ? nB : nA;
                const double *gr = R.gmInv[X], *gi = gr + (size_t)m * m;
                double sr = 0.0, sim = 0.0;
#pragma unroll 8
                for (int k = 0; k < m; ++k) {
                    const size_t og = (size_t)k + (size_t)(q - off) * m;
                    double tr, ti;
                    cmul(zr[t * NX + off + k], zi[t * NX + off + k], gr[og], gi[og], tr, ti);
                    sr += tr;
                    sim += ti;
                }
                br_[t * NX + q] = sr;
                bi_[t * NX + q] = sim;
            }
            __syncthreads();
            double v[8] = {0, 0, 0, 0, 0, 0, 0, 0}, w8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int q = tid; q < nX; q += BT) {
                const bool inA = q < nA;
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int uu = 0; uu < 2; ++uu) {
                        double tr, ti;
                        cmul(br_[t * NX + q], bi_[t * NX + q], ar_[uu * NX + q], ai_[uu * NX + q], tr, ti);
                        const int k = 2 * (t * 2 + uu);
                        v[k] += inA ? tr : 0.0;
                        v[k + 1] += inA ? ti : 0.0;
                        w8[k] += inA ? 0.0 : tr;
                        w8[k + 1] += inA ? 0.0 : ti;
                    }
            }
            block_sum8<BT>(v, red);
            block_sum8<BT>(w8, red + 8 * (BT / 64));
            if (tid == 0) {
                const int m[2] = {nA, nB};
                double dX[2];
#pragma unroll
                for (int X = 0; X < 2; ++X) {
                    double tr[2][2], ti[2][2], ir[2][2], ii[2][2], dr = 1.0, di = 0.0;
#pragma unroll
                    for (int t = 0; t < 2; ++t)
#pragma unroll
                        for (int uu = 0; uu < 2; ++uu) {
                            const int k = 2 * (t * 2 + uu);
                            tr[t][uu] = (X ? w8[k] : v[k]) + (t == uu ? 1.0 : 0.0);
                            ti[t][uu] = X ? w8[k + 1] : v[k + 1];
                        }
                    if (m[X]) c22_inv(tr, ti, ir, ii, dr, di);
                    else {
                        ir[0][0] = ir[1][1] = 1.0;
                        ir[0][1] = ir[1][0] = 0.0;
                        ii[0][0] = ii[0][1] = ii[1][0] = ii[1][1] = 0.0;
                    }
#pragma unroll
                    for (int t = 0; t < 2; ++t)
#pragma unroll
                        for (int uu = 0; uu < 2; ++uu) {
                            tinv[X][0][t][uu] = ir[t][uu];
                            tinv[X][1][t][uu] = ii[t][uu];
                        }
                    dX[X] = m[X] ? (a.kind == 1 ? dr : dr * dr + di * di) : 1.0;
                }
                const double u = pu[r][qb];
                const double p = prob_s * pow(dX[0], R.lambda) * pow(dX[1], 1.0 - R.lambda);
                const int ok = u < p;
                if (fabs(u - p) < KDQ_NEAR_TIE) ++nties[r];
                acc_s = ok;
                if (ok) {
                    R.detg[0] *= dX[0];
                    R.detg[1] *= dX[1];
                }
            }
            __syncthreads();
            if (!acc_s) continue;   // block-uniform
            const int n0 = na[r];
            // ---- accepted: GMupdate! operands, the working rows / columns at the element's
            // sites (before the update), and the bordering vectors x, y from the block start
            for (int it = tid; it < 2 * nX; it += BT) {
                const int uu = it / nX, q = it % nX;
                const int X = q < nA ? 0 : 1, off = X ? nA : 0, m = X ? nB : nA;
                const double *gr = R.gmInv[X], *gi = gr + (size_t)m * m;
                double sr = 0.0, sim = 0.0;
#pragma unroll 8
                for (int k = 0; k < m; ++k) {
                    const size_t og = (size_t)(q - off) + (size_t)k * m;
                    double tr, ti;
                    cmul(gr[og], gi[og], ar_[uu * NX + off + k], ai_[uu * NX + off + k], tr, ti);
                    sr += tr;
                    sim += ti;
                }
                nr_[uu * NX + q] = sr;
                ni_[uu * NX + q] = sim;
                double wr = 0.0, wi = 0.0;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    double tr, ti;
                    cmul(tinv[X][0][uu][k], tinv[X][1][uu][k], br_[k * NX + q], bi_[k * NX + q], tr, ti);
                    wr += tr;
                    wi += ti;
                }
                wr_[uu * NX + q] = wr;
                wi_[uu * NX + q] = wi;
                cxr[uu * NX + q] = o0r[O0(r, q, ls(uu))];
                cxi[uu * NX + q] = o0i[O0(r, q, ls(uu))];
                rxr[uu * NX + q] = t0r[T0(r, ls(uu), q)];
                rxi[uu * NX + q] = t0i[T0(r, ls(uu), q)];
            }
            if (tid < 2 * KS) {
                const int t = tid / KS, i = tid % KS;
                csr[t * KS + i] = gtr[GT(r, i, ls(t))];
                csi[t * KS + i] = gti[GT(r, i, ls(t))];
                rsr[t * KS + i] = gtr[GT(r, ls(t), i)];
                rsi[t * KS + i] = gti[GT(r, ls(t), i)];
            }
            if (tid < 2 * n0) {
                // x[t][k] = Σ_u W[t][u] (δ(A_u, a_k) − G0[A_u, a_k]); y[k][u] = Σ_t G0[a_k, A_t] W[t][u]
                const int t = tid / 2, k = tid % 2;
                double sxr = 0.0, sxi = 0.0, syr = 0.0, syi = 0.0;
                for (int u = 0; u < n0; ++u) {
                    const int au = accl[r][u];
                    double tr, ti;
                    cmul(Wr[GT(r, t, u)], Wi[GT(r, t, u)], (au == ls(k) ? 1.0 : 0.0) - g0r[GT(r, au, ls(k))],
                         -g0i[GT(r, au, ls(k))], tr, ti);
                    sxr += tr;
                    sxi += ti;
                    // y[k][t]: here t plays the column index of W
                    cmul(g0r[GT(r, ls(k), au)], g0i[GT(r, ls(k), au)], Wr[GT(r, u, t)], Wi[GT(r, u, t)], tr, ti);
                    syr += tr;
                    syi += ti;
                }
                xr[t * 2 + k] = sxr;
                xi[t * 2 + k] = sxi;
                yr[k * KS + t] = syr;
                yi[k * KS + t] = syi;
            }
            __syncthreads();
            region_update(R, r, nA, nB, NX, nr_, ni_, wr_, wi_, cxr, cxi, rxr, rxi, rmr, rmi);
            // working blocks: with f = w (E_s - Gt[s, S]) and g = w Gt0[s, X],
            // Gt[S,S] -= Gt[S,s] f, G0t[X,S] -= G0t[X,s] f, Gt0[S,X] += Gt[S,s] g
            for (int e = tid; e < KS * KS; e += BT) {
                const int i = e / KS, j = e % KS;
                double sr = 0.0, sim = 0.0;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    double fr = 0.0, fi = 0.0;
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(rmr[t][k], rmi[t][k], (j == ls(k) ? 1.0 : 0.0) - rsr[k * KS + j], -rsi[k * KS + j], tr, ti);
                        fr += tr;
                        fi += ti;
                    }
                    double tr, ti;
                    cmul(csr[t * KS + i], csi[t * KS + i], fr, fi, tr, ti);
                    sr += tr;
                    sim += ti;
                }
                gtr[GT(r, i, j)] -= sr;
                gti[GT(r, i, j)] -= sim;
            }
            for (int e = tid; e < nX * KS; e += BT) {
                const int x = e / KS, j = e % KS;
                double sr = 0.0, sim = 0.0;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    double fr = 0.0, fi = 0.0;
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(rmr[t][k], rmi[t][k], (j == ls(k) ? 1.0 : 0.0) - rsr[k * KS + j], -rsi[k * KS + j], tr, ti);
                        fr += tr;
                        fi += ti;
                    }
                    double tr, ti;
                    cmul(cxr[t * NX + x], cxi[t * NX + x], fr, fi, tr, ti);
                    sr += tr;
                    sim += ti;
                }
                o0r[O0(r, x, j)] -= sr;
                o0i[O0(r, x, j)] -= sim;
            }
            for (int e = tid; e < KS * nX; e += BT) {
                const int i = e / nX, x = e % nX;
                double sr = 0.0, sim = 0.0;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    double gr = 0.0, gi = 0.0;
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(rmr[t][k], rmi[t][k], rxr[k * NX + x], rxi[k * NX + x], tr, ti);
                        gr += tr;
                        gi += ti;
                    }
                    double tr, ti;
                    cmul(csr[t * KS + i], csi[t * KS + i], gr, gi, tr, ti);
                    sr += tr;
                    sim += ti;
                }
                t0r[T0(r, i, x)] += sr;
                t0i[T0(r, i, x)] += sim;
            }
            // W' = [[W − x w y, −x w], [w y, w]] (rows / columns n0, n0 + 1 new)
            for (int e = tid; e < (n0 + 2) * (n0 + 2); e += BT) {
                const int t = e % (n0 + 2), u = e / (n0 + 2);
                double vr = 0.0, vi = 0.0;
                if (t < n0 && u < n0) {
                    vr = Wr[GT(r, t, u)];
                    vi = Wi[GT(r, t, u)];
#pragma unroll
                    for (int k = 0; k < 2; ++k)
#pragma unroll
                        for (int k2 = 0; k2 < 2; ++k2) {
                            double pr, pi, tr, ti;
                            cmul(xr[t * 2 + k], xi[t * 2 + k], rmr[k][k2], rmi[k][k2], pr, pi);
                            cmul(pr, pi, yr[k2 * KS + u], yi[k2 * KS + u], tr, ti);
                            vr -= tr;
                            vi -= ti;
                        }
                } else if (t < n0) {
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(xr[t * 2 + k], xi[t * 2 + k], rmr[k][u - n0], rmi[k][u - n0], tr, ti);
                        vr -= tr;
                        vi -= ti;
                    }
                } else if (u < n0) {
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        double tr, ti;
                        cmul(rmr[t - n0][k], rmi[t - n0][k], yr[k * KS + u], yi[k * KS + u], tr, ti);
                        vr += tr;
                        vi += ti;
                    }
                } else {
                    vr = rmr[t - n0][u - n0];
                    vi = rmi[t - n0][u - n0];
                }
                Wr[GT(r, t, u)] = vr;
                Wi[GT(r, t, u)] = vi;
            }
            if (tid == 0) {
                accl[r][n0] = l0;
                accl[r][n0 + 1] = l1;
                sv[r][qb] |= 1 << 16;
            }
            __syncthreads();
            na[r] += 2;
        }
    }
    for (int r = 0; r < 2; ++r) {
        const ChainDev &chr = a.chains[c + r * nch];
        if (tid == 0) {
            *chr.acc_cnt = na[r];
            chr.counters[0] += kb;
            chr.counters[1] += na[r] / 2;
            chr.counters[2] += nties[r];
        }
        if (tid < na[r]) chr.acc_idx[tid] = site[accl[r][tid]];
        for (int e = tid; e < na[r] * na[r]; e += BT) {
            const int t = e % na[r], u = e / na[r];
            chr.W[t + u * 64] = Wr[r * KS * KS + t * KS + u];
            chr.W[64 * 64 + t + u * 64] = Wi[r * KS * KS + t * KS + u];
        }
        if (tid < kb) {
            const int v = sv[r][tid];
            chr.s[(size_t)(e0 + tid) + (size_t)(lt - 1) * chr.n_elem] =
                (uint8_t)((v >> 16) ? (v >> 8) & 0xff : v & 0xff);
        }
    }
#undef GT
#undef T0
#undef O0
}

__global__ void k_scee_detg_signed(const SceeRegion *reg, int nch, const double *la, const double *lb,
                                   const double *pa, const double *pb) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    const SceeRegion &R = reg[c];
    R.detg[0] = R.nA ? exp(la[c]) * pa[2 * c] : 1.0;
    R.detg[1] = R.nB ? exp(lb[c]) * pb[2 * c] : 1.0;
}

}  // namespace

hipError_t launch_scee_accum(hipStream_t st, int nch, const SceeRegion *reg) {
    hipLaunchKernelGGL(k_scee_accum, dim3((nch + 63) / 64), dim3(64), 0, st, reg, nch);
    return hipGetLastError();
}

hipError_t launch_diag(hipStream_t st, int nchains, int max_n, const ChainDev *chains, const Opnd *X, int ndim,
                       double v, int set) {
    dim3 grid((max_n + 63) / 64, (max_n + 3) / 4, nchains);
    hipLaunchKernelGGL(k_diag, grid, dim3(256), 0, st, chains, X, ndim, v, set);
    return hipGetLastError();
}

size_t scee_r2_lds_bytes(int nx_max, int max_ns) {
    const size_t nxm = nx_max > 1 ? nx_max : 1;
    return (20 * nxm + 16 * (size_t)max_ns + 8 * 2 * (RT / 64)) * sizeof(double);
}

hipError_t launch_scee_detg_signed(hipStream_t st, int nch, const SceeRegion *reg, const double *logdet_a,
                                   const double *logdet_b, const double *phase_a, const double *phase_b) {
    hipLaunchKernelGGL(k_scee_detg_signed, dim3((nch + 63) / 64), dim3(64), 0, st, reg, nch, logdet_a, logdet_b,
                       phase_a, phase_b);
    return hipGetLastError();
}

size_t scee_r2_block_lds_bytes(int kb, int nx_max) {
    const size_t ks = 2 * (size_t)kb, nx = nx_max > 1 ? nx_max : 1;
    return (8 * ks * ks + 8 * ks * nx + 28 * nx + 8 * ks + 16 * (BT / 64)) * sizeof(double);
}

hipError_t launch_scee_r2_block(hipStream_t st, const SceeR2Args &a, int kb) {
    if (kb != 4 && kb != 8) return hipErrorInvalidValue;
    const size_t bytes = scee_r2_block_lds_bytes(kb, a.nx_max);
    if (bytes > 150 * 1024) return hipErrorInvalidValue;
    const void *f = kb == 8 ? reinterpret_cast<const void *>(k_scee_r2_block<8>)
                            : reinterpret_cast<const void *>(k_scee_r2_block<4>);
    if (hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)) return e;
    if (kb == 8) hipLaunchKernelGGL(k_scee_r2_block<8>, dim3(a.nch), dim3(BT), bytes, st, a);
    else hipLaunchKernelGGL(k_scee_r2_block<4>, dim3(a.nch), dim3(BT), bytes, st, a);
    return hipGetLastError();
}

size_t scee_r2_block_uv_lds_bytes(int kb, int nx_max) {
    const size_t ks = 2 * (size_t)kb, nx = nx_max > 1 ? nx_max : 1;
    return (12 * ks * ks + 8 * ks * nx + 28 * nx + 16 * ks + 16 * (BT / 64)) * sizeof(double);
}

hipError_t launch_scee_r2_block_uv(hipStream_t st, const SceeR2Args &a, int kb) {
    if (kb != 4 && kb != 8) return hipErrorInvalidValue;
    const size_t bytes = scee_r2_block_uv_lds_bytes(kb, a.nx_max);
    if (bytes > 150 * 1024) return hipErrorInvalidValue;
    const void *f = kb == 8 ? reinterpret_cast<const void *>(k_scee_r2_block_uv<8>)
                            : reinterpret_cast<const void *>(k_scee_r2_block_uv<4>);
    if (hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)) return e;
    if (kb == 8) hipLaunchKernelGGL(k_scee_r2_block_uv<8>, dim3(a.nch), dim3(BT), bytes, st, a);
    else hipLaunchKernelGGL(k_scee_r2_block_uv<4>, dim3(a.nch), dim3(BT), bytes, st, a);
    return hipGetLastError();
}

hipError_t launch_scee_r2(hipStream_t st, const SceeR2Args &a) {
    const size_t bytes = scee_r2_lds_bytes(a.nx_max, a.max_ns);
    if (bytes > 163840) return hipErrorInvalidValue;
    if (hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_scee_r2),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes))
        return e;
    hipLaunchKernelGGL(k_scee_r2, dim3(a.nch), dim3(RT), bytes, st, a);
    return hipGetLastError();
}
