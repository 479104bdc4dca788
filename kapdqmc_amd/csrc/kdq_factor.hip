// FP64 factorizations of the stabilisation step (complex, planar storage, column-major).
//
// Orthonormalisation of the N x Np propagators (the role of geqrf+orgqr / gerqf+orgrq at
// src/tU/phy_update.jl:44-45,85-86) by two-pass Cholesky QR: S = Uᴴ U (MFMA GEMM),
// S + σI = L Lᴴ (k_chol), U <- U L⁻ᴴ (k_trsm_rlh), then once more without the shift. Only
// span(U) enters G = I - U_R (U_L U_R)⁻¹ U_L (src/public/Gupdate.jl:139-157), and U L⁻ᴴ has
// exactly the span of U for any invertible L, so the result equals the Householder one up
// to rounding. The first pass is shifted (σ = c·u·‖U‖_F², Fukaya et al., SIAM J. Sci.
// Comput. 42 (2020) A477) so that its Cholesky cannot break down for cond(U) < ~1e12; the
// second pass restores orthonormality to O(u).
// Overlap solve of get_G!: blocked LU with partial pivoting (k_lu, getrf semantics) and
// X = M⁻¹ U_L (k_lu_solve, getrs semantics), log|det M| from the diagonal of U.
//
// Each factorization runs in one workgroup per chain with its current 16-wide panel in LDS;
// the triangular solves run over 16-row (16-column) slabs held in LDS, one workgroup each.
#include "kdq_kernels.h"

namespace {

constexpr int NT = 256;     // threads of the triangular solves
constexpr int NF = 1024;    // threads of the one-workgroup Cholesky (4 waves per SIMD hide the
                            // FP64 dependency latency that one wave per SIMD exposes)
constexpr int NB = 16;      // panel / slab width
constexpr int PS = NB + 1;  // padded LDS row stride of a panel (doubles)
constexpr int RB = 16;      // rows (right-hand sides) per workgroup of the solves
constexpr int TB = 8;       // trailing-update entries per thread whose loads are issued together
constexpr int NL = 1024;    // threads of the one-workgroup LU
constexpr int LB = 4;       // its trailing-update entries per thread per batch

template <int T>
KDQ_DEV double blk_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < T / 64; ++w) s += red[w];
    return s;
}

// Inverse of the panel's kb x kb diagonal block (LDS rows [r * PS + q]) by row sweeps: thread
// (r, q) = (tid / NB, tid % NB), tid < NB², keeps X[r][q] in registers; row i is final after
// step i and is published in LDS (X: real [i * NB + q], then imaginary) to the rows it feeds.
// LU = false: the lower Cholesky factor (real positive diagonal), to columns 0..NB-1 of D.
// LU = true: the unit-lower factor (forward, X, columns 0..NB-1) and the upper factor
// (backward in the same steps, X + 2 NB², columns NB..2NB-1). Entries outside the triangle
// come out as exact zeros. Every thread of the block calls it (one barrier per step).
template <bool LU>
KDQ_DEV void diag_inv(const double *Pr, const double *Pi, double *X, int kb, int k0, const Opnd &D) {
    const int tid = threadIdx.x, r = tid / NB, q = tid % NB;
    const bool act = r < kb && q < kb;
    double *Xi = X + NB * NB, *Yr = X + 2 * NB * NB, *Yi = X + 3 * NB * NB;
    double lr = r == q ? 1.0 : 0.0, li = 0.0, ur = lr, ui = 0.0;
    for (int s = 0; s < kb; ++s) {
        const int i = kb - 1 - s;
        if (act && r == s) {
            double xr = lr, xi = li;
            if (!LU) {
                const double id = 1.0 / Pr[s * PS + s];
                xr *= id;
                xi *= id;
            }
            X[s * NB + q] = xr;
            Xi[s * NB + q] = xi;
            const size_t o = (size_t)(k0 + s) + (size_t)q * D.ld;
            D.re[o] = xr;
            D.im[o] = xi;
        }
        if (LU && act && r == i) {
            const double dr = Pr[i * PS + i], di = Pi[i * PS + i], d2 = dr * dr + di * di;
            double xr, xi;
            cmul(ur, ui, dr / d2, -di / d2, xr, xi);
            Yr[i * NB + q] = xr;
            Yi[i * NB + q] = xi;
            const size_t o = (size_t)(k0 + i) + (size_t)(NB + q) * D.ld;
            D.re[o] = xr;
            D.im[o] = xi;
        }
        __syncthreads();
        double tr_, ti_;
        if (act && r > s) {
            cmul(Pr[r * PS + s], Pi[r * PS + s], X[s * NB + q], Xi[s * NB + q], tr_, ti_);
            lr -= tr_;
            li -= ti_;
        }
        if (LU && act && r < i) {
            cmul(Pr[r * PS + i], Pi[r * PS + i], Yr[i * NB + q], Yi[i * NB + q], tr_, ti_);
            ur -= tr_;
            ui -= ti_;
        }
    }
}

// Inverses of the 16 x 16 diagonal blocks of a factored n x n matrix (lower Cholesky factor,
// or the unit-lower and upper LU factors) into D: one workgroup per block and chain.
template <bool LU>
__global__ __launch_bounds__(NT) void k_diag_inv(const ChainDev *chains, const Opnd *A_, int ndim, const Opnd *D_) {
    const int c = blockIdx.y, k0 = blockIdx.x * NB;
    const int n = dim_of(chains[c], ndim);
    if (k0 >= n) return;   // uniform per workgroup
    const int kb = min(NB, n - k0);
    const Opnd A = A_[c];
    __shared__ double Pr[NB * PS], Pi[NB * PS], X[4 * NB * NB];
    const int tid = threadIdx.x, r = tid % NB, q = tid / NB;
    if (r < kb && q < kb) {
        const size_t o = (size_t)(k0 + r) + (size_t)(k0 + q) * A.ld;
        Pr[r * PS + q] = A.re[o];
        Pi[r * PS + q] = A.im[o];
    }
    __syncthreads();
    diag_inv<LU>(Pr, Pi, X, kb, k0, D_[c]);
}

// In-place Cholesky S + σI = L Lᴴ of the Hermitian n x n matrix (lower triangle read and
// written), σ = shift · trace(S). info (sticky) receives j+1 for a non-positive pivot j.
__global__ __launch_bounds__(NF) void k_chol(CholArgs a) {
    const int c = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = dim_of(ch, a.ndim);
    const Opnd A = a.A[c];
    const size_t ld = A.ld;
    double *Re = A.re, *Im = A.im;
    extern __shared__ double lds[];
    double *Pr = lds, *Pi = lds + (size_t)n * PS;  // panel rows k0.., [r][q]
    __shared__ double red[NF / 64];
    __shared__ int bad;
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    double tr = 0.0;
    for (int i = tid; i < n; i += NF) tr += Re[i + (size_t)i * ld];
    const double sigma = a.shift * blk_sum<NF>(tr, red);
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0), m = n - k0;
        __syncthreads();
        for (int e = tid; e < m * kb; e += NF) {
            const int r = e % m, q = e / m;
            const size_t o = (k0 + r) + (size_t)(k0 + q) * ld;
            Pr[r * PS + q] = Re[o] + (r == q ? sigma : 0.0);
            Pi[r * PS + q] = Im[o];
        }
        __syncthreads();
        for (int j = 0; j < kb; ++j) {
            const double djj = Pr[j * PS + j];
            __syncthreads();
            if (!(djj > 0.0) && tid == 0 && !bad) bad = k0 + j + 1;
            const double d = sqrt(fmax(djj, 1e-300)), id = 1.0 / d;
            for (int r = j + tid; r < m; r += NF) {
                if (r == j) {
                    Pr[j * PS + j] = d;
                    Pi[j * PS + j] = 0.0;
                } else {
                    Pr[r * PS + j] *= id;
                    Pi[r * PS + j] *= id;
                }
            }
            __syncthreads();
            const int nq = kb - j - 1;
            for (int e = tid; e < (m - j - 1) * nq; e += NF) {
                const int r = j + 1 + e / nq, q = j + 1 + e % nq;
                if (r < q) continue;
                double tr_, ti_;
                cmul(Pr[r * PS + j], Pi[r * PS + j], Pr[q * PS + j], -Pi[q * PS + j], tr_, ti_);
                Pr[r * PS + q] -= tr_;
                Pi[r * PS + q] -= ti_;
            }
            __syncthreads();
        }
        for (int e = tid; e < m * kb; e += NF) {
            const int r = e % m, q = e / m;
            if (r < q) continue;
            const size_t o = (k0 + r) + (size_t)(k0 + q) * ld;
            Re[o] = Pr[r * PS + q];
            Im[o] = Pi[r * PS + q];
        }
        // trailing Hermitian update (lower): S[i][j] -= Σ_q P[i][q] conj(P[j][q]), i >= j.
        // TB entries per thread at a time, their loads issued before any store (the stores
        // may alias later loads as far as the compiler knows, which would serialise them)
        const int mt = m - kb;
        for (int e0 = tid; e0 < mt * mt; e0 += NF * TB) {
            double gr[TB], gi[TB];
            size_t off[TB];
            bool ok[TB];
#pragma unroll
            for (int b = 0; b < TB; ++b) {
                const int e = e0 + b * NF, i = kb + e % mt, j = kb + e / mt;
                ok[b] = e < mt * mt && i >= j;
                off[b] = ok[b] ? (k0 + i) + (size_t)(k0 + j) * ld : 0;
                gr[b] = ok[b] ? Re[off[b]] : 0.0;
                gi[b] = ok[b] ? Im[off[b]] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < TB; ++b) {
                if (!ok[b]) continue;
                const int e = e0 + b * NF, i = kb + e % mt, j = kb + e / mt;
                double sr = 0.0, si = 0.0;
#pragma unroll 4
                for (int q = 0; q < kb; ++q) {
                    const double ar = Pr[i * PS + q], ai = Pi[i * PS + q];
                    const double br = Pr[j * PS + q], bi = -Pi[j * PS + q];
                    sr += ar * br - ai * bi;
                    si += ar * bi + ai * br;
                }
                gr[b] -= sr;
                gi[b] -= si;
            }
#pragma unroll
            for (int b = 0; b < TB; ++b)
                if (ok[b]) {
                    Re[off[b]] = gr[b];
                    Im[off[b]] = gi[b];
                }
        }
    }
    __syncthreads();
    if (tid == 0 && bad && a.info && !a.info[c]) a.info[c] = bad;
    if (a.logdiag) {
        double lsum = 0.0;
        for (int i = tid; i < n; i += NF) lsum += log(Re[i + (size_t)i * ld]);
        lsum = blk_sum<NF>(lsum, red);
        if (tid == 0) a.logdiag[c] += lsum;
    }
}

// X <- X L⁻ᴴ for the m x n matrix X (in place; L lower, real positive diagonal):
// X[r][j] = (X[r][j] - Σ_{i<j} X[r][i] conj(L[j][i])) / L[j][j]. One workgroup per 16 rows;
// the rows live in LDS, the current 16-column slab of L is staged per step.
__global__ __launch_bounds__(NT) void k_trsm_rlh(TrsmArgs t) {
    const int c = blockIdx.y;
    const ChainDev &ch = t.chains[c];
    const int m = dim_of(ch, t.mdim), n = dim_of(ch, t.ndim);
    const int r0 = blockIdx.x * RB;
    if (r0 >= m) return;
    const int rb = min(RB, m - r0);
    const Opnd X = t.X[c], L = t.L[c];
    extern __shared__ double lds[];
    double *Xr = lds, *Xi = lds + (size_t)n * RB;                      // [j][r]
    double *Lr = lds + 2 * (size_t)n * RB, *Li = Lr + (size_t)n * PS;  // [jj][i]
    double *Dr_ = Li + (size_t)n * PS, *Di_ = Dr_ + NB * PS;           // block of L⁻¹ [jj][i]
    const int tid = threadIdx.x;
    for (int e = tid; e < n * RB; e += NT) {
        const int r = e % RB, j = e / RB;
        double vr = 0.0, vi = 0.0;
        if (r < rb) {
            const size_t o = (size_t)(r0 + r) + (size_t)j * X.ld;
            vr = X.re[o];
            vi = X.im[o];
        }
        Xr[e] = vr;
        Xi[e] = vi;
    }
    for (int b0 = 0; b0 < n; b0 += NB) {
        const int bw = min(NB, n - b0), rest = n - b0;
        __syncthreads();
        for (int e = tid; e < rest * bw; e += NT) {
            const int jj = e % rest, i = e / rest;
            const size_t o = (size_t)(b0 + jj) + (size_t)(b0 + i) * L.ld;
            Lr[jj * PS + i] = L.re[o];
            Li[jj * PS + i] = L.im[o];
        }
        if (t.Dinv)   // the slab's block of L⁻¹ (lower, written by k_chol)
            for (int e = tid; e < bw * bw; e += NT) {
                const int jj = e % bw, i = e / bw;
                const Opnd D = t.Dinv[c];
                const size_t o = (size_t)(b0 + jj) + (size_t)i * D.ld;
                Dr_[jj * PS + i] = D.re[o];
                Di_[jj * PS + i] = D.im[o];
            }
        __syncthreads();
        if (t.Dinv) {
            // X_slab <- X_slab (L_ss⁻¹)ᴴ: X[r][j] = Σ_{i <= j} X[r][i] conj(L⁻¹[j][i]), all (r, j) at once
            const int r = tid % RB, j = tid / RB;
            double sr = 0.0, si = 0.0;
            if (j < bw)
                for (int i = 0; i <= j; ++i) {
                    const double xr = Xr[(b0 + i) * RB + r], xi = Xi[(b0 + i) * RB + r];
                    const double lr = Dr_[j * PS + i], li = -Di_[j * PS + i];
                    sr += xr * lr - xi * li;
                    si += xr * li + xi * lr;
                }
            __syncthreads();
            if (j < bw) {
                Xr[(b0 + j) * RB + r] = sr;
                Xi[(b0 + j) * RB + r] = si;
            }
        } else if (tid < RB) {
            const int r = tid;
            for (int j = 0; j < bw; ++j) {
                double sr = Xr[(b0 + j) * RB + r], si = Xi[(b0 + j) * RB + r];
                for (int i = 0; i < j; ++i) {
                    const double xr = Xr[(b0 + i) * RB + r], xi = Xi[(b0 + i) * RB + r];
                    const double lr = Lr[j * PS + i], li = -Li[j * PS + i];
                    sr -= xr * lr - xi * li;
                    si -= xr * li + xi * lr;
                }
                const double id = 1.0 / Lr[j * PS + j];
                Xr[(b0 + j) * RB + r] = sr * id;
                Xi[(b0 + j) * RB + r] = si * id;
            }
        }
        __syncthreads();
        const int nrest = rest - bw;
        for (int e = tid; e < nrest * RB; e += NT) {
            const int r = e % RB, jj = bw + e / RB;
            double sr = 0.0, si = 0.0;
#pragma unroll 4
            for (int i = 0; i < bw; ++i) {
                const double xr = Xr[(b0 + i) * RB + r], xi = Xi[(b0 + i) * RB + r];
                const double lr = Lr[jj * PS + i], li = -Li[jj * PS + i];
                sr += xr * lr - xi * li;
                si += xr * li + xi * lr;
            }
            Xr[(b0 + jj) * RB + r] -= sr;
            Xi[(b0 + jj) * RB + r] -= si;
        }
    }
    __syncthreads();
    for (int e = tid; e < n * RB; e += NT) {
        const int r = e % RB, j = e / RB;
        if (r >= rb) continue;
        const size_t o = (size_t)(r0 + r) + (size_t)j * X.ld;
        X.re[o] = Xr[e];
        X.im[o] = Xi[e];
    }
}

// In-place LU with partial pivoting of the n x n matrix, P A = L U (getrf semantics: the
// interchanges are applied to whole rows; pivots stored 0-based, row k swapped with piv[k]).
// logdet receives log|det A|; info (sticky) receives j+1 for an exactly zero pivot j.
__global__ __launch_bounds__(NL) void k_lu(LuArgs a) {
    const int c = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = dim_of(ch, a.ndim);
    const Opnd A = a.A[c];
    const size_t ld = A.ld;
    double *Re = A.re, *Im = A.im;
    int *piv = a.piv + (size_t)c * a.piv_ld;
    extern __shared__ double lds[];
    double *Pr = lds, *Pi = lds + (size_t)n * PS;                      // panel [r][q]
    double *Ur = lds + 2 * (size_t)n * PS, *Ui = Ur + (size_t)n * PS;  // U12 [col][j]
    __shared__ int lpiv[NB], bad;
    __shared__ double s_logdet, s_phr, s_phi;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        bad = 0;
        s_logdet = 0.0;
        s_phr = 1.0;
        s_phi = 0.0;
    }
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0), m = n - k0;
        __syncthreads();
        for (int e = tid; e < m * kb; e += NL) {
            const int r = e % m, q = e / m;
            const size_t o = (k0 + r) + (size_t)(k0 + q) * ld;
            Pr[r * PS + q] = Re[o];
            Pi[r * PS + q] = Im[o];
        }
        __syncthreads();
        // panel factorisation by wave 0 alone (wave-synchronous: no workgroup barriers in the
        // column loop; LDS accesses of one wave execute in program order)
        if (wave == 0) {
            for (int j = 0; j < kb; ++j) {
                // pivot: first index of max |P[r][j]|², r >= j
                double best = -1.0;
                int bi = m;
                for (int r = j + lane; r < m; r += 64) {
                    const double v = Pr[r * PS + j] * Pr[r * PS + j] + Pi[r * PS + j] * Pi[r * PS + j];
                    if (v > best) {
                        best = v;
                        bi = r;
                    }
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const double ob = __shfl_xor(best, off, 64);
                    const int oi = __shfl_xor(bi, off, 64);
                    if (ob > best || (ob == best && oi < bi)) {
                        best = ob;
                        bi = oi;
                    }
                }
                const int p = bi;   // the same in every lane after the butterfly
                if (lane == 0) {
                    lpiv[j] = p;
                    piv[k0 + j] = k0 + p;
                }
                if (p != j && lane < kb) {
                    const double ar = Pr[j * PS + lane], ai = Pi[j * PS + lane];
                    Pr[j * PS + lane] = Pr[p * PS + lane];
                    Pi[j * PS + lane] = Pi[p * PS + lane];
                    Pr[p * PS + lane] = ar;
                    Pi[p * PS + lane] = ai;
                }
                __builtin_amdgcn_wave_barrier();
                const double pr = Pr[j * PS + j], pim = Pi[j * PS + j];
                const double p2 = pr * pr + pim * pim;
                if (p2 == 0.0) {   // singular: leave the column, flag (uniform branch)
                    if (lane == 0 && !bad) bad = k0 + j + 1;
                    continue;
                }
                const double ir = pr / p2, ii = -pim / p2;
                for (int r = j + 1 + lane; r < m; r += 64) {
                    double tr_, ti_;
                    cmul(Pr[r * PS + j], Pi[r * PS + j], ir, ii, tr_, ti_);
                    Pr[r * PS + j] = tr_;
                    Pi[r * PS + j] = ti_;
                }
                __builtin_amdgcn_wave_barrier();
                const int nq = kb - j - 1;
                for (int e = lane; e < (m - j - 1) * nq; e += 64) {
                    const int r = j + 1 + e / nq, q = j + 1 + e % nq;
                    double tr_, ti_;
                    cmul(Pr[r * PS + j], Pi[r * PS + j], Pr[j * PS + q], Pi[j * PS + q], tr_, ti_);
                    Pr[r * PS + q] -= tr_;
                    Pi[r * PS + q] -= ti_;
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
        // panel (L21, U11 and L11) back to memory; log|det| of this panel's pivots
        for (int e = tid; e < m * kb; e += NL) {
            const int r = e % m, q = e / m;
            const size_t o = (k0 + r) + (size_t)(k0 + q) * ld;
            Re[o] = Pr[r * PS + q];
            Im[o] = Pi[r * PS + q];
        }
        if (tid == 0) {
            double s = 0.0, phr = s_phr, phi = s_phi;
            for (int j = 0; j < kb; ++j) {
                const double ur = Pr[j * PS + j], ui = Pi[j * PS + j], m2 = ur * ur + ui * ui;
                s += 0.5 * log(m2);
                // phase of det: the pivots' phases, a sign per row interchange
                const double im = m2 > 0.0 ? 1.0 / sqrt(m2) : 0.0;
                const double nr = phr * ur * im - phi * ui * im, ni = phr * ui * im + phi * ur * im;
                phr = lpiv[j] != j ? -nr : nr;
                phi = lpiv[j] != j ? -ni : ni;
            }
            s_logdet += s;
            s_phr = phr;
            s_phi = phi;
        }
        // the panel's interchanges on every column outside it, as one row permutation of the
        // (at most 2 kb) rows they touch: src[t] is the original row (panel-relative) that ends in
        // row rows[t]; every entry is loaded before any is stored
        // The rows: 0 .. kb-1 (thread t < kb), then each distinct pivot row >= kb (thread kb + j
        // for the first j that names it), compacted in order; every thread of the first 2 kb
        // follows its row back through the swaps with the pivots in registers.
        __shared__ int rows[2 * NB], src[2 * NB], nrows;
        if (tid < 2 * NB) {
            int lp[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) lp[j] = j < kb ? lpiv[j] : j;
            int row = -1, pos = -1;
            if (tid < kb) {
                row = tid;
                pos = tid;
            } else if (tid < 2 * kb) {
                const int j = tid - kb, p = lpiv[j];
                bool rep = p >= kb;
                int before = 0;   // representatives among the earlier pivots
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    if (i < j) {
                        rep = rep && lp[i] != p;
                        bool ri = lp[i] >= kb;
#pragma unroll
                        for (int i2 = 0; i2 < NB; ++i2)
                            if (i2 < i) ri = ri && lp[i2] != lp[i];
                        before += ri ? 1 : 0;
                    }
                }
                if (rep) {
                    row = p;
                    pos = kb + before;
                }
            }
            if (row >= 0) {
                int r = row;
#pragma unroll
                for (int j = NB - 1; j >= 0; --j) {
                    if (j >= kb) continue;
                    if (r == j) r = lp[j];
                    else if (r == lp[j]) r = j;
                }
                rows[pos] = row;
                src[pos] = r;
            }
            if (tid == 0) {   // the count of distinct pivot rows >= kb
                int cnt = kb;
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    bool ri = i < kb && lp[i] >= kb;
#pragma unroll
                    for (int i2 = 0; i2 < NB; ++i2)
                        if (i2 < i) ri = ri && lp[i2] != lp[i];
                    cnt += ri ? 1 : 0;
                }
                nrows = cnt;
            }
        }
        __syncthreads();
        {
            // whole columns per round (cpi of them, nr rows each), so that every row of a column
            // is read before any of its rows is written; the same number of rounds in every thread
            const int nr = nrows, nout = n - kb;   // columns outside the panel
            const int cpi = nr > 0 ? (NL * LB) / nr : nout;
            for (int c0 = 0; c0 < nout && nr > 0; c0 += cpi) {
                double vr[LB], vi[LB];
#pragma unroll
                for (int b = 0; b < LB; ++b) {
                    const int e = tid + b * NL, t = e % nr, cq = c0 + e / nr, col = cq < k0 ? cq : cq + kb;
                    const bool ok = e < cpi * nr && cq < nout;
                    const size_t o = ok ? (k0 + src[t]) + (size_t)col * ld : 0;
                    vr[b] = ok ? Re[o] : 0.0;
                    vi[b] = ok ? Im[o] : 0.0;
                }
                __syncthreads();
#pragma unroll
                for (int b = 0; b < LB; ++b) {
                    const int e = tid + b * NL, t = e % nr, cq = c0 + e / nr, col = cq < k0 ? cq : cq + kb;
                    if (e >= cpi * nr || cq >= nout) continue;
                    const size_t o = (k0 + rows[t]) + (size_t)col * ld;
                    Re[o] = vr[b];
                    Im[o] = vi[b];
                }
                __syncthreads();
            }
        }
        __syncthreads();
        // U12 = L11⁻¹ A12 (unit lower), one thread per column: its kb entries loaded together
        // into the thread's LDS column (no global store between them), solved there, stored
        const int ncol = n - k0 - kb;
        for (int q = tid; q < ncol; q += NL) {
            const size_t col = (size_t)(k0 + kb + q) * ld;
#pragma unroll
            for (int j = 0; j < NB; ++j)
                if (j < kb) {
                    Ur[q * PS + j] = Re[k0 + j + col];
                    Ui[q * PS + j] = Im[k0 + j + col];
                }
            for (int j = 0; j < kb; ++j) {
                double sr = Ur[q * PS + j], si = Ui[q * PS + j];
                for (int i = 0; i < j; ++i) {
                    double tr_, ti_;
                    cmul(Pr[j * PS + i], Pi[j * PS + i], Ur[q * PS + i], Ui[q * PS + i], tr_, ti_);
                    sr -= tr_;
                    si -= ti_;
                }
                Ur[q * PS + j] = sr;
                Ui[q * PS + j] = si;
                Re[k0 + j + col] = sr;
                Im[k0 + j + col] = si;
            }
        }
        __syncthreads();
        // trailing update A22 -= L21 U12, LB entries per thread at a time with their loads issued
        // before any store (the stores may alias later loads as far as the compiler knows)
        const int mt = m - kb;
        for (int e0 = tid; e0 < mt * ncol; e0 += NL * LB) {
            double gr[LB], gi[LB];
            size_t off[LB];
#pragma unroll
            for (int b = 0; b < LB; ++b) {
                const int e = e0 + b * NL, i = kb + e % mt, q = e / mt;
                const bool ok = e < mt * ncol;
                off[b] = ok ? (k0 + i) + (size_t)(k0 + kb + q) * ld : 0;
                gr[b] = ok ? Re[off[b]] : 0.0;
                gi[b] = ok ? Im[off[b]] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < LB; ++b) {
                const int e = e0 + b * NL, i = kb + e % mt, q = e / mt;
                if (e >= mt * ncol) continue;
                double sr = 0.0, si = 0.0;
#pragma unroll 4
                for (int j = 0; j < kb; ++j) {
                    const double ar = Pr[i * PS + j], ai = Pi[i * PS + j];
                    const double br = Ur[q * PS + j], bi = Ui[q * PS + j];
                    sr += ar * br - ai * bi;
                    si += ar * bi + ai * br;
                }
                Re[off[b]] = gr[b] - sr;
                Im[off[b]] = gi[b] - si;
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (a.logdet) a.logdet[c] = s_logdet;
        if (a.phase) {
            a.phase[2 * c] = s_phr;
            a.phase[2 * c + 1] = s_phi;
        }
        if (bad && a.info && !a.info[c]) a.info[c] = bad;
    }
}

// X = M⁻¹ Cᴴ with the LU factors of M (n x n, from k_lu) and the nr x n matrix C (so X is
// n x nr): one workgroup per 16 right-hand sides held in LDS. Row interchanges, unit-lower
// forward and upper backward substitution by 16-column slabs of the factors.
__global__ __launch_bounds__(NT) void k_lu_solve(LuSolveArgs s) {
    const int c = blockIdx.y;
    const ChainDev &ch = s.chains[c];
    const int n = dim_of(ch, s.ndim), nr = dim_of(ch, s.rdim);
    const int r0 = blockIdx.x * RB;
    if (r0 >= nr) return;
    const int rb = min(RB, nr - r0);
    const Opnd F = s.LU[c], Cm = s.C[c], X = s.X[c];
    const int *piv = s.piv + (size_t)c * s.piv_ld;
    extern __shared__ double lds[];
    double *Br = lds, *Bi = lds + (size_t)n * RB;                      // [i][q]
    double *Fr = lds + 2 * (size_t)n * RB, *Fi = Fr + (size_t)n * PS;  // factor slab [row][i]
    double *Dr_ = Fi + (size_t)n * PS, *Di_ = Dr_ + NB * PS;           // block inverse [jj][i]
    const int tid = threadIdx.x;
    const size_t ld = F.ld;
    for (int e = tid; e < n * RB; e += NT) {
        const int q = e % RB, i = e / RB;
        double vr = 0.0, vi = 0.0;
        if (q < rb) {
            const size_t o = (size_t)(r0 + q) + (size_t)i * Cm.ld;
            vr = Cm.re[o];
            vi = -Cm.im[o];
        }
        Br[e] = vr;
        Bi[e] = vi;
    }
    __syncthreads();
    if (tid < RB) {
        const int q = tid;
        for (int k = 0; k < n; ++k) {
            const int p = piv[k];
            if (p != k) {
                const double ar = Br[k * RB + q], ai = Bi[k * RB + q];
                Br[k * RB + q] = Br[p * RB + q];
                Bi[k * RB + q] = Bi[p * RB + q];
                Br[p * RB + q] = ar;
                Bi[p * RB + q] = ai;
            }
        }
    }
    // forward substitution with the unit lower factor
    for (int b0 = 0; b0 < n; b0 += NB) {
        const int bw = min(NB, n - b0), rest = n - b0;
        __syncthreads();
        for (int e = tid; e < rest * bw; e += NT) {
            const int jj = e % rest, i = e / rest;
            const size_t o = (size_t)(b0 + jj) + (size_t)(b0 + i) * ld;
            Fr[jj * PS + i] = F.re[o];
            Fi[jj * PS + i] = F.im[o];
        }
        if (s.Dinv)   // L11⁻¹ of this slab (unit lower, written by k_lu)
            for (int e = tid; e < bw * bw; e += NT) {
                const int jj = e % bw, i = e / bw;
                const Opnd D = s.Dinv[c];
                const size_t o = (size_t)(b0 + jj) + (size_t)i * D.ld;
                Dr_[jj * PS + i] = D.re[o];
                Di_[jj * PS + i] = D.im[o];
            }
        __syncthreads();
        if (s.Dinv) {   // B_slab <- L11⁻¹ B_slab, every (j, q) at once
            const int q = tid % RB, j = tid / RB;
            double sr = 0.0, si = 0.0;
            if (j < bw)
                for (int i = 0; i <= j; ++i) {
                    double tr_, ti_;
                    cmul(Dr_[j * PS + i], Di_[j * PS + i], Br[(b0 + i) * RB + q], Bi[(b0 + i) * RB + q], tr_, ti_);
                    sr += tr_;
                    si += ti_;
                }
            __syncthreads();
            if (j < bw) {
                Br[(b0 + j) * RB + q] = sr;
                Bi[(b0 + j) * RB + q] = si;
            }
        } else if (tid < RB) {
            const int q = tid;
            for (int j = 1; j < bw; ++j) {
                double sr = Br[(b0 + j) * RB + q], si = Bi[(b0 + j) * RB + q];
                for (int i = 0; i < j; ++i) {
                    double tr_, ti_;
                    cmul(Fr[j * PS + i], Fi[j * PS + i], Br[(b0 + i) * RB + q], Bi[(b0 + i) * RB + q], tr_, ti_);
                    sr -= tr_;
                    si -= ti_;
                }
                Br[(b0 + j) * RB + q] = sr;
                Bi[(b0 + j) * RB + q] = si;
            }
        }
        __syncthreads();
        for (int e = tid; e < (rest - bw) * RB; e += NT) {
            const int q = e % RB, jj = bw + e / RB;
            double sr = 0.0, si = 0.0;
#pragma unroll 4
            for (int i = 0; i < bw; ++i) {
                const double ar = Fr[jj * PS + i], ai = Fi[jj * PS + i];
                const double br = Br[(b0 + i) * RB + q], bi = Bi[(b0 + i) * RB + q];
                sr += ar * br - ai * bi;
                si += ar * bi + ai * br;
            }
            Br[(b0 + jj) * RB + q] -= sr;
            Bi[(b0 + jj) * RB + q] -= si;
        }
    }
    // backward substitution with the upper factor, last slab first
    const int nslab = (n + NB - 1) / NB;
    for (int sb = nslab - 1; sb >= 0; --sb) {
        const int b0 = sb * NB, bw = min(NB, n - b0), top = b0 + bw;
        __syncthreads();
        for (int e = tid; e < top * bw; e += NT) {
            const int jj = e % top, i = e / top;
            const size_t o = (size_t)jj + (size_t)(b0 + i) * ld;
            Fr[jj * PS + i] = F.re[o];
            Fi[jj * PS + i] = F.im[o];
        }
        if (s.Dinv)   // U11⁻¹ of this slab (columns 16..31 of the inverse blocks)
            for (int e = tid; e < bw * bw; e += NT) {
                const int jj = e % bw, i = e / bw;
                const Opnd D = s.Dinv[c];
                const size_t o = (size_t)(b0 + jj) + (size_t)(NB + i) * D.ld;
                Dr_[jj * PS + i] = D.re[o];
                Di_[jj * PS + i] = D.im[o];
            }
        __syncthreads();
        if (s.Dinv) {   // B_slab <- U11⁻¹ B_slab
            const int q = tid % RB, j = tid / RB;
            double sr = 0.0, si = 0.0;
            if (j < bw)
                for (int i = j; i < bw; ++i) {
                    double tr_, ti_;
                    cmul(Dr_[j * PS + i], Di_[j * PS + i], Br[(b0 + i) * RB + q], Bi[(b0 + i) * RB + q], tr_, ti_);
                    sr += tr_;
                    si += ti_;
                }
            __syncthreads();
            if (j < bw) {
                Br[(b0 + j) * RB + q] = sr;
                Bi[(b0 + j) * RB + q] = si;
            }
        } else if (tid < RB) {
            const int q = tid;
            for (int j = bw - 1; j >= 0; --j) {
                double sr = Br[(b0 + j) * RB + q], si = Bi[(b0 + j) * RB + q];
                for (int i = j + 1; i < bw; ++i) {
                    double tr_, ti_;
                    cmul(Fr[(b0 + j) * PS + i], Fi[(b0 + j) * PS + i], Br[(b0 + i) * RB + q],
                         Bi[(b0 + i) * RB + q], tr_, ti_);
                    sr -= tr_;
                    si -= ti_;
                }
                const double ur = Fr[(b0 + j) * PS + j], ui = Fi[(b0 + j) * PS + j];
                const double u2 = ur * ur + ui * ui;
                double xr, xi;
                cmul(sr, si, ur / u2, -ui / u2, xr, xi);
                Br[(b0 + j) * RB + q] = xr;
                Bi[(b0 + j) * RB + q] = xi;
            }
        }
        __syncthreads();
        for (int e = tid; e < b0 * RB; e += NT) {
            const int q = e % RB, jj = e / RB;
            double sr = 0.0, si = 0.0;
#pragma unroll 4
            for (int i = 0; i < bw; ++i) {
                const double ar = Fr[jj * PS + i], ai = Fi[jj * PS + i];
                const double br = Br[(b0 + i) * RB + q], bi = Bi[(b0 + i) * RB + q];
                sr += ar * br - ai * bi;
                si += ar * bi + ai * br;
            }
            Br[jj * RB + q] -= sr;
            Bi[jj * RB + q] -= si;
        }
    }
    __syncthreads();
    for (int e = tid; e < n * RB; e += NT) {
        const int i = e % n, q = e / n;
        if (q >= rb) continue;
        const size_t o = (size_t)i + (size_t)(r0 + q) * X.ld;
        X.re[o] = Br[i * RB + q];
        X.im[o] = Bi[i * RB + q];
    }
}

template <class K>
hipError_t allow_lds(K kern, size_t bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

}  // namespace

size_t factor_lds_bytes(int which, int max_n) {
    const size_t n = (size_t)max_n;
    switch (which) {
        case 0: return 2 * n * PS * sizeof(double);                 // k_chol
        case 1: return (2 * n * RB + 2 * n * PS + 2 * NB * PS) * sizeof(double);  // k_trsm_rlh, k_lu_solve
        case 2: return 4 * n * PS * sizeof(double);                 // k_lu
        default: return 0;
    }
}

hipError_t launch_chol(hipStream_t st, int nchains, int max_n, const CholArgs &a) {
    const size_t b = factor_lds_bytes(0, max_n);
    if (hipError_t e = allow_lds(k_chol, b)) return e;
    hipLaunchKernelGGL(k_chol, dim3(nchains), dim3(NF), b, st, a);
    if (a.Dinv) hipLaunchKernelGGL(k_diag_inv<false>, dim3((max_n + NB - 1) / NB, nchains), dim3(NT), 0, st,
                                   a.chains, a.A, a.ndim, a.Dinv);
    return hipGetLastError();
}

hipError_t launch_trsm_rlh(hipStream_t st, int nchains, int max_m, int max_n, const TrsmArgs &t) {
    const size_t b = factor_lds_bytes(1, max_n);
    if (hipError_t e = allow_lds(k_trsm_rlh, b)) return e;
    hipLaunchKernelGGL(k_trsm_rlh, dim3((max_m + RB - 1) / RB, nchains), dim3(NT), b, st, t);
    return hipGetLastError();
}

hipError_t launch_lu(hipStream_t st, int nchains, int max_n, const LuArgs &a) {
    const size_t b = factor_lds_bytes(2, max_n);
    if (hipError_t e = allow_lds(k_lu, b)) return e;
    hipLaunchKernelGGL(k_lu, dim3(nchains), dim3(NL), b, st, a);
    if (a.Dinv) hipLaunchKernelGGL(k_diag_inv<true>, dim3((max_n + NB - 1) / NB, nchains), dim3(NT), 0, st,
                                   a.chains, a.A, a.ndim, a.Dinv);
    return hipGetLastError();
}

hipError_t launch_lu_solve(hipStream_t st, int nchains, int max_n, int max_r, const LuSolveArgs &s) {
    const size_t b = factor_lds_bytes(1, max_n);
    if (hipError_t e = allow_lds(k_lu_solve, b)) return e;
    hipLaunchKernelGGL(k_lu_solve, dim3((max_r + RB - 1) / RB, nchains), dim3(NT), b, st, s);
    return hipGetLastError();
}
