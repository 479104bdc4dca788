// Register-resident factorisations of the stabilisation, one 512-thread workgroup per chain
// with the whole n x n matrix (n <= 144) in registers:
//   k_inv      the inverse of the overlap M = U_L U_R of get_G! (src/public/Gupdate.jl:139-157,
//              where the reference inverts it through LAPACK's LU) by Gauss-Jordan elimination
//              with partial pivoting; log|det M|;
//   k_cholinv  the Cholesky factor of the Gram matrix UᴴU of a Cholesky-QR pass (the role of
//              the QR at src/tU/phy_update.jl:41-47,81-87) and its inverse in the same pass.
//
// Layout: thread t is column class tc = t / 32 (0..15) and row class tr = t % 32; it holds
// M[tr + 32 i][tc + 16 k] for i < RS, k < CS (n <= 32 RS, n <= 16 CS), so the 32 holders of a
// column are one half of a wave64 and the pivot search is a shuffle butterfly without LDS.
// Step j (column j): the half holding column j finds the pivot row p among the rows not yet
// pivoted (largest |m|², first index on ties), publishes column j (0 at row p), p and 1/m_pj;
// one barrier; the lane holding row p passes u = row p / m_pj to its half through a per-half
// LDS row; the column owners set column j to 0 outside row p (1 at row p) and every thread
// applies the rank-1 update m_rc -= m_rj u_c. Rows are not interchanged (the pivot order is
// kept instead) and row p keeps the factor m_pj until the store: the registers then hold
// m_pj M⁻¹[q(r)][p(c)] at (r, c) for the pivot row r = p(j), with p(j) the pivot row of step j
// and q its inverse, and are stored scaled and permuted. log|det M| = Σ log|pivot|.
// n³ complex multiply-adds, 2 waves per SIMD; every wave64 VALU instruction (FP64 FMA
// included) takes 4 cycles, so no register is written under a per-lane condition on the pivot.
#include "kdq_kernels.h"

#include <utility>

namespace {

// f(std::integral_constant<int, 0>{}), ..., f(.. N - 1 ..): a compile-time column slot per
// block of 16 steps (a loop the compiler might not unroll would turn the register arrays into
// scratch memory)
template <class F, int... I>
KDQ_DEV void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
KDQ_DEV void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

constexpr int IT = 512;
constexpr int SENT = 1 << 30;

template <int RS, int CS>
struct RegLds {
    double colr[2][32 * RS], coli[2][32 * RS];   // the step's multiplier column (by row)
    double rowr[16][CS], rowi[16][CS];           // a row of the matrix passed to its half
    double s1[2], s2[2];                          // the step's scalars
    int sp[2];                                    // the step's pivot row
    double pd[16 * CS];                           // per step: |pivot|² or the Cholesky diagonal
    double ir[16 * CS], ii[16 * CS];              // per step: 1/pivot (1/d for the Cholesky pass)
    int pvt[16 * CS], qv[32 * RS];                // pivot row of step j, step of pivot row r
    double red[IT / 64];
    int bad;
};

// Loads the n x n matrix into the register layout (zeros outside it); returns the trace.
// LOWER: only the slots holding some entry on or below the diagonal (32 i + 31 >= 16 k); the
// others are 0 and stay 0 (the Cholesky pass never needs them, so they take no registers).
KDQ_DEV constexpr bool low_slot(int i, int k) { return 32 * i + 31 >= 16 * k; }

template <int RS, int CS, bool LOWER>
KDQ_DEV double load_regs(const Opnd &A, int n, int tr, int tc, double (&xr)[RS][CS], double (&xi)[RS][CS]) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < RS; ++i) {
        const int r = tr + 32 * i;
#pragma unroll
        for (int k = 0; k < CS; ++k) {
            const int col = tc + 16 * k;
            const bool ok = r < n && col < n && (!LOWER || low_slot(i, k));
            const size_t o = ok ? (size_t)r + (size_t)col * A.ld : 0;
            xr[i][k] = ok ? A.re[o] : 0.0;
            xi[i][k] = ok ? A.im[o] : 0.0;
            if (ok && r == col) t += xr[i][k];
        }
    }
    return t;
}

// x[i][K] -= d_i u for the row slots i >= i0 (i0 uniform)
template <int RS, int CS, bool LOWER>
KDQ_DEV void upd_slot(double (&xr)[RS][CS], double (&xi)[RS][CS], const double (&dr)[RS], const double (&di)[RS],
                      int K, double ur, double ui, int i0) {
#pragma unroll
    for (int i = 0; i < RS; ++i)
        if (i >= i0 && (!LOWER || low_slot(i, K))) {
            xr[i][K] = fma(-dr[i], ur, fma(di[i], ui, xr[i][K]));
            xi[i][K] = fma(-dr[i], ui, fma(-di[i], ur, xi[i][K]));
        }
}

// Gauss-Jordan pivot search of step j, called by the half holding column j (slot K): the
// pivot row p among the rows not pivoted yet; the column to LDS (0 at row p); 1/piv.
template <int RS, int CS>
KDQ_DEV void inv_pivot(RegLds<RS, CS> &S, const double (&xr)[RS][CS], const double (&xi)[RS][CS], unsigned elim,
                       int K, int tr, int hb, int j) {
    const int b = j & 1;
    double best = -1.0;
    int bi = SENT;
#pragma unroll
    for (int i = 0; i < RS; ++i)
        if (!((elim >> i) & 1u)) {
            const double v = xr[i][K] * xr[i][K] + xi[i][K] * xi[i][K];
            if (bi == SENT || v > best) {
                best = v;
                bi = tr + 32 * i;
            }
        }
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (oi != SENT && (bi == SENT || ob > best || (ob == best && oi < bi))) {
            best = ob;
            bi = oi;
        }
    }
    const int p = __shfl(bi, hb, 64), ps = p >> 5, pl = p & 31;
    double vr = xr[0][K], vi = xi[0][K];
#pragma unroll
    for (int i = 1; i < RS; ++i)
        if (ps == i) {
            vr = xr[i][K];
            vi = xi[i][K];
        }
    vr = __shfl(vr, hb + pl, 64);
    vi = __shfl(vi, hb + pl, 64);
#pragma unroll
    for (int i = 0; i < RS; ++i) {
        const int r = tr + 32 * i;
        S.colr[b][r] = r == p ? 0.0 : xr[i][K];
        S.coli[b][r] = r == p ? 0.0 : xi[i][K];
    }
    if (tr == 0) {
        const double p2 = vr * vr + vi * vi;
        const bool ok = p2 > 0.0;   // an exactly singular step leaves the matrix as it is
        S.sp[b] = p;
        S.pvt[j] = p;
        S.qv[p] = j;
        S.pd[j] = p2;
        S.s1[b] = S.ir[j] = ok ? vr / p2 : 0.0;
        S.s2[b] = S.ii[j] = ok ? -vi / p2 : 0.0;
    }
}

// Cholesky column of step j, called by the half holding column j (slot K): d = sqrt of the
// current diagonal, the multipliers l_r = x_rj / d (rows > j, else 0) to LDS, 1/d
template <int RS, int CS>
KDQ_DEV void chol_pivot(RegLds<RS, CS> &S, const double (&xr)[RS][CS], const double (&xi)[RS][CS], int K, int tr,
                        int hb, int j) {
    const int b = j & 1, js = j >> 5, jl = j & 31;
    double v = xr[0][K];
#pragma unroll
    for (int i = 1; i < RS; ++i)
        if (js == i) v = xr[i][K];
    const double dj = __shfl(v, hb + jl, 64);
    const double d = sqrt(fmax(dj, 1e-300)), id = 1.0 / d;
#pragma unroll
    for (int i = 0; i < RS; ++i) {
        const int r = tr + 32 * i;
        S.colr[b][r] = r > j ? xr[i][K] * id : 0.0;
        S.coli[b][r] = r > j ? xi[i][K] * id : 0.0;
    }
    if (tr == 0) {
        S.s1[b] = id;
        S.ir[j] = id;
        S.pd[j] = dj;
    }
}

// block sum of v (every thread calls it)
template <int RS, int CS>
KDQ_DEV double block_sum(RegLds<RS, CS> &S, double v, int tid) {
    v = wave_sum(v);
    __syncthreads();
    if ((tid & 63) == 0) S.red[tid >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < IT / 64; ++w) t += S.red[w];
    return t;
}

// In-place Gauss-Jordan with the pivot row kept unnormalised: at step j (pivot p, piv = m_pj)
// the column owners set column j to 0 outside row p and to 1 at row p, and every entry takes
// m_rc -= d_r u_c with u = row p / piv (u_j = 1/piv, formed by the lane holding row p) and
// d = column j (d_p = 0). Rows r != p get the Gauss-Jordan update; row p stays piv · u (its
// column j: 1 = piv / piv), and keeps that factor through later steps (its multipliers carry
// it too), so the store scales it by 1/piv. No register is written under a per-lane condition
// on the pivot row (the compiler would evaluate such code for every slot).
template <int RS, int CS>
__global__ __launch_bounds__(IT) void k_inv(InvArgs a) {
    const int c = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = dim_of(ch, a.ndim);
    const Opnd A = a.A[c], X = a.X[c];
    const int tid = threadIdx.x, tr = tid & 31, tc = tid >> 5, hb = tid & 32;
    __shared__ RegLds<RS, CS> S;
    double xr[RS][CS], xi[RS][CS];
    load_regs<RS, CS, false>(A, n, tr, tc, xr, xi);
    unsigned elim = 0;   // bit i: row tr + 32 i pivoted already (or outside the matrix)
#pragma unroll
    for (int i = 0; i < RS; ++i)
        if (tr + 32 * i >= n) elim |= 1u << i;
#pragma unroll
    for (int kj = 0; kj < CS; ++kj) {
        for (int jj = 0; jj < 16; ++jj) {
            const int j = 16 * kj + jj;
            if (j >= n) break;
            const int b = j & 1;
            if (tc == jj) inv_pivot<RS, CS>(S, xr, xi, elim, kj, tr, hb, j);
            __syncthreads();
            const int p = __builtin_amdgcn_readfirstlane(S.sp[b]), ps = p >> 5, pl = p & 31;
            const double ir = S.s1[b], ii = S.s2[b];
            if (tr == pl) elim |= 1u << ps;
            // u = row p / piv (u_j = 1/piv) from the lane holding row p, to its half's LDS row
#pragma unroll
            for (int i = 0; i < RS; ++i)
                if (ps == i) {
                    if (tr == pl) {
#pragma unroll
                        for (int k = 0; k < CS; ++k) {
                            double nr = ir, ni = ii;
                            if (tc + 16 * k != j) cmul(xr[i][k], xi[i][k], ir, ii, nr, ni);
                            S.rowr[tc][k] = nr;
                            S.rowi[tc][k] = ni;
                        }
                    }
                }
            __builtin_amdgcn_wave_barrier();
            if (tc == jj) {
#pragma unroll
                for (int i = 0; i < RS; ++i) {
                    xr[i][kj] = tr + 32 * i == p ? 1.0 : 0.0;
                    xi[i][kj] = 0.0;
                }
            }
            double dr[RS], di[RS];
#pragma unroll
            for (int i = 0; i < RS; ++i) {
                dr[i] = S.colr[b][tr + 32 * i];
                di[i] = S.coli[b][tr + 32 * i];
            }
#pragma unroll
            for (int k = 0; k < CS; ++k)
                upd_slot<RS, CS, false>(xr, xi, dr, di, k, S.rowr[tc][k], S.rowi[tc][k], 0);
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS; ++i) {
        const int r = tr + 32 * i;
        if (r < n) {
            const int q = S.qv[r];
            const double sr = S.ir[q], si = S.ii[q];
#pragma unroll
            for (int k = 0; k < CS; ++k) {
                const int col = tc + 16 * k;
                if (col < n) {
                    double vr, vi;
                    cmul(xr[i][k], xi[i][k], sr, si, vr, vi);
                    const size_t o = (size_t)q + (size_t)S.pvt[col] * X.ld;
                    X.re[o] = vr;
                    X.im[o] = vi;
                }
            }
        }
    }
    if (tid == 0) S.bad = 0x7fffffff;
    __syncthreads();
    double l = 0.0;
    for (int j = tid; j < n; j += IT) {
        const double p2 = S.pd[j];
        if (p2 > 0.0) l += 0.5 * log(p2);
        else atomicMin(&S.bad, j + 1);
    }
    const double ld = block_sum<RS, CS>(S, l, tid);
    if (tid == 0) {
        if (a.logdet) a.logdet[c] = ld;
        if (S.bad != 0x7fffffff && a.info && !a.info[c]) a.info[c] = S.bad;
    }
}

// X = L⁻¹ for the Cholesky factor of S + σI = L Lᴴ (S Hermitian n x n, n <= 144, σ = shift ·
// trace S), both in one register pass of the layout above: step j takes the column
// l = S'[:, j] / d (rows > j, d = sqrt(S'[j][j]) of the current Schur complement) and applies
// one rank-1 update to the rows below j with the vector u = (row j of the partial inverse
// scaled by 1/d for columns <= j, conj(l) for columns > j): the right-hand columns continue the
// Cholesky factorisation, the left-hand ones the forward elimination of [L | I], which needs
// L's column j just when the factorisation makes it. Row j is left unscaled in the registers
// (later steps neither read nor change it) and scaled by 1/d when stored. X is lower
// triangular (zeros above); log d_j summed into logdiag, info (sticky) j+1 for a non-positive
// pivot.
template <int RS, int CS>
__global__ __launch_bounds__(IT) void k_cholinv(CholInvArgs a) {
    const int c = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = dim_of(ch, a.ndim);
    const Opnd A = a.A[c], X = a.X[c];
    const int tid = threadIdx.x, tr = tid & 31, tc = tid >> 5, hb = tid & 32;
    __shared__ RegLds<RS, CS> S;
    double xr[RS][CS], xi[RS][CS];
    const double tr0 = load_regs<RS, CS, true>(A, n, tr, tc, xr, xi);
    {
        const double sigma = a.shift * block_sum<RS, CS>(S, tr0, tid);
#pragma unroll
        for (int i = 0; i < RS; ++i)
#pragma unroll
            for (int k = 0; k < CS; ++k)
                if (tr + 32 * i == tc + 16 * k) xr[i][k] += sigma;
    }
    static_for<CS>([&](auto kjc) {
        constexpr int kj = decltype(kjc)::value;
        for (int jj = 0; jj < 16; ++jj) {
            const int j = 16 * kj + jj;
            if (j >= n) break;
            const int b = j & 1, js = j >> 5, jl = j & 31;
            if (tc == jj) chol_pivot<RS, CS>(S, xr, xi, kj, tr, hb, j);
            __syncthreads();
            const double id = S.s1[b];
            // row j of the partial inverse scaled by 1/d (columns <= j; 1/d at j) to its half's LDS row
#pragma unroll
            for (int i = 0; i < RS; ++i)
                if (js == i) {
                    if (tr == jl) {
#pragma unroll
                        for (int k = 0; k < CS; ++k) {
                            if (!low_slot(i, k)) continue;
                            const int col = tc + 16 * k;
                            S.rowr[tc][k] = col < j ? xr[i][k] * id : (col == j ? id : 0.0);
                            S.rowi[tc][k] = col < j ? xi[i][k] * id : 0.0;
                        }
                    }
                }
            __builtin_amdgcn_wave_barrier();
            if (tc == jj) {
#pragma unroll
                for (int i = 0; i < RS; ++i)
                    if (tr + 32 * i > j) {
                        xr[i][kj] = 0.0;
                        xi[i][kj] = 0.0;
                    }
            }
            double dr[RS], di[RS];
#pragma unroll
            for (int i = 0; i < RS; ++i) {
                dr[i] = S.colr[b][tr + 32 * i];
                di[i] = S.coli[b][tr + 32 * i];
            }
            const int i0 = (j + 1) >> 5;   // row slots holding rows > j
#pragma unroll
            for (int k = 0; k < CS; ++k) {
                const int col = tc + 16 * k;
                const double ur = col > j ? S.colr[b][col] : S.rowr[tc][k];
                const double ui = col > j ? -S.coli[b][col] : S.rowi[tc][k];
                upd_slot<RS, CS, true>(xr, xi, dr, di, k, ur, ui, i0);
            }
            __builtin_amdgcn_wave_barrier();
        }
    });
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS; ++i) {
        const int r = tr + 32 * i;
        if (r < n) {
            const double sc = S.ir[r];   // row r was left unscaled: 1/d_r
#pragma unroll
            for (int k = 0; k < CS; ++k) {
                const int col = tc + 16 * k;
                if (col < n) {
                    const size_t o = (size_t)r + (size_t)col * X.ld;
                    const bool low = low_slot(i, k) && r >= col;
                    X.re[o] = low ? (r == col ? sc : xr[i][k] * sc) : 0.0;
                    X.im[o] = low && r != col ? xi[i][k] * sc : 0.0;
                }
            }
        }
    }
    if (tid == 0) S.bad = 0x7fffffff;
    __syncthreads();
    double l = 0.0;
    for (int j = tid; j < n; j += IT) {
        const double dj = S.pd[j];
        l += log(sqrt(fmax(dj, 1e-300)));
        if (!(dj > 0.0)) atomicMin(&S.bad, j + 1);
    }
    const double ls = block_sum<RS, CS>(S, l, tid);
    if (tid == 0) {
        if (a.logdiag) a.logdiag[c] += ls;
        if (S.bad != 0x7fffffff && a.info && !a.info[c]) a.info[c] = S.bad;
    }
}

}  // namespace

bool inv_supported(int max_n) { return max_n >= 1 && max_n <= 144; }

hipError_t launch_inv(hipStream_t st, int nchains, int max_n, const InvArgs &a) {
    if (max_n <= 32) hipLaunchKernelGGL((k_inv<1, 2>), dim3(nchains), dim3(IT), 0, st, a);
    else if (max_n <= 64) hipLaunchKernelGGL((k_inv<2, 4>), dim3(nchains), dim3(IT), 0, st, a);
    else if (max_n <= 96) hipLaunchKernelGGL((k_inv<3, 6>), dim3(nchains), dim3(IT), 0, st, a);
    else if (max_n <= 128) hipLaunchKernelGGL((k_inv<4, 8>), dim3(nchains), dim3(IT), 0, st, a);
    else if (max_n <= 144) hipLaunchKernelGGL((k_inv<5, 9>), dim3(nchains), dim3(IT), 0, st, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_cholinv(hipStream_t st, int nchains, int max_n, const CholInvArgs &a) {
    if (max_n <= 32) hipLaunchKernelGGL((k_cholinv<1, 2>), dim3(nchains), dim3(IT), 0, st, a);
    else if (max_n <= 64) hipLaunchKernelGGL((k_cholinv<2, 4>), dim3(nchains), dim3(IT), 0, st, a);
    else if (max_n <= 96) hipLaunchKernelGGL((k_cholinv<3, 6>), dim3(nchains), dim3(IT), 0, st, a);
    else if (max_n <= 128) hipLaunchKernelGGL((k_cholinv<4, 8>), dim3(nchains), dim3(IT), 0, st, a);
    else if (max_n <= 144) hipLaunchKernelGGL((k_cholinv<5, 9>), dim3(nchains), dim3(IT), 0, st, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}
