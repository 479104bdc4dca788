// Inverse of the per-chain n x n overlap M = U_L U_R of get_G! (src/public/Gupdate.jl:139-157,
// where the reference inverts it through LAPACK's LU) by Gauss-Jordan elimination with partial
// pivoting, the whole matrix held in registers of one 512-thread workgroup per chain.
//
// Layout: thread t is column class tc = t / 32 (0..15) and row class tr = t % 32; it holds
// M[tr + 32 i][tc + 16 k] for i < RS, k < CS (n <= 32 RS, n <= 16 CS), so the 32 holders of a
// column are one half of a wave64 and the pivot search is a shuffle butterfly without LDS.
// Step j (column j): the half holding column j finds the pivot row p among the rows not yet
// pivoted (largest |m|², first index on ties), publishes column j (0 at row p), p and 1/m_pj;
// one barrier; then every thread normalises its part of row p if it holds it (row p lives in
// one lane of each half, which passes it to its half through a per-half LDS row), sets its
// column-j entries outside row p to 0 and applies the rank-1 update m_rc -= m_rj m'_pc. Rows
// are not interchanged (the pivot order is kept instead): the registers then hold
// M⁻¹[q(r)][p(c)] at (r, c), with p(j) the pivot row of step j and q its inverse, and are
// stored that way. log|det M| = Σ log|pivot|. n³ complex multiply-adds, 2 waves per SIMD.
#include "kdq_kernels.h"

#include <utility>

namespace {

// f(std::integral_constant<int, 0>{}), ..., f(.. N - 1 ..): the column slot of a step is then a
// compile-time index into the register arrays (a loop the compiler might not unroll would turn
// them into scratch memory)
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
__global__ __launch_bounds__(IT) void k_inv(InvArgs a) {
    const int c = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = dim_of(ch, a.ndim);
    const Opnd A = a.A[c], X = a.X[c];
    const int tid = threadIdx.x, tr = tid & 31, tc = tid >> 5, hb = tid & 32;
    __shared__ double colr[2][32 * RS], coli[2][32 * RS];   // column j of step j, 0 at the pivot row
    __shared__ double rowr[16][CS], rowi[16][CS];           // the normalised pivot row, per half
    __shared__ double sir[2], sii[2];                       // 1 / pivot
    __shared__ int sp[2], pvt[16 * CS], qv[32 * RS], bad;
    __shared__ double red[IT / 64];
    double xr[RS][CS], xi[RS][CS];
    unsigned elim = 0;   // bit i: row tr + 32 i pivoted already (or outside the matrix)
#pragma unroll
    for (int i = 0; i < RS; ++i) {
        const int r = tr + 32 * i;
        if (r >= n) elim |= 1u << i;
#pragma unroll
        for (int k = 0; k < CS; ++k) {
            const int col = tc + 16 * k;
            const bool ok = r < n && col < n;
            const size_t o = ok ? (size_t)r + (size_t)col * A.ld : 0;
            xr[i][k] = ok ? A.re[o] : 0.0;
            xi[i][k] = ok ? A.im[o] : 0.0;
        }
    }
    if (tid == 0) bad = 0;
    double ldet = 0.0;
#pragma unroll
    for (int kj = 0; kj < CS; ++kj) {
        for (int jj = 0; jj < 16; ++jj) {
            const int j = 16 * kj + jj;
            if (j >= n) break;
            const int b = j & 1;
            if (tc == jj) {
                double best = -1.0;
                int bi = SENT;
#pragma unroll
                for (int i = 0; i < RS; ++i)
                    if (!((elim >> i) & 1u)) {
                        const double v = xr[i][kj] * xr[i][kj] + xi[i][kj] * xi[i][kj];
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
                double vr = xr[0][kj], vi = xi[0][kj];
#pragma unroll
                for (int i = 1; i < RS; ++i)
                    if (ps == i) {
                        vr = xr[i][kj];
                        vi = xi[i][kj];
                    }
                vr = __shfl(vr, hb + pl, 64);
                vi = __shfl(vi, hb + pl, 64);
#pragma unroll
                for (int i = 0; i < RS; ++i) {
                    const int r = tr + 32 * i;
                    colr[b][r] = r == p ? 0.0 : xr[i][kj];
                    coli[b][r] = r == p ? 0.0 : xi[i][kj];
                }
                if (tr == 0) {
                    const double p2 = vr * vr + vi * vi;
                    sp[b] = p;
                    pvt[j] = p;
                    qv[p] = j;
                    if (p2 > 0.0) {
                        sir[b] = vr / p2;
                        sii[b] = -vi / p2;
                        ldet += 0.5 * log(p2);
                    } else {   // exactly singular: flagged, the step leaves the matrix as it is
                        sir[b] = 0.0;
                        sii[b] = 0.0;
                        if (!bad) bad = j + 1;
                    }
                }
            }
            __syncthreads();
            const int p = sp[b], ps = p >> 5, pl = p & 31;
            const double ir = sir[b], ii = sii[b];
            if (tr == pl) elim |= 1u << ps;
#pragma unroll
            for (int i = 0; i < RS; ++i)
                if (ps == i && tr == pl) {
#pragma unroll
                    for (int k = 0; k < CS; ++k) {
                        double nr, ni;
                        if (tc + 16 * k == j) {
                            nr = ir;
                            ni = ii;
                        } else {
                            cmul(xr[i][k], xi[i][k], ir, ii, nr, ni);
                        }
                        xr[i][k] = nr;
                        xi[i][k] = ni;
                        rowr[tc][k] = nr;
                        rowi[tc][k] = ni;
                    }
                }
            __builtin_amdgcn_wave_barrier();
            if (tc == jj) {
#pragma unroll
                for (int i = 0; i < RS; ++i)
                    if (tr + 32 * i != p) {
                        xr[i][kj] = 0.0;
                        xi[i][kj] = 0.0;
                    }
            }
            double dr[RS], di[RS];
#pragma unroll
            for (int i = 0; i < RS; ++i) {
                dr[i] = colr[b][tr + 32 * i];
                di[i] = coli[b][tr + 32 * i];
            }
#pragma unroll
            for (int k = 0; k < CS; ++k) {
                const double ur = rowr[tc][k], ui = rowi[tc][k];
#pragma unroll
                for (int i = 0; i < RS; ++i) {
                    xr[i][k] = fma(-dr[i], ur, fma(di[i], ui, xr[i][k]));
                    xi[i][k] = fma(-dr[i], ui, fma(-di[i], ur, xi[i][k]));
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS; ++i) {
        const int r = tr + 32 * i;
#pragma unroll
        for (int k = 0; k < CS; ++k) {
            const int col = tc + 16 * k;
            if (r < n && col < n) {
                const size_t o = (size_t)qv[r] + (size_t)pvt[col] * X.ld;
                X.re[o] = xr[i][k];
                X.im[o] = xi[i][k];
            }
        }
    }
    if (a.logdet) {
        double s = wave_sum(ldet);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < IT / 64; ++w) t += red[w];
            a.logdet[c] = t;
        }
    }
    if (tid == 0 && bad && a.info && !a.info[c]) a.info[c] = bad;
}

// X = L⁻¹ for the Cholesky factor of S + σI = L Lᴴ (S Hermitian n x n, n <= 144, σ = shift ·
// trace S), both in one register pass of the layout above: step j takes the column
// l = S'[:, j] / d (rows > j, d = sqrt(S'[j][j]) of the current Schur complement) and applies
// one rank-1 update to the rows below j with the vector u = (row j of the partial inverse
// scaled by 1/d for columns <= j, conj(l) for columns > j): the right-hand columns continue the
// Cholesky factorisation, the left-hand ones the forward elimination of [L | I], which needs
// L's column j just when the factorisation makes it. X is lower triangular (zeros above);
// log d_j summed into logdiag, info (sticky) j+1 for a non-positive pivot.
template <int RS, int CS>
__global__ __launch_bounds__(IT) void k_cholinv(CholInvArgs a) {
    const int c = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int n = dim_of(ch, a.ndim);
    const Opnd A = a.A[c], X = a.X[c];
    const int tid = threadIdx.x, tr = tid & 31, tc = tid >> 5, hb = tid & 32;
    __shared__ double colr[2][32 * RS + 16 * CS], coli[2][32 * RS + 16 * CS];   // l (0 at rows <= j)
    __shared__ double rowr[16][CS], rowi[16][CS];   // row j of the partial inverse, per half
    __shared__ double sid[2];                       // 1 / d
    __shared__ int bad;
    __shared__ double red[IT / 64];
    double xr[RS][CS], xi[RS][CS];
    double tr0 = 0.0;
#pragma unroll
    for (int i = 0; i < RS; ++i) {
        const int r = tr + 32 * i;
#pragma unroll
        for (int k = 0; k < CS; ++k) {
            const int col = tc + 16 * k;
            const bool ok = r < n && col < n;
            const size_t o = ok ? (size_t)r + (size_t)col * A.ld : 0;
            xr[i][k] = ok ? A.re[o] : 0.0;
            xi[i][k] = ok ? A.im[o] : 0.0;
            if (ok && r == col) tr0 += xr[i][k];
        }
    }
    if (tid == 0) bad = 0;
    {
        const double s = wave_sum(tr0);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < IT / 64; ++w) t += red[w];
        const double sigma = a.shift * t;
#pragma unroll
        for (int i = 0; i < RS; ++i)
#pragma unroll
            for (int k = 0; k < CS; ++k)
                if (tr + 32 * i == tc + 16 * k) xr[i][k] += sigma;
    }
    double lsum = 0.0;
    static_for<CS>([&](auto kjc) {
        constexpr int kj = decltype(kjc)::value;
        for (int jj = 0; jj < 16; ++jj) {
            const int j = 16 * kj + jj;
            if (j >= n) break;
            const int b = j & 1, js = j >> 5, jl = j & 31;
            if (tc == jj) {
                double v = xr[0][kj];
#pragma unroll
                for (int i = 1; i < RS; ++i)
                    if (js == i) v = xr[i][kj];
                const double dj = __shfl(v, hb + jl, 64);
                const double d = sqrt(fmax(dj, 1e-300)), id = 1.0 / d;
#pragma unroll
                for (int i = 0; i < RS; ++i) {
                    const int r = tr + 32 * i;
                    colr[b][r] = r > j ? xr[i][kj] * id : 0.0;
                    coli[b][r] = r > j ? xi[i][kj] * id : 0.0;
                }
                if (tr == 0) {
                    sid[b] = id;
                    lsum += log(d);
                    if (!(dj > 0.0) && !bad) bad = j + 1;
                }
            }
            __syncthreads();
            const double id = sid[b];
#pragma unroll
            for (int i = 0; i < RS; ++i)
                if (js == i && tr == jl) {
#pragma unroll
                    for (int k = 0; k < CS; ++k) {
                        const int col = tc + 16 * k;
                        double nr = 0.0, ni = 0.0;
                        if (col < j) {
                            nr = xr[i][k] * id;
                            ni = xi[i][k] * id;
                        } else if (col == j) {
                            nr = id;
                        }
                        xr[i][k] = nr;
                        xi[i][k] = ni;
                        rowr[tc][k] = nr;
                        rowi[tc][k] = ni;
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
                dr[i] = colr[b][tr + 32 * i];
                di[i] = coli[b][tr + 32 * i];
            }
#pragma unroll
            for (int k = 0; k < CS; ++k) {
                const int col = tc + 16 * k;
                double ur, ui;
                if (col > j) {
                    ur = colr[b][col];
                    ui = -coli[b][col];
                } else {
                    ur = rowr[tc][k];
                    ui = rowi[tc][k];
                }
#pragma unroll
                for (int i = 0; i < RS; ++i)
                    if (32 * i + 31 > j) {   // rows below j only (the multiplier is 0 elsewhere)
                        xr[i][k] = fma(-dr[i], ur, fma(di[i], ui, xr[i][k]));
                        xi[i][k] = fma(-dr[i], ui, fma(-di[i], ur, xi[i][k]));
                    }
            }
            __builtin_amdgcn_wave_barrier();
        }
    });
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS; ++i) {
        const int r = tr + 32 * i;
#pragma unroll
        for (int k = 0; k < CS; ++k) {
            const int col = tc + 16 * k;
            if (r < n && col < n) {
                const size_t o = (size_t)r + (size_t)col * X.ld;
                X.re[o] = r >= col ? xr[i][k] : 0.0;
                X.im[o] = r >= col ? xi[i][k] : 0.0;
            }
        }
    }
    if (a.logdiag) {
        const double s = wave_sum(lsum);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < IT / 64; ++w) t += red[w];
            a.logdiag[c] += t;
        }
    }
    if (tid == 0 && bad && a.info && !a.info[c]) a.info[c] = bad;
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
