// FP64 factorizations of the stabilisation step, one workgroup per chain.
//
//  k_qr : Householder QR of the N x Np propagator with the explicit orthonormal factor
//         formed in place (LAPACK geqrf + orgqr as called at src/tU/phy_update.jl:85-86;
//         the RQ of U_L at :44-45 is the QR of U_Lᴴ, which is what the library stores).
//  k_inv: Gauss-Jordan inverse with partial pivoting of the Np x Np overlap U_L U_R,
//         with log|det| from the pivots (get_G!, src/public/Gupdate.jl:139-157, solves
//         with getrf/getrs; G = I - U_R (U_L U_R)⁻¹ U_L is the same matrix).
#include "kdq_kernels.h"

namespace {

constexpr int NT = 256;

KDQ_DEV double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) s += red[w];
    return s;
}

__global__ __launch_bounds__(NT) void k_qr(QRArgs q) {
    const int c = blockIdx.x;
    const ChainDev &ch = q.chains[c];
    const int m = dim_of(ch, q.mdim), n = dim_of(ch, q.ndim);
    const Opnd A = q.A[c];
    const size_t ld = A.ld;
    __shared__ double taur[256], taui[256];
    __shared__ double red[NT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *Re = A.re, *Im = A.im;

    for (int j = 0; j < n; ++j) {
        double part = 0.0;
        for (int i = j + 1 + tid; i < m; i += NT) {
            const double a = Re[i + j * ld], b = Im[i + j * ld];
            part += a * a + b * b;
        }
        const double xnorm2 = block_sum(part, red);
        const double alr = Re[j + j * ld], ali = Im[j + j * ld];
        double tr = 0.0, ti = 0.0;
        if (!(xnorm2 == 0.0 && ali == 0.0)) {
            const double beta = -copysign(sqrt(alr * alr + ali * ali + xnorm2), alr);
            tr = (beta - alr) / beta;
            ti = -ali / beta;
            // scal = 1 / (alpha - beta)
            const double dr = alr - beta, di = ali;
            const double d2 = dr * dr + di * di;
            const double sr = dr / d2, si = -di / d2;
            for (int i = j + 1 + tid; i < m; i += NT) {
                double vr, vi;
                cmul(Re[i + j * ld], Im[i + j * ld], sr, si, vr, vi);
                Re[i + j * ld] = vr;
                Im[i + j * ld] = vi;
            }
            __syncthreads();
            if (tid == 0) {
                Re[j + j * ld] = beta;
                Im[j + j * ld] = 0.0;
            }
        }
        if (tid == 0) {
            taur[j] = tr;
            taui[j] = ti;
        }
        __syncthreads();
        if (tr == 0.0 && ti == 0.0) continue;
        // apply Hᴴ = I - conj(tau) v vᴴ to columns j+1..n-1 (one wave per column)
        for (int cc = j + 1 + wave; cc < n; cc += NT / 64) {
            double wr = 0.0, wi = 0.0;
            for (int i = j + 1 + lane; i < m; i += 64) {
                const double vr = Re[i + j * ld], vi = -Im[i + j * ld];
                double pr, pi;
                cmul(vr, vi, Re[i + cc * ld], Im[i + cc * ld], pr, pi);
                wr += pr;
                wi += pi;
            }
            wr = wave_sum(wr) + Re[j + cc * ld];
            wi = wave_sum(wi) + Im[j + cc * ld];
            double fr, fi;
            cmul(wr, wi, tr, -ti, fr, fi);
            for (int i = j + 1 + lane; i < m; i += 64) {
                double pr, pi;
                cmul(Re[i + j * ld], Im[i + j * ld], fr, fi, pr, pi);
                Re[i + cc * ld] -= pr;
                Im[i + cc * ld] -= pi;
            }
            if (lane == 0) {
                Re[j + cc * ld] -= fr;
                Im[j + cc * ld] -= fi;
            }
        }
        __syncthreads();
    }
    // orgqr (zung2r order): Q = H_0 ... H_{n-1} [I; 0] in place
    for (int j = n - 1; j >= 0; --j) {
        const double tr = taur[j], ti = taui[j];
        if (j < n - 1) {
            for (int cc = j + 1 + wave; cc < n; cc += NT / 64) {
                double wr = 0.0, wi = 0.0;
                for (int i = j + 1 + lane; i < m; i += 64) {
                    double pr, pi;
                    cmul(Re[i + j * ld], -Im[i + j * ld], Re[i + cc * ld], Im[i + cc * ld], pr, pi);
                    wr += pr;
                    wi += pi;
                }
                wr = wave_sum(wr) + Re[j + cc * ld];
                wi = wave_sum(wi) + Im[j + cc * ld];
                double fr, fi;
                cmul(wr, wi, tr, ti, fr, fi);
                for (int i = j + 1 + lane; i < m; i += 64) {
                    double pr, pi;
                    cmul(Re[i + j * ld], Im[i + j * ld], fr, fi, pr, pi);
                    Re[i + cc * ld] -= pr;
                    Im[i + cc * ld] -= pi;
                }
                if (lane == 0) {
                    Re[j + cc * ld] -= fr;
                    Im[j + cc * ld] -= fi;
                }
            }
        }
        __syncthreads();
        for (int i = tid; i < m; i += NT) {
            if (i > j) {
                double pr, pi;
                cmul(Re[i + j * ld], Im[i + j * ld], -tr, -ti, pr, pi);
                Re[i + j * ld] = pr;
                Im[i + j * ld] = pi;
            } else if (i == j) {
                Re[i + j * ld] = 1.0 - tr;
                Im[i + j * ld] = -ti;
            } else {
                Re[i + j * ld] = 0.0;
                Im[i + j * ld] = 0.0;
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(NT) void k_inv(InvArgs q) {
    const int c = blockIdx.x;
    const ChainDev &ch = q.chains[c];
    const int n = dim_of(ch, q.ndim);
    const Opnd A = q.A[c];
    const size_t ld = A.ld;
    double *Re = A.re, *Im = A.im;
    __shared__ int piv[512];
    __shared__ double colr[512], coli[512], rowr[512], rowi[512];
    __shared__ double redv[NT / 64];
    __shared__ int redi[NT / 64];
    __shared__ int s_p;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double logdet = 0.0;
    int info = 0;
    for (int k = 0; k < n; ++k) {
        // pivot search: argmax |A[i,k]|, i >= k (first index on ties)
        double best = -1.0;
        int bi = n;
        for (int i = k + tid; i < n; i += NT) {
            const double a = Re[i + k * ld], b = Im[i + k * ld];
            const double v = a * a + b * b;
            if (v > best) { best = v; bi = i; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        if (lane == 0) { redv[wave] = best; redi[wave] = bi; }
        __syncthreads();
        if (tid == 0) {
            double b = redv[0];
            int p = redi[0];
            for (int w = 1; w < NT / 64; ++w)
                if (redv[w] > b || (redv[w] == b && redi[w] < p)) { b = redv[w]; p = redi[w]; }
            s_p = p;
            piv[k] = p;
        }
        __syncthreads();
        const int p = s_p;
        if (p != k) {
            for (int j = tid; j < n; j += NT) {
                const double ar = Re[k + j * ld], ai = Im[k + j * ld];
                Re[k + j * ld] = Re[p + j * ld];
                Im[k + j * ld] = Im[p + j * ld];
                Re[p + j * ld] = ar;
                Im[p + j * ld] = ai;
            }
        }
        __syncthreads();
        const double pr = Re[k + k * ld], pim = Im[k + k * ld];
        const double p2 = pr * pr + pim * pim;
        if (p2 == 0.0) {
            if (!info) info = k + 1;
            continue;
        }
        logdet += 0.5 * log(p2);
        const double ir = pr / p2, ii = -pim / p2;   // 1 / pivot
        for (int j = tid; j < n; j += NT) {
            double vr, vi;
            if (j == k) { vr = ir; vi = ii; }
            else cmul(Re[k + j * ld], Im[k + j * ld], ir, ii, vr, vi);
            rowr[j] = vr;
            rowi[j] = vi;
            colr[j] = Re[j + k * ld];
            coli[j] = Im[j + k * ld];
        }
        __syncthreads();
        for (int j = tid; j < n; j += NT) {
            Re[k + j * ld] = rowr[j];
            Im[k + j * ld] = rowi[j];
        }
        // eliminate: A[i,j] -= A[i,k] * rowk[j] (j != k), A[i,k] = -A[i,k] / pivot
        for (int e = tid; e < n * n; e += NT) {
            const int i = e % n, j = e / n;
            if (i == k) continue;
            double tr, ti;
            if (j == k) {
                cmul(colr[i], coli[i], ir, ii, tr, ti);
                Re[i + j * ld] = -tr;
                Im[i + j * ld] = -ti;
            } else {
                cmul(colr[i], coli[i], rowr[j], rowi[j], tr, ti);
                Re[i + j * ld] -= tr;
                Im[i + j * ld] -= ti;
            }
        }
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, in reverse order
    for (int k = n - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k)
            for (int i = tid; i < n; i += NT) {
                const double ar = Re[i + k * ld], ai = Im[i + k * ld];
                Re[i + k * ld] = Re[i + p * ld];
                Im[i + k * ld] = Im[i + p * ld];
                Re[i + p * ld] = ar;
                Im[i + p * ld] = ai;
            }
        __syncthreads();
    }
    if (tid == 0) {
        if (q.logdet) q.logdet[c] = logdet;
        if (q.info && info) q.info[c] = info;
    }
}

__global__ void k_copy(const ChainDev *chains, const Opnd *src, const Opnd *dst, int mdim, int ndim,
                       int cplx, int to_cplx) {
    const int c = blockIdx.z;
    const ChainDev &ch = chains[c];
    const int M = dim_of(ch, mdim), N = dim_of(ch, ndim);
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= M || j >= N) return;
    const Opnd s = src[c], d = dst[c];
    d.re[i + (size_t)j * d.ld] = s.re[i + (size_t)j * s.ld];
    if (to_cplx) d.im[i + (size_t)j * d.ld] = 0.0;
    else if (cplx) d.im[i + (size_t)j * d.ld] = s.im[i + (size_t)j * s.ld];
}

__global__ void k_initial_s(const ChainDev *chains, int nt) {
    const int c = blockIdx.y;
    const ChainDev &ch = chains[c];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ch.n_elem * nt) return;
    const int i = e % ch.n_elem, lt = e / ch.n_elem;
    ch.s[e] = (uint8_t)init_field_value(ch.seed, ch.chain_id, 0, (uint32_t)lt, (uint32_t)i);
}

__global__ void k_zero_meas(const ChainDev *chains) {
    const ChainDev &ch = chains[blockIdx.x];
    if (threadIdx.x < KDQ_MEAS_SLOTS) ch.meas[threadIdx.x] = 0.0;
}

}  // namespace

hipError_t launch_qr(hipStream_t st, int nchains, const QRArgs &q) {
    hipLaunchKernelGGL(k_qr, dim3(nchains), dim3(NT), 0, st, q);
    return hipGetLastError();
}

hipError_t launch_inv(hipStream_t st, int nchains, const InvArgs &q) {
    hipLaunchKernelGGL(k_inv, dim3(nchains), dim3(NT), 0, st, q);
    return hipGetLastError();
}

hipError_t launch_copy_mat(hipStream_t st, int nchains, int max_m, int max_n, const ChainDev *chains,
                           const Opnd *src, const Opnd *dst, int mdim, int ndim, bool cplx) {
    dim3 grid((max_m + 63) / 64, (max_n + 3) / 4, nchains);
    hipLaunchKernelGGL(k_copy, grid, dim3(256), 0, st, chains, src, dst, mdim, ndim, cplx ? 1 : 0, 0);
    return hipGetLastError();
}

hipError_t launch_real_to_cplx(hipStream_t st, int nchains, int max_m, int max_n, const ChainDev *chains,
                               const Opnd *src, const Opnd *dst, int mdim, int ndim) {
    dim3 grid((max_m + 63) / 64, (max_n + 3) / 4, nchains);
    hipLaunchKernelGGL(k_copy, grid, dim3(256), 0, st, chains, src, dst, mdim, ndim, 0, 1);
    return hipGetLastError();
}

hipError_t launch_initial_s(hipStream_t st, int nchains, int max_elem, int nt, const ChainDev *chains) {
    dim3 grid((max_elem * nt + 255) / 256, nchains);
    hipLaunchKernelGGL(k_initial_s, grid, dim3(256), 0, st, chains, nt);
    return hipGetLastError();
}

hipError_t launch_zero_meas(hipStream_t st, int nchains, const ChainDev *chains) {
    hipLaunchKernelGGL(k_zero_meas, dim3(nchains), dim3(64), 0, st, chains);
    return hipGetLastError();
}
