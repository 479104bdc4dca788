// Small device helpers of the engine: strided matrix copies (optionally real -> complex),
// the initial HS field (Initial_s, src/tU/GreenMatrix.jl:1-8, drawn from the counter-based
// generator) and the reset of the per-chain measurement accumulators.
#include "kdq_kernels.h"

namespace {

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

__global__ void k_bump(uint32_t *counter) { *counter += 1; }

// FP64 -> FP32 copy of an n x n complex matrix; optionally max |dst_before - src| per chain
// (non-negative doubles order like their bit patterns, so an unsigned 64-bit max works)
__global__ void k_to_f32(const ChainDev *chains, const Opnd *src, const OpndF *dst, int ndim, double *maxerr) {
    const int c = blockIdx.z;
    const int n = dim_of(chains[c], ndim);
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    double err = 0.0;
    if (i < n && j < n) {
        const Opnd s = src[c];
        const OpndF d = dst[c];
        const size_t os = (size_t)i + (size_t)j * s.ld, od = (size_t)i + (size_t)j * d.ld;
        const double sr = s.re[os], si = s.im[os];
        if (maxerr) err = fmax(fabs((double)d.re[od] - sr), fabs((double)d.im[od] - si));
        d.re[od] = (float)sr;
        d.im[od] = (float)si;
    }
    if (!maxerr) return;
    err = wave_max(err);
    if ((threadIdx.x & 63) == 0 && err > 0.0)
        atomicMax(reinterpret_cast<unsigned long long *>(maxerr + c), (unsigned long long)__double_as_longlong(err));
}

// Health check at a stabilisation node: max |wrapped - fresh| per chain into maxerr (the
// reference's commented wrap-error check, src/tU/phy_update.jl:54-58,92-97) and a sticky flag
// when either matrix holds a non-finite entry. Also used for max |G² - G| (wrapped = G², fresh = G).
// One workgroup per 64 x 32 tile (each wave walks 8 columns), one atomic per workgroup.
constexpr int NC_COLS = 32;
__global__ void k_node_check(const ChainDev *chains, const Opnd *wrapped, const Opnd *fresh, int ndim,
                             double *maxerr, int *poison) {
    const int c = blockIdx.z;
    const int n = dim_of(chains[c], ndim);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    double err = 0.0;
    int bad = 0;
    if (i < n) {
        const Opnd w = wrapped[c], f = fresh[c];
        for (int jj = wave; jj < NC_COLS; jj += 4) {
            const int j = blockIdx.y * NC_COLS + jj;
            if (j >= n) break;
            const size_t ow = (size_t)i + (size_t)j * w.ld, of = (size_t)i + (size_t)j * f.ld;
            const double wr = w.re[ow], wi = w.im[ow], fr = f.re[of], fi = f.im[of];
            const bool b = !(isfinite(wr) && isfinite(wi) && isfinite(fr) && isfinite(fi));
            bad |= b;
            if (!b) err = fmax(err, fmax(fabs(wr - fr), fabs(wi - fi)));
        }
    }
    err = wave_max(err);
    const unsigned long long anybad = __ballot(bad);
    __shared__ double se[4];
    __shared__ int sb[4];
    if (lane == 0) {
        se[wave] = err;
        sb[wave] = anybad != 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double e = fmax(fmax(se[0], se[1]), fmax(se[2], se[3]));
        if (e > 0.0)
            atomicMax(reinterpret_cast<unsigned long long *>(maxerr + c), (unsigned long long)__double_as_longlong(e));
        if ((sb[0] | sb[1] | sb[2] | sb[3]) && poison) atomicOr(poison + c, 1);
    }
}

// out[c] = Re tr G of the current G (projector check: N - Np)
__global__ void k_trace(const ChainDev *chains, const Opnd *G, double *out) {
    const int c = blockIdx.x;
    const int n = chains[c].ns;
    const Opnd g = G[c];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) v += g.re[(size_t)i + (size_t)i * g.ld];
    v = wave_sum(v);
    if (threadIdx.x == 0) out[c] = v;
}

__global__ void k_to_f64(const ChainDev *chains, const OpndF *src, const Opnd *dst, int ndim, int cplx) {
    const int c = blockIdx.z;
    const int n = dim_of(chains[c], ndim);
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    const OpndF s = src[c];
    const Opnd d = dst[c];
    const size_t os = (size_t)i + (size_t)j * s.ld, od = (size_t)i + (size_t)j * d.ld;
    d.re[od] = s.re[os];
    if (cplx) d.im[od] = s.im[os];
}

__global__ void k_zero_meas(const ChainDev *chains) {
    const ChainDev &ch = chains[blockIdx.x];
    if (threadIdx.x < KDQ_MEAS_SLOTS) ch.meas[threadIdx.x] = 0.0;
}

}  // namespace

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

hipError_t launch_bump(hipStream_t st, uint32_t *counter) {
    hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, st, counter);
    return hipGetLastError();
}

hipError_t launch_to_f32(hipStream_t st, int nchains, int max_n, const ChainDev *chains, const Opnd *src,
                         const OpndF *dst, int ndim, double *maxerr) {
    dim3 grid((max_n + 63) / 64, (max_n + 3) / 4, nchains);
    hipLaunchKernelGGL(k_to_f32, grid, dim3(256), 0, st, chains, src, dst, ndim, maxerr);
    return hipGetLastError();
}

hipError_t launch_node_check(hipStream_t st, int nchains, int max_n, const ChainDev *chains, const Opnd *wrapped,
                             const Opnd *fresh, int ndim, double *maxerr, int *poison) {
    dim3 grid((max_n + 63) / 64, (max_n + NC_COLS - 1) / NC_COLS, nchains);
    hipLaunchKernelGGL(k_node_check, grid, dim3(256), 0, st, chains, wrapped, fresh, ndim, maxerr, poison);
    return hipGetLastError();
}

hipError_t launch_trace(hipStream_t st, int nchains, const ChainDev *chains, const Opnd *G, double *out) {
    hipLaunchKernelGGL(k_trace, dim3(nchains), dim3(64), 0, st, chains, G, out);
    return hipGetLastError();
}

hipError_t launch_to_f64(hipStream_t st, int nchains, int max_n, const ChainDev *chains, const OpndF *src,
                         const Opnd *dst, int ndim, int cplx) {
    dim3 grid((max_n + 63) / 64, (max_n + 3) / 4, nchains);
    hipLaunchKernelGGL(k_to_f64, grid, dim3(256), 0, st, chains, src, dst, ndim, cplx);
    return hipGetLastError();
}
