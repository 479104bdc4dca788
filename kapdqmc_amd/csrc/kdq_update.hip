// Local HS-field update of the t-U model as delayed (block) Sherman-Morrison updates.
//
// Reference: UpdatePhyLayer! (src/tU/phy_update.jl:256-266) visits the sites of slice lt
// in order; each proposal needs the current G_ii (get_r!, src/tU/GreenMatrix.jl:49-54)
// and each accepted flip applies the rank-1 update of Gupdate! (src/public/Gupdate.jl:5-11)
//     G <- G - G[:,i] r (e_i - G[i,:]),   r = Δ / (1 + Δ (1 - G_ii)).
// Here the sites are processed in blocks S of k consecutive sites:
//   k_decide_tu (one workgroup per chain) loads G[S,S], runs the k sequential proposals
//   on that k x k submatrix (kept in registers, rank-1 updates only touch S x S), and
//   builds by bordering the k_a x k_a matrix W with
//       G_new = G - G[:,A] W (E_A - G[A,:])          (A = accepted sites, in order)
//   which equals the product of the sequential rank-1 updates exactly.
//   k_flush applies that rank-|A| correction to the whole N x N matrix (all workgroups).
#include "kdq_kernels.h"

namespace {

constexpr int KB = 64;   // maximum block size
constexpr int LP = KB + 1;

__global__ __launch_bounds__(256) void k_decide_tu(DecideArgs a) {
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
    __shared__ double G0r[KB][LP], G0i[KB][LP];   // original G[S,S]
    __shared__ double Wr[KB][LP], Wi[KB][LP];     // Woodbury factor (a x a)
    __shared__ double colr[KB], coli[KB], rowr[KB], rowi[KB];
    __shared__ double dgr[KB], dgi[KB];
    __shared__ double xr[KB], xi[KB], yr[KB], yi[KB];
    __shared__ double us[KB];
    __shared__ int chs[KB], accl[KB];
    __shared__ int sv[KB];
    __shared__ double dtr[4][4], dti[4][4];

    const Opnd G = a.G[c];
    const int lt = a.lt;
    // stage G[S,S]
    for (int e = tid; e < kb * kb; e += 256) {
        const int r = e % kb, q = e / kb;
        const size_t off = (size_t)(s0 + r) + (size_t)(s0 + q) * G.ld;
        G0r[r][q] = G.re[off];
        G0i[r][q] = G.im[off];
    }
    const uint32_t tag = (uint32_t)lt | ((uint32_t)a.dir << 31);
    if (tid < kb) {
        int chv;
        double u;
        proposal_draw(ch.seed, ch.chain_id, 0, a.sweep, tag, (uint32_t)(s0 + tid), chv, u);
        chs[tid] = chv;
        us[tid] = u;
        sv[tid] = ch.s[(size_t)(s0 + tid) + (size_t)(lt - 1) * ch.n_elem];
    }
    if (tid < 16) {
        const int si = tid >> 2, sx = tid & 3;
        const double x = ch.alpha[lt - 1] * (ch.eta[sx] - ch.eta[si]);
        double sn, cs;
        sincos(x, &sn, &cs);
        dtr[si][sx] = cs - 1.0;
        dti[si][sx] = sn;
    }
    __syncthreads();
    // working copy: thread owns row r, columns g*16 .. g*16+15
    const int r = tid & 63, g = tid >> 6;
    double wr[16], wi[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int cc = g * 16 + q;
        wr[q] = (r < kb && cc < kb) ? G0r[r][cc] : 0.0;
        wi[q] = (r < kb && cc < kb) ? G0i[r][cc] : 0.0;
    }
    if (tid < kb) {
        dgr[tid] = G0r[tid][tid];
        dgi[tid] = G0i[tid][tid];
    }
    __syncthreads();

    int na = 0;
    for (int m = 0; m < kb; ++m) {
        const int si = sv[m];
        const int cidx = chs[m];
        const int sx = cidx + 1 + ((cidx + 1) >= si ? 1 : 0);
        const double Dr = dtr[si - 1][sx - 1], Di = dti[si - 1][sx - 1];
        // p = 1 + Δ (1 - G_mm)
        const double omr = 1.0 - dgr[m], omi = -dgi[m];
        double pr, pi;
        cmul(Dr, Di, omr, omi, pr, pi);
        pr += 1.0;
        const double p2 = pr * pr + pi * pi;
        const double prob = p2 * (ch.gamma[sx - 1] / ch.gamma[si - 1]);
        if (!(us[m] < prob)) continue;
        // r = Δ / p
        const double rr = (Dr * pr + Di * pi) / p2, ri = (Di * pr - Dr * pi) / p2;
        // publish column m and row m of the working submatrix
        if ((m >> 4) == g && r < kb) {
            colr[r] = wr[m & 15];
            coli[r] = wi[m & 15];
        }
        if (r == m) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                rowr[g * 16 + q] = wr[q];
                rowi[g * 16 + q] = wi[q];
            }
        }
        __syncthreads();
        // rank-1 update of the working submatrix
        if (r < kb) {
            double fr, fi;
            cmul(colr[r], coli[r], rr, ri, fr, fi);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int cc = g * 16 + q;
                const double vr = (cc == m ? 1.0 : 0.0) - rowr[cc], vi = -rowi[cc];
                double tr, ti;
                cmul(fr, fi, vr, vi, tr, ti);
                wr[q] -= tr;
                wi[q] -= ti;
            }
        }
        // bordering vectors: x = W G0[A, m], y = G0[m, A] W
        if (tid < na) {
            double sr = 0.0, sim = 0.0;
            for (int u = 0; u < na; ++u) {
                double tr, ti;
                cmul(Wr[tid][u], Wi[tid][u], G0r[accl[u]][m], G0i[accl[u]][m], tr, ti);
                sr += tr;
                sim += ti;
            }
            xr[tid] = sr;
            xi[tid] = sim;
        } else if (tid >= 128 && tid < 128 + na) {
            const int u = tid - 128;
            double sr = 0.0, sim = 0.0;
            for (int t = 0; t < na; ++t) {
                double tr, ti;
                cmul(G0r[m][accl[t]], G0i[m][accl[t]], Wr[t][u], Wi[t][u], tr, ti);
                sr += tr;
                sim += ti;
            }
            yr[u] = sr;
            yi[u] = sim;
        }
        __syncthreads();
        // diagonal of the working matrix
        if (r < kb && (r >> 4) == g) {
            dgr[r] = wr[r & 15];
            dgi[r] = wi[r & 15];
        }
        // W' = [[W + x r y, x r], [r y, r]]
        const int na1 = na + 1;
        for (int e = tid; e < na1 * na1; e += 256) {
            const int t = e % na1, u = e / na1;
            if (t < na && u < na) {
                double tr, ti, vr2, vi2;
                cmul(xr[t], xi[t], rr, ri, tr, ti);
                cmul(tr, ti, yr[u], yi[u], vr2, vi2);
                Wr[t][u] += vr2;
                Wi[t][u] += vi2;
            } else if (t < na) {
                cmul(xr[t], xi[t], rr, ri, Wr[t][u], Wi[t][u]);
            } else if (u < na) {
                cmul(rr, ri, yr[u], yi[u], Wr[t][u], Wi[t][u]);
            } else {
                Wr[t][u] = rr;
                Wi[t][u] = ri;
            }
        }
        if (tid == 0) {
            accl[na] = m;
            sv[m] = sx;
        }
        __syncthreads();
        na = na1;
    }
    // outputs
    if (tid == 0) {
        *ch.acc_cnt = na;
        ch.counters[0] += kb;
        ch.counters[1] += na;
    }
    if (tid < na) ch.acc_idx[tid] = s0 + accl[tid];
    for (int e = tid; e < na * na; e += 256) {
        const int t = e % na, u = e / na;
        ch.W[t + u * KB] = Wr[t][u];
        ch.W[KB * KB + t + u * KB] = Wi[t][u];
    }
    if (tid < kb) ch.s[(size_t)(s0 + tid) + (size_t)(lt - 1) * ch.n_elem] = (uint8_t)sv[tid];
}

// Gout[I,J] = Gin[I,J] - Gin[I,A] · (W · Y[:,J]),  Y[t, j] = δ(A_t, j) - Gin[A_t, j].
__global__ __launch_bounds__(256) void k_flush(FlushArgs f) {
    const int c = blockIdx.z;
    const ChainDev &ch = f.chains[c];
    const int n = ch.ns;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    if (i0 >= n || j0 >= n) return;
    const Opnd Gi = f.Gin[c], Go = f.Gout[c];
    const int na = *ch.acc_cnt;
    const int tid = threadIdx.x;
    const int ii = tid & 63, jg = tid >> 6;
    if (na == 0) {
        for (int q = 0; q < 16; ++q) {
            const int i = i0 + ii, j = j0 + jg * 16 + q;
            if (i < n && j < n) {
                Go.re[(size_t)i + (size_t)j * Go.ld] = Gi.re[(size_t)i + (size_t)j * Gi.ld];
                Go.im[(size_t)i + (size_t)j * Go.ld] = Gi.im[(size_t)i + (size_t)j * Gi.ld];
            }
        }
        return;
    }
    __shared__ double Ar[KB][LP], Ai[KB][LP];   // W, later Gin[I, A]
    __shared__ double Pr[KB][LP], Pi[KB][LP];   // Y[:, J], later Z
    __shared__ int aidx[KB];
    if (tid < na) aidx[tid] = ch.acc_idx[tid];
    for (int e = tid; e < na * na; e += 256) {
        const int t = e % na, u = e / na;
        Ar[t][u] = ch.W[t + u * KB];
        Ai[t][u] = ch.W[KB * KB + t + u * KB];
    }
    __syncthreads();
    for (int e = tid; e < na * 64; e += 256) {
        const int jj = e & 63, t = e >> 6;
        const int j = j0 + jj;
        double yr = 0.0, yi = 0.0;
        if (j < n) {
            const size_t off = (size_t)aidx[t] + (size_t)j * Gi.ld;
            yr = (aidx[t] == j ? 1.0 : 0.0) - Gi.re[off];
            yi = -Gi.im[off];
        }
        Pr[t][jj] = yr;
        Pi[t][jj] = yi;
    }
    __syncthreads();
    // Z = W · Y  (thread: column jj = ii, rows t = jg + 4 q)
    double zr[16], zi[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int t = jg + 4 * q;
        double sr = 0.0, sim = 0.0;
        if (t < na) {
            for (int u = 0; u < na; ++u) {
                double tr, ti;
                cmul(Ar[t][u], Ai[t][u], Pr[u][ii], Pi[u][ii], tr, ti);
                sr += tr;
                sim += ti;
            }
        }
        zr[q] = sr;
        zi[q] = sim;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int t = jg + 4 * q;
        if (t < na) {
            Pr[t][ii] = zr[q];
            Pi[t][ii] = zi[q];
        }
    }
    // Gin[I, A] into the W space
    for (int e = tid; e < na * 64; e += 256) {
        const int r = e & 63, t = e >> 6;
        const int i = i0 + r;
        double vr = 0.0, vi = 0.0;
        if (i < n) {
            const size_t off = (size_t)i + (size_t)aidx[t] * Gi.ld;
            vr = Gi.re[off];
            vi = Gi.im[off];
        }
        Ar[r][t] = vr;
        Ai[r][t] = vi;
    }
    __syncthreads();
    const int i = i0 + ii;
    if (i >= n) return;
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
        const int jj = jg * 16 + q, j = j0 + jj;
        if (j >= n) break;
        double sr = 0.0, sim = 0.0;
        for (int t = 0; t < na; ++t) {
            double tr, ti;
            cmul(Ar[ii][t], Ai[ii][t], Pr[t][jj], Pi[t][jj], tr, ti);
            sr += tr;
            sim += ti;
        }
        const size_t oi = (size_t)i + (size_t)j * Gi.ld, oo = (size_t)i + (size_t)j * Go.ld;
        Go.re[oo] = Gi.re[oi] - sr;
        Go.im[oo] = Gi.im[oi] - sim;
    }
}

}  // namespace

hipError_t launch_decide_tu(hipStream_t st, int nchains, const DecideArgs &a) {
    if (a.kb < 1 || a.kb > KB) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_decide_tu, dim3(nchains), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_flush(hipStream_t st, int nchains, int max_ns, const FlushArgs &f) {
    const int t = (max_ns + 63) / 64;
    hipLaunchKernelGGL(k_flush, dim3(t, t, nchains), dim3(256), 0, st, f);
    return hipGetLastError();
}
