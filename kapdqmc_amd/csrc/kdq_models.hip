// t-V and t-U-V kernels: rank-2 block decisions, 2x2 bond rotations and observables.
//
// Rank-2 local update (t-V: src/tV/phy_update.jl:190-205 with get_r! at
// src/tV/GreenMatrix.jl:76-89; t-U-V: src/tUV/phy_update.jl:188-203 with get_r! at
// src/tUV/GreenMatrix.jl:73-89). A flip of field element e touches the two sites
// (x_e, y_e); with the 2x2 matrix Δ of the flip
//     r = I + Δ (I - G[sub,sub]),   p = det r (t-V) or |det r|² (t-U-V),
//     G <- G - G[:,sub] (r⁻¹Δ) (E_sub - G[sub,:]).
// As for t-U, a block of up to 32 consecutive elements is decided on the submatrix of the
// ≤ 64 sites it touches (in registers), and the accepted flips are folded into one
// Woodbury factor W built by bordering:
//     x = W (E_A[:,sub] - G[A,sub]),  y = G[sub,A] W,  w = r⁻¹Δ,
//     W' = [[W - x w y, -x w], [w y, w]],  A' = A ∪ sub   (sites may repeat in A),
// then k_flush applies G <- G - G[:,A] W (E_A - G[A,:]).
#include "kdq_kernels.h"

namespace {

constexpr int KS = 64;        // maximum sites per block
constexpr int KE = 32;        // maximum elements per block
constexpr int LP = KS + 1;

__global__ __launch_bounds__(256) void k_decide_r2(DecideR2Args a) {
    const int c = blockIdx.x;
    const ChainDev &ch = a.chains[c];
    const int tid = threadIdx.x;
    if (a.block >= a.nblocks[c]) {
        if (tid == 0) *ch.acc_cnt = 0;
        return;
    }
    const BlockDesc &bd = a.blocks[a.block_offset[c] + a.family * a.nblocks[c] + a.block];
    const int ns = bd.nslot, ne = bd.ne, e0 = bd.e0;
    const int n = ch.ns;
    __shared__ double G0r[KS][LP], G0i[KS][LP];
    __shared__ double Wr[KS][LP], Wi[KS][LP];
    __shared__ double colr[2][KS], coli[2][KS], rowr[2][KS], rowi[2][KS];
    __shared__ double blkr[KE][4], blki[KE][4];   // current 2x2 block of every element
    __shared__ double xr[KS][2], xi[KS][2], yr[2][KS], yi[2][KS];
    __shared__ double us[KE];
    __shared__ int chs[KE], sv[KE], acc[KS], gsite[KS];
    __shared__ int sx0[KE], sx1[KE];
    __shared__ double wsh[8];
    // per (s, s') = 4 si + sx: Δ (2x2) and the weight factor (extra · γ'/γ)
    __shared__ double dtab_r[16][4], dtab_i[16][4], ftab[16];

    const Opnd G = a.G[c];
    const int lt = a.lt;
    if (tid < ns) gsite[tid] = bd.sites[tid];
    if (tid < ne) {
        sx0[tid] = bd.sl0[tid];
        sx1[tid] = bd.sl1[tid];
    }
    __syncthreads();
    for (int e = tid; e < ns * ns; e += 256) {
        const int r = e % ns, q = e / ns;
        const size_t off = (size_t)gsite[r] + (size_t)gsite[q] * G.ld;
        G0r[r][q] = G.re[off];
        G0i[r][q] = G.im[off];
    }
    const uint32_t tag = (uint32_t)lt | ((uint32_t)a.dir << 31);
    const int kind = ch.kind;
    const double h = 0.7071067811865476;   // 2^0.5 / 2
    if (tid >= 128 && tid < 144) {
        const int si = (tid - 128) >> 2, sx = (tid - 128) & 3;   // 0-based old / new values
        double Dr[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, Di[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
        double extra = 1.0;
        if (kind == 1) {   // t-V: Δ = uv diag(e^{v}-1, e^{-v}-1) uv
            const double v = ch.alpha[lt - 1] * (ch.eta[sx] - ch.eta[si]);
            const double t0 = exp(v) - 1.0, t1 = exp(-v) - 1.0;
            const double uv[2][2] = {{-h, -h}, {-h, h}};
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int q = 0; q < 2; ++q) Dr[p][q] = (uv[p][0] * t0) * uv[0][q] + (uv[p][1] * t1) * uv[1][q];
        } else {           // t-U-V: Δ = diag(f(Δs) - 1, f(a Δs) - 1)
            const double ds = ch.eta[sx] - ch.eta[si];
            if (ch.is_complex) {
                double s0, c0, s1, c1;
                sincos(ds, &s0, &c0);
                sincos(ch.a * ds, &s1, &c1);
                Dr[0][0] = c0 - 1.0; Di[0][0] = s0;
                Dr[1][1] = c1 - 1.0; Di[1][1] = s1;
            } else {
                Dr[0][0] = exp(ds) - 1.0;
                Dr[1][1] = exp(ch.a * ds) - 1.0;
                extra = exp((ch.eta[si] - ch.eta[sx]) * (ch.a + 1.0));
            }
        }
        const int k = tid - 128;
        dtab_r[k][0] = Dr[0][0]; dtab_r[k][1] = Dr[0][1]; dtab_r[k][2] = Dr[1][0]; dtab_r[k][3] = Dr[1][1];
        dtab_i[k][0] = Di[0][0]; dtab_i[k][1] = Di[0][1]; dtab_i[k][2] = Di[1][0]; dtab_i[k][3] = Di[1][1];
        ftab[k] = extra * (ch.gamma[sx] / ch.gamma[si]);
    }
    if (tid < ne) {
        int chv;
        double u;
        const uint32_t sweep = a.sweep_dev ? *a.sweep_dev : a.sweep;
        sweep_draw(ch, sweep, lt, a.dir, (uint32_t)(e0 + tid), chv, u);
        chs[tid] = chv;
        us[tid] = u;
        sv[tid] = ch.s[(size_t)(e0 + tid) + (size_t)(lt - 1) * ch.n_elem];
    }
    __syncthreads();
    const int r = tid & 63, g = tid >> 6;
    double wr[16], wi[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int cc = g * 16 + q;
        wr[q] = (r < ns && cc < ns) ? G0r[r][cc] : 0.0;
        wi[q] = (r < ns && cc < ns) ? G0i[r][cc] : 0.0;
    }
    if (tid < ne) {
        const int x = sx0[tid], y = sx1[tid];
        blkr[tid][0] = G0r[x][x]; blki[tid][0] = G0i[x][x];
        blkr[tid][1] = G0r[x][y]; blki[tid][1] = G0i[x][y];
        blkr[tid][2] = G0r[y][x]; blki[tid][2] = G0i[y][x];
        blkr[tid][3] = G0r[y][y]; blki[tid][3] = G0i[y][y];
    }
    __syncthreads();

    int na = 0, nties = 0;
    for (int m = 0; m < ne; ++m) {
        const int si = sv[m];
        const int cidx = chs[m];
        const int sx = cidx + 1 + ((cidx + 1) >= si ? 1 : 0);
        const int k = (si - 1) * 4 + (sx - 1);
        double Dr[2][2], Di[2][2];
        Dr[0][0] = dtab_r[k][0]; Dr[0][1] = dtab_r[k][1]; Dr[1][0] = dtab_r[k][2]; Dr[1][1] = dtab_r[k][3];
        Di[0][0] = dtab_i[k][0]; Di[0][1] = dtab_i[k][1]; Di[1][0] = dtab_i[k][2]; Di[1][1] = dtab_i[k][3];
        // r = I + Δ (I - G_ss)
        double Omr[2][2], Omi[2][2];
        Omr[0][0] = 1.0 - blkr[m][0]; Omi[0][0] = -blki[m][0];
        Omr[0][1] = -blkr[m][1];      Omi[0][1] = -blki[m][1];
        Omr[1][0] = -blkr[m][2];      Omi[1][0] = -blki[m][2];
        Omr[1][1] = 1.0 - blkr[m][3]; Omi[1][1] = -blki[m][3];
        double Rr[2][2], Ri[2][2];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                double t0r, t0i, t1r, t1i;
                cmul(Dr[p][0], Di[p][0], Omr[0][q], Omi[0][q], t0r, t0i);
                cmul(Dr[p][1], Di[p][1], Omr[1][q], Omi[1][q], t1r, t1i);
                Rr[p][q] = t0r + t1r + (p == q ? 1.0 : 0.0);
                Ri[p][q] = t0i + t1i;
            }
        double detr, deti, tr, ti;
        cmul(Rr[0][0], Ri[0][0], Rr[1][1], Ri[1][1], detr, deti);
        cmul(Rr[0][1], Ri[0][1], Rr[1][0], Ri[1][0], tr, ti);
        detr -= tr;
        deti -= ti;
        double prob = (kind == 1) ? detr : (detr * detr + deti * deti);
        prob *= ftab[k];
        nties += fabs(us[m] - prob) < KDQ_NEAR_TIE;
        if (!(us[m] < prob)) continue;
        // w = r⁻¹ Δ
        const double d2 = detr * detr + deti * deti;
        const double idr = detr / d2, idi = -deti / d2;
        double Ir[2][2], Ii[2][2];
        cmul(Rr[1][1], Ri[1][1], idr, idi, Ir[0][0], Ii[0][0]);
        cmul(-Rr[0][1], -Ri[0][1], idr, idi, Ir[0][1], Ii[0][1]);
        cmul(-Rr[1][0], -Ri[1][0], idr, idi, Ir[1][0], Ii[1][0]);
        cmul(Rr[0][0], Ri[0][0], idr, idi, Ir[1][1], Ii[1][1]);
        double Wwr[2][2], Wwi[2][2];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                double t0r, t0i, t1r, t1i;
                cmul(Ir[p][0], Ii[p][0], Dr[0][q], Di[0][q], t0r, t0i);
                cmul(Ir[p][1], Ii[p][1], Dr[1][q], Di[1][q], t1r, t1i);
                Wwr[p][q] = t0r + t1r;
                Wwi[p][q] = t0i + t1i;
            }
        const int sxl[2] = {sx0[m], sx1[m]};
        // publish columns and rows x, y of the working submatrix
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int sl = sxl[k];
            if ((sl >> 4) == g && r < ns) {
                colr[k][r] = wr[sl & 15];
                coli[k][r] = wi[sl & 15];
            }
            if (r == sl) {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    rowr[k][g * 16 + q] = wr[q];
                    rowi[k][g * 16 + q] = wi[q];
                }
            }
        }
        __syncthreads();
        // working update: Gw[p][q] -= Σ_k Gw[p][sub_k] (w (E_sub - Gw[sub,:]))[k][q]
        if (r < ns) {
            double fr[2], fi[2];   // f_k = Σ_l Gw[r][sub_l] w[l][k]
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double t0r, t0i, t1r, t1i;
                cmul(colr[0][r], coli[0][r], Wwr[0][k], Wwi[0][k], t0r, t0i);
                cmul(colr[1][r], coli[1][r], Wwr[1][k], Wwi[1][k], t1r, t1i);
                fr[k] = t0r + t1r;
                fi[k] = t0i + t1i;
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int cc = g * 16 + q;
                double acr = 0.0, aci = 0.0;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const double er = (cc == sxl[k] ? 1.0 : 0.0) - rowr[k][cc], ei = -rowi[k][cc];
                    double t0r, t0i;
                    cmul(fr[k], fi[k], er, ei, t0r, t0i);
                    acr += t0r;
                    aci += t0i;
                }
                wr[q] -= acr;
                wi[q] -= aci;
            }
        }
        // bordering vectors x = W (E_A[:,sub] - G0[A,sub]) (na x 2), y = G0[sub,A] W (2 x na)
        if (tid < na) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double sr = 0.0, sim = 0.0;
                for (int u = 0; u < na; ++u) {
                    const double er = (acc[u] == sxl[k] ? 1.0 : 0.0) - G0r[acc[u]][sxl[k]];
                    const double ei = -G0i[acc[u]][sxl[k]];
                    double t0r, t0i;
                    cmul(Wr[tid][u], Wi[tid][u], er, ei, t0r, t0i);
                    sr += t0r;
                    sim += t0i;
                }
                xr[tid][k] = sr;
                xi[tid][k] = sim;
            }
        } else if (tid >= 128 && tid < 128 + na) {
            const int u = tid - 128;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double sr = 0.0, sim = 0.0;
                for (int t = 0; t < na; ++t) {
                    double t0r, t0i;
                    cmul(G0r[sxl[k]][acc[t]], G0i[sxl[k]][acc[t]], Wr[t][u], Wi[t][u], t0r, t0i);
                    sr += t0r;
                    sim += t0i;
                }
                yr[k][u] = sr;
                yi[k][u] = sim;
            }
        }
        __syncthreads();
        // refresh the 2x2 blocks of the remaining elements from the working registers
        for (int mm = m + 1; mm < ne; ++mm) {
            const int px = sx0[mm], py = sx1[mm];
            const int rows[4] = {px, px, py, py}, cols[4] = {px, py, px, py};
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (rows[q] == r && (cols[q] >> 4) == g) {
                    blkr[mm][q] = wr[cols[q] & 15];
                    blki[mm][q] = wi[cols[q] & 15];
                }
        }
        // W' = [[W - x w y, -x w], [w y, w]]
        const int na2 = na + 2;
        for (int e = tid; e < na2 * na2; e += 256) {
            const int t = e % na2, u = e / na2;
            double vr, vi;
            if (t < na && u < na) {
                double xw0r, xw0i, xw1r, xw1i, t0r, t0i, t1r, t1i;   // (x w)[t][k]
                cmul(xr[t][0], xi[t][0], Wwr[0][0], Wwi[0][0], t0r, t0i);
                cmul(xr[t][1], xi[t][1], Wwr[1][0], Wwi[1][0], t1r, t1i);
                xw0r = t0r + t1r; xw0i = t0i + t1i;
                cmul(xr[t][0], xi[t][0], Wwr[0][1], Wwi[0][1], t0r, t0i);
                cmul(xr[t][1], xi[t][1], Wwr[1][1], Wwi[1][1], t1r, t1i);
                xw1r = t0r + t1r; xw1i = t0i + t1i;
                cmul(xw0r, xw0i, yr[0][u], yi[0][u], t0r, t0i);
                cmul(xw1r, xw1i, yr[1][u], yi[1][u], t1r, t1i);
                vr = Wr[t][u] - (t0r + t1r);
                vi = Wi[t][u] - (t0i + t1i);
            } else if (t < na) {
                const int k = u - na;
                double t0r, t0i, t1r, t1i;
                cmul(xr[t][0], xi[t][0], Wwr[0][k], Wwi[0][k], t0r, t0i);
                cmul(xr[t][1], xi[t][1], Wwr[1][k], Wwi[1][k], t1r, t1i);
                vr = -(t0r + t1r);
                vi = -(t0i + t1i);
            } else if (u < na) {
                const int k = t - na;
                double t0r, t0i, t1r, t1i;
                cmul(Wwr[k][0], Wwi[k][0], yr[0][u], yi[0][u], t0r, t0i);
                cmul(Wwr[k][1], Wwi[k][1], yr[1][u], yi[1][u], t1r, t1i);
                vr = t0r + t1r;
                vi = t0i + t1i;
            } else {
                vr = Wwr[t - na][u - na];
                vi = Wwi[t - na][u - na];
            }
            Wr[t][u] = vr;
            Wi[t][u] = vi;
        }
        if (tid == 0) {
            acc[na] = sxl[0];
            acc[na + 1] = sxl[1];
            sv[m] = sx;
        }
        __syncthreads();
        na = na2;
    }
    if (tid == 0) {
        *ch.acc_cnt = na;
        ch.counters[0] += ne;
        ch.counters[1] += na / 2;
        ch.counters[2] += nties;
    }
    if (tid < na) ch.acc_idx[tid] = gsite[acc[tid]];
    for (int e = tid; e < na * na; e += 256) {
        const int t = e % na, u = e / na;
        ch.W[t + u * KS] = Wr[t][u];
        ch.W[KS * KS + t + u * KS] = Wi[t][u];
    }
    if (tid < ne) ch.s[(size_t)(e0 + tid) + (size_t)(lt - 1) * ch.n_elem] = (uint8_t)sv[tid];
}

// t-V bond rotations of family j at slice lt: M = blockdiag over bonds (x, y) of
// m(v) = uv diag(e^{v}, e^{-v}) uv, v = sign α[lt] η[s]; m(v)⁻¹ = m(-v).
// mode 0: X <- M X M⁻¹ (both sides, N x N), mode 1: X <- M X (rows only, N x ncols),
// mode 2: X <- X M⁻¹ (columns only, nrows x N).
KDQ_DEV void bond_m(const ChainDev &ch, int e, int lt, int sign, double m[2][2]) {
    const double h = 0.7071067811865476;
    const double v = sign * ch.alpha[lt - 1] * ch.eta[ch.s[(size_t)e + (size_t)(lt - 1) * ch.n_elem] - 1];
    const double a = exp(v), b = exp(-v);
    const double uv[2][2] = {{-h, -h}, {-h, h}};
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) m[p][q] = (uv[p][0] * a) * uv[0][q] + (uv[p][1] * b) * uv[1][q];
}

__global__ void k_bondrot(BondRotArgs a) {
    const int c = blockIdx.z;
    const ChainDev &ch = a.chains[c];
    const int nb = ch.nb;
    const Opnd X = a.X[c];
    const int ncols = a.mode == 0 ? nb : dim_of(ch, a.ncol_dim);
    const int bi = blockIdx.x * 64 + (threadIdx.x & 63);   // row bond
    const int cj = blockIdx.y * 4 + (threadIdx.x >> 6);    // column bond (mode 0) or column
    if (bi >= nb || cj >= ncols) return;
    const int e = a.family * nb + bi;
    const int x = ch.ex[e], y = ch.ey[e];
    double ml[2][2];
    bond_m(ch, e, a.lt, a.sign, ml);
    if (a.mode == 2) {   // row cj, columns x and y: [X_x X_y] <- [X_x X_y] m(-v)
        double mr[2][2];
        bond_m(ch, e, a.lt, -a.sign, mr);
        const size_t o0 = (size_t)cj + (size_t)x * X.ld, o1 = (size_t)cj + (size_t)y * X.ld;
        const double ar = X.re[o0], ai = X.im[o0], br = X.re[o1], bi2 = X.im[o1];
        X.re[o0] = ar * mr[0][0] + br * mr[1][0];
        X.im[o0] = ai * mr[0][0] + bi2 * mr[1][0];
        X.re[o1] = ar * mr[0][1] + br * mr[1][1];
        X.im[o1] = ai * mr[0][1] + bi2 * mr[1][1];
        return;
    }
    if (a.mode == 1) {
        const size_t o0 = (size_t)x + (size_t)cj * X.ld, o1 = (size_t)y + (size_t)cj * X.ld;
        const double ar = X.re[o0], ai = X.im[o0], br = X.re[o1], bi2 = X.im[o1];
        X.re[o0] = ml[0][0] * ar + ml[0][1] * br;
        X.im[o0] = ml[0][0] * ai + ml[0][1] * bi2;
        X.re[o1] = ml[1][0] * ar + ml[1][1] * br;
        X.im[o1] = ml[1][0] * ai + ml[1][1] * bi2;
        return;
    }
    const int e2 = a.family * nb + cj;
    const int x2 = ch.ex[e2], y2 = ch.ey[e2];
    double mr[2][2];
    bond_m(ch, e2, a.lt, -a.sign, mr);
    const int rr[2] = {x, y}, cc[2] = {x2, y2};
    double Br[2][2], Bi[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const size_t o = (size_t)rr[p] + (size_t)cc[q] * X.ld;
            Br[p][q] = X.re[o];
            Bi[p][q] = X.im[o];
        }
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int l = 0; l < 2; ++l) {
                    sr += ml[p][k] * Br[k][l] * mr[l][q];
                    si += ml[p][k] * Bi[k][l] * mr[l][q];
                }
            const size_t o = (size_t)rr[p] + (size_t)cc[q] * X.ld;
            X.re[o] = sr;
            X.im[o] = si;
        }
}

}  // namespace

hipError_t launch_decide_r2(hipStream_t st, int nchains, const DecideR2Args &a) {
    hipLaunchKernelGGL(k_decide_r2, dim3(nchains), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_bondrot(hipStream_t st, int nchains, int max_nb, int max_cols, const BondRotArgs &a) {
    dim3 grid((max_nb + 63) / 64, ((a.mode == 0 ? max_nb : max_cols) + 3) / 4, nchains);
    hipLaunchKernelGGL(k_bondrot, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

