// Kinetic products by cell momenta (block-circulant eK, eK⁻¹).
//
// The hopping matrix of every lattice the models build (nnK_Matrix, src/public/Geometry.jl:
// 178-207: HoneyComb120/60, SQUARE90/45 with periodic L x L cells and uniform bond phases) is
// invariant under cell translations, so eK = exp(-Δτ K) and eK⁻¹ are block-circulant: with
// site s = 2·(x + L·y) + b, E[(c,a), (c',b)] = E_ab(c' - c). A product E·X then factors as
//     X̂_b(k) = Σ_c e^{-i k·c} X[(c,b)],   Ŷ_a(k) = Σ_b Λ_ab(k) X̂_b(k),   Y[(c,a)] = Σ_k e^{+i k·c} Ŷ_a(k)
// with Λ(k) = (1/L²) Σ_d E(d) e^{+i k·d} (ChainDev::circ). The cell transforms are two passes of
// length-L DFTs (along x, then y). Each workgroup holds LPW lines of X in LDS (columns for a
// left product E·X, rows for a right product X·E = (Eᵀ Xᵀ)ᵀ), transforms them, applies the 2 x 2
// blocks and transforms back: one read and one write of X, and 4 passes of 2L² outputs with L
// complex terms each per line (8L³ complex MACs) against (2L²)² for the dense product.
// op selects E (0), Eᵀ (1: Λ_ba(-k)) or Eᴴ (2: conj Λ_ba(k)). The diagonal factors of the GEMM
// path are applied in the same places: `in` scales the line entries before the product,
// `out` the products, `line` the whole line (per column of a left product, per row of a right
// product). Every line is read completely before it is written, so Y may be X.
#include "kdq_kernels.h"

namespace {

constexpr int CT = 256;   // threads per workgroup

// One length-L DFT of `base[m * stride]`, m < L, in place: forward (FWD) with e^{-2πi km/L},
// else e^{+2πi km/L}; twiddle w[e] = e^{-2πi e/L} (wr, wi).
template <int L, bool FWD>
KDQ_DEV void dft1(double2 *base, int stride, const double (&wr)[L], const double (&wi)[L]) {
    double xr[L], xi[L];
#pragma unroll
    for (int m = 0; m < L; ++m) {
        const double2 v = base[m * stride];
        xr[m] = v.x;
        xi[m] = v.y;
    }
#pragma unroll
    for (int k = 0; k < L; ++k) {
        double sr = xr[0], si = xi[0];
#pragma unroll
        for (int m = 1; m < L; ++m) {
            const int e = (k * m) % L;
            const double cr = wr[e], ci = FWD ? wi[e] : -wi[e];
            sr = fma(xr[m], cr, fma(-xi[m], ci, sr));
            si = fma(xr[m], ci, fma(xi[m], cr, si));
        }
        base[k * stride] = make_double2(sr, si);
    }
}

// all length-L DFTs of one pass over the LPW lines in LDS (line stride NPAD): along x
// (elements 2·(x + L·y) + b, stride 2) or along y (stride 2L)
template <int L, bool FWD, bool ALONG_Y>
KDQ_DEV void dft_pass(double2 *buf, int nlines, int npad, const double (&wr)[L], const double (&wi)[L]) {
    const int per_line = 2 * L;   // (y or x, b) pairs
    for (int q = threadIdx.x; q < nlines * per_line; q += CT) {
        const int l = q / per_line, r = q % per_line, b = r & 1, o = r >> 1;
        double2 *base = buf + (size_t)l * npad + (ALONG_Y ? 2 * o + b : 2 * L * o + b);
        dft1<L, FWD>(base, ALONG_Y ? 2 * L : 2, wr, wi);
    }
}

template <int L, int LPW>
KDQ_DEV void circ_body(const CircArgs &a, const ChainDev &ch, int c, int line0, unsigned char *lds) {
    constexpr int NC = L * L, NS = 2 * NC, NPAD = NS + 1;
    const int nl = dim_of(ch, a.nline_dim);
    const int nlw = min(LPW, nl - line0);
    double2 *buf = reinterpret_cast<double2 *>(lds);   // [LPW][NPAD]
    double2 *lam = buf + LPW * NPAD;                    // [NC][2][2]
    double2 *pin = lam + 4 * NC, *pout = pin + NS;      // position factors
    double2 *pline = pout + NS;                         // [LPW] line factors
    const int tid = threadIdx.x;
    double wr[L], wi[L];
#pragma unroll
    for (int e = 0; e < L; ++e) {
        double sn, cs;
        sincospi(2.0 * e / L, &sn, &cs);
        wr[e] = cs;
        wi[e] = -sn;
    }
    // the 2 x 2 blocks of E, Eᵀ or Eᴴ acting on a line; a right product X·E acts on the rows
    // of X with Eᵀ
    const int eop = a.rows ? 1 : a.op;
    const double *tab = ch.circ + (size_t)a.mat * NC * 8;
    for (int e = tid; e < 4 * NC; e += CT) {
        const int k = e >> 2, ab = e & 3, ia = ab >> 1, ib = ab & 1;
        int ks = k, sa = ia, sb = ib;
        if (eop != 0) {   // Λᵀ: (b, a) at -k; Λᴴ: conj (b, a) at k
            sa = ib;
            sb = ia;
            if (eop == 1) {
                const int kx = k % L, ky = k / L;
                ks = ((L - kx) % L) + L * ((L - ky) % L);
            }
        }
        const double *t = tab + (size_t)ks * 8 + 2 * (2 * sa + sb);
        lam[e] = make_double2(t[0], eop == 2 ? -t[1] : t[1]);
    }
    for (int p = tid; p < NS; p += CT) {
        double r, i;
        diag_value(ch, a.in, p, r, i);
        pin[p] = make_double2(r, i);
        diag_value(ch, a.out, p, r, i);
        pout[p] = make_double2(r, i);
    }
    if (tid < nlw) {
        double r, i;
        diag_value(ch, a.line, line0 + tid, r, i);
        pline[tid] = make_double2(r, i);
    }
    __syncthreads();
    const Opnd X = a.X[c], Y = a.Y[c];
    // load (and scale by `in`)
    if (!a.rows) {   // lines are columns: positions fastest (contiguous)
        for (int e = tid; e < LPW * NS; e += CT) {
            const int l = e / NS, p = e % NS;
            double2 v = make_double2(0.0, 0.0);
            if (l < nlw) {
                const size_t o = (size_t)p + (size_t)(line0 + l) * X.ld;
                const double xr = X.re[o], xi = X.im[o], fr = pin[p].x, fi = pin[p].y;
                v = make_double2(xr * fr - xi * fi, xr * fi + xi * fr);
            }
            buf[l * NPAD + p] = v;
        }
    } else {         // lines are rows: the LPW rows of one column together
        for (int e = tid; e < LPW * NS; e += CT) {
            const int l = e % LPW, p = e / LPW;
            double2 v = make_double2(0.0, 0.0);
            if (l < nlw) {
                const size_t o = (size_t)(line0 + l) + (size_t)p * X.ld;
                const double xr = X.re[o], xi = X.im[o], fr = pin[p].x, fi = pin[p].y;
                v = make_double2(xr * fr - xi * fi, xr * fi + xi * fr);
            }
            buf[l * NPAD + p] = v;
        }
    }
    __syncthreads();
    dft_pass<L, true, false>(buf, nlw, NPAD, wr, wi);
    __syncthreads();
    dft_pass<L, true, true>(buf, nlw, NPAD, wr, wi);
    __syncthreads();
    for (int e = tid; e < nlw * NC; e += CT) {
        const int l = e / NC, k = e % NC;
        double2 *u = buf + (size_t)l * NPAD + 2 * k;
        const double2 u0 = u[0], u1 = u[1];
        const double2 l00 = lam[4 * k], l01 = lam[4 * k + 1], l10 = lam[4 * k + 2], l11 = lam[4 * k + 3];
        u[0] = make_double2(l00.x * u0.x - l00.y * u0.y + l01.x * u1.x - l01.y * u1.y,
                            l00.x * u0.y + l00.y * u0.x + l01.x * u1.y + l01.y * u1.x);
        u[1] = make_double2(l10.x * u0.x - l10.y * u0.y + l11.x * u1.x - l11.y * u1.y,
                            l10.x * u0.y + l10.y * u0.x + l11.x * u1.y + l11.y * u1.x);
    }
    __syncthreads();
    dft_pass<L, false, true>(buf, nlw, NPAD, wr, wi);
    __syncthreads();
    dft_pass<L, false, false>(buf, nlw, NPAD, wr, wi);
    __syncthreads();
    // store (scaled by `out` and the line factor)
    if (!a.rows) {
        for (int e = tid; e < LPW * NS; e += CT) {
            const int l = e / NS, p = e % NS;
            if (l >= nlw) continue;
            const double2 v = buf[l * NPAD + p], f = pout[p], g = pline[l];
            const double tr = v.x * f.x - v.y * f.y, ti = v.x * f.y + v.y * f.x;
            const size_t o = (size_t)p + (size_t)(line0 + l) * Y.ld;
            Y.re[o] = tr * g.x - ti * g.y;
            Y.im[o] = tr * g.y + ti * g.x;
        }
    } else {
        for (int e = tid; e < LPW * NS; e += CT) {
            const int l = e % LPW, p = e / LPW;
            if (l >= nlw) continue;
            const double2 v = buf[l * NPAD + p], f = pout[p], g = pline[l];
            const double tr = v.x * f.x - v.y * f.y, ti = v.x * f.y + v.y * f.x;
            const size_t o = (size_t)(line0 + l) + (size_t)p * Y.ld;
            Y.re[o] = tr * g.x - ti * g.y;
            Y.im[o] = tr * g.y + ti * g.x;
        }
    }
}

template <int LPW>
__global__ __launch_bounds__(CT) void k_circ(CircArgs a) {
    const int c = blockIdx.y;
    const ChainDev &ch = a.chains[c];
    const int line0 = blockIdx.x * LPW;
    if (line0 >= dim_of(ch, a.nline_dim)) return;   // uniform per workgroup
    extern __shared__ __align__(16) unsigned char lds_raw[];
    switch (ch.cL) {
        case 2: circ_body<2, LPW>(a, ch, c, line0, lds_raw); break;
        case 3: circ_body<3, LPW>(a, ch, c, line0, lds_raw); break;
        case 4: circ_body<4, LPW>(a, ch, c, line0, lds_raw); break;
        case 6: circ_body<6, LPW>(a, ch, c, line0, lds_raw); break;
        case 8: circ_body<8, LPW>(a, ch, c, line0, lds_raw); break;
        case 9: circ_body<9, LPW>(a, ch, c, line0, lds_raw); break;
        case 12: circ_body<12, LPW>(a, ch, c, line0, lds_raw); break;
        case 15: circ_body<15, LPW>(a, ch, c, line0, lds_raw); break;
        default: break;   // not reached: the host takes the dense path unless every cL is listed
    }
}

}  // namespace

bool circ_supported(int L) { return L == 2 || L == 3 || L == 4 || L == 6 || L == 8 || L == 9 || L == 12 || L == 15; }

size_t circ_lds_bytes(int L, int lpw) {
    const size_t nc = (size_t)L * L, ns = 2 * nc;
    return sizeof(double2) * ((size_t)lpw * (ns + 1) + 4 * nc + 2 * ns + lpw);
}

hipError_t launch_circ(hipStream_t st, int nchains, int max_lines, int max_L, const CircArgs &a) {
    if (!circ_supported(max_L) || a.op < 0 || a.op > 2 || (a.rows && a.op != 0)) return hipErrorInvalidValue;
    constexpr int LPW = 8;
    const size_t bytes = circ_lds_bytes(max_L, LPW);
    if (hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_circ<LPW>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes))
        return e;
    hipLaunchKernelGGL(k_circ<LPW>, dim3((max_lines + LPW - 1) / LPW, nchains), dim3(CT), bytes, st, a);
    return hipGetLastError();
}
