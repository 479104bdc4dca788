// Kinetic products by cell momenta (block-circulant eK, eK⁻¹).
//
// The hopping matrix of every lattice the models build (nnK_Matrix, src/public/Geometry.jl:
// 178-207: HoneyComb120/60, SQUARE90/45 with periodic L x L cells and uniform bond phases) is
// invariant under cell translations, so eK = exp(-Δτ K) and eK⁻¹ are block-circulant: with
// site s = 2·(x + L·y) + b, E[(c,a), (c',b)] = E_ab(c' - c). A product E·X then factors as
//     X̂_b(k) = Σ_c e^{-i k·c} X[(c,b)],   Ŷ_a(k) = Σ_b Λ_ab(k) X̂_b(k),   Y[(c,a)] = Σ_k e^{+i k·c} Ŷ_a(k)
// with Λ(k) = (1/L²) Σ_d E(d) e^{+i k·d} (ChainDev::circ). The cell transforms are length-L FFTs
// (radix 2, 3, 4, 5 butterflies and one Cooley-Tukey step for L = 6, 8, 9, 12, 15) along y, then
// x. A workgroup takes LPW = 96 / L lines of X (columns for a left product E·X, rows for a right
// product X·E = (Eᵀ Xᵀ)ᵀ) with its 192 threads, one per (line, x, b): each thread loads its L
// elements of a line straight from X, transforms them along y in registers and puts them in
// LDS; one pass over LDS takes the x transforms, the 2 x 2 blocks and the inverse x transforms
// (one thread per (line, ky, b)); each thread then takes its elements back, transforms them
// backward along y and stores them. One read and one write of X and O(L² log L) flops per line,
// against (2L²)² for the dense product.
// op selects E (0), Eᵀ (1: Λ_ba(-k)) or Eᴴ (2: conj Λ_ba(k)). The diagonal factors of the GEMM
// path (the t-U HS phases, DiagSpec kind 0 or 1) are applied in the same places: `in` scales the
// line entries before the product, `out` the products, `line` the whole line (per column of a
// left product, per row of a right product). Every line is read completely before it is
// written, so Y may be X.
#include "kdq_kernels.h"

// Multiply-adds fuse only within one source expression (not across statements, as the
// backend's default would), so the single-product kernel and the chained kernel, which inline
// the same transforms into different surroundings, round every step alike.
#pragma clang fp contract(on)

namespace {

constexpr int CT = 192;   // threads per workgroup: LPW(L) = 96 / L lines x 2L DFTs, one thread each

// Length-L DFTs in registers, in place and in natural order: forward (FWD) X_k = Σ_n x_n
// e^{-2πi kn/L}, backward with e^{+2πi kn/L}. Radices 2, 3, 4 and 5 by their butterflies;
// composite L = P·Q by one Cooley-Tukey step (Q transforms of length P over n = Q n1 + n2, the
// twiddles e^{∓2πi n2 k1/L}, P transforms of length Q; X_{k1 + P k2}) with the twiddle table
// w[e] = e^{-2πi e/L} (wr, wi) of the full length.
template <bool FWD>
KDQ_DEV void mul_i(double &re, double &im) {   // z <- ∓i z
    const double t = re;
    if (FWD) {
        re = im;
        im = -t;
    } else {
        re = -im;
        im = t;
    }
}

template <int L, bool FWD>
KDQ_DEV void fft(double *xr, double *xi, const double *wr, const double *wi);

template <int P, int Q, bool FWD>
KDQ_DEV void fft_pq(double *xr, double *xi, const double *wr, const double *wi) {
    constexpr int L = P * Q;
    double yr[L], yi[L];   // y[n2 P + k1]
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
        double ar[P], ai[P], sr[P], si[P];
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) {
            ar[n1] = xr[Q * n1 + n2];
            ai[n1] = xi[Q * n1 + n2];
        }
        // the length-P twiddles are every Q-th entry of the length-L table
#pragma unroll
        for (int e = 0; e < P; ++e) {
            sr[e] = wr[Q * e];
            si[e] = wi[Q * e];
        }
        fft<P, FWD>(ar, ai, sr, si);
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) {
            const int e = (n2 * k1) % L;
            if (e == 0) {
                yr[n2 * P + k1] = ar[k1];
                yi[n2 * P + k1] = ai[k1];
            } else {
                const double cr = wr[e], ci = FWD ? wi[e] : -wi[e];
                yr[n2 * P + k1] = ar[k1] * cr - ai[k1] * ci;
                yi[n2 * P + k1] = ar[k1] * ci + ai[k1] * cr;
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
        double br[Q], bi[Q], sr[Q], si[Q];
#pragma unroll
        for (int n2 = 0; n2 < Q; ++n2) {
            br[n2] = yr[n2 * P + k1];
            bi[n2] = yi[n2 * P + k1];
        }
#pragma unroll
        for (int e = 0; e < Q; ++e) {
            sr[e] = wr[P * e];
            si[e] = wi[P * e];
        }
        fft<Q, FWD>(br, bi, sr, si);
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) {
            xr[k1 + P * k2] = br[k2];
            xi[k1 + P * k2] = bi[k2];
        }
    }
}

template <int L, bool FWD>
KDQ_DEV void fft(double *xr, double *xi, const double *wr, const double *wi) {
    if constexpr (L == 1) {
    } else if constexpr (L == 2) {
        const double ar = xr[0], ai = xi[0];
        xr[0] = ar + xr[1];
        xi[0] = ai + xi[1];
        xr[1] = ar - xr[1];
        xi[1] = ai - xi[1];
    } else if constexpr (L == 3) {
        // x0 + x1 + x2 and (x0 - (x1 + x2)/2) ± (∓i) sin(2π/3) (x1 - x2), the last as fmas
        constexpr double h = FWD ? 0.86602540378443864676 : -0.86602540378443864676;   // ±sin(2π/3)
        const double s1r = xr[1] + xr[2], s1i = xi[1] + xi[2];
        const double dr = xr[1] - xr[2], di = xi[1] - xi[2];
        const double tr = xr[0] - 0.5 * s1r, ti = xi[0] - 0.5 * s1i;
        xr[0] += s1r;
        xi[0] += s1i;
        xr[1] = fma(h, di, tr);
        xi[1] = fma(-h, dr, ti);
        xr[2] = fma(-h, di, tr);
        xi[2] = fma(h, dr, ti);
    } else if constexpr (L == 4) {
        const double t0r = xr[0] + xr[2], t0i = xi[0] + xi[2], t1r = xr[0] - xr[2], t1i = xi[0] - xi[2];
        const double t2r = xr[1] + xr[3], t2i = xi[1] + xi[3];
        double t3r = xr[1] - xr[3], t3i = xi[1] - xi[3];
        mul_i<FWD>(t3r, t3i);
        xr[0] = t0r + t2r;
        xi[0] = t0i + t2i;
        xr[2] = t0r - t2r;
        xi[2] = t0i - t2i;
        xr[1] = t1r + t3r;
        xi[1] = t1i + t3i;
        xr[3] = t1r - t3r;
        xi[3] = t1i - t3i;
    } else if constexpr (L == 5) {
        constexpr double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;   // cos(2π/5), cos(4π/5)
        constexpr double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;    // sin(2π/5), sin(4π/5)
        const double b1r = xr[1] + xr[4], b1i = xi[1] + xi[4], b2r = xr[2] + xr[3], b2i = xi[2] + xi[3];
        const double d1r = xr[1] - xr[4], d1i = xi[1] - xi[4], d2r = xr[2] - xr[3], d2i = xi[2] - xi[3];
        const double t1r = xr[0] + c1 * b1r + c2 * b2r, t1i = xi[0] + c1 * b1i + c2 * b2i;
        const double t2r = xr[0] + c2 * b1r + c1 * b2r, t2i = xi[0] + c2 * b1i + c1 * b2i;
        double u1r = s1 * d1r + s2 * d2r, u1i = s1 * d1i + s2 * d2i;
        double u2r = s2 * d1r - s1 * d2r, u2i = s2 * d1i - s1 * d2i;
        mul_i<FWD>(u1r, u1i);
        mul_i<FWD>(u2r, u2i);
        xr[0] += b1r + b2r;
        xi[0] += b1i + b2i;
        xr[1] = t1r + u1r;
        xi[1] = t1i + u1i;
        xr[4] = t1r - u1r;
        xi[4] = t1i - u1i;
        xr[2] = t2r + u2r;
        xi[2] = t2i + u2i;
        xr[3] = t2r - u2r;
        xi[3] = t2i - u2i;
    } else if constexpr (L == 6) {
        fft_pq<2, 3, FWD>(xr, xi, wr, wi);
    } else if constexpr (L == 8) {
        fft_pq<2, 4, FWD>(xr, xi, wr, wi);
    } else if constexpr (L == 9) {
        fft_pq<3, 3, FWD>(xr, xi, wr, wi);
    } else if constexpr (L == 12) {
        fft_pq<3, 4, FWD>(xr, xi, wr, wi);
    } else if constexpr (L == 15) {
        fft_pq<3, 5, FWD>(xr, xi, wr, wi);
    } else {
        static_assert(L == 0, "no transform of this length");
    }
}

// The middle pass of a product, in place on the LPW lines in LDS (line stride npad; element
// 2·(x + L·y) + b of a line at (2L + 1)·y + 2x + b, rows of 2L padded to 2L + 1), after the y
// transforms: thread (line, ky = o, b) transforms its row forward along x, takes the other
// component (b ^ 1) of every momentum from the neighbouring lane, forms y_b(k) = M_bb(k) x_b(k)
// + M_b,1-b(k) x_1-b(k) and transforms back. M = Λ (eop 0), Λ(-k)ᵀ (eop 1) or Λᴴ (eop 2), from
// the chain's table staged in LDS (Λ_ij(k) at tab + 8k + 2(2i + j), k = kx + L·ky). Every
// thread of the workgroup calls it (the lane exchange).
template <int L>
KDQ_DEV void x_lambda_pass(double2 *buf, int nlines, int npad, const double (&wr)[L], const double (&wi)[L],
                           const double *tab, int eop) {
    const int q = threadIdx.x, l = q / (2 * L), r = q % (2 * L), b = r & 1, o = r >> 1;
    const bool act = l < nlines;
    double2 *base = buf + (size_t)l * npad + (2 * L + 1) * o + b;
    double xr[L], xi[L];
#pragma unroll
    for (int m = 0; m < L; ++m) {
        double2 v = make_double2(0.0, 0.0);
        if (act) v = base[2 * m];
        xr[m] = v.x;
        xi[m] = v.y;
    }
    fft<L, true>(xr, xi, wr, wi);
    const double cj = eop == 2 ? -1.0 : 1.0;
    const int own = 2 * (2 * b + b), oth = eop == 0 ? 2 * (2 * b + (1 - b)) : 2 * (2 * (1 - b) + b);
#pragma unroll
    for (int m = 0; m < L; ++m) {
        const double pr = __shfl_xor(xr[m], 1), pi = __shfl_xor(xi[m], 1);
        int ks = m + L * o;
        if (eop == 1) ks = ((L - m) % L) + L * ((L - o) % L);
        const double *t = tab + (size_t)ks * 8;
        const double ar = t[own], ai = cj * t[own + 1], br = t[oth], bi = cj * t[oth + 1];
        const double yr = ar * xr[m] - ai * xi[m] + br * pr - bi * pi;
        const double yi = ar * xi[m] + ai * xr[m] + br * pi + bi * pr;
        xr[m] = yr;
        xi[m] = yi;
    }
    fft<L, false>(xr, xi, wr, wi);
    if (act) {
#pragma unroll
        for (int m = 0; m < L; ++m) base[2 * m] = make_double2(xr[m], xi[m]);
    }
}

KDQ_HD int lpw_of(int L) { return L >= 96 ? 1 : 96 / L; }

// complex product with a fixed fma order (the load, the store and the steps of a chained
// product must round alike)
KDQ_DEV double2 cmul_f(double2 v, double2 f) {
    return make_double2(fma(v.x, f.x, -v.y * f.y), fma(v.x, f.y, v.y * f.x));
}

// the four values cis(±α η_v) of the `in`, `out` and `line` factors of step `step` (threads 32..43)
KDQ_DEV void set_fac(const CircArgs &a, const ChainDev &ch, double2 (*f)[4], int step, int tid) {
    if (tid >= 32 && tid < 44) {
        const int w = (tid - 32) >> 2, v = (tid - 32) & 3;
        const DiagSpec &d = w == 0 ? a.in : (w == 1 ? a.out : a.line);
        double r = 1.0, i = 0.0;
        if (d.kind == 1) {
            const int lt = d.lt + step * a.dlt;
            sincos(d.sign * ch.alpha[lt - 1] * ch.eta[v], &i, &r);
            if (d.conj) i = -i;
        }
        f[w][v] = make_double2(r, i);
    }
}

// the fields of a factor's slice at step `step`
KDQ_DEV const uint8_t *step_fields(const CircArgs &a, const ChainDev &ch, const DiagSpec &d, int step) {
    const int lt = d.lt > 0 ? d.lt + step * a.dlt : 1;
    return ch.s + (size_t)(lt - 1) * ch.n_elem;
}

// a plane element at a byte offset from the plane's (workgroup-uniform) base
KDQ_DEV double ld_at(const double *base, uint32_t off) {
    return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + off);
}
KDQ_DEV void st_at(double *base, uint32_t off, double v) {
    *reinterpret_cast<double *>(reinterpret_cast<char *>(base) + off) = v;
}

// Thread -> (line, x, b) of the y transforms: for columns the 2L threads of a line together (a
// line's 2L elements at one y are contiguous in memory), for rows b and the lines fastest (the
// lines of one column are contiguous). Threads past 2L·LPW have x >= L or l >= LPW.
template <int L>
KDQ_DEV void y_thread(bool rows, int &l, int &x, int &b) {
    constexpr int LPW = 96 / L;
    const int q = threadIdx.x;
    if (rows) {
        b = q & 1;
        l = (q >> 1) % LPW;
        x = (q >> 1) / LPW;
    } else {
        l = q / (2 * L);
        x = (q % (2 * L)) >> 1;
        b = q & 1;
    }
}

// One product (or a chain of nsteps products) on a group of LPW lines. Each thread loads its
// L elements of a line (fixed x and b, every y) from X, applies `in` and transforms them along
// y in registers, and writes them to LDS; the x transforms with the 2 x 2 blocks run as one
// pass over LDS; each thread reads its elements back, transforms them backward along y and
// stores them (scaled by `out` and `line`) — or, between the steps of a chain, applies this
// step's `out` and `line` and the next step's `in` in registers and starts the next step.
template <int L, bool MULTI>
KDQ_DEV void circ_body(const CircArgs &a, const ChainDev &ch, int c, int grp, unsigned char *lds) {
    constexpr int NC = L * L, NPAD = (2 * L + 1) * L + 1, LPW = 96 / L, YS = 2 * L + 1;
    const int nl = dim_of(ch, a.nline_dim);
    const int line0 = grp * LPW;
    if (line0 >= nl) return;   // uniform per workgroup
    const int nlw = min(LPW, nl - line0);
    const bool use_in = a.in.kind != 0, use_out = a.out.kind != 0, use_line = a.line.kind != 0;
    const int nsteps = MULTI && a.nsteps > 1 ? a.nsteps : 1;
    double2 *buf = reinterpret_cast<double2 *>(lds);   // [LPW][NPAD]
    double *lamt = reinterpret_cast<double *>(buf + LPW * NPAD);   // the chain's Λ table of E [k][a][b]
    const int tid = threadIdx.x;
    int l, x, b;
    y_thread<L>(a.rows, l, x, b);
    const bool act = l < nlw && x < L;
    const Opnd X = a.X[c], Y = a.Y[c];
    // byte offsets of a thread's elements from the planes' base (32 bits: a matrix plane holds at
    // most a few MB), element m at off + m·step
    auto off0 = [&](const Opnd &M) {
        const uint32_t p = 2 * x + b, ln = line0 + l;
        return 8u * (a.rows ? ln + p * (uint32_t)M.ld : p + ln * (uint32_t)M.ld);
    };
    auto step = [&](const Opnd &M) { return 16u * L * (a.rows ? (uint32_t)M.ld : 1u); };
    // the t-U HS factors take four values per slice (one per field value): cis(±α η_v) of the
    // `in`, `out` and `line` factors of a step, evaluated once per workgroup and step (kind 0 = 1)
    __shared__ double twr[16], twi[16];   // w[e] = e^{-2πi e/L}, read from LDS by the butterflies
    __shared__ double2 fac[2][3][4];
    const double(&wr)[L] = *reinterpret_cast<const double(*)[L]>(twr);
    const double(&wi)[L] = *reinterpret_cast<const double(*)[L]>(twi);
    double2 *ybase = buf + (size_t)l * NPAD + 2 * x + b;
    auto tables = [&]() {
        if (tid < L) {
            double sn, cs;
            sincospi(2.0 * tid / L, &sn, &cs);
            twr[tid] = cs;
            twi[tid] = -sn;
        }
        set_fac(a, ch, fac[0], 0, tid);
    };
    // the twiddles and the first step's factors: after the loads are issued where the registers
    // allow it (L <= 12), before them at L = 15
    if constexpr (L > 12) tables();
    {
        double xr[L], xi[L];
        if (act) {   // the loads in flight while the tables are built and staged
            const uint32_t o0 = off0(X), sx = step(X);
#pragma unroll
            for (int m = 0; m < L; ++m) {
                xr[m] = ld_at(X.re, o0 + m * sx);
                xi[m] = ld_at(X.im, o0 + m * sx);
            }
        }
        if constexpr (L <= 12) tables();
        {
            const int m0 = MULTI && a.plain_last && nsteps == 1 ? a.mat_last : a.mat;
            const double *tab = ch.circ + (size_t)m0 * NC * 8;
            for (int e = tid; e < NC * 8; e += CT) lamt[e] = tab[e];
        }
        __syncthreads();
        if (act) {   // `in` (unless the only step is the plain one), forward along y, to LDS
            if (use_in && !(MULTI && a.plain_last && nsteps == 1)) {
                const uint8_t *sin_ = step_fields(a, ch, a.in, 0);
#pragma unroll
                for (int m = 0; m < L; ++m) {
                    const double2 w = cmul_f(make_double2(xr[m], xi[m]), fac[0][0][sin_[2 * (x + L * m) + b] - 1]);
                    xr[m] = w.x;
                    xi[m] = w.y;
                }
            }
            fft<L, true>(xr, xi, wr, wi);
#pragma unroll
            for (int m = 0; m < L; ++m) ybase[m * YS] = make_double2(xr[m], xi[m]);
        }
    }
    for (int st = 0; st < nsteps; ++st) {
        const int cur = st & 1;
        if (MULTI && a.plain_last && st == nsteps - 1 && st > 0 && a.mat_last != a.mat) {
            // the closing matrix's table in place of the step matrix's (every thread has left the
            // previous step's x pass: a barrier followed it)
            const double *tab = ch.circ + (size_t)a.mat_last * NC * 8;
            for (int e = tid; e < NC * 8; e += CT) lamt[e] = tab[e];
        }
        __syncthreads();
        // forward x, the 2 x 2 blocks of E, Eᵀ or Eᴴ acting on a line (a right product X·E acts on
        // the rows of X with Eᵀ), backward x
        x_lambda_pass<L>(buf, nlw, NPAD, wr, wi, lamt, a.rows ? 1 : a.op);
        // the next step's factors (every thread has used fac[cur ^ 1] before the barrier above)
        const bool more = MULTI && st + 1 < nsteps;
        if (more) set_fac(a, ch, fac[cur ^ 1], st + 1, tid);
        __syncthreads();
        if (!act) continue;
        double xr[L], xi[L];
#pragma unroll
        for (int m = 0; m < L; ++m) {
            const double2 v = ybase[m * YS];
            xr[m] = v.x;
            xi[m] = v.y;
        }
        fft<L, false>(xr, xi, wr, wi);
        if (more) {   // this step's `out` and `line`, the next step's `in`, forward along y, to LDS
            const bool next_plain = a.plain_last && st + 1 == nsteps - 1;
            const uint8_t *so = step_fields(a, ch, a.out, st), *sl = step_fields(a, ch, a.line, st),
                          *si = step_fields(a, ch, a.in, st + 1);
            const double2 fl = use_line ? fac[cur][2][sl[line0 + l] - 1] : make_double2(1.0, 0.0);
#pragma unroll
            for (int m = 0; m < L; ++m) {
                const int p = 2 * (x + L * m) + b;
                double2 v = make_double2(xr[m], xi[m]);
                if (use_out) v = cmul_f(v, fac[cur][1][so[p] - 1]);
                v = cmul_f(v, fl);
                if (use_in && !next_plain) v = cmul_f(v, fac[cur ^ 1][0][si[p] - 1]);
                xr[m] = v.x;
                xi[m] = v.y;
            }
            fft<L, true>(xr, xi, wr, wi);
#pragma unroll
            for (int m = 0; m < L; ++m) ybase[m * YS] = make_double2(xr[m], xi[m]);
        } else {   // store (scaled by `out` and the line factor of the last step, unless it is plain)
            const bool plain = MULTI && a.plain_last;
            const uint8_t *sout = step_fields(a, ch, a.out, st), *sline = step_fields(a, ch, a.line, st);
            const double2 fl = use_line && !plain ? fac[cur][2][sline[line0 + l] - 1] : make_double2(1.0, 0.0);
            const uint32_t o0 = off0(Y), sy = step(Y);
#pragma unroll
            for (int m = 0; m < L; ++m) {
                const int p = 2 * (x + L * m) + b;
                double2 v = make_double2(xr[m], xi[m]);
                if (use_out && !plain) v = cmul_f(v, fac[cur][1][sout[p] - 1]);
                v = cmul_f(v, fl);
                st_at(Y.re, o0 + m * sy, v.x);
                st_at(Y.im, o0 + m * sy, v.y);
            }
        }
    }
}

// Workgroup -> (chain, line group). Workgroups go to the 8 XCDs round-robin by linear id; with
// xcd_map set (batches of at least 8 chains) every group of chain c runs on XCD c % 8, and
// neighbouring groups one after another there: the row pass reads half of each 128-byte line
// per group, and the neighbouring group finds the other half in the same L2.
constexpr int NXCD = 8;
// BIG: the instantiation for cell edge 15 too (batches whose largest edge is 15; it needs more
// registers than the smaller edges, so those launch without it)
template <bool MULTI, bool BIG>
KDQ_DEV void circ_dispatch(const CircArgs &a, int nch, int gx, int xcd_map) {
    int c, grp;
    if (xcd_map) {
        const int L = blockIdx.x, x = L % NXCD, k = L / NXCD;
        c = x + NXCD * (k / gx);
        grp = k % gx;
        if (c >= nch) return;
    } else {
        c = blockIdx.y;
        grp = blockIdx.x;
    }
    const ChainDev &ch = a.chains[c];
    extern __shared__ __align__(16) unsigned char lds_raw[];
    switch (ch.cL) {
        case 2: circ_body<2, MULTI>(a, ch, c, grp, lds_raw); break;
        case 3: circ_body<3, MULTI>(a, ch, c, grp, lds_raw); break;
        case 4: circ_body<4, MULTI>(a, ch, c, grp, lds_raw); break;
        case 6: circ_body<6, MULTI>(a, ch, c, grp, lds_raw); break;
        case 8: circ_body<8, MULTI>(a, ch, c, grp, lds_raw); break;
        case 9: circ_body<9, MULTI>(a, ch, c, grp, lds_raw); break;
        case 12: circ_body<12, MULTI>(a, ch, c, grp, lds_raw); break;
        case 15:
            if constexpr (BIG) circ_body<15, MULTI>(a, ch, c, grp, lds_raw);
            break;
        default: break;   // not reached: the host takes the dense path unless every cL is listed
    }
}

template <bool BIG>
__global__ __launch_bounds__(CT, 3) void k_circ(CircArgs a, int nch, int gx, int xcd_map) {
    circ_dispatch<false, BIG>(a, nch, gx, xcd_map);
}

// a product of several factors (a.nsteps > 1) in one launch, the lines kept in LDS between them
template <bool BIG>
__global__ __launch_bounds__(CT, 2) void k_circ_chain(CircArgs a, int nch, int gx, int xcd_map) {
    circ_dispatch<true, BIG>(a, nch, gx, xcd_map);
}

}  // namespace

bool circ_supported(int L) { return L == 2 || L == 3 || L == 4 || L == 6 || L == 8 || L == 9 || L == 12 || L == 15; }

int circ_lines_per_wg(int L) { return lpw_of(L); }

size_t circ_lds_bytes(int L) {
    const size_t lpw = (size_t)lpw_of(L), npad = (2 * (size_t)L + 1) * L + 1;
    return sizeof(double2) * lpw * npad + sizeof(double) * 8 * (size_t)L * L;
}

hipError_t launch_circ(hipStream_t st, int nchains, int grid_x, size_t lds_bytes, int lmax, const CircArgs &a) {
    if (a.op < 0 || a.op > 2 || (a.rows && a.op != 0) || grid_x < 1 || a.mat < 0 || a.mat > 3 || a.mat_last < 0 ||
        a.mat_last > 3 || lmax < 1 || lmax > 15)
        return hipErrorInvalidValue;
    const bool big = lmax > 12, multi = a.nsteps > 1 || a.plain_last;
    auto kern = multi ? (big ? k_circ_chain<true> : k_circ_chain<false>) : (big ? k_circ<true> : k_circ<false>);
    if (hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes))
        return e;
    if (nchains >= NXCD) {
        const int groups = (nchains + NXCD - 1) / NXCD;
        hipLaunchKernelGGL(kern, dim3(NXCD * groups * grid_x), dim3(CT), lds_bytes, st, a, nchains, grid_x, 1);
    } else {
        hipLaunchKernelGGL(kern, dim3(grid_x, nchains), dim3(CT), lds_bytes, st, a, nchains, grid_x, 0);
    }
    return hipGetLastError();
}
