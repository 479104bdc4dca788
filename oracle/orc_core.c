/*
 * ORACLE — test infrastructure only (see orc.h). Random numbers, lattice geometry
 * and dense FP64 linear algebra used by the model restatements.
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================= Philox4x32-10 (Salmon et al., SC'11) ======================== */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Draw layout (identical in the HIP library):
 *   key = {seed_lo, seed_hi}
 *   ctr = {elem, slice_tag, sweep, (stream << 24) | chain}
 *   choice = floor(3 * x0 / 2^32), u = ((x1 >> 5) * 2^26 + (x2 >> 6)) / 2^53
 * slice_tag = lt | (dir << 31), dir 0 forward, 1 backward. */
void orc_draw(uint64_t seed, uint32_t chain, uint32_t stream, uint32_t sweep, uint32_t slice_tag,
              uint32_t elem, int *choice, double *u) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {elem, slice_tag, sweep, (stream << 24) | (chain & 0xFFFFFFu)};
    uint32_t x[4];
    orc_philox4x32_10(ctr, key, x);
    *choice = (int)(((uint64_t)x[0] * 3u) >> 32);
    *u = ((double)(x[1] >> 5) * 67108864.0 + (double)(x[2] >> 6)) * (1.0 / 9007199254740992.0);
}

/* One proposal of the plain sweep: the replayed draw when a buffer is set, else Philox
 * stream 0. A sweep outside the buffer reads as a rejected proposal (u = +inf). */
void orc_sweep_draw(const orc_replay *r, int nt, int n_elem, uint64_t seed, uint32_t chain, uint32_t sweep,
                    int lt, int dir, uint32_t elem, int *choice, double *u) {
    if (r && r->u) {
        uint32_t s = sweep - r->sweep0;
        if (s < r->nsweeps) {
            size_t q = (((size_t)s * 2 + (size_t)dir) * (size_t)nt + (size_t)(lt - 1)) * (size_t)n_elem + elem;
            *choice = r->choice[q] % 3;
            *u = r->u[q];
        } else {
            *choice = 0;
            *u = INFINITY;
        }
        return;
    }
    orc_draw(seed, chain, 0, sweep, (uint32_t)lt | ((uint32_t)dir << 31), elem, choice, u);
}

/* Initial_s replacement (src/tU/GreenMatrix.jl:1-8 draws i.i.d. uniform over {1,2,3,4}):
 * stream tag 15 is reserved for initial fields. */
int orc_init_value(uint64_t seed, uint32_t chain, uint32_t stream, uint32_t slice, uint32_t elem) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {elem, slice, 0xFFFFFFFFu, (stream << 24) | (chain & 0xFFFFFFu)};
    uint32_t x[4];
    orc_philox4x32_10(ctr, key, x);
    return 1 + (int)(((uint64_t)x[0] * 4u) >> 32);
}

/* =============================== geometry ========================================== */
static int mod1(int a, int n) { int r = a % n; if (r <= 0) r += n; return r; }

/* Geometry.jl:8-13 — 1-based even index of cell (x,y) */
int orc_xy_i(const int site[2], int x, int y) { return 2 * (x + (y - 1) * site[0]); }

/* Geometry.jl:18-21 */
void orc_i_xy(const int site[2], int i, int *x, int *y) {
    int j = (i + 1) / 2;               /* ceil(i/2) */
    *x = mod1(j, site[0]);
    *y = (j + site[0] - 1) / site[0];  /* ceil(j/site[1]) */
}

/* Geometry.jl:74-134 — 1-based neighbour list; returns coordination number */
int orc_nn2idx(int lattice, const int site[2], int idx, int nn[4]) {
    int x, y;
    orc_i_xy(site, idx, &x, &y);
    int L1 = site[0], L2 = site[1];
    if (lattice == ORC_SQ90) {
        if (idx % 2 == 1) {
            nn[0] = idx + 1;
            nn[1] = orc_xy_i(site, mod1(x - 1, L1), mod1(y - 1, L2));
            nn[2] = orc_xy_i(site, x, mod1(y - 1, L2));
            nn[3] = orc_xy_i(site, mod1(x - 1, L1), y);
        } else {
            nn[0] = orc_xy_i(site, mod1(x + 1, L1), y) - 1;
            nn[1] = orc_xy_i(site, x, mod1(y + 1, L2)) - 1;
            nn[2] = idx - 1;
            nn[3] = orc_xy_i(site, mod1(x + 1, L1), mod1(y + 1, L2)) - 1;
        }
        return 4;
    } else if (lattice == ORC_SQ45) {
        if (idx % 2 == 1) {
            nn[0] = orc_xy_i(site, mod1(x - 1, L1), y);
            nn[1] = idx + 1;
            nn[2] = orc_xy_i(site, mod1(x - 1, L1), mod1(y + 1, L2));
            nn[3] = orc_xy_i(site, x, mod1(y - 1, L2));
        } else {
            nn[0] = orc_xy_i(site, x, mod1(y + 1, L2)) - 1;
            nn[1] = orc_xy_i(site, mod1(x + 1, L1), mod1(y - 1, L2)) - 1;
            nn[2] = idx - 1;
            nn[3] = orc_xy_i(site, mod1(x + 1, L1), mod1(y, L2)) - 1;
        }
        return 4;
    } else if (lattice == ORC_HC120) {
        if (idx % 2 == 1) {
            nn[0] = idx + 1;
            nn[1] = orc_xy_i(site, mod1(x + 1, L1), y);
            nn[2] = orc_xy_i(site, x, mod1(y - 1, L2));
        } else {
            nn[0] = idx - 1;
            nn[1] = orc_xy_i(site, x, mod1(y + 1, L2)) - 1;
            nn[2] = orc_xy_i(site, mod1(x - 1, L1), y) - 1;
        }
        return 3;
    } else { /* ORC_HC60 */
        if (idx % 2 == 1) {
            nn[0] = idx + 1;
            nn[1] = orc_xy_i(site, mod1(x + 1, L1), mod1(y - 1, L2));
            nn[2] = orc_xy_i(site, x, mod1(y - 1, L2));
        } else {
            nn[0] = idx - 1;
            nn[1] = orc_xy_i(site, x, mod1(y + 1, L2)) - 1;
            nn[2] = orc_xy_i(site, mod1(x - 1, L1), mod1(y + 1, L2)) - 1;
        }
        return 3;
    }
}

/* Geometry.jl:178-207, flux = 0, t = (1,1,1). K is ns x ns column-major. */
void orc_nnK(int lattice, const int site[2], double *K) {
    int ns = 2 * site[0] * site[1];
    memset(K, 0, sizeof(double) * (size_t)ns * ns);
    for (int i = 1; i <= ns; ++i) {
        int nn[4];
        int z = orc_nn2idx(lattice, site, i, nn);
        for (int j = 0; j < z; ++j) K[(size_t)(i - 1) + (size_t)(nn[j] - 1) * ns] = 1.0;
    }
}

/* Geometry.jl:24-66 honeycomb branch: nnidx[count, j] = (i, nn[j]) for odd i.
 * Output: nnidx[(count-1) + (j-1)*(ns/2)] pairs stored as 2 ints, 1-based. */
void orc_nnidx_F(int lattice, const int site[2], int *nnidx) {
    int ns = 2 * site[0] * site[1], half = ns / 2, count = 0;
    for (int i = 1; i <= ns; i += 2) {
        int nn[4];
        int z = orc_nn2idx(lattice, site, i, nn);
        for (int j = 0; j < z; ++j) {
            nnidx[2 * (count + j * half) + 0] = i;
            nnidx[2 * (count + j * half) + 1] = nn[j];
        }
        ++count;
    }
}

/* Geometry.jl:209-307, rectangular areas on HoneyComb / SQUARE90 (1-based output). */
int orc_area_index(int lattice, const int site[2], const int a0[2], const int a1[2], int *index) {
    (void)lattice;
    int c = 0;
    for (int lx = a0[0]; lx <= a1[0]; ++lx)
        for (int ly = a0[1]; ly <= a1[1]; ++ly) {
            index[c] = orc_xy_i(site, lx, ly) - 1;
            index[c + 1] = index[c] + 1;
            c += 2;
        }
    return c;
}

/* ============================ dense linear algebra ================================== */
/* C = alpha*op(A)*op(B) + beta*C; op in {'N','T','C'} (Julia mul!, BLAS zgemm semantics) */
void orc_zgemm(char ta, char tb, int m, int n, int k, zdouble alpha, const zdouble *A, int lda,
               const zdouble *B, int ldb, zdouble beta, zdouble *C, int ldc) {
    zdouble *col = NULL;
#pragma omp parallel for schedule(static) if ((long long)m * n * k > 200000)
    for (int j = 0; j < n; ++j) {
        zdouble acc_local[1];
        (void)acc_local;
        zdouble *cj = C + (size_t)j * ldc;
        /* accumulate into a private column to keep BLAS beta semantics exact */
        zdouble *tmp = (zdouble *)malloc(sizeof(zdouble) * (size_t)m);
        for (int i = 0; i < m; ++i) tmp[i] = 0;
        for (int p = 0; p < k; ++p) {
            zdouble b;
            if (tb == 'N') b = B[p + (size_t)j * ldb];
            else if (tb == 'T') b = B[j + (size_t)p * ldb];
            else b = conj(B[j + (size_t)p * ldb]);
            if (b == 0) continue;
            if (ta == 'N') {
                const zdouble *ap = A + (size_t)p * lda;
                for (int i = 0; i < m; ++i) tmp[i] += ap[i] * b;
            } else if (ta == 'T') {
                for (int i = 0; i < m; ++i) tmp[i] += A[p + (size_t)i * lda] * b;
            } else {
                for (int i = 0; i < m; ++i) tmp[i] += conj(A[p + (size_t)i * lda]) * b;
            }
        }
        for (int i = 0; i < m; ++i) cj[i] = (beta == 0 ? 0 : beta * cj[i]) + alpha * tmp[i];
        free(tmp);
    }
    (void)col;
}

/* Cyclic Jacobi eigensolver for a real symmetric matrix (Golub & Van Loan, Alg. 8.5.3).
 * Replaces LAPACK.syevd! in the model constructors (src/tU/model.jl:61,73).
 * Eigenvalues ascending in w, eigenvectors in the columns of V. */
void orc_jacobi_eigh(int n, const double *Ain, double *w, double *V) {
    double *A = (double *)malloc(sizeof(double) * (size_t)n * n);
    memcpy(A, Ain, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n * n; ++i) V[i] = 0;
    for (int i = 0; i < n; ++i) V[i + (size_t)i * n] = 1;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double a = A[i + (size_t)j * n];
                tot += a * a;
                if (i != j) off += a * a;
            }
        if (off <= 1e-30 * tot || off == 0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p + (size_t)q * n];
                if (fabs(apq) < 1e-300) continue;
                double app = A[p + (size_t)p * n], aqq = A[q + (size_t)q * n];
                double theta = (aqq - app) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) { /* A <- A J */
                    double akp = A[k + (size_t)p * n], akq = A[k + (size_t)q * n];
                    A[k + (size_t)p * n] = c * akp - s * akq;
                    A[k + (size_t)q * n] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) { /* A <- J^T A */
                    double apk = A[p + (size_t)k * n], aqk = A[q + (size_t)k * n];
                    A[p + (size_t)k * n] = c * apk - s * aqk;
                    A[q + (size_t)k * n] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k + (size_t)p * n], vkq = V[k + (size_t)q * n];
                    V[k + (size_t)p * n] = c * vkp - s * vkq;
                    V[k + (size_t)q * n] = s * vkp + c * vkq;
                }
            }
    }
    /* sort ascending */
    for (int i = 0; i < n; ++i) w[i] = A[i + (size_t)i * n];
    for (int i = 0; i < n; ++i) {
        int mi = i;
        for (int j = i + 1; j < n; ++j) if (w[j] < w[mi]) mi = j;
        if (mi != i) {
            double t = w[i]; w[i] = w[mi]; w[mi] = t;
            for (int k = 0; k < n; ++k) {
                double v = V[k + (size_t)i * n];
                V[k + (size_t)i * n] = V[k + (size_t)mi * n];
                V[k + (size_t)mi * n] = v;
            }
        }
    }
    free(A);
}

/* Householder QR of the m x n matrix A (m >= n), replaced in place by the explicit
 * orthonormal factor Q (LAPACK geqrf + orgqr as called at src/tU/phy_update.jl:85-86).
 * Reflector H = I - tau v v^H with v_0 = 1 and H^H x = beta e_1 (zlarfg convention).
 * Optionally returns sum_j log|R_jj| (for the log-determinant of the weight). */
void orc_zqr_q(int m, int n, zdouble *A, int lda, double *logabsr) {
    zdouble *tau = (zdouble *)malloc(sizeof(zdouble) * (size_t)n);
    double lr = 0;
    for (int j = 0; j < n; ++j) {
        zdouble *a = A + (size_t)j * lda;
        double xnorm2 = 0;
        for (int i = j + 1; i < m; ++i) xnorm2 += creal(a[i] * conj(a[i]));
        zdouble alpha = a[j];
        if (xnorm2 == 0 && cimag(alpha) == 0) {
            tau[j] = 0;
            lr += log(fabs(creal(alpha)));
            continue;
        }
        double beta = -copysign(sqrt(creal(alpha * conj(alpha)) + xnorm2), creal(alpha));
        tau[j] = (beta - alpha) / beta;
        zdouble scal = 1.0 / (alpha - beta);
        for (int i = j + 1; i < m; ++i) a[i] *= scal;
        a[j] = beta;
        lr += log(fabs(beta));
        /* apply H^H = I - conj(tau) v v^H to trailing columns */
        for (int c = j + 1; c < n; ++c) {
            zdouble *ac = A + (size_t)c * lda;
            zdouble wsum = ac[j];
            for (int i = j + 1; i < m; ++i) wsum += conj(a[i]) * ac[i];
            wsum *= conj(tau[j]);
            ac[j] -= wsum;
            for (int i = j + 1; i < m; ++i) ac[i] -= a[i] * wsum;
        }
    }
    /* orgqr: Q = H_0 H_1 ... H_{n-1} [I; 0] (backward accumulation) */
    zdouble *Q = (zdouble *)calloc((size_t)m * n, sizeof(zdouble));
    for (int j = 0; j < n; ++j) Q[j + (size_t)j * m] = 1;
    for (int j = n - 1; j >= 0; --j) {
        const zdouble *v = A + (size_t)j * lda;
        for (int c = j; c < n; ++c) {
            zdouble *qc = Q + (size_t)c * m;
            zdouble wsum = qc[j];
            for (int i = j + 1; i < m; ++i) wsum += conj(v[i]) * qc[i];
            wsum *= tau[j];
            qc[j] -= wsum;
            for (int i = j + 1; i < m; ++i) qc[i] -= v[i] * wsum;
        }
    }
    for (int j = 0; j < n; ++j) memcpy(A + (size_t)j * lda, Q + (size_t)j * m, sizeof(zdouble) * m);
    free(Q);
    free(tau);
    if (logabsr) *logabsr = lr;
}

/* LU with partial pivoting (LAPACK getrf semantics, 0-based ipiv). Returns 0 or k+1 if
 * U(k,k) is exactly zero. */
int orc_zgetrf(int n, zdouble *A, int lda, int *ipiv) {
    int info = 0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = cabs(A[k + (size_t)k * lda]);
        for (int i = k + 1; i < n; ++i) {
            double v = cabs(A[i + (size_t)k * lda]);
            if (v > best) { best = v; p = i; }
        }
        ipiv[k] = p;
        if (p != k)
            for (int c = 0; c < n; ++c) {
                zdouble t = A[k + (size_t)c * lda];
                A[k + (size_t)c * lda] = A[p + (size_t)c * lda];
                A[p + (size_t)c * lda] = t;
            }
        zdouble piv = A[k + (size_t)k * lda];
        if (piv == 0) { if (!info) info = k + 1; continue; }
        for (int i = k + 1; i < n; ++i) A[i + (size_t)k * lda] /= piv;
        for (int c = k + 1; c < n; ++c) {
            zdouble akc = A[k + (size_t)c * lda];
            if (akc == 0) continue;
            for (int i = k + 1; i < n; ++i) A[i + (size_t)c * lda] -= A[i + (size_t)k * lda] * akc;
        }
    }
    return info;
}

/* getrs('N'): solve A X = B with the factors of orc_zgetrf */
void orc_zgetrs(int n, int nrhs, const zdouble *A, int lda, const int *ipiv, zdouble *B, int ldb) {
    for (int c = 0; c < nrhs; ++c) {
        zdouble *b = B + (size_t)c * ldb;
        for (int k = 0; k < n; ++k)
            if (ipiv[k] != k) { zdouble t = b[k]; b[k] = b[ipiv[k]]; b[ipiv[k]] = t; }
        for (int k = 0; k < n; ++k)
            for (int i = k + 1; i < n; ++i) b[i] -= A[i + (size_t)k * lda] * b[k];
        for (int k = n - 1; k >= 0; --k) {
            b[k] /= A[k + (size_t)k * lda];
            for (int i = 0; i < k; ++i) b[i] -= A[i + (size_t)k * lda] * b[k];
        }
    }
}

zdouble orc_zdet(int n, const zdouble *Ain, int lda) {
    zdouble *A = (zdouble *)malloc(sizeof(zdouble) * (size_t)n * n);
    int *ipiv = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; ++j) memcpy(A + (size_t)j * n, Ain + (size_t)j * lda, sizeof(zdouble) * n);
    orc_zgetrf(n, A, n, ipiv);
    zdouble d = 1;
    for (int k = 0; k < n; ++k) { d *= A[k + (size_t)k * n]; if (ipiv[k] != k) d = -d; }
    free(A); free(ipiv);
    return d;
}

/* getrf + getri: in-place inverse */
int orc_zinv(int n, zdouble *A, int lda) {
    zdouble *F = (zdouble *)malloc(sizeof(zdouble) * (size_t)n * n);
    int *ipiv = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; ++j) memcpy(F + (size_t)j * n, A + (size_t)j * lda, sizeof(zdouble) * n);
    int info = orc_zgetrf(n, F, n, ipiv);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) A[i + (size_t)j * lda] = (i == j);
    /* solve column by column into A (lda stride) */
    zdouble *b = (zdouble *)malloc(sizeof(zdouble) * (size_t)n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) b[i] = (i == c);
        orc_zgetrs(n, 1, F, n, ipiv, b, n);
        for (int i = 0; i < n; ++i) A[i + (size_t)c * lda] = b[i];
    }
    free(b); free(F); free(ipiv);
    return info;
}

/* get_G! (src/public/Gupdate.jl:139-157): G = I - BR (BL BR)^{-1} BL via LU + getrs.
 * BL is np x ns, BR is ns x np. Optionally returns log|det(BL BR)|. */
void orc_get_G(int ns, int np, const zdouble *BL, int ldbl, const zdouble *BR, int ldbr,
               zdouble *G, double *logabsdet) {
    zdouble *M = (zdouble *)malloc(sizeof(zdouble) * (size_t)np * np);
    zdouble *X = (zdouble *)malloc(sizeof(zdouble) * (size_t)np * ns);
    int *ipiv = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    orc_zgemm('N', 'N', np, np, ns, 1.0, BL, ldbl, BR, ldbr, 0.0, M, np);
    orc_zgetrf(np, M, np, ipiv);
    if (logabsdet) {
        double l = 0;
        for (int k = 0; k < np; ++k) l += log(cabs(M[k + (size_t)k * np]));
        *logabsdet = l;
    }
    for (int j = 0; j < ns; ++j)
        for (int i = 0; i < np; ++i) X[i + (size_t)j * np] = BL[i + (size_t)j * ldbl];
    orc_zgetrs(np, ns, M, np, ipiv, X, np);
    orc_zgemm('N', 'N', ns, ns, np, -1.0, BR, ldbr, X, np, 0.0, G, ns);
    for (int i = 0; i < ns; ++i) G[i + (size_t)i * ns] += 1.0;
    free(M); free(X); free(ipiv);
}
