// Host-side launchers of the HIP kernels (one translation unit per kernel family).
#pragma once
#include "kdq_common.h"

#include <vector>

// C = rowS ⊙ (op(A) · diag(inS) · op(B)) ⊙ colS, with
//   mode 0: C = product, mode 1: C = I - product, mode 2: C = Cin - product,
//   mode 3: C = -product.
// op codes: 0 = N, 1 = T, 2 = C (conjugate transpose). inS scales the complex operand
// along the contraction index. Dimensions are per-chain codes (DIM_NS / DIM_NP).
template <class T>
struct GemmArgsT {
    const ChainDev *chains;
    const typename OpndOf<T>::type *A;
    const typename OpndOf<T>::type *B;
    const typename OpndOf<T>::type *C;
    const typename OpndOf<T>::type *Cin;
    int mdim, ndim, kdim, mode;
    DiagSpec rowS, colS, inS;
};
typedef GemmArgsT<double> GemmArgs;
typedef GemmArgsT<float> GemmArgsF;

hipError_t launch_gemm(hipStream_t st, int nchains, int max_m, int max_n, int max_k, bool a_cplx, bool b_cplx,
                       int opa, int opb, const GemmArgs &g, int tile = 0);   // tile: 0 auto, 32, 64
// FP32 variant (FP32 MFMA, FP32 accumulation; diagonal factors evaluated in FP64, rounded)
hipError_t launch_gemm_f32(hipStream_t st, int nchains, int max_m, int max_n, int max_k, bool a_cplx, bool b_cplx,
                           int opa, int opb, const GemmArgsF &g, int tile = 0);

// Kinetic product through cell momenta (block-circulant eK / eK⁻¹, ChainDev::circ, cL > 0 for
// every chain): per chain, Y = line ⊙ (out ⊙ op(E) (in ⊙ X)) over the nline_dim columns of X
// (rows = 0), or Y = line ⊙ ((X ⊙ in) op(E) ⊙ out) over its rows (rows = 1, op(E) acting from
// the right), E = eK (mat 0), eK⁻¹ (mat 1), HalfeK (2) or HalfeK⁻¹ (3), op 0: E, 1: Eᵀ, 2: Eᴴ. The diagonal factors are
// those of the GEMM path, t-U HS phases only (DiagSpec kind 0 = none, 1). Y may be X.
struct CircArgs {
    const ChainDev *chains;
    const Opnd *X;
    const Opnd *Y;
    int mat, op, rows, nline_dim;
    DiagSpec in, out, line;
    // nsteps > 1 (lines = columns): the product of nsteps such factors, step s with the
    // diagonal factors of slice lt + s·dlt, the lines held in LDS between the steps (the same
    // arithmetic as nsteps single launches, Y_s = X_{s+1}); 0 or 1: one factor
    int nsteps, dlt;
    // plain_last: the last step is the product with matrix mat_last alone (no diagonal factors),
    // e.g. HalfeK (2) or HalfeK⁻¹ (3) closing the measurement's propagation
    int plain_last, mat_last;
};
bool circ_supported(int L);   // cell edges with a kernel instantiation
int circ_lines_per_wg(int L);  // lines of X per workgroup for cell edge L
size_t circ_lds_bytes(int L);  // its dynamic LDS
// grid_x: the largest workgroup count over the chains (lines / lines per workgroup),
// lds_bytes: the largest circ_lds_bytes over the chains, lmax: the largest cell edge
hipError_t launch_circ(hipStream_t st, int nchains, int grid_x, size_t lds_bytes, int lmax, const CircArgs &a);

// Local update of one block of sites [s0, s0 + kb) at slice lt (t-U model).
struct DecideArgs {
    const ChainDev *chains;
    const Opnd *G;
    int lt, dir, s0, kb;
    uint32_t sweep;
    int pad;
    const uint32_t *sweep_dev;   // when set, the sweep counter is read here (graph replays)
    const OpndF *Gf;             // when set, G[S,S] is read from this FP32 copy instead of G
    // deferred flush: G is the slice's G0 and the current G is G0 - U V with the factors'
    // rank rk[blk]; the block reads G0[S,S] - U[S,:] V[:,S] and records rk[blk + 1]
    int deferred, blk;
    const Opnd *U, *V;           // FP64 factors (N x N each; V stored transposed)
    const OpndF *Uf, *Vf;        // FP32 factors (mixed precision)
    // slot: the block's accepted count, sites and W go to slot `slot` of acc_cnt, acc_idx and W.
    // pre: the previous block's update G - G[:,A] W (E_A - G[A,:]) (A, W, count in slot
    // 1 - slot) has not been applied to G yet (its flush runs concurrently); the block reads
    // G[S,S] + G[S,A] W G[A,S]
    int slot, pre;
    // prewrap (with pre = 0, FP64): G is not read; the block's G[S,S] of the wrapped G is formed
    // as line ⊙ (T E) ⊙ out on S x S from the wrap's first product T and the dense kinetic
    // factor E of its second (real, or complex when E.im is set), so the decisions can run
    // beside that second product
    int prewrap, pw_nkc;         // pw_nkc: chunks of 32 of the contraction index (N / 32, rounded up)
    const Opnd *T, *E;
    DiagSpec pw_line, pw_out;
    double *pw_part;             // partial products per chain and chunk: [c][chunk][re, im][32][32]
};
hipError_t launch_decide_tu(hipStream_t st, int nchains, const DecideArgs &a);
// prewrap: the partial products T[S, K] E[K, S] per chunk K of 32 (launch before the decisions,
// same stream)
hipError_t launch_prewrap_part(hipStream_t st, int nchains, const DecideArgs &a);

// Deferred flush, per block after its decisions: with r = rk[blk], A the accepted set and W
// its factor, append X = G[:,A] (columns r.. of U) and Y = W (E_A - G[A,:]) (rows r.. of V),
// where G = G0 - U[:, :r] V[:r, :] is the current matrix. V is stored transposed (N x rank,
// V[t, j] at j + t·ld) so that its rows are contiguous. T = double or float.
template <class T>
struct DeferArgsT {
    const ChainDev *chains;
    const typename OpndOf<T>::type *G0, *U, *V;
    int blk, kmax;
};
hipError_t launch_defer_xy(hipStream_t st, int nchains, int max_ns, const DeferArgsT<double> &a);
hipError_t launch_defer_xy(hipStream_t st, int nchains, int max_ns, const DeferArgsT<float> &a);

// Delayed-update flush with the block's accepted set A and factor W (per chain):
//   mode 0: Xout = Xin - Cin[:,A] · W · (E_A - Rin[A,:])
//   mode 1: Xout = Xin + Cin[:,A] · W · Rin[A,:]
// The plain sweep uses mode 0 with Cin = Rin = Xin = G. Null Cin / Rin mean Xin.
// Default (one_launch = 0): two launches; Z = W · Y (|A| x N) and the gathered columns
// ±Cin[:,A] (N x |A|) into the scratch operands Zs (ld 64) and Cs, then Xout = Xin - Cs · Zs
// of rank |A|. one_launch = 1: each workgroup owns a strip of columns J of one chain, forms
// Z[:,J] = W · Y[:,J] in LDS and streams the strip's rows through Xout = Xin - Cin[:,A] · Z[:,J]
// (the gathered columns read straight from Cin); Xout must alias none of Xin, Cin, Rin (every
// caller writes a spare buffer). Both forms read the accepted count on the device and give
// identical results (same sums, same order).
template <class T>
struct FlushArgsT {
    const ChainDev *chains;
    const typename OpndOf<T>::type *Gin;    // Xin
    const typename OpndOf<T>::type *Gout;   // Xout
    int kmax, mode;     // largest accepted set the preceding decisions can produce (<= 64)
    const typename OpndOf<T>::type *Cin;
    const typename OpndOf<T>::type *Rin;
    const typename OpndOf<T>::type *Zs;     // scratch: 64 x N per chain (two-launch form)
    const typename OpndOf<T>::type *Cs;     // scratch: N x 64 per chain (two-launch form)
    int slot;           // the rank, accepted sites and W of slot `slot`
    int one_launch;     // 1: the strip form
    int commit_lt, commit_s0;   // > 0: write back the field values a prewrapped block 0 left in
                                // chains[c].pw_fld (slice commit_lt, sites commit_s0..), two-launch form
};
typedef FlushArgsT<double> FlushArgs;
typedef FlushArgsT<float> FlushArgsF;
hipError_t launch_flush(hipStream_t st, int nchains, int max_ns, const FlushArgs &f);
hipError_t launch_flush_f32(hipStream_t st, int nchains, int max_ns, const FlushArgsF &f);   // FP32 G

// Mixed precision: dst = FP32(src) for per-chain n x n complex matrices; when maxerr is set,
// max |dst_before - src| is folded into maxerr[c] first (FP32 propagation error against the
// FP64 recompute). to_f64 is the reverse copy.
hipError_t launch_node_check(hipStream_t st, int nchains, int max_n, const ChainDev *chains, const Opnd *wrapped,
                             const Opnd *fresh, int ndim, double *maxerr, int *poison);
hipError_t launch_trace(hipStream_t st, int nchains, const ChainDev *chains, const Opnd *G, double *out);
hipError_t launch_to_f32(hipStream_t st, int nchains, int max_n, const ChainDev *chains, const Opnd *src,
                         const OpndF *dst, int ndim, double *maxerr);
hipError_t launch_to_f64(hipStream_t st, int nchains, int max_n, const ChainDev *chains, const OpndF *src,
                         const Opnd *dst, int ndim, int cplx);

// Cholesky S + shift·trace(S)·I = L Lᴴ in place (lower triangle) of the per-chain n x n
// Hermitian matrix; info (sticky, may be null) receives j+1 on a non-positive pivot.
struct CholArgs {
    const ChainDev *chains;
    const Opnd *A;
    int ndim, pad;
    double shift;
    int *info;
    double *logdiag;   // when set, logdiag[c] += Σ_j log L_jj (the log|det| of the factor)
    const Opnd *Dinv;  // when set: the inverses of the 16 x 16 diagonal blocks of L, block b at
                       // rows 16b.., columns 0..15
};
hipError_t launch_chol(hipStream_t st, int nchains, int max_n, const CholArgs &a);

// X <- X L⁻ᴴ in place for the m x n matrix X and the lower n x n factor L.
struct TrsmArgs {
    const ChainDev *chains;
    const Opnd *X;
    const Opnd *L;
    int mdim, ndim;
    const Opnd *Dinv;  // when set: the diagonal-block inverses written by k_chol
};
hipError_t launch_trsm_rlh(hipStream_t st, int nchains, int max_m, int max_n, const TrsmArgs &t);

// In-place LU with partial pivoting (getrf semantics, 0-based pivots at piv + chain·piv_ld);
// logdet (may be null) receives log|det|, info (sticky) j+1 on an exactly zero pivot.
struct LuArgs {
    const ChainDev *chains;
    const Opnd *A;
    int ndim, piv_ld;
    int *piv;
    double *logdet;
    int *info;
    double *phase;   // [2 per chain] det A / |det A| (re, im), or null
    const Opnd *Dinv;  // when set: inverses of the 16 x 16 diagonal blocks, unit lower L at
                       // columns 0..15 and upper U at columns 16..31 (block b at rows 16b..)
};
hipError_t launch_lu(hipStream_t st, int nchains, int max_n, const LuArgs &a);

// X = M⁻¹ Cᴴ from the LU factors of M (n x n) for the per-chain nr x n matrix C.
struct LuSolveArgs {
    const ChainDev *chains;
    const Opnd *LU;
    const Opnd *C;
    const Opnd *X;
    const int *piv;
    int ndim, rdim, piv_ld, pad;
    const Opnd *Dinv;  // when set: the diagonal-block inverses written by k_lu
};
hipError_t launch_lu_solve(hipStream_t st, int nchains, int max_n, int max_r, const LuSolveArgs &s);

// X = A⁻¹ for the per-chain n x n complex A (n <= 144) by Gauss-Jordan elimination with partial
// pivoting in registers; X may be A. logdet (if set) receives log|det A|, info (sticky) j+1 for
// an exactly zero pivot at step j.
struct InvArgs {
    const ChainDev *chains;
    const Opnd *A;
    const Opnd *X;
    int ndim, pad;
    double *logdet;
    int *info;
};
bool inv_supported(int max_n);
hipError_t launch_inv(hipStream_t st, int nchains, int max_n, const InvArgs &a);

// X = L⁻¹ (lower, zeros above) for S + σI = L Lᴴ, S Hermitian n x n (n <= 144), σ = shift ·
// trace S, in one register pass; X may be S. logdiag (if set) += Σ log L_jj, info (sticky) j+1
// for a non-positive pivot.
struct CholInvArgs {
    const ChainDev *chains;
    const Opnd *A;
    const Opnd *X;
    int ndim, pad;
    double shift;
    int *info;
    double *logdiag;
};
hipError_t launch_cholinv(hipStream_t st, int nchains, int max_n, const CholInvArgs &a);
size_t factor_lds_bytes(int which, int max_n);

// Observables of phy_measure on the sandwiched G0 (t-U); accumulates into chains[c].meas.
struct MeasArgs {
    const ChainDev *chains;
    const Opnd *G;
    double *meas;   // nullptr: accumulate into chains[c].meas; else into meas + c * KDQ_MEAS_SLOTS
    double *part;   // tile partial sums, meas_part_doubles(nchains, max_ns) doubles
    int T;          // tiles per dimension, meas_tiles(max_ns)
};
// phy_measure of every chain (any model) on the sandwiched G0: tile reduction + per-chain finish.
hipError_t launch_measure(hipStream_t st, int nchains, const MeasArgs &m);
int meas_tiles(int max_ns);
size_t meas_part_doubles(int nchains, int max_ns);

// Rank-2 local update (t-V bond flips, t-U-V directed-bond flips) of one block.
struct BlockDesc {
    int e0, ne, nslot, pad;   // field elements [e0, e0 + ne) within the slice, touched sites
    int sites[64];            // global site of each local slot
    uint8_t sl0[32], sl1[32]; // local slots of the two sites of each element
};
struct DecideR2Args {
    const ChainDev *chains;
    const Opnd *G;
    const BlockDesc *blocks;
    const int *block_offset;  // [nchains] first block of the chain
    const int *nblocks;       // [nchains] blocks per family of the chain
    int lt, dir, family, block;
    uint32_t sweep;
    int pad;
    const uint32_t *sweep_dev;   // when set, the sweep counter is read here (graph replays)
};
hipError_t launch_decide_r2(hipStream_t st, int nchains, const DecideR2Args &a);

// t-V bond rotations of one family (mode 0: X <- M X M⁻¹, mode 1: X <- M X rows only).
struct BondRotArgs {
    const ChainDev *chains;
    const Opnd *X;
    int family, lt, sign, mode, ncol_dim, pad;   // mode 2: ncol_dim gives the row count
};
hipError_t launch_bondrot(hipStream_t st, int nchains, int max_nb, int max_cols, const BondRotArgs &a);
// Observables of t-V (kind 1) and t-U-V (kind 2).

// Small helpers.
hipError_t launch_copy_mat(hipStream_t st, int nchains, int max_m, int max_n, const ChainDev *chains,
                           const Opnd *src, const Opnd *dst, int mdim, int ndim, bool cplx);
hipError_t launch_real_to_cplx(hipStream_t st, int nchains, int max_m, int max_n, const ChainDev *chains,
                               const Opnd *src, const Opnd *dst, int mdim, int ndim);
hipError_t launch_initial_s(hipStream_t st, int nchains, int max_elem, int nt, const ChainDev *chains);
hipError_t launch_zero_meas(hipStream_t st, int nchains, const ChainDev *chains);
hipError_t launch_bump(hipStream_t st, uint32_t *counter);   // *counter += 1 (in-graph sweep counter)

// ---------------------------------------------------------------- swap-replica Rényi-2 (SCEE)
// Per chain: regions A and B (0-based sites, A first), the replicas' region blocks of G(Θ)
// M[r][X] = G0_r[X,X] (kept current inside a block of decisions), the inverse Grover
// matrices gmInv[X], |det g_X| products detg[X] and the running Rényi accumulator.
// All matrices planar complex, column-major, ld = region size.
struct SceeRegion {
    int nA, nB, nlam;
    int dop;               // 1: disorder operator (one replica, X = e^{iα}I + (1-e^{iα})G0[X,X])
    double lambda;
    const int *idx;        // [nA + nB]
    double *M[2][2];       // [replica][region]
    double *gmInv[2];      // [region]
    double *C[2];          // [region] product, then Grover matrix
    double *detg;          // [2]
    double *osum;          // [2]: Σ (detg_A / detg_B)^(1/Nλ), count
    double ea[2];          // e^{iα} of the disorder operator
};
// Entries: replica r of chain c is entry c + r·nch of the chain tables.
struct SceeArgs {
    const ChainDev *chains;
    const Opnd *Gt, *Gt0, *G0t, *G0;   // current buffers, per entry
    const SceeRegion *reg;             // [nch]
    int nch, lt, dir, s0, kb, nx_max;
    uint32_t sweep;
    int pad;
};
hipError_t launch_scee_decide(hipStream_t st, const SceeArgs &a);
size_t scee_lds_bytes(int ks, int nx);   // dynamic LDS of the decide kernel for block size ks
hipError_t launch_scee_gather(hipStream_t st, int nch, int nx_max, const SceeArgs &a);
hipError_t launch_scee_grover_form(hipStream_t st, int nch, int total, const SceeRegion *reg);
// Disorder operator (get_ABX!, src/tU/disorder_operate.jl:202-223): C = e^{iα}I + (1-e^{iα})M[0]
hipError_t launch_dop_form(hipStream_t st, int nch, int total, const SceeRegion *reg);
hipError_t launch_scee_detg(hipStream_t st, int nch, const SceeRegion *reg, const double *logdet_a,
                            const double *logdet_b);
hipError_t launch_scee_accum(hipStream_t st, int nch, const SceeRegion *reg);
// det g of the regions from the LU log|det| and phase: |det|² (t-U, t-U-V) or the signed
// real det (t-V, signed = 1)
hipError_t launch_scee_detg_signed(hipStream_t st, int nch, const SceeRegion *reg, const double *logdet_a,
                                   const double *logdet_b, const double *phase_a, const double *phase_b);
// Decisions of the rank-2 models (t-V bonds of one family, t-U-V directed bonds) of slice lt,
// replica 1 then replica 2 per element, every accepted change applied at once to the
// replica's four matrices and to both regions.
struct SceeR2Args {
    const ChainDev *chains;
    const Opnd *Gt, *Gt0, *G0t, *G0;   // per entry (replica r of chain c at c + r·nch)
    const SceeRegion *reg;
    int nch, lt, dir, family, kind;    // family: t-V bond family (elements family·nb ..), -1: every element; kind 1 t-V, 2 t-U-V
    uint32_t sweep;
    int nx_max, max_ns, s0;            // s0: first bond of the block within the family (block kernel)
    const uint32_t *sweep_dev;
};
hipError_t launch_scee_r2(hipStream_t st, const SceeR2Args &a);
size_t scee_r2_lds_bytes(int nx_max, int max_ns);
// t-V delayed blocks of kb (4 or 8) bonds of one family: decisions with LDS working sets, W and
// the accepted sites for the four-matrix flush; returns hipErrorInvalidValue if the block's
// LDS does not fit (then the element kernel is used)
hipError_t launch_scee_r2_block(hipStream_t st, const SceeR2Args &a, int kb);
size_t scee_r2_block_lds_bytes(int kb, int nx_max);
hipError_t launch_scee_r2_block_uv(hipStream_t st, const SceeR2Args &a, int kb);
size_t scee_r2_block_uv_lds_bytes(int kb, int nx_max);
// X[i][i] += v for the per-chain n x n matrices (dimension code ndim); set = 1 writes I.
hipError_t launch_diag(hipStream_t st, int nchains, int max_n, const ChainDev *chains, const Opnd *X, int ndim,
                       double v, int set);

// ---------------------------------------------------------------- self-learning moves
typedef struct kdq_net_s *kdq_net;
// Inner chain of a global move on exp(-E) of the network for nch chains whose trial fields
// (device arrays; d_field_ptrs is the device table of them) change in place: nsteps local
// steps; E0_host receives the starting energies, d_E1 the final ones.
int slmc_inner(kdq_net n, hipStream_t st, int nch, uint8_t **d_field_ptrs, const uint64_t *d_seed,
               const uint32_t *d_chain, double *d_E1, int nsteps, int move, double *E0_host);
int slmc_geometry(kdq_net n, int *ns, int *nt);
