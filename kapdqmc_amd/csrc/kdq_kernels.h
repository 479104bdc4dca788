// Host-side launchers of the HIP kernels (one translation unit per kernel family).
#pragma once
#include "kdq_common.h"

// C = rowS ⊙ (op(A) · diag(inS) · op(B)) ⊙ colS, with
//   mode 0: C = product, mode 1: C = I - product, mode 2: C = Cin - product.
// op codes: 0 = N, 1 = T, 2 = C (conjugate transpose). inS scales the complex operand
// along the contraction index. Dimensions are per-chain codes (DIM_NS / DIM_NP).
struct GemmArgs {
    const ChainDev *chains;
    const Opnd *A;
    const Opnd *B;
    const Opnd *C;
    const Opnd *Cin;
    int mdim, ndim, kdim, mode;
    DiagSpec rowS, colS, inS;
};

hipError_t launch_gemm(hipStream_t st, int nchains, int max_m, int max_n, bool a_cplx, bool b_cplx,
                       int opa, int opb, const GemmArgs &g);

// Local update of one block of sites [s0, s0 + kb) at slice lt (t-U model).
struct DecideArgs {
    const ChainDev *chains;
    const Opnd *G;
    int lt, dir, s0, kb;
    uint32_t sweep;
    int pad;
};
hipError_t launch_decide_tu(hipStream_t st, int nchains, const DecideArgs &a);

// Gout = Gin - Gin[:,A] · W · (E_A - Gin[A,:]) with the block's accepted set A.
struct FlushArgs {
    const ChainDev *chains;
    const Opnd *Gin;
    const Opnd *Gout;
};
hipError_t launch_flush(hipStream_t st, int nchains, int max_ns, const FlushArgs &f);

// In-place Householder QR with explicit Q of the per-chain m x n complex matrix.
struct QRArgs {
    const ChainDev *chains;
    const Opnd *A;
    int mdim, ndim;
};
hipError_t launch_qr(hipStream_t st, int nchains, const QRArgs &q);

// In-place Gauss-Jordan inverse of the per-chain n x n complex matrix (partial pivoting);
// logdet (may be null) receives log|det| per chain, info receives 0 or k+1 on a zero pivot.
struct InvArgs {
    const ChainDev *chains;
    const Opnd *A;
    int ndim, pad;
    double *logdet;
    int *info;
};
hipError_t launch_inv(hipStream_t st, int nchains, const InvArgs &q);

// Observables of phy_measure on the sandwiched G0 (t-U); accumulates into chains[c].meas.
struct MeasArgs {
    const ChainDev *chains;
    const Opnd *G;
};
hipError_t launch_measure_tu(hipStream_t st, int nchains, const MeasArgs &m);

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
};
hipError_t launch_decide_r2(hipStream_t st, int nchains, const DecideR2Args &a);

// t-V bond rotations of one family (mode 0: X <- M X M⁻¹, mode 1: X <- M X rows only).
struct BondRotArgs {
    const ChainDev *chains;
    const Opnd *X;
    int family, lt, sign, mode, ncol_dim, pad;
};
hipError_t launch_bondrot(hipStream_t st, int nchains, int max_nb, int max_cols, const BondRotArgs &a);
// Observables of t-V (kind 1) and t-U-V (kind 2).
hipError_t launch_measure_model(hipStream_t st, int nchains, int kind, const MeasArgs &m);

// Small helpers.
hipError_t launch_copy_mat(hipStream_t st, int nchains, int max_m, int max_n, const ChainDev *chains,
                           const Opnd *src, const Opnd *dst, int mdim, int ndim, bool cplx);
hipError_t launch_real_to_cplx(hipStream_t st, int nchains, int max_m, int max_n, const ChainDev *chains,
                               const Opnd *src, const Opnd *dst, int mdim, int ndim);
hipError_t launch_initial_s(hipStream_t st, int nchains, int max_elem, int nt, const ChainDev *chains);
hipError_t launch_zero_meas(hipStream_t st, int nchains, const ChainDev *chains);
