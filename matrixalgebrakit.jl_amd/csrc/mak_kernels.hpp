// Argument structs and host launchers shared by the library's translation
// units.  Every pointer is a device pointer; matrices are column-major.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mak_common.hpp"

namespace mak {

// ---------------------------------------------------------------- panel / TSQR
struct Geqr2NodeArgs {
  void* X;          // stacked input rows of this level (factored in place)
  int64_t ldx;
  int H;            // rows of the stacked input
  int nodes;        // workgroups (tree nodes) of this level
  int chunk;        // rows per node (> 0), or 0 for a balanced split of H
  int nbc;          // active columns (<= 32)
  void* Tout;       // T factor of node i at Tout + i * tstride (32 x 32, ld 32)
  int64_t tstride;
  void* Rnext;      // R of node i -> rows [i*nbc, i*nbc + nbc) of the next level
  int64_t ldr;
  void* V1out;      // copy of the top 32x32 of V per node (unit lower), or null
  const int* gate;  // when non-null the launch runs only if *gate == gate_want
  int gate_want;
  int variant;      // node kernel: 0 = default per element type, 1..3 = a fixed one (tests)
};

// Device-side predication of a launch (all threads of a workgroup agree).
__device__ inline bool gate_closed(const int* gate, int want) { return gate != nullptr && *gate != want; }

// One independent panel node of a batch (geqr2_node2_tasks_kernel).
struct NodeTask {
  void* X;          // H x nbc panel, factored in place
  int64_t ldx;
  void* Tout;       // 32 x 32 T factor (ld 32)
  void* Rn;         // optional: R (nbc x nbc) of the node, ld ldr (TSQR leaves)
  int64_t ldr;
  void* V1;         // optional: top 32 x 32 of V (unit lower), ld 32
  const int* gate;  // optional: the task runs only if *gate == 1 (its panel's CholeskyQR2 was rejected)
  int H;
  int nbc;
};

// One panel of a batch factored by CholeskyQR2 with Householder
// reconstruction in one workgroup (cholqr2_panel_tasks_kernel).
struct CholQrTask {
  void* X;          // H x nbc panel (H >= 64): compact WY form on success, untouched otherwise
  int64_t ldx;
  void* Q1;         // H x nbc scratch (Q1 = X R1^{-1}), ld ldq
  int64_t ldq;
  void* Tout;       // 32 x 32 T factor (ld 32)
  int* status;      // out: 0 accepted, 1 rejected (the Householder launches gated on it run)
  int H;
  int nbc;
};
template <typename T> void launch_cholqr2_panel_tasks(const CholQrTask* tasks, int ntasks, hipStream_t s);

constexpr int kMaxTreeLevels = 8;

struct TsqrRootArgs {
  const void* Xtop;  // factored top node (V strictly below, R on/above)
  int64_t ldxtop;
  int htop;
  int nbc;
  const void* Ttop;
  int nchain;        // levels below the top
  const void* chainV1[kMaxTreeLevels];
  int64_t chainLd[kMaxTreeLevels];
  const void* chainT[kMaxTreeLevels];
  void* Mtop;        // out: T_top V_top1^H (32 x 32)
  void* Tout;        // out: reconstructed T of the panel (32 x 32)
  void* Uout;        // out: U of the LU of Q - D (32 x 32)
  void* Apanel;      // out: top nbc rows of the panel <- [D R ; L1]
  int64_t lda;
  const int* gate;   // run only if *gate == gate_want (when non-null)
  int gate_want;
};

struct TsqrExpandArgs {
  int H;
  int nodes;
  int chunk;
  int nbc;
  void* X;                 // V of this level's nodes (leaves: the panel itself)
  int64_t ldx;
  const void* Tnodes;      // T of node i at Tnodes + i * 32 * 32
  const void* V1base;      // top 32x32 of V of node i at V1base + i * v1stride
  int64_t v1stride;
  int64_t v1ld;
  const void* Qparent;     // parent level explicit Q (used when Mtop is null)
  int64_t ldqp;
  const void* Mtop;        // parent is the top node: Q block = E - V_top Mtop
  const void* Xtop;
  int64_t ldxtop;
  void* Qout;              // inner levels: explicit Q rows of this level
  int64_t ldqo;
  const void* Umat;        // leaves: U of the LU; rows >= nbc become V = Q U^{-1}
  const int* gate;         // run only if *gate == gate_want (when non-null)
  int gate_want;
};

template <typename T> int geqr2_node_capacity();
// Node capacity of size class cls (0: 256, 1: 512, 2: 1024 rows; class 2 is
// real only) for geqr2_node_tasks launches.
template <typename T> int node_class_rows(int cls);
template <typename T> void launch_geqr2_node_tasks(const NodeTask* tasks, int ntasks, int cls, hipStream_t s);
template <typename T> void launch_tsqr_root_finish_tasks(const TsqrRootArgs* tasks, int ntasks, hipStream_t s);
template <typename T>
void launch_tsqr_expand_tasks(const TsqrExpandArgs* tasks, int ntasks, int nodes, int max_node_rows, hipStream_t s);
template <typename T> void launch_geqr2_node(const Geqr2NodeArgs& a, int leaf_rows_max, hipStream_t s);
template <typename T> void launch_tsqr_root_finish(const TsqrRootArgs& a, hipStream_t s);
template <typename T> void launch_tsqr_expand(const TsqrExpandArgs& a, int max_node_rows, hipStream_t s);

// ---------------------------------------------------------------- GEMM
enum Op { OP_N = 0, OP_T = 1, OP_C = 2 };
enum TriHint { TRI_A_LOWER = 1, TRI_A_UPPER = 2, TRI_B_UPPER = 4, TRI_B_LOWER = 8 };

struct GemmArgs {
  int M, N, K;
  const void* A;
  int64_t lda, strideA;   // strideA: element offset between batch entries
  const void* B;
  int64_t ldb, strideB;
  void* C;
  int64_t ldc, strideC;
  double alpha_re, alpha_im, beta_re, beta_im;
  int ksplit;             // >1: slice z of K writes C + z * splitStride (beta must be 0)
  int kchunk;
  int64_t splitStride;
  int batch;
  int unitLowerA;         // stored A is unit lower trapezoidal (strict lower part valid)
  // Structure hints (FP64, complex and FP32 / BF16 kernels; kernels that ignore
  // them are also correct): the K range of a tile is clipped to the terms that can be
  // nonzero when op(A) is lower (tri & TRI_A_LOWER: k <= i) or upper
  // (TRI_A_UPPER: k >= i) triangular and op(B) is upper (TRI_B_UPPER: k <= j)
  // or lower (TRI_B_LOWER: k >= j) triangular; upperC skips tiles strictly
  // below the diagonal (their C is left as it was), for a consumer that reads
  // only the upper triangle.
  int tri;
  int upperC;
  const int* gate;        // run only if *gate == gate_want (when non-null)
  int gate_want;
  // Carried first Gram of the next QR panel (real C = C - op(A) op(B) with
  // op N, N): when cGw is not null, the workgroup of tile (0, 0) also forms
  // G1 = cGw - Rt^H Rt from Rt = C(0:32, 0:cNbc) after its update (rows /
  // columns beyond cNbc: identity), factors G1 = W^H W and writes W and W^{-1}
  // (32 x 32, ld 32) to cR1 / cR1i -- or, when G1 is not positive definite,
  // sets cStatus[2] = cJ0 and cStatus[0] = 1.  Skipped when cStatus[0] is set.
  const void* cGw;
  void* cR1;
  void* cR1i;
  int* cStatus;
  int cJ0, cNbc;
  // Side job of a real split-K product (M <= 32 tiles): when rPart is not
  // null, one extra workgroup sums the rN 32 x 32 partials at rPart (in order)
  // into rOut (32 x 32), concurrently with the product's tiles.
  const void* rPart;
  void* rOut;
  int rN;
  // Tile size (square) forced for the real kernels: 0 = chosen by the
  // dispatcher from the launch size, else 32, 64 or 128.
  int tile;
  // FP32 / BF16 products only (gemm_prec with prec 32 or 16): when Cin is not
  // null, C = alpha op(A) op(B) + beta Cin (Cin m x n, ld ldcin; it may be C).
  const void* Cin;
  int64_t ldcin;
};

// C = alpha op(A) op(B) + beta C on FP64 MFMA (real) or 4 real MFMAs (complex).
template <typename T> void gemm(int opA, int opB, GemmArgs g, hipStream_t s);

// Split-K partial count and chunk for an (M x N x K) product with small M:
// about kSplitTargetWgs workgroups, slices of at least 256 rows.  ksplit * N
// stays below kSplitTargetWgs * bn + N.
constexpr int kSplitTargetWgs = 512;
void splitk_plan(int N, int K, int bn, int& ksplit, int& kchunk);

// W2 = op(Tm) * sum_z Wp[z]  (nbc x ncols, ld nbc), op = conj-transpose if transT;
// skipped when gate is given and *gate != 0
template <typename T>
void wy_reduce(const void* Wp, int ksplit, int64_t slice, const void* Tm, int transT, int nbc,
               int ncols, void* W2, hipStream_t s, const int* gate = nullptr);

// Precision-dispatched GEMM on FP64 data: prec 64 (FP64 MFMA), 32 (FP32 MFMA),
// 16 (BF16 MFMA), FP32 accumulation for 32 / 16.  Complex data runs in FP64.
// ws: gemm_prec_workspace_bytes(M, N, K) bytes (the rounded operand copies;
// unused for prec 64).
size_t gemm_prec_workspace_bytes(int M, int N, int K);
template <typename T> void gemm_prec(int prec, int opA, int opB, GemmArgs g, void* ws, hipStream_t s);

// C = alpha op(A) op(B) + beta Cin (Cin as in gemm_prec) on BF16 MFMA when
// *g16 == 1, else FP32 MFMA when *g32 == 1, else nothing (g.gate is
// ignored), FP32 accumulation: the precision is chosen on the device, and
// both forms take the same two launches (the packs, the product).  Real data
// (cplx false; any ops; ws: gemm_prec_workspace_bytes(M, N, K)) or complex
// data (cplx true; A M x K and B K x N both op N, real alpha and beta,
// TRI_B_UPPER the only hint; through the real embedding [A.re A.im]
// [[B.re B.im], [-B.im B.re]] with the parts interleaved along k; ws:
// gemm_lo_complex_workspace_bytes(M, N, K)).  Returns -1 for arguments
// outside those forms (nothing enqueued), else 0.
size_t gemm_lo_complex_workspace_bytes(int M, int N, int K);
int gemm_lo_reduced(const GemmArgs& g, int opA, int opB, const int* g16, const int* g32, void* ws, hipStream_t s,
                    bool cplx);

// ---------------------------------------------------------------- dense blocks
// Cholesky with inverse by the persistent tile kernel (mak_chol.hip):
// G (n x n, upper triangle read) -> R (upper, strict lower part zeroed) with
// G = R^H R, and X = R^{-1} (upper, strict lower part zeroed), n <= kCholTileMaxN.
// flags: chol_tile_flag_ints(n) ints of device scratch.  info (device, may be
// null): set to info_offset + j when the leading j x j minor is not positive
// definite (if still 0).  grid <= 0: default workgroup count.  Returns 0 or a
// HIP launch error (1); -3 when n is too large.
//
// rp (may be null): also Y = B X (B, Y m x n, column-major, elements of T),
// computed in the same launch as X's column blocks become final.  Xc: n
// columns of scratch (elements of T), leading dimension ldxc >= n a multiple
// of 16, 128-byte aligned (a copy of X's final tiles that the product reads
// through the caches: with whole 128-byte lines per column, no cached line
// holds a value that is not final yet).
constexpr int kCholTileMaxNt = 64;
constexpr int kCholTileMaxN = 32 * kCholTileMaxNt;
struct CholRightProduct {
  const void* B;
  int64_t ldb;
  void* Y;
  int64_t ldy;
  int m;
  void* Xc;
  int64_t ldxc;
};
size_t chol_tile_flag_ints(int n);
template <typename T>
int chol_tile_factor(T* G, int64_t ldg, int n, T* X, int64_t ldx, int* flags, int* info, int info_offset,
                     hipStream_t s, int grid = 0, long long* trace = nullptr, const CholRightProduct* rp = nullptr);
int chol_tile_order_check(int nt, int ny, int ysplit);
template <typename T>
void potrf_upper(T* Z, int64_t ldz, int n, T* Winv, int64_t ldw, T* tmp, int* dinfo, hipStream_t s);
template <typename T>
void triangular_right_inplace(T* X, int64_t ldx, int m, int kb, const T* U, int64_t ldu, double alpha,
                              hipStream_t s);
template <typename T>
void sumsq(const T* X, int64_t ldx, const T* Y, int64_t ldy, int m, int n, double* partial, double* out,
           hipStream_t s);
template <typename T>
void axpby(int m, int n, double a, const T* X, int64_t ldx, double b, T* Y, int64_t ldy, hipStream_t s);
template <typename T> void herm_part(T* C, int64_t ldc, int n, hipStream_t s);
template <typename T> void add_diag(T* A, int64_t lda, int n, double v, hipStream_t s);
template <typename T> void fill_zero(T* A, int64_t lda, int m, int n, hipStream_t s);
template <typename T> void copy_mat(int m, int n, const T* X, int64_t ldx, T* Y, int64_t ldy, hipStream_t s);

// ---------------------------------------------------------------- block reflectors
// Vc (rows x kg) = unit lower trapezoidal V stored below A's diagonal.
template <typename T> void copy_unit_lower(const T* A, int64_t lda, int rows, int kg, T* Vc, int64_t ldvc,
                                           hipStream_t s);
// out (M x N, ld ldo) = sum of the ks slices P + z * slice (each M x N, ld M).
template <typename T> void sum_slices(const T* P, int ks, int64_t slice, int M, int N, T* out, int64_t ldo,
                                      hipStream_t s);
// T (kg x kg, ld kg) of the block reflector of np <= 4 consecutive 32-column
// panels, from S = V^H V (kg x kg, ld kg) and the panels' 32 x 32 T blocks.
template <typename T> void wy_merge(const T* S, const T* Tp, int np, int kg, T* Tout, hipStream_t s);
// Every group of four panels at once (workgroup g: panels 4g .. 4g + 3 of
// npan, kg = min(minmn - 128 g, 128) columns): S and Tout hold one 128 x 128
// slot per group (ld kg), Tp the panels' 32 x 32 T blocks; groups of one
// panel are copied through.
template <typename T> void wy_merge_groups(const T* S, const T* Tp, int npan, int minmn, T* Tout, hipStream_t s);

// ---------------------------------------------------------------- elementwise
template <typename T> void set_identity(void* Q, int64_t ldq, int m, int n, hipStream_t s);
// R <- triu(A[0:minmn, :]) row-scaled by conj(sign_safe(diag)) when positive
template <typename T>
void qr_extract_r(const void* A, int64_t lda, void* R, int64_t ldr, int rrows, int n, int minmn,
                  int positive, hipStream_t s);
// Q[:, j] *= sign_safe(A[j, j]) for j < minmn
template <typename T>
void qr_gauge_q(const void* A, int64_t lda, void* Q, int64_t ldq, int m, int minmn, hipStream_t s);

}  // namespace mak
