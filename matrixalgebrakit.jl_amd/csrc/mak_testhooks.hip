// Internal entry points used only by the test suite to check kernels in
// isolation (GEMM, one panel node, the blocked factorisation).  They are not
// part of the public C ABI in include/mak_mi355x.h.
#include "mak_common.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

extern "C" {

int maktest_gemm(void* stream, int dtype, int opA, int opB, int M, int N, int K, const void* A, int lda,
                 const void* B, int ldb, void* C, int ldc, double alpha, double beta, int unitLowerA, int ksplit,
                 int kchunk) {
  mak::GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha_re = alpha; g.beta_re = beta;
  g.ksplit = ksplit; g.kchunk = kchunk; g.splitStride = (int64_t)ldc * N; g.batch = 1;
  g.unitLowerA = unitLowerA;
  if (dtype == MAK_F64) mak::gemm<double>(opA, opB, g, (hipStream_t)stream);
  else mak::gemm<mak::zd>(opA, opB, g, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Real FP64 GEMM with the structure hints (tri, upperC), split-K slices
// (ksplit > 1: slice z at C + z * ldc * N) and a forced tile size (0 = auto).
int maktest_gemm_hint(void* stream, int opA, int opB, int M, int N, int K, const double* A, int lda,
                      const double* B, int ldb, double* C, int ldc, double alpha, double beta, int tri, int upperC,
                      int ksplit, int tile) {
  mak::GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha_re = alpha; g.beta_re = beta;
  g.ksplit = ksplit > 1 ? ksplit : 1;
  g.kchunk = (K + g.ksplit - 1) / g.ksplit;
  g.kchunk = ((g.kchunk + 15) / 16) * 16;
  if (g.kchunk < 16) g.kchunk = 16;
  g.ksplit = (K + g.kchunk - 1) / g.kchunk;
  if (g.ksplit < 1) g.ksplit = 1;
  g.splitStride = (int64_t)ldc * N; g.batch = 1;
  g.tri = tri; g.upperC = upperC; g.tile = tile;
  mak::gemm<double>(opA, opB, g, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// FP32 / BF16 MFMA product of FP64 data (prec 32 / 16; 64 runs the FP64
// kernel); ws: maktest_gemm_prec_ws(M, N, K) bytes.
size_t maktest_gemm_prec_ws(int M, int N, int K) { return mak::gemm_prec_workspace_bytes(M, N, K); }
int maktest_gemm_prec(void* stream, int prec, int opA, int opB, int M, int N, int K, const double* A, int lda,
                      const double* B, int ldb, double* C, int ldc, double alpha, double beta, int tri, void* ws) {
  mak::GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha_re = alpha; g.beta_re = beta;
  g.ksplit = 1; g.kchunk = K; g.batch = 1; g.tri = tri;
  mak::gemm_prec<double>(prec, opA, opB, g, ws, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Status words a qr call left in its workspace `work` (synchronous read):
// out[0] 1 when the whole-matrix CholeskyQR2 result was taken, out[1] the
// first 32-column panel handed to Householder TSQR (-1: none), out[2] the
// whole-matrix path's Cholesky info, out[3] 1 when that path exists for the
// shape, out[4..5] its ||Q1^H Q1 - I||_F (a double), out[6 / 7 / 8] 1 when
// its final product took the FP32 correction / ran on FP64 / took the BF16
// correction (nine ints).
int maktest_qr_record(int dtype, int full, int m, int n, const void* work, int* out) {
  if (dtype == MAK_F64) mak::qr_record<double>((const char*)work, m, n, full, out);
  else mak::qr_record<mak::zd>((const char*)work, m, n, full, out);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int maktest_geqr2(void* stream, int dtype, int H, int nbc, void* X, int ldx, void* Tout) {
  const int cap = dtype == MAK_F64 ? mak::geqr2_node_capacity<double>() : mak::geqr2_node_capacity<mak::zd>();
  if (H > cap || nbc > 32) return -3;
  mak::Geqr2NodeArgs g{};
  g.X = X; g.ldx = ldx; g.H = H; g.nodes = 1; g.chunk = 0; g.nbc = nbc;
  g.Tout = Tout; g.tstride = 32 * 32;
  if (dtype == MAK_F64) mak::launch_geqr2_node<double>(g, 0, (hipStream_t)stream);
  else mak::launch_geqr2_node<mak::zd>(g, 0, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}


// One TSQR level: `nodes` workgroups over H stacked rows (balanced split),
// each node factored in place; node i's T at Tout + i * 32 * 32.  variant:
// Geqr2NodeArgs::variant.
int maktest_geqr2_level(void* stream, int dtype, int H, int nodes, int nbc, void* X, int ldx, void* Tout,
                        int variant) {
  const int cap = dtype == MAK_F64 ? mak::geqr2_node_capacity<double>() : mak::geqr2_node_capacity<mak::zd>();
  if (nodes < 1 || H < nodes || nbc < 1 || nbc > 32) return -4;
  if ((H + nodes - 1) / nodes > cap) return -3;
  mak::Geqr2NodeArgs g{};
  g.X = X; g.ldx = ldx; g.H = H; g.nodes = nodes; g.chunk = 0; g.nbc = nbc;
  g.Tout = Tout; g.tstride = 32 * 32;
  g.variant = variant;
  if (dtype == MAK_F64) mak::launch_geqr2_node<double>(g, 0, (hipStream_t)stream);
  else mak::launch_geqr2_node<mak::zd>(g, 0, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Cholesky with inverse: Z (n x n, upper part) -> W (upper) with Z = W^H W and
// Y = W^{-1}; info receives j > 0 when the leading j x j minor is not positive
// definite.  tmp: n * n / 4 + 64 * n elements.
int maktest_potrf(void* stream, int dtype, int n, void* Z, int ldz, void* Y, int ldy, void* tmp, int* info) {
  if (dtype == MAK_F64)
    mak::potrf_upper<double>((double*)Z, ldz, n, (double*)Y, ldy, (double*)tmp, info, (hipStream_t)stream);
  else
    mak::potrf_upper<mak::zd>((mak::zd*)Z, ldz, n, (mak::zd*)Y, ldy, (mak::zd*)tmp, info, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The persistent tile Cholesky alone (n <= 2048): G -> R, X = R^{-1}; grid <= 0
// selects the default workgroup count.  flags: maktest_chol_tile_flag_ints(n) ints.
size_t maktest_chol_tile_flag_ints(int n) { return mak::chol_tile_flag_ints(n); }
int maktest_chol_tile(void* stream, int dtype, int n, void* G, int ldg, void* X, int ldx, int* flags, int* info,
                      int grid) {
  if (dtype == MAK_F64)
    return mak::chol_tile_factor<double>((double*)G, ldg, n, (double*)X, ldx, flags, info, 0, (hipStream_t)stream,
                                         grid);
  return mak::chol_tile_factor<mak::zd>((mak::zd*)G, ldg, n, (mak::zd*)X, ldx, flags, info, 0, (hipStream_t)stream,
                                        grid);
}
// As maktest_chol_tile (real), also writing 8 wall-clock stamps (100 MHz) per
// panel iteration to trace (8 * ceil(n / 32) long longs).
int maktest_chol_tile_trace(void* stream, int n, double* G, int ldg, double* X, int ldx, int* flags, int* info,
                            int grid, long long* trace) {
  return mak::chol_tile_factor<double>(G, ldg, n, X, ldx, flags, info, 0, (hipStream_t)stream, grid, trace);
}
// As maktest_chol_tile (real), also Y = B X (B, Y m x n) in the same launch;
// Xc: ldxc * n doubles of scratch (ldxc >= n a multiple of 16, 128-byte aligned).
// trace (may be null): as maktest_chol_tile_trace, and from entry 256 the
// start and end stamps of every right-product task (by queue index: 256 + 2 *
// the total task count entries).
int maktest_chol_tile_product(void* stream, int n, double* G, int ldg, double* X, int ldx, int* flags, int* info,
                              int grid, int m, const double* B, int ldb, double* Y, int ldy, double* Xc, int ldxc,
                              long long* trace) {
  const mak::CholRightProduct rp{B, ldb, Y, ldy, m, Xc, ldxc};
  return mak::chol_tile_factor<double>(G, ldg, n, X, ldx, flags, info, 0, (hipStream_t)stream, grid, trace, &rp);
}
// Host-only check of the tile kernel's task order for nt tiles per side, ny
// right-product row blocks per pair step and pairs from ysplit on splitting
// them in two: the task count when every dependency precedes its task, else -1.
int maktest_chol_order_check(int nt, int ny, int ysplit) { return mak::chol_tile_order_check(nt, ny, ysplit); }

// Blocked factorisation only: A <- compact WY (V below, R above), one 32x32 T
// per panel written to Tpanels (32 x 32 x npanels); ws from the QR query.
int maktest_geqrf(void* stream, int dtype, int m, int n, void* A, int lda, void* Tpanels, void* ws) {
  const int minmn = m < n ? m : n;
  const int npan = (minmn + 31) / 32;
  if (dtype == MAK_F64) {
    mak::geqrf_blocked<double>((double*)A, lda, m, n, (char*)ws, minmn, (hipStream_t)stream);
    hipMemcpyAsync(Tpanels, ws, sizeof(double) * 32 * 32 * npan, hipMemcpyDeviceToDevice, (hipStream_t)stream);
  } else {
    mak::geqrf_blocked<mak::zd>((mak::zd*)A, lda, m, n, (char*)ws, minmn, (hipStream_t)stream);
    hipMemcpyAsync(Tpanels, ws, sizeof(mak::zd) * 32 * 32 * npan, hipMemcpyDeviceToDevice, (hipStream_t)stream);
  }
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // extern "C"

// One wave runs a single MFMA on per-lane operands and writes the 4 result
// registers of every lane: d[lane * 4 + r].  Used to check fragment layouts.
__global__ void maktest_mfma_kernel(int which, const double* a, const double* b, double* d) {
  const int l = threadIdx.x;
  if (which == 0) {
    mak::v4d c = {0.0, 0.0, 0.0, 0.0};
    c = mak::mfma_f64(a[l], b[l], c);
    for (int r = 0; r < 4; ++r) d[l * 4 + r] = c[r];
  } else if (which == 1) {
    mak::v4f c = {0.f, 0.f, 0.f, 0.f};
    c = mak::mfma_f32((float)a[l], (float)b[l], c);
    for (int r = 0; r < 4; ++r) d[l * 4 + r] = c[r];
  } else {
    mak::v8bf av, bv;
    for (int e = 0; e < 8; ++e) {
      av[e] = (__bf16)(float)a[l * 8 + e];
      bv[e] = (__bf16)(float)b[l * 8 + e];
    }
    mak::v4f c = {0.f, 0.f, 0.f, 0.f};
    c = mak::mfma_bf16(av, bv, c);
    for (int r = 0; r < 4; ++r) d[l * 4 + r] = c[r];
  }
}

extern "C" int maktest_mfma(void* stream, int which, const double* a, const double* b, double* d) {
  hipLaunchKernelGGL(maktest_mfma_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, which, a, b, d);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}
