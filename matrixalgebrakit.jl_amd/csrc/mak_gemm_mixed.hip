// Reduced-precision GEMM on FP64 data for the mixed-precision polar bulk:
//   C = alpha op(A) op(B) + beta C with A, B, C in FP64 memory, the product
//   computed on FP32 MFMA (v_mfma_f32_16x16x4_f32) or BF16 MFMA
//   (v_mfma_f32_16x16x32_bf16), both accumulating in FP32.  Operands are
//   rounded to the compute type when they are staged into LDS; alpha, beta and
//   the accumulation with C happen in FP64.
//
// Result layout of the FP32-accumulating 16x16 MFMAs: lane l, register r holds
// row 4 (l / 16) + r, column l % 16.  As in the FP64 kernel the operands are
// swapped so that lane l holds ROW l % 16 of C (128-byte column stores).
#include <algorithm>

#include "mak_common.hpp"
#include "mak_kernels.hpp"

namespace mak {

template <int OPA>
__device__ inline double mload_a(const double* A, int64_t lda, int gi, int gk) {
  return (OPA == OP_N) ? A[gi + (int64_t)gk * lda] : A[gk + (int64_t)gi * lda];
}
template <int OPB>
__device__ inline double mload_b(const double* B, int64_t ldb, int gk, int gj) {
  return (OPB == OP_N) ? B[gk + (int64_t)gj * ldb] : B[gj + (int64_t)gk * ldb];
}

__device__ inline void mixed_tile_coords(int& tm, int& tn) {
  const int gx = gridDim.x, gy = gridDim.y;
  const int total = gx * gy;
  int L = blockIdx.x + blockIdx.y * gx;
  if ((total & 7) == 0) L = (L & 7) * (total >> 3) + (L >> 3);
  tm = L % gx;
  tn = L / gx;
}

// FP32 compute.  128 x 128 tile, BK = 16, 4 waves of 64 x 64.
template <int OPA, int OPB>
__global__ void __launch_bounds__(256, 2) gemm_mixed_f32_kernel(GemmArgs g) {
  constexpr int BM = 128, BN = 128, BK = 16, WTM = 64, WTN = 64, TM = 4, TN = 4;
  __shared__ float As[BK][BM + 4];
  __shared__ float Bs[BK][BN + 4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  int tmi, tni;
  mixed_tile_coords(tmi, tni);
  const int m0 = tmi * BM, n0 = tni * BN;
  const double* A = reinterpret_cast<const double*>(g.A);
  const double* B = reinterpret_cast<const double*>(g.B);
  double* C = reinterpret_cast<double*>(g.C);
  const int ia0 = (OPA == OP_N) ? (tid % BM) : (tid / BK);
  const int ka0 = (OPA == OP_N) ? (tid / BM) : (tid % BK);
  constexpr int dA = (OPA == OP_N) ? (256 / BM) : (256 / BK);
  const int jb0 = (OPB == OP_N) ? (tid / BK) : (tid % BN);
  const int kb0 = (OPB == OP_N) ? (tid % BK) : (tid / BN);
  constexpr int dB = (OPB == OP_N) ? (256 / BK) : (256 / BN);
  float pa[8], pb[8];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = (OPA == OP_N) ? ia0 : ia0 + u * dA, k = (OPA == OP_N) ? ka0 + u * dA : ka0;
      const int gi = m0 + i, gk = k0 + k;
      pa[u] = (gi < g.M && gk < g.K) ? (float)mload_a<OPA>(A, g.lda, gi, gk) : 0.f;
      const int j = (OPB == OP_N) ? jb0 + u * dB : jb0, kk = (OPB == OP_N) ? kb0 : kb0 + u * dB;
      const int gj = n0 + j, gk2 = k0 + kk;
      pb[u] = (gj < g.N && gk2 < g.K) ? (float)mload_b<OPB>(B, g.ldb, gk2, gj) : 0.f;
    }
  };
  v4f acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  if (g.K > 0) fetch(0);
  for (int k0 = 0; k0 < g.K; k0 += BK) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = (OPA == OP_N) ? ia0 : ia0 + u * dA, k = (OPA == OP_N) ? ka0 + u * dA : ka0;
      As[k][i] = pa[u];
      const int j = (OPB == OP_N) ? jb0 + u * dB : jb0, kk = (OPB == OP_N) ? kb0 : kb0 + u * dB;
      Bs[kk][j] = pb[u];
    }
    __syncthreads();
    if (k0 + BK < g.K) fetch(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = As[kk + (lane >> 4)][wm * WTM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = Bs[kk + (lane >> 4)][wn * WTN + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = mfma_f32(bf[j], af[i], acc[i][j]);
    }
    __syncthreads();
  }
  const int rowb = m0 + wm * WTM + (lane & 15), colb = n0 + wn * WTN;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rowb + i * 16, col = colb + j * 16 + mfma_f32_row(lane, r);
        if (row < g.M && col < g.N) {
          double* cp = C + row + (int64_t)col * g.ldc;
          const double v = g.alpha_re * (double)acc[i][j][r];
          *cp = (g.beta_re != 0.0) ? fma(g.beta_re, *cp, v) : v;
        }
      }
}

// BF16 compute.  128 x 128 tile, BK = 32 (one MFMA k-step), 4 waves of 64 x 64.
// LDS holds rows of op(A) and columns of op(B) with k contiguous so a lane's
// 8 consecutive k values are one 16-byte read.
template <int OPA, int OPB>
__global__ void __launch_bounds__(256, 2) gemm_mixed_bf16_kernel(GemmArgs g) {
  constexpr int BM = 128, BN = 128, BK = 32, WTM = 64, WTN = 64, TM = 4, TN = 4;
  __shared__ __attribute__((aligned(16))) __bf16 As[BM][BK + 8];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[BN][BK + 8];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  int tmi, tni;
  mixed_tile_coords(tmi, tni);
  const int m0 = tmi * BM, n0 = tni * BN;
  const double* A = reinterpret_cast<const double*>(g.A);
  const double* B = reinterpret_cast<const double*>(g.B);
  double* C = reinterpret_cast<double*>(g.C);
  // 128 x 32 tile = 4096 elements = 16 per thread
  const int ia0 = (OPA == OP_N) ? (tid % BM) : (tid / BK);
  const int ka0 = (OPA == OP_N) ? (tid / BM) : (tid % BK);
  constexpr int dA = (OPA == OP_N) ? (256 / BM) : (256 / BK);
  const int jb0 = (OPB == OP_N) ? (tid / BK) : (tid % BN);
  const int kb0 = (OPB == OP_N) ? (tid % BK) : (tid / BN);
  constexpr int dB = (OPB == OP_N) ? (256 / BK) : (256 / BN);
  float pa[16], pb[16];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int i = (OPA == OP_N) ? ia0 : ia0 + u * dA, k = (OPA == OP_N) ? ka0 + u * dA : ka0;
      const int gi = m0 + i, gk = k0 + k;
      pa[u] = (gi < g.M && gk < g.K) ? (float)mload_a<OPA>(A, g.lda, gi, gk) : 0.f;
      const int j = (OPB == OP_N) ? jb0 + u * dB : jb0, kk = (OPB == OP_N) ? kb0 : kb0 + u * dB;
      const int gj = n0 + j, gk2 = k0 + kk;
      pb[u] = (gj < g.N && gk2 < g.K) ? (float)mload_b<OPB>(B, g.ldb, gk2, gj) : 0.f;
    }
  };
  v4f acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  if (g.K > 0) fetch(0);
  for (int k0 = 0; k0 < g.K; k0 += BK) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int i = (OPA == OP_N) ? ia0 : ia0 + u * dA, k = (OPA == OP_N) ? ka0 + u * dA : ka0;
      As[i][k] = (__bf16)pa[u];
      const int j = (OPB == OP_N) ? jb0 + u * dB : jb0, kk = (OPB == OP_N) ? kb0 : kb0 + u * dB;
      Bs[j][kk] = (__bf16)pb[u];
    }
    __syncthreads();
    if (k0 + BK < g.K) fetch(k0 + BK);
    v8bf af[TM], bf[TN];
    const int kg = 8 * (lane >> 4);
#pragma unroll
    for (int i = 0; i < TM; ++i) af[i] = *reinterpret_cast<const v8bf*>(&As[wm * WTM + i * 16 + (lane & 15)][kg]);
#pragma unroll
    for (int j = 0; j < TN; ++j) bf[j] = *reinterpret_cast<const v8bf*>(&Bs[wn * WTN + j * 16 + (lane & 15)][kg]);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = mfma_bf16(bf[j], af[i], acc[i][j]);
    __syncthreads();
  }
  const int rowb = m0 + wm * WTM + (lane & 15), colb = n0 + wn * WTN;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rowb + i * 16, col = colb + j * 16 + mfma_f32_row(lane, r);
        if (row < g.M && col < g.N) {
          double* cp = C + row + (int64_t)col * g.ldc;
          const double v = g.alpha_re * (double)acc[i][j][r];
          *cp = (g.beta_re != 0.0) ? fma(g.beta_re, *cp, v) : v;
        }
      }
}

template <int OPA, int OPB> static void launch_mixed(int prec, const GemmArgs& g, hipStream_t s) {
  dim3 grid(cdiv(g.M, 128), cdiv(g.N, 128), 1);
  if (prec == 16)
    hipLaunchKernelGGL((gemm_mixed_bf16_kernel<OPA, OPB>), grid, dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL((gemm_mixed_f32_kernel<OPA, OPB>), grid, dim3(256), 0, s, g);
}

// Precision-dispatched GEMM: 64 -> FP64 MFMA, 32 / 16 -> FP32 / BF16 MFMA
// (real data only; complex data always runs in FP64).
template <typename T> void gemm_prec(int prec, int opA, int opB, GemmArgs g, hipStream_t s) {
  if constexpr (ScalarTraits<T>::is_complex) {
    gemm<T>(opA, opB, g, s);
  } else {
    if (prec == 64 || g.M <= 0 || g.N <= 0) { gemm<T>(opA, opB, g, s); return; }
    const int a = (opA == OP_N) ? 0 : 1, b = (opB == OP_N) ? 0 : 1;
    switch (a * 2 + b) {
      case 0: launch_mixed<OP_N, OP_N>(prec, g, s); break;
      case 1: launch_mixed<OP_N, OP_T>(prec, g, s); break;
      case 2: launch_mixed<OP_T, OP_N>(prec, g, s); break;
      default: launch_mixed<OP_T, OP_T>(prec, g, s); break;
    }
  }
}
template void gemm_prec<double>(int, int, int, GemmArgs, hipStream_t);
template void gemm_prec<zd>(int, int, int, GemmArgs, hipStream_t);

}  // namespace mak
