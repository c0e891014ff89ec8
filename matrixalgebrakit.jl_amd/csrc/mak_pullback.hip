// Reverse-mode rules of qr_compact! and left_polar! on MI355X, for a batch of
// equally sized matrices (strided storage, every product a batched MFMA GEMM).
//
// qr_pullback! (MatrixAlgebraKit src/pullbacks/qr.jl:73-120), rank p:
//   Q is re-formed from the compact WY factors kept by the primal
//   (V strictly below the diagonal of A_wy, one 32x32 T per panel), with the
//   positive gauge of R;
//   dQ1 -= Q1 (dR12 R12^H), dA2 += Q1 dR12                    (n > p)
//   M = triu(dR11) R11^H - Q1^H dQ1 ; M <- copyltu(M) (real diagonal)
//   dA1 += (dQ1 + Q1 M) R11^{-H}
// left_polar_pullback! (src/pullbacks/polar.jl:7-29):
//   dP <- herm(dP) ; M = W^H dW - dP P ; solve P C + C P = M - M^H
//   dA += W (C + dP) + (dW - W W^H dW) P^{-1}
// The Sylvester equation (LinearAlgebra.sylvester in the reference,
// src/common/pullbacks.jl:15) is solved by preconditioned conjugate gradients
// on the SPD operator C -> P C + C P (diagonal preconditioner
// 1 / (P_ii + P_jj)); P^{-1} comes from a batched Cholesky factorisation.
// CG needs about sqrt(cond(P)) steps: when it has not converged after
// kPcgFirst steps, the batch is solved exactly through the eigendecomposition
// P = Z diag(lambda) Z^H (rocSOLVER syevd / heevd, batched):
//   C = Z [(Z^H F Z)_ij / (lambda_i + lambda_j)] Z^H,  P^{-1} = Z diag(1/lambda) Z^H,
// whose accuracy does not depend on cond(P).
#include <algorithm>
#include <cmath>
#include <vector>

#include "mak_common.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

bool eigh_available();
size_t eigh_workspace_bytes(int n, int batch);
template <typename T>
int eigh_batched(hipStream_t s, int n, T* A, int64_t lda, int64_t strideA, double* D, int64_t strideD, int batch,
                 char* ws);

namespace {

constexpr int NB = 32;
constexpr int DBK = 64;
constexpr int kPcgFirst = 48;  // CG steps before the eigendecomposition takes over

inline size_t al256(size_t b) { return ((b + 255) / 256) * 256; }

template <typename T> GemmArgs bga(int batch, int M, int N, int K, const T* A, int64_t lda, int64_t sA, const T* B,
                                   int64_t ldb, int64_t sB, T* C, int64_t ldc, int64_t sC, double alpha,
                                   double beta) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.strideA = sA;
  g.B = B; g.ldb = ldb; g.strideB = sB;
  g.C = C; g.ldc = ldc; g.strideC = sC;
  g.alpha_re = alpha; g.beta_re = beta;
  g.ksplit = 1; g.kchunk = K; g.batch = batch;
  return g;
}

inline dim3 egrid(int64_t per, int batch) { return dim3((unsigned)cdiv64(std::max<int64_t>(per, 1), 256), 1, batch); }

// ------------------------------------------------------------------ helpers
template <typename T>
__global__ void bcopy_kernel(int m, int n, const T* X, int64_t ldx, int64_t sx, T* Y, int64_t ldy, int64_t sy) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  Y[blockIdx.z * sy + r + c * ldy] = X[blockIdx.z * sx + r + c * ldx];
}
template <typename T>
__global__ void bfill_kernel(int m, int n, T* Y, int64_t ldy, int64_t sy, int ident) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  Y[blockIdx.z * sy + r + c * ldy] = (ident && r == c) ? s_one(T()) : s_zero(T());
}
// Y = triu(X) (p x n block)
template <typename T>
__global__ void btriu_kernel(int m, int n, const T* X, int64_t ldx, int64_t sx, T* Y, int64_t ldy, int64_t sy) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  Y[blockIdx.z * sy + r + c * ldy] = (r <= c) ? X[blockIdx.z * sx + r + c * ldx] : s_zero(T());
}
// M <- copyltu(M): lower part <- conj(upper part), real diagonal
template <typename T> __global__ void bcopyltu_kernel(int n, T* M, int64_t ldm, int64_t sm) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  T* Mb = M + blockIdx.z * sm;
  if (r > c) Mb[r + c * ldm] = s_conj(Mb[c + r * ldm]);
  else if (r == c) Mb[r + c * ldm] = s_from_real<T>(s_re(Mb[r + c * ldm]));
}
// Y = (X + X^H) / 2
template <typename T>
__global__ void bherm_kernel(int n, const T* X, int64_t ldx, int64_t sx, T* Y, int64_t ldy, int64_t sy) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  const T* Xb = X + blockIdx.z * sx;
  Y[blockIdx.z * sy + r + c * ldy] = s_scale(s_add(Xb[r + c * ldx], s_conj(Xb[c + r * ldx])), 0.5);
}
// F = M - M^H
template <typename T>
__global__ void bskew2_kernel(int n, const T* M, int64_t ldm, int64_t sm, T* F, int64_t ldf, int64_t sf) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  const T* Mb = M + blockIdx.z * sm;
  F[blockIdx.z * sf + r + c * ldf] = s_sub(Mb[r + c * ldm], s_conj(Mb[c + r * ldm]));
}
// Y += a X
template <typename T>
__global__ void baxpy_kernel(int m, int n, double a, const T* X, int64_t ldx, int64_t sx, T* Y, int64_t ldy,
                             int64_t sy) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  T* y = Y + blockIdx.z * sy + r + c * ldy;
  *y = s_add(*y, s_scale(X[blockIdx.z * sx + r + c * ldx], a));
}
// Q[:, j] *= sign_safe(Awy[j, j])  (positive gauge of the re-formed Q)
template <typename T>
__global__ void bgauge_kernel(int m, int k, const T* A, int64_t lda, int64_t sa, T* Q, int64_t ldq, int64_t sq) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)m * k) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  T* q = Q + blockIdx.z * sq + r + c * ldq;
  *q = s_mul(*q, sign_safe(A[blockIdx.z * sa + c + c * lda]));
}

// Inverse of a kb x kb (kb <= 64) upper-triangular diagonal block of U into Y
// (rank-1 steps in LDS).  Block b of the grid's z dimension handles matrix b;
// the grid's x dimension enumerates the diagonal blocks.
template <typename T>
__global__ void __launch_bounds__(256) btrtri_diag_kernel(int n, const T* U, int64_t ldu, int64_t su, T* Y,
                                                          int64_t ldy, int64_t sy) {
  __shared__ T Ul[DBK][DBK + 1];
  __shared__ T S[DBK][DBK + 1];
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * DBK;
  const int kb = min(DBK, n - k0);
  const T* Ub = U + blockIdx.z * su + k0 + (int64_t)k0 * ldu;
  T* Yb = Y + blockIdx.z * sy + k0 + (int64_t)k0 * ldy;
  for (int i = tid; i < DBK * DBK; i += 256) {
    const int r = i % DBK, c = i / DBK;
    Ul[r][c] = (r <= c && c < kb) ? Ub[r + c * ldu] : s_zero(T());
  }
  __syncthreads();
  for (int i = tid; i < DBK * DBK; i += 256) {
    const int r = i % DBK, c = i / DBK;
    S[r][c] = (r == c && r < kb) ? s_div(s_one(T()), Ul[r][r]) : s_zero(T());
  }
  __syncthreads();
  for (int jj = kb - 1; jj > 0; --jj) {
    for (int i = tid; i < DBK * DBK; i += 256) {
      const int r = i % DBK, c = i / DBK;
      if (r < jj && c >= jj && c < kb) S[r][c] = s_sub(S[r][c], s_mul(s_div(Ul[r][jj], Ul[r][r]), S[jj][c]));
    }
    __syncthreads();
  }
  for (int i = tid; i < DBK * DBK; i += 256) {
    const int r = i % DBK, c = i / DBK;
    if (r < kb && c < kb) Yb[r + c * ldy] = (r <= c) ? S[r][c] : s_zero(T());
  }
}

// X (m x kb) <- alpha X U for a kb x kb upper-triangular U (in place, batched)
template <typename T>
__global__ void __launch_bounds__(256) btrmm_kernel(T* X, int64_t ldx, int64_t sx, int m, int kb, const T* U,
                                                    int64_t ldu, int64_t su, double alpha) {
  __shared__ T Us[DBK][DBK + 1];
  __shared__ T Xs[64][DBK + 1];
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x * 64;
  T* Xb = X + blockIdx.z * sx;
  const T* Ub = U + blockIdx.z * su;
  for (int i = tid; i < DBK * DBK; i += 256) {
    const int r = i % DBK, c = i / DBK;
    Us[r][c] = (r <= c && c < kb) ? Ub[r + c * ldu] : s_zero(T());
  }
  for (int i = tid; i < 64 * DBK; i += 256) {
    const int r = i % 64, c = i / 64;
    Xs[r][c] = (r0 + r < m && c < kb) ? Xb[(r0 + r) + c * ldx] : s_zero(T());
  }
  __syncthreads();
  for (int i = tid; i < 64 * DBK; i += 256) {
    const int r = i % 64, c = i / 64;
    if (r0 + r < m && c < kb) {
      T s = s_zero(T());
      for (int k = 0; k <= c; ++k) s = s_fma(Xs[r][k], Us[k][c], s);
      Xb[(r0 + r) + c * ldx] = s_scale(s, alpha);
    }
  }
}

// Y = U^{-1} for batched n x n upper-triangular U (Y fully overwritten).
template <typename T>
void btrtri_upper(int batch, int n, const T* U, int64_t ldu, int64_t su, T* Y, int64_t ldy, int64_t sy,
                  hipStream_t s) {
  hipLaunchKernelGGL((bfill_kernel<T>), egrid((int64_t)n * n, batch), dim3(256), 0, s, n, n, Y, ldy, sy, 0);
  hipLaunchKernelGGL((btrtri_diag_kernel<T>), dim3(cdiv(n, DBK), 1, batch), dim3(256), 0, s, n, U, ldu, su, Y, ldy,
                     sy);
  for (int k = DBK; k < n; k += DBK) {
    const int kb = std::min(DBK, n - k);
    T* Yc = Y + (int64_t)k * ldy;
    gemm<T>(OP_N, OP_N, bga<T>(batch, k, kb, k, Y, ldy, sy, U + (int64_t)k * ldu, ldu, su, Yc, ldy, sy, 1.0, 0.0), s);
    hipLaunchKernelGGL((btrmm_kernel<T>), dim3(cdiv(k, 64), 1, batch), dim3(256), 0, s, Yc, ldy, sy, k, kb,
                       Y + k + (int64_t)k * ldy, ldy, sy, -1.0);
  }
}

// Batched Cholesky factor kernel for one 64x64 diagonal block (upper, Z = W^H W)
template <typename T>
__global__ void __launch_bounds__(256) bpotrf_diag_kernel(T* Z, int64_t ldz, int64_t sz, int kb, int* info) {
  __shared__ T S[DBK][DBK + 1];
  __shared__ T Wl[DBK][DBK + 1];
  const int tid = threadIdx.x;
  T* Zb = Z + blockIdx.z * sz;
  for (int i = tid; i < DBK * DBK; i += 256) {
    const int r = i % DBK, c = i / DBK;
    S[r][c] = (r <= c && c < kb) ? Zb[r + c * ldz] : s_zero(T());
    Wl[r][c] = s_zero(T());
  }
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    const double d2 = s_re(S[j][j]);
    if (!(d2 > 0.0)) {
      if (tid == 0) info[0] = 1;
      break;
    }
    const double d = sqrt(d2), rd = 1.0 / d, rd2 = 1.0 / d2;
    for (int c = j + tid; c < kb; c += 256) Wl[j][c] = (c == j) ? s_from_real<T>(d) : s_scale(S[j][c], rd);
    for (int i = tid; i < DBK * DBK; i += 256) {
      const int r = i % DBK, c = i / DBK;
      if (r > j && r <= c && c < kb) S[r][c] = s_sub(S[r][c], s_scale(s_mul(s_conj(S[j][r]), S[j][c]), rd2));
    }
    __syncthreads();
  }
  __syncthreads();
  for (int i = tid; i < DBK * DBK; i += 256) {
    const int r = i % DBK, c = i / DBK;
    if (r < kb && c < kb) Zb[r + c * ldz] = (r <= c) ? Wl[r][c] : s_zero(T());
  }
}

// Batched Cholesky Z = W^H W (W upper, in place; strict lower part zeroed on the
// diagonal blocks only -- callers read the upper part).  tmp: DBK x n per matrix.
template <typename T>
void bpotrf_upper(int batch, int n, T* Z, int64_t ldz, int64_t sz, T* Ydiag, int64_t ldy, int64_t sy, T* tmp,
                  int64_t stmp, int* dinfo, hipStream_t s) {
  for (int k = 0; k < n; k += DBK) {
    const int kb = std::min(DBK, n - k);
    T* Zkk = Z + k + (int64_t)k * ldz;
    hipLaunchKernelGGL((bpotrf_diag_kernel<T>), dim3(1, 1, batch), dim3(256), 0, s, Zkk, ldz, sz, kb, dinfo);
    const int rest = n - k - kb;
    if (rest > 0) {
      // Y_kk = W_kk^{-1}
      hipLaunchKernelGGL((btrtri_diag_kernel<T>), dim3(1, 1, batch), dim3(256), 0, s, kb, Zkk, ldz, sz,
                         Ydiag, ldy, sy);
      T* Zrow = Z + k + (int64_t)(k + kb) * ldz;
      gemm<T>(OP_C, OP_N, bga<T>(batch, kb, rest, kb, Ydiag, ldy, sy, Zrow, ldz, sz, tmp, DBK, stmp, 1.0, 0.0), s);
      hipLaunchKernelGGL((bcopy_kernel<T>), egrid((int64_t)kb * rest, batch), dim3(256), 0, s, kb, rest, tmp,
                         (int64_t)DBK, stmp, Zrow, ldz, sz);
      T* Z22 = Z + (k + kb) + (int64_t)(k + kb) * ldz;
      gemm<T>(OP_C, OP_N, bga<T>(batch, rest, rest, kb, Zrow, ldz, sz, Zrow, ldz, sz, Z22, ldz, sz, -1.0, 1.0), s);
    }
  }
}

// Batched Q re-formation from compact WY (qr_compact shape, Q m x k).
template <typename T>
void borgqr(int batch, int m, int n, const T* A, int64_t lda, int64_t sa, const T* Tp, int64_t st, T* Q, int64_t ldq,
            int64_t sq, T* Wb, T* W2b, int positive, hipStream_t s) {
  const int k = std::min(m, n);
  hipLaunchKernelGGL((bfill_kernel<T>), egrid((int64_t)m * k, batch), dim3(256), 0, s, m, k, Q, ldq, sq, 1);
  const int npan = cdiv(k, NB);
  const int64_t sw = (int64_t)NB * k;
  for (int p = npan - 1; p >= 0; --p) {
    const int j0 = p * NB, nbc = std::min(NB, k - j0), rows = m - j0, ncols = k - j0;
    const T* Vp = A + j0 + (int64_t)j0 * lda;
    T* C = Q + j0 + (int64_t)j0 * ldq;
    // W = V^H C ; W2 = T W ; C -= V W2
    GemmArgs g1 = bga<T>(batch, nbc, ncols, rows, Vp, lda, sa, C, ldq, sq, Wb, nbc, sw, 1.0, 0.0);
    g1.unitLowerA = 1;
    gemm<T>(OP_C, OP_N, g1, s);
    gemm<T>(OP_N, OP_N, bga<T>(batch, nbc, ncols, nbc, Tp + (int64_t)p * NB * NB, NB, st, Wb, nbc, sw, W2b, nbc, sw,
                               1.0, 0.0), s);
    GemmArgs g3 = bga<T>(batch, rows, ncols, nbc, Vp, lda, sa, W2b, nbc, sw, C, ldq, sq, -1.0, 1.0);
    g3.unitLowerA = 1;
    gemm<T>(OP_N, OP_N, g3, s);
  }
  if (positive)
    hipLaunchKernelGGL((bgauge_kernel<T>), egrid((int64_t)m * k, batch), dim3(256), 0, s, m, k, A, lda, sa, Q, ldq,
                       sq);
}

// ---- batched reductions for CG: out[b] = Re sum conj(X_b) Y_b
constexpr int kRedBlocks = 64;
template <typename T>
__global__ void __launch_bounds__(256) bdot_kernel(int m, int n, const T* X, int64_t ldx, int64_t sx, const T* Y,
                                                   int64_t ldy, int64_t sy, double* partial) {
  __shared__ double red[4];
  const int64_t tot = (int64_t)m * n;
  const T* Xb = X + blockIdx.z * sx;
  const T* Yb = Y + blockIdx.z * sy;
  double acc = 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * 256) {
    const int r = (int)(idx % m), c = (int)(idx / m);
    acc += s_re(s_cfma(Xb[r + c * ldx], Yb[r + c * ldy], s_zero(T())));
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.z * kRedBlocks + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ void bsum_kernel(const double* partial, double* out) {
  const int b = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < kRedBlocks; i += 64) acc += partial[b * kRedBlocks + i];
  acc = wave_sum(acc);
  if (threadIdx.x == 0) out[b] = acc;
}
template <typename T>
void bdot(int batch, int m, int n, const T* X, int64_t ldx, int64_t sx, const T* Y, int64_t ldy, int64_t sy,
          double* partial, double* out, hipStream_t s) {
  hipLaunchKernelGGL((bdot_kernel<T>), dim3(kRedBlocks, 1, batch), dim3(256), 0, s, m, n, X, ldx, sx, Y, ldy, sy,
                     partial);
  hipLaunchKernelGGL(bsum_kernel, dim3(batch), dim3(64), 0, s, partial, out);
}
// Z = R ./ (P_ii + P_jj)
template <typename T>
__global__ void bprecond_kernel(int n, const T* R, const T* P, int64_t ldp, int64_t sp, T* Z) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  const T* Pb = P + blockIdx.z * sp;
  const double d = s_re(Pb[r + r * ldp]) + s_re(Pb[c + c * ldp]);
  const int64_t o = blockIdx.z * (int64_t)n * n + r + (int64_t)c * n;
  Z[o] = s_scale(R[o], 1.0 / d);
}
// X += a D ; R -= a Q   with a = rz[b] / dq[b]
template <typename T>
__global__ void bcg_step_kernel(int nn, const double* rz, const double* dq, const T* D, const T* Qm, T* X, T* R) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= nn) return;
  const int b = blockIdx.z;
  const double a = (dq[b] != 0.0) ? rz[b] / dq[b] : 0.0;
  const int64_t o = (int64_t)b * nn + idx;
  X[o] = s_add(X[o], s_scale(D[o], a));
  R[o] = s_sub(R[o], s_scale(Qm[o], a));
}
// D = Z + (rz_new[b] / rz[b]) D
template <typename T>
__global__ void bcg_dir_kernel(int nn, const double* rznew, const double* rz, const T* Z, T* D) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= nn) return;
  const int b = blockIdx.z;
  const double beta = (rz[b] != 0.0) ? rznew[b] / rz[b] : 0.0;
  const int64_t o = (int64_t)b * nn + idx;
  D[o] = s_add(Z[o], s_scale(D[o], beta));
}

}  // namespace

// ---------------------------------------------------------------------------
// workspace layouts
// ---------------------------------------------------------------------------
template <typename T> size_t qr_pullback_ws(int batch, int m, int n) {
  const int k = std::min(m, n);
  size_t b = 0;
  b += al256((size_t)batch * m * k * sizeof(T));      // Q
  b += al256((size_t)batch * m * k * sizeof(T));      // X = dQ1 (+ Q M)
  b += al256((size_t)batch * k * k * sizeof(T)) * 3;  // M, Rinv, triu(dR11)
  b += al256((size_t)batch * NB * k * sizeof(T)) * 2; // WY scratch
  b += al256((size_t)batch * k * k * sizeof(T));      // triu(R11)
  return b;
}
// out[b] = max(out[b], max |X_ij|) over a rows x cols block of matrix b
// (upper = 1: only i <= j), by a grid-stride loop per batch member and an
// integer atomic max on the bit patterns (monotone for non-negative doubles).
template <typename T>
__global__ void bmaxabs_kernel(int rows, int cols, const T* X, int64_t ldx, int64_t sx, int upper, double* out) {
  __shared__ double red[256];
  const T* x = X + (int64_t)blockIdx.z * sx;
  const int64_t tot = (int64_t)rows * cols;
  double v = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) {
    const int r = (int)(e % rows), c = (int)(e / rows);
    if (!upper || r <= c) v = fmax(v, sqrt(s_abs2(x[r + (int64_t)c * ldx])));
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    atomicMax(reinterpret_cast<unsigned long long*>(out + blockIdx.z), (unsigned long long)__double_as_longlong(red[0]));
}
__global__ void bzero_doubles_kernel(double* p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
template <typename T>
void bmaxabs(int batch, int rows, int cols, const T* X, int64_t ldx, int64_t sx, int upper, double* out,
             hipStream_t s) {
  if (batch <= 0 || rows <= 0 || cols <= 0 || X == nullptr) return;
  const int64_t tot = (int64_t)rows * cols;
  const unsigned gx = (unsigned)std::min<int64_t>(cdiv64(tot, 256), 64);
  hipLaunchKernelGGL((bmaxabs_kernel<T>), dim3(gx, 1, batch), dim3(256), 0, s, rows, cols, X, ldx, sx, upper, out);
}

template <typename T>
__global__ void bsylv_divide_kernel(int n, T* X, int64_t ldx, int64_t sx, const double* lam, int64_t sl) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  const double* l = lam + (int64_t)blockIdx.z * sl;
  T* x = X + (int64_t)blockIdx.z * sx + r + (int64_t)c * ldx;
  *x = s_scale(*x, 1.0 / (l[r] + l[c]));
}
template <typename T>
__global__ void bscale_cols_inv_kernel(int m, int n, T* X, int64_t ldx, int64_t sx, const double* lam, int64_t sl) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  T* x = X + (int64_t)blockIdx.z * sx + r + (int64_t)c * ldx;
  *x = s_scale(*x, 1.0 / lam[(int64_t)blockIdx.z * sl + c]);
}

template <typename T> size_t polar_pullback_ws(int batch, int m, int n) {
  size_t b = 0;
  const size_t nn = al256((size_t)batch * n * n * sizeof(T));
  b += al256((size_t)batch * n * sizeof(double)) + al256(eigh_workspace_bytes(n, std::max(batch, 1)));  // lambda
  b += nn * 10;                                        // dPh, M, F, X, R, Z, D, Qm, Pc, Pinv
  b += al256((size_t)batch * m * n * sizeof(T));      // Zw = dW - W W^H dW
  b += al256((size_t)batch * DBK * n * sizeof(T));    // Cholesky scratch
  b += al256((size_t)batch * DBK * DBK * sizeof(T) * cdiv(n, DBK) + 256);
  b += al256((size_t)batch * kRedBlocks * sizeof(double)) + al256((size_t)batch * 4 * sizeof(double));
  b += 256;
  return b;
}

template <typename T>
int qr_pullback(hipStream_t s, int batch, int m, int n, int p, int positive, const T* A, int64_t lda, int64_t sa,
                const T* Tp, int64_t st, const T* R, int64_t ldr, int64_t sr, const T* dQ, int64_t lddq, int64_t sdq,
                const T* dR, int64_t lddr, int64_t sdr, T* dA, int64_t ldda, int64_t sda, void* work, size_t wbytes,
                double* gauge_out) {
  if (batch < 0) return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  const int k = std::min(m, n);
  if (p < 0 || p > k) return -5;
  if (lda < std::max(1, m)) return -8;
  if (ldr < std::max(1, k)) return -13;
  if (dQ && lddq < std::max(1, m)) return -16;
  if (dR && lddr < std::max(1, k)) return -19;
  if (ldda < std::max(1, m)) return -22;
  if (work == nullptr && batch > 0 && p > 0) return -24;
  if (wbytes < qr_pullback_ws<T>(batch, m, n)) return -25;
  if (gauge_out && batch > 0) {
    // the gauge-dependent part of the cotangents (src/pullbacks/qr.jl:27-29,
    // :34-47): dQ[:, p:k] and the upper triangle of dR[p:k, p:n]
    hipLaunchKernelGGL(bzero_doubles_kernel, dim3(cdiv(batch, 256)), dim3(256), 0, s, gauge_out, batch);
    if (dQ) bmaxabs<T>(batch, m, k - p, dQ + (int64_t)p * lddq, lddq, sdq, 0, gauge_out, s);
    if (dR) bmaxabs<T>(batch, k - p, n - p, dR + p + (int64_t)p * lddr, lddr, sdr, 1, gauge_out, s);
  }
  if (batch == 0 || p == 0) return 0;
  char* w = reinterpret_cast<char*>(work);
  auto take = [&](size_t bytes) { char* q = w; w += al256(bytes); return reinterpret_cast<T*>(q); };
  const int64_t smk = (int64_t)m * k, skk = (int64_t)k * k;
  T* Q = take((size_t)batch * smk * sizeof(T));
  T* X = take((size_t)batch * smk * sizeof(T));
  T* M = take((size_t)batch * skk * sizeof(T));
  T* Rinv = take((size_t)batch * skk * sizeof(T));
  T* dR11u = take((size_t)batch * skk * sizeof(T));
  T* Wb = take((size_t)batch * NB * k * sizeof(T));
  T* W2b = take((size_t)batch * NB * k * sizeof(T));
  T* Rt = take((size_t)batch * skk * sizeof(T));
  // Q = orgqr(V, T) (times the gauge signs when the primal used positive = true)
  borgqr<T>(batch, m, n, A, lda, sa, Tp, st, Q, m, smk, Wb, W2b, positive, s);
  // Rt = triu(R11): only the upper triangle of R is read (UpperTriangular(R11), src/pullbacks/qr.jl:92)
  hipLaunchKernelGGL((btriu_kernel<T>), egrid((int64_t)p * p, batch), dim3(256), 0, s, p, p, R, ldr, sr, Rt,
                     (int64_t)k, skk);
  // X = dQ1
  if (dQ) hipLaunchKernelGGL((bcopy_kernel<T>), egrid((int64_t)m * p, batch), dim3(256), 0, s, m, p, dQ, lddq, sdq, X,
                             (int64_t)m, smk);
  else hipLaunchKernelGGL((bfill_kernel<T>), egrid((int64_t)m * p, batch), dim3(256), 0, s, m, p, X, (int64_t)m, smk, 0);
  if (dR && n > p) {
    // X -= Q1 (dR12 R12^H) ; dA2 += Q1 dR12
    gemm<T>(OP_N, OP_C, bga<T>(batch, p, p, n - p, dR + (int64_t)p * lddr, lddr, sdr, R + (int64_t)p * ldr, ldr, sr,
                               M, k, skk, 1.0, 0.0), s);
    gemm<T>(OP_N, OP_N, bga<T>(batch, m, p, p, Q, m, smk, M, k, skk, X, m, smk, -1.0, 1.0), s);
    gemm<T>(OP_N, OP_N, bga<T>(batch, m, n - p, p, Q, m, smk, dR + (int64_t)p * lddr, lddr, sdr,
                               dA + (int64_t)p * ldda, ldda, sda, 1.0, 1.0), s);
  }
  // M = triu(dR11) R11^H - Q1^H X
  if (dR) {
    hipLaunchKernelGGL((btriu_kernel<T>), egrid((int64_t)p * p, batch), dim3(256), 0, s, p, p, dR, lddr, sdr, dR11u,
                       (int64_t)k, skk);
    gemm<T>(OP_N, OP_C, bga<T>(batch, p, p, p, dR11u, k, skk, Rt, k, skk, M, k, skk, 1.0, 0.0), s);
    gemm<T>(OP_C, OP_N, bga<T>(batch, p, p, m, Q, m, smk, X, m, smk, M, k, skk, -1.0, 1.0), s);
  } else {
    gemm<T>(OP_C, OP_N, bga<T>(batch, p, p, m, Q, m, smk, X, m, smk, M, k, skk, -1.0, 0.0), s);
  }
  hipLaunchKernelGGL((bcopyltu_kernel<T>), egrid((int64_t)p * p, batch), dim3(256), 0, s, p, M, (int64_t)k, skk);
  // X += Q1 M
  gemm<T>(OP_N, OP_N, bga<T>(batch, m, p, p, Q, m, smk, M, k, skk, X, m, smk, 1.0, 1.0), s);
  // dA1 += X R11^{-H}
  btrtri_upper<T>(batch, p, Rt, k, skk, Rinv, k, skk, s);
  gemm<T>(OP_N, OP_C, bga<T>(batch, m, p, p, X, m, smk, Rinv, k, skk, dA, ldda, sda, 1.0, 1.0), s);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

template <typename T>
int polar_pullback(hipStream_t s, int batch, int m, int n, const T* W, int64_t ldw, int64_t sw, const T* P,
                   int64_t ldp, int64_t sp, const T* dW, int64_t lddw, int64_t sdw, const T* dP, int64_t lddp,
                   int64_t sdp, T* dA, int64_t ldda, int64_t sda, int maxiter, double tol, void* work, size_t wbytes,
                   int* iters_out) {
  if (batch < 0) return -2;
  if (m < n) return -3;
  if (n < 0) return -4;
  if (ldw < std::max(1, m)) return -6;
  if (ldp < std::max(1, n)) return -9;
  if (dW && lddw < std::max(1, m)) return -12;
  if (dP && lddp < std::max(1, n)) return -15;
  if (ldda < std::max(1, m)) return -18;
  if (work == nullptr && batch > 0 && n > 0) return -22;
  if (wbytes < polar_pullback_ws<T>(batch, m, n)) return -23;
  if (batch == 0 || n == 0) return 0;
  if (maxiter <= 0) maxiter = 200;
  if (!(tol > 0.0)) tol = 1e-14;
  char* w = reinterpret_cast<char*>(work);
  auto take = [&](size_t bytes) { char* q = w; w += al256(bytes); return q; };
  const int64_t snn = (int64_t)n * n, smn = (int64_t)m * n;
  const size_t nnb = (size_t)batch * snn * sizeof(T);
  double* lam = (double*)take((size_t)batch * n * sizeof(double));
  char* eigws = take(eigh_workspace_bytes(n, batch));
  T* dPh = (T*)take(nnb);
  T* M = (T*)take(nnb);
  T* F = (T*)take(nnb);
  T* X = (T*)take(nnb);
  T* R = (T*)take(nnb);
  T* Z = (T*)take(nnb);
  T* D = (T*)take(nnb);
  T* Qm = (T*)take(nnb);
  T* Pc = (T*)take(nnb);
  T* Pinv = (T*)take(nnb);
  T* Zw = (T*)take((size_t)batch * smn * sizeof(T));
  T* tmp = (T*)take((size_t)batch * DBK * n * sizeof(T));
  T* Yd = (T*)take((size_t)batch * DBK * DBK * sizeof(T) * cdiv(n, DBK) + 256);
  double* partial = (double*)take((size_t)batch * kRedBlocks * sizeof(double));
  double* sc = (double*)take((size_t)batch * 4 * sizeof(double));
  int* dinfo = (int*)take(256);
  double* rz = sc;
  double* rznew = sc + batch;
  double* dq = sc + 2 * batch;
  double* ff = sc + 3 * batch;
  const dim3 gnn = egrid(snn, batch);
  // dPh = herm(dP) ; M = W^H dW - dPh P
  if (dP) hipLaunchKernelGGL((bherm_kernel<T>), gnn, dim3(256), 0, s, n, dP, lddp, sdp, dPh, (int64_t)n, snn);
  if (dW) gemm<T>(OP_C, OP_N, bga<T>(batch, n, n, m, W, ldw, sw, dW, lddw, sdw, M, n, snn, 1.0, 0.0), s);
  else hipLaunchKernelGGL((bfill_kernel<T>), gnn, dim3(256), 0, s, n, n, M, (int64_t)n, snn, 0);
  if (dP) gemm<T>(OP_N, OP_N, bga<T>(batch, n, n, n, dPh, n, snn, P, ldp, sp, M, n, snn, -1.0, 1.0), s);
  // F = M - M^H ; solve P C + C P = F  (PCG, C in X)
  hipLaunchKernelGGL((bskew2_kernel<T>), gnn, dim3(256), 0, s, n, M, (int64_t)n, snn, F, (int64_t)n, snn);
  hipLaunchKernelGGL((bfill_kernel<T>), gnn, dim3(256), 0, s, n, n, X, (int64_t)n, snn, 0);
  hipLaunchKernelGGL((bcopy_kernel<T>), gnn, dim3(256), 0, s, n, n, F, (int64_t)n, snn, R, (int64_t)n, snn);
  hipLaunchKernelGGL((bprecond_kernel<T>), gnn, dim3(256), 0, s, n, R, P, ldp, sp, Z);
  hipLaunchKernelGGL((bcopy_kernel<T>), gnn, dim3(256), 0, s, n, n, Z, (int64_t)n, snn, D, (int64_t)n, snn);
  bdot<T>(batch, n, n, R, n, snn, Z, n, snn, partial, rz, s);
  bdot<T>(batch, n, n, F, n, snn, F, n, snn, partial, ff, s);
  std::vector<double> hff(batch), hrr(batch);
  hipMemcpyAsync(hff.data(), ff, batch * sizeof(double), hipMemcpyDeviceToHost, s);
  hipStreamSynchronize(s);
  int it = 0;
  bool done = true;
  for (int b = 0; b < batch; ++b) done = done && (hff[b] == 0.0);
  const bool can_eig = eigh_available();
  const int pcg_steps = can_eig ? std::min(maxiter, kPcgFirst) : maxiter;
  for (it = 0; it < pcg_steps && !done; ++it) {
    // Qm = P D + D P
    gemm<T>(OP_N, OP_N, bga<T>(batch, n, n, n, P, ldp, sp, D, n, snn, Qm, n, snn, 1.0, 0.0), s);
    gemm<T>(OP_N, OP_N, bga<T>(batch, n, n, n, D, n, snn, P, ldp, sp, Qm, n, snn, 1.0, 1.0), s);
    bdot<T>(batch, n, n, D, n, snn, Qm, n, snn, partial, dq, s);
    hipLaunchKernelGGL((bcg_step_kernel<T>), gnn, dim3(256), 0, s, (int)snn, rz, dq, D, Qm, X, R);
    hipLaunchKernelGGL((bprecond_kernel<T>), gnn, dim3(256), 0, s, n, R, P, ldp, sp, Z);
    bdot<T>(batch, n, n, R, n, snn, Z, n, snn, partial, rznew, s);
    hipLaunchKernelGGL((bcg_dir_kernel<T>), gnn, dim3(256), 0, s, (int)snn, rznew, rz, Z, D);
    std::swap(rz, rznew);
    if ((it + 1) % 4 == 0 || it + 1 == pcg_steps) {
      bdot<T>(batch, n, n, R, n, snn, R, n, snn, partial, dq, s);
      hipMemcpyAsync(hrr.data(), dq, batch * sizeof(double), hipMemcpyDeviceToHost, s);
      hipStreamSynchronize(s);
      done = true;
      for (int b = 0; b < batch; ++b) done = done && (hrr[b] <= tol * tol * hff[b]);
    }
  }
  if (iters_out) *iters_out = it;
  // CG has not converged: solve exactly through P = Z diag(lambda) Z^H (Z in Pc)
  bool eig = false;
  if (!done && can_eig) {
    hipLaunchKernelGGL((bcopy_kernel<T>), gnn, dim3(256), 0, s, n, n, P, ldp, sp, Pc, (int64_t)n, snn);
    const int est = eigh_batched<T>(s, n, Pc, n, snn, lam, n, batch, eigws);
    if (est != 0) return est;
    std::vector<double> hl((size_t)batch * n);
    hipMemcpyAsync(hl.data(), lam, hl.size() * sizeof(double), hipMemcpyDeviceToHost, s);
    hipStreamSynchronize(s);
    for (int b = 0; b < batch; ++b)
      if (!(hl[(size_t)b * n] > 0.0)) return 2;  // P not positive definite
    gemm<T>(OP_N, OP_N, bga<T>(batch, n, n, n, F, n, snn, Pc, n, snn, Qm, n, snn, 1.0, 0.0), s);   // F Z
    gemm<T>(OP_C, OP_N, bga<T>(batch, n, n, n, Pc, n, snn, Qm, n, snn, R, n, snn, 1.0, 0.0), s);   // Z^H F Z
    hipLaunchKernelGGL((bsylv_divide_kernel<T>), gnn, dim3(256), 0, s, n, R, (int64_t)n, snn, (const double*)lam,
                       (int64_t)n);
    gemm<T>(OP_N, OP_N, bga<T>(batch, n, n, n, Pc, n, snn, R, n, snn, Qm, n, snn, 1.0, 0.0), s);   // Z Y
    gemm<T>(OP_N, OP_C, bga<T>(batch, n, n, n, Qm, n, snn, Pc, n, snn, X, n, snn, 1.0, 0.0), s);   // Z Y Z^H
    done = true;
    eig = true;
  }
  // C = X + dPh ; dA += W C
  if (dP) hipLaunchKernelGGL((baxpy_kernel<T>), gnn, dim3(256), 0, s, n, n, 1.0, dPh, (int64_t)n, snn, X, (int64_t)n, snn);
  gemm<T>(OP_N, OP_N, bga<T>(batch, m, n, n, W, ldw, sw, X, n, snn, dA, ldda, sda, 1.0, 1.0), s);
  if (dW) {
    // Zw = dW - W (W^H dW) ; dA += Zw P^{-1}
    gemm<T>(OP_C, OP_N, bga<T>(batch, n, n, m, W, ldw, sw, dW, lddw, sdw, M, n, snn, 1.0, 0.0), s);
    hipLaunchKernelGGL((bcopy_kernel<T>), egrid(smn, batch), dim3(256), 0, s, m, n, dW, lddw, sdw, Zw, (int64_t)m, smn);
    gemm<T>(OP_N, OP_N, bga<T>(batch, m, n, n, W, ldw, sw, M, n, snn, Zw, m, smn, -1.0, 1.0), s);
    if (eig) {
      // P^{-1} = (Z diag(1/lambda)) Z^H
      hipLaunchKernelGGL((bcopy_kernel<T>), gnn, dim3(256), 0, s, n, n, Pc, (int64_t)n, snn, Qm, (int64_t)n, snn);
      hipLaunchKernelGGL((bscale_cols_inv_kernel<T>), gnn, dim3(256), 0, s, n, n, Qm, (int64_t)n, snn,
                         (const double*)lam, (int64_t)n);
      gemm<T>(OP_N, OP_C, bga<T>(batch, n, n, n, Qm, n, snn, Pc, n, snn, Pinv, n, snn, 1.0, 0.0), s);
      gemm<T>(OP_N, OP_N, bga<T>(batch, m, n, n, Zw, m, smn, Pinv, n, snn, dA, ldda, sda, 1.0, 1.0), s);
    } else {
      // P = Wp^H Wp ; P^{-1} = Wp^{-1} Wp^{-H}
      hipLaunchKernelGGL((bcopy_kernel<T>), gnn, dim3(256), 0, s, n, n, P, ldp, sp, Pc, (int64_t)n, snn);
      hipMemsetAsync(dinfo, 0, sizeof(int), s);
      bpotrf_upper<T>(batch, n, Pc, n, snn, Yd, DBK, (int64_t)DBK * DBK, tmp, (int64_t)DBK * n, dinfo, s);
      btrtri_upper<T>(batch, n, Pc, n, snn, Qm, n, snn, s);
      gemm<T>(OP_N, OP_C, bga<T>(batch, n, n, n, Qm, n, snn, Qm, n, snn, Pinv, n, snn, 1.0, 0.0), s);
      gemm<T>(OP_N, OP_N, bga<T>(batch, m, n, n, Zw, m, smn, Pinv, n, snn, dA, ldda, sda, 1.0, 1.0), s);
      int hinfo = 0;
      hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, s);
      hipStreamSynchronize(s);
      if (hinfo != 0) return 2;
    }
  }
  if (hipGetLastError() != hipSuccess) return 1;
  return done ? 0 : 6;
}

size_t qr_full_cotangent_ws_bytes(size_t elem, int batch, int m, int n) {
  const int k = std::min(m, n), r = m - k;
  return al256((size_t)batch * k * std::max(r, 0) * elem) + al256((size_t)batch * m * std::max(r, 0) * elem) + 256;
}
template <typename T> size_t qr_full_cotangent_ws(int batch, int m, int n) {
  return qr_full_cotangent_ws_bytes(sizeof(T), batch, m, n);
}

// check_and_prepare_qr_cotangents for a qr_full! cotangent (src/pullbacks/
// qr.jl:4-53): dQ (m x m, as the full Q) -> dQ1 (m x k, the compact
// cotangent the pullback takes) and gauge_out[b] (device) = max |.| of the
// gauge-dependent part: for p = k, dQ3 - Q1 Q1^H dQ3 (dQ3 = dQ[:, k:m]), with
// dQ1 = dQ[:, :k] - Q3 (Q1^H dQ3)^H; for p < k, dQ[:, p:m], with dQ1 =
// dQ[:, :k].  Q: the full Q (m x m).
template <typename T>
int qr_full_cotangent(hipStream_t s, int batch, int m, int n, int p, const T* Q, int64_t ldq, int64_t sq, const T* dQ,
                      int64_t lddq, int64_t sdq, T* dQ1, int64_t ldd1, int64_t sd1, double* gauge_out, void* work,
                      size_t wbytes) {
  if (batch < 0) return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  const int k = std::min(m, n), r = m - k;
  if (p < 0 || p > k) return -5;
  if (ldq < std::max(1, m)) return -7;
  if (lddq < std::max(1, m)) return -10;
  if (ldd1 < std::max(1, m)) return -13;
  if (wbytes < qr_full_cotangent_ws<T>(batch, m, n) || (work == nullptr && batch > 0 && r > 0)) return -17;
  if (batch == 0 || m == 0) return 0;
  if (Q == nullptr || dQ == nullptr || dQ1 == nullptr) return -6;
  const int64_t smk = (int64_t)m * k;
  hipLaunchKernelGGL((bcopy_kernel<T>), egrid(smk, batch), dim3(256), 0, s, m, k, dQ, lddq, sdq, dQ1, ldd1, sd1);
  if (gauge_out) hipLaunchKernelGGL(bzero_doubles_kernel, dim3(cdiv(batch, 256)), dim3(256), 0, s, gauge_out, batch);
  if (r == 0) return hipGetLastError() == hipSuccess ? 0 : 1;
  if (p < k) {
    if (gauge_out) bmaxabs<T>(batch, m, m - p, dQ + (int64_t)p * lddq, lddq, sdq, 0, gauge_out, s);
    return hipGetLastError() == hipSuccess ? 0 : 1;
  }
  char* w = reinterpret_cast<char*>(work);
  T* X = reinterpret_cast<T*>(w);                                                  // Q1^H dQ3 (k x r)
  T* T3 = reinterpret_cast<T*>(w + al256((size_t)batch * k * r * sizeof(T)));     // dQ3 - Q1 X (m x r)
  const int64_t skr = (int64_t)k * r, smr = (int64_t)m * r;
  const T* dQ3 = dQ + (int64_t)k * lddq;
  const T* Q3 = Q + (int64_t)k * ldq;
  gemm<T>(OP_C, OP_N, bga<T>(batch, k, r, m, Q, ldq, sq, dQ3, lddq, sdq, X, k, skr, 1.0, 0.0), s);
  if (gauge_out) {
    hipLaunchKernelGGL((bcopy_kernel<T>), egrid(smr, batch), dim3(256), 0, s, m, r, dQ3, lddq, sdq, T3, (int64_t)m,
                       smr);
    gemm<T>(OP_N, OP_N, bga<T>(batch, m, r, k, Q, ldq, sq, X, k, skr, T3, m, smr, -1.0, 1.0), s);
    bmaxabs<T>(batch, m, r, T3, m, smr, 0, gauge_out, s);
  }
  gemm<T>(OP_N, OP_C, bga<T>(batch, m, k, r, Q3, ldq, sq, X, k, skr, dQ1, ldd1, sd1, -1.0, 1.0), s);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

#define MAK_INST_PB(T)                                                                                            \
  template size_t qr_pullback_ws<T>(int, int, int);                                                               \
  template size_t polar_pullback_ws<T>(int, int, int);                                                            \
  template int qr_pullback<T>(hipStream_t, int, int, int, int, int, const T*, int64_t, int64_t, const T*, int64_t,\
                              const T*, int64_t, int64_t, const T*, int64_t, int64_t, const T*, int64_t, int64_t, \
                              T*, int64_t, int64_t, void*, size_t, double*);                                      \
  template int qr_full_cotangent<T>(hipStream_t, int, int, int, int, const T*, int64_t, int64_t, const T*,        \
                                    int64_t, int64_t, T*, int64_t, int64_t, double*, void*, size_t);              \
  template int polar_pullback<T>(hipStream_t, int, int, int, const T*, int64_t, int64_t, const T*, int64_t,       \
                                 int64_t, const T*, int64_t, int64_t, const T*, int64_t, int64_t, T*, int64_t,    \
                                 int64_t, int, double, void*, size_t, int*);
MAK_INST_PB(double)
MAK_INST_PB(zd)

}  // namespace mak

extern "C" {

int mak_pullback_workspace_size(int kind, int dtype, int batch, int m, int n, size_t* bytes) {
  if (bytes == nullptr) return -6;
  if (batch < 0) return -3;
  if (m < 0) return -4;
  if (n < 0) return -5;
  if (kind == 0) *bytes = dtype == MAK_F64 ? mak::qr_pullback_ws<double>(batch, m, n)
                                           : mak::qr_pullback_ws<mak::zd>(batch, m, n);
  else if (kind == 1) *bytes = dtype == MAK_F64 ? mak::polar_pullback_ws<double>(batch, m, n)
                                                : mak::polar_pullback_ws<mak::zd>(batch, m, n);
  else if (kind == 2) *bytes = dtype == MAK_F64 ? mak::qr_full_cotangent_ws<double>(batch, m, n)
                                                : mak::qr_full_cotangent_ws<mak::zd>(batch, m, n);
  else return -1;
  if (dtype != MAK_F64 && dtype != MAK_C64) return -2;
  return 0;
}

#define MAK_QRPB(SUFFIX, T, CT)                                                                                     \
  int mak_qr_pullback_##SUFFIX(void* s, int batch, int m, int n, int p, int positive, const CT* A, int lda,      \
                               long long sa,                                                                       \
                               const CT* Tp, long long st, const CT* R, int ldr, long long sr, const CT* dQ,       \
                               int lddq, long long sdq, const CT* dR, int lddr, long long sdr, CT* dA, int ldda,   \
                               long long sda, void* work, size_t wbytes, double* gauge_out) {                      \
    return mak::qr_pullback<T>((hipStream_t)s, batch, m, n, p, positive, (const T*)A, lda, sa, (const T*)Tp, st,   \
                               (const T*)R, ldr, sr, (const T*)dQ, lddq, sdq, (const T*)dR, lddr, sdr, (T*)dA,    \
                               ldda, sda, work, wbytes, gauge_out);                                                \
  }
MAK_QRPB(d, double, double)
MAK_QRPB(z, mak::zd, void)

#define MAK_QRFC(SUFFIX, T, CT)                                                                                     \
  int mak_qr_full_cotangent_##SUFFIX(void* s, int batch, int m, int n, int p, const CT* Q, int ldq, long long sq,  \
                                     const CT* dQ, int lddq, long long sdq, CT* dQ1, int ldd1, long long sd1,      \
                                     double* gauge_out, void* work, size_t wbytes) {                                \
    return mak::qr_full_cotangent<T>((hipStream_t)s, batch, m, n, p, (const T*)Q, ldq, sq, (const T*)dQ, lddq,    \
                                     sdq, (T*)dQ1, ldd1, sd1, gauge_out, work, wbytes);                            \
  }
MAK_QRFC(d, double, double)
MAK_QRFC(z, mak::zd, void)

#define MAK_PPB(SUFFIX, T, CT)                                                                                      \
  int mak_left_polar_pullback_##SUFFIX(void* s, int batch, int m, int n, const CT* W, int ldw, long long sw,        \
                                       const CT* P, int ldp, long long sp, const CT* dW, int lddw, long long sdw,  \
                                       const CT* dP, int lddp, long long sdp, CT* dA, int ldda, long long sda,     \
                                       int maxiter, double tol, void* work, size_t wbytes, int* iters) {           \
    return mak::polar_pullback<T>((hipStream_t)s, batch, m, n, (const T*)W, ldw, sw, (const T*)P, ldp, sp,         \
                                  (const T*)dW, lddw, sdw, (const T*)dP, lddp, sdp, (T*)dA, ldda, sda, maxiter,    \
                                  tol, work, wbytes, iters);                                                        \
  }
MAK_PPB(d, double, double)
MAK_PPB(z, mak::zd, void)

}  // extern "C"
