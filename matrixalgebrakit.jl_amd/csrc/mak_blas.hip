// Dense building blocks for the polar iterations (FP64 / ComplexF64):
//   * potrf_upper: recursive Cholesky Z = W^H W (W upper) that also returns
//     W^{-1}; 64 x 64 leaves are factored and inverted in registers by one
//     workgroup, every off-diagonal product is an MFMA GEMM;
//   * Frobenius norms (deterministic two-pass reductions);
//   * elementwise helpers (identity fill, axpby, Hermitian part).
#include <algorithm>

#include "mak_common.hpp"
#include "mak_dense32.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

constexpr int DB = 64;  // diagonal block of the blocked Cholesky

template <typename T> void fill_zero(T* A, int64_t lda, int m, int n, hipStream_t s);

// Factor the kb x kb (kb <= 64) diagonal block Z (upper part read) into W
// (upper, written over Z; strict lower part set to zero) and write W^{-1}
// (upper, zero below) to Y.  info[0] receives offset + j + 1 when the leading
// (j+1) x (j+1) minor is not positive definite.
//
// 2 x 2 blocks of 32, one workgroup of 4 waves, everything in LDS:
//   W11, Y11 = chol_inv(Z11)                 (wave 0, no workgroup barrier)
//   W12 = Y11^H Z12 ; Z22 -= W12^H W12       (MFMA, 4 waves)
//   W22, Y22 = chol_inv(Z22)                 (wave 0)
//   Y12 = -(Y11 W12) Y22                     (MFMA)
// Rows/columns >= kb are padded with the identity.
template <typename T>
__global__ void __launch_bounds__(256) potrf_inv64_kernel(T* Z, int64_t ldz, int kb, T* Y, int64_t ldy,
                                                          int* info, int offset) {
  __shared__ T Z11[32][33], Z12[32][33], Z22[32][33];
  __shared__ T Y11[32][33], Y12[32][33], Y22[32][33];
  __shared__ int bad_s;
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  for (int e = tid; e < 32 * 32; e += 256) {
    const int r = e % 32, c = e / 32;
    auto ld = [&](int rr, int cc) {
      return (rr <= cc && cc < kb) ? Z[rr + (int64_t)cc * ldz] : ((rr == cc) ? s_one(T()) : s_zero(T()));
    };
    Z11[r][c] = ld(r, c);
    Z12[r][c] = ld(r, c + 32);
    Z22[r][c] = ld(r + 32, c + 32);
  }
  if (tid == 0) bad_s = 0;
  __syncthreads();
  if (w == 0) {
    const int b = wave_chol_inv32<T>(Z11, Y11);
    if (lane_id() == 0) bad_s = b;
  }
  __syncthreads();
  if (bad_s) {
    if (tid == 0 && info && info[0] == 0) info[0] = offset + bad_s;
    return;
  }
  mm32<T>(Z12, Y11, true, Z12, false, 1.0, false);    // W12 = Y11^H Z12
  mm32<T>(Z22, Z12, true, Z12, false, -1.0, true);    // Z22 -= W12^H W12
  if (w == 0) {
    const int b = wave_chol_inv32<T>(Z22, Y22);
    if (lane_id() == 0) bad_s = b;
  }
  __syncthreads();
  if (bad_s) {
    if (tid == 0 && info && info[0] == 0) info[0] = offset + 32 + bad_s;
    return;
  }
  mm32<T>(Y12, Y11, false, Z12, false, 1.0, false);   // Y11 W12
  mm32<T>(Y12, Y12, false, Y22, false, -1.0, false);  // -(Y11 W12) Y22
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e % 64, c = e / 64;
    if (r < kb && c < kb) {
      T wv = s_zero(T()), yv = s_zero(T());
      if (r <= c) {
        if (c < 32) { wv = Z11[r][c]; yv = Y11[r][c]; }
        else if (r < 32) { wv = Z12[r][c - 32]; yv = Y12[r][c - 32]; }
        else { wv = Z22[r - 32][c - 32]; yv = Y22[r - 32][c - 32]; }
      }
      Z[r + (int64_t)c * ldz] = wv;
      Y[r + (int64_t)c * ldy] = yv;
    }
  }
}

template <typename T> static GemmArgs gargs(int M, int N, int K, const T* A, int64_t lda, const T* B, int64_t ldb,
                                            T* C, int64_t ldc, double alpha, double beta) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha_re = alpha; g.beta_re = beta;
  g.ksplit = 1; g.kchunk = K; g.batch = 1;
  return g;
}

template <typename T>
__global__ void zero_lower_kernel(T* A, int64_t lda, int n) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  if (r > c) A[r + c * lda] = s_zero(T());
}

template <typename T>
void triangular_right_inplace(T* X, int64_t ldx, int m, int kb, const T* U, int64_t ldu, double alpha,
                              hipStream_t s);

template <typename T>
__global__ void copy_kernel(int m, int n, const T* X, int64_t ldx, T* Y, int64_t ldy) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  Y[r + c * ldy] = X[r + c * ldx];
}
template <typename T> void copy_mat(int m, int n, const T* X, int64_t ldx, T* Y, int64_t ldy, hipStream_t s) {
  const int64_t tot = (int64_t)m * n;
  if (tot <= 0) return;
  hipLaunchKernelGGL((copy_kernel<T>), dim3((unsigned)cdiv64(tot, 256)), dim3(256), 0, s, m, n, X, ldx, Y, ldy);
}

// Recursive Cholesky with inverse on an n x n block (splits at multiples of
// DB, so the leaves are the DB x DB diagonal blocks):
//   Z = [Z11 Z12; . Z22] = W^H W,  W = [W11 W12; 0 W22],  Y = W^{-1}:
//   W11, Y11 <- recursion;  W12 = Y11^H Z12;  Z22 -= W12^H W12;
//   W22, Y22 <- recursion;  Y12 = -(Y11 W12) Y22.
// Every product is one MFMA GEMM; the DB x DB leaves run potrf_inv64_kernel.
// W12 is written conjugate-transposed into the free lower block Z21 (the
// product that forms it cannot run in place), and potrf_finish_kernel moves
// every such block to the upper part at the end.  Y's strict lower part is
// zeroed once before the recursion (the triangular-hint GEMMs read the zeros
// inside their diagonal tiles); nothing below Y's diagonal is written after.
template <typename T>
static void potrf_inv_rec(T* Z, int64_t ldz, T* Y, int64_t ldy, int n, int offset, T* tmp, int* dinfo,
                          hipStream_t s) {
  if (n <= DB) {
    hipLaunchKernelGGL((potrf_inv64_kernel<T>), dim3(1), dim3(256), 0, s, Z, ldz, n, Y, ldy, dinfo, offset);
    return;
  }
  const int n1 = ((n / 2 + DB - 1) / DB) * DB;
  const int n2 = n - n1;
  T* Z12 = Z + (int64_t)n1 * ldz;
  T* Z21 = Z + n1;
  T* Z22 = Z + n1 + (int64_t)n1 * ldz;
  T* Y12 = Y + (int64_t)n1 * ldy;
  T* Y22 = Y + n1 + (int64_t)n1 * ldy;
  potrf_inv_rec<T>(Z, ldz, Y, ldy, n1, offset, tmp, dinfo, s);
  // Z21 = W12^H = Z12^H Y11  (Y11 upper)
  GemmArgs g = gargs<T>(n2, n1, n1, Z12, ldz, Y, ldy, Z21, ldz, 1.0, 0.0);
  g.tri = TRI_B_UPPER;
  gemm<T>(OP_C, OP_N, g, s);
  // Z22 -= W12^H W12 = Z21 Z21^H  (upper triangle only: the recursion reads nothing else)
  g = gargs<T>(n2, n2, n1, Z21, ldz, Z21, ldz, Z22, ldz, -1.0, 1.0);
  g.upperC = 1;
  gemm<T>(OP_N, OP_C, g, s);
  potrf_inv_rec<T>(Z22, ldz, Y22, ldy, n2, offset + n1, tmp, dinfo, s);
  // tmp = Y11 W12 = Y11 Z21^H;  Y12 = -tmp Y22
  g = gargs<T>(n1, n2, n1, Y, ldy, Z21, ldz, tmp, n1, 1.0, 0.0);
  g.tri = TRI_A_UPPER;
  gemm<T>(OP_N, OP_C, g, s);
  g = gargs<T>(n1, n2, n2, tmp, n1, Y22, ldy, Y12, ldy, -1.0, 0.0);
  g.tri = TRI_B_UPPER;
  gemm<T>(OP_N, OP_N, g, s);
}

// After the recursion: every 32 x 32 block strictly below the diagonal that
// lies outside the DB x DB leaves holds W^H there; move it (conjugate
// transposed) to the mirrored upper block and zero it.  Blocks inside a leaf
// were written by the leaf kernel.  One workgroup per lower block pair, the
// transpose staged through LDS.
template <typename T>
__global__ void __launch_bounds__(256) potrf_finish_kernel(T* Z, int64_t ldz, int n) {
  __shared__ T tile[32][33];
  const int bc = blockIdx.x, br = blockIdx.y;  // lower block (rows br, cols bc), br > bc
  if (br <= bc || (br * 32) / DB == (bc * 32) / DB) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int r = br * 32 + tx, c = bc * 32 + k;
    if (r < n && c < n) {
      tile[k][tx] = Z[r + (int64_t)c * ldz];
      Z[r + (int64_t)c * ldz] = s_zero(T());
    }
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int r = bc * 32 + tx, c = br * 32 + k;  // upper element (r, c) = conj(lower (c, r))
    if (r < n && c < n) Z[r + (int64_t)c * ldz] = s_conj(tile[tx][k]);
  }
}

// Z (n x n, upper part) -> W (upper, in place, strict lower zeroed) with
// Z = W^H W, and Winv = W^{-1} (upper, n x n, fully overwritten).  tmp holds
// potrf_tmp_elems(n) elements.  dinfo (device) receives j > 0 when the leading
// j x j minor is not positive definite.
template <typename T>
void potrf_upper(T* Z, int64_t ldz, int n, T* Winv, int64_t ldw, T* tmp, int* dinfo, hipStream_t s) {
  hipMemsetAsync(dinfo, 0, sizeof(int), s);
  if (n <= 0) return;
  hipLaunchKernelGGL((zero_lower_kernel<T>), dim3((unsigned)cdiv64((int64_t)n * n, 256)), dim3(256), 0, s, Winv, ldw,
                     n);
  potrf_inv_rec<T>(Z, ldz, Winv, ldw, n, 0, tmp, dinfo, s);
  if (n > DB) {
    const int nb = cdiv(n, 32);
    hipLaunchKernelGGL((potrf_finish_kernel<T>), dim3(nb, nb), dim3(256), 0, s, Z, ldz, n);
  }
}

// ---------------------------------------------------------------------------
// In-place right multiplication by a small upper-triangular block:
//   X (m x kb) <- alpha X U, U upper (kb x kb, kb <= 64).
// Each workgroup owns 64 rows of X; column c of the result only needs columns
// 0..c of X, so rows are processed in registers.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) trmm_right_upper_kernel(T* X, int64_t ldx, int m, int kb, const T* U,
                                                               int64_t ldu, double alpha) {
  __shared__ T Us[DB][DB + 1];
  __shared__ T Xs[64][DB + 1];
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x * 64;
  for (int i = tid; i < DB * DB; i += 256) {
    const int r = i % DB, c = i / DB;
    Us[r][c] = (r <= c && c < kb) ? U[r + c * ldu] : s_zero(T());
  }
  for (int i = tid; i < 64 * DB; i += 256) {
    const int r = i % 64, c = i / 64;
    Xs[r][c] = (r0 + r < m && c < kb) ? X[(r0 + r) + c * ldx] : s_zero(T());
  }
  __syncthreads();
  for (int i = tid; i < 64 * DB; i += 256) {
    const int r = i % 64, c = i / 64;
    if (r0 + r < m && c < kb) {
      T s = s_zero(T());
      for (int k = 0; k <= c; ++k) s = s_fma(Xs[r][k], Us[k][c], s);
      X[(r0 + r) + c * ldx] = s_scale(s, alpha);
    }
  }
}

template <typename T>
void triangular_right_inplace(T* X, int64_t ldx, int m, int kb, const T* U, int64_t ldu, double alpha,
                              hipStream_t s) {
  if (m <= 0 || kb <= 0) return;
  hipLaunchKernelGGL((trmm_right_upper_kernel<T>), dim3(cdiv(m, 64)), dim3(256), 0, s, X, ldx, m, kb, U, ldu, alpha);
}

// ---------------------------------------------------------------------------
// Frobenius norms: ||X||_F and ||X - Y||_F (two-pass, deterministic)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) sumsq_kernel(const T* X, int64_t ldx, const T* Y, int64_t ldy, int m, int n,
                                                    double* partial) {
  __shared__ double red[4];
  const int64_t tot = (int64_t)m * n;
  double s = 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * 256) {
    const int r = (int)(idx % m), c = (int)(idx / m);
    T v = X[r + c * ldx];
    if (Y) v = s_sub(v, Y[r + c * ldy]);
    s += s_abs2(v);
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ void sum_partials_kernel(const double* partial, int n, double* out) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = red[0] + red[1] + red[2] + red[3];
}

constexpr int kNormBlocks = 512;

template <typename T>
void sumsq(const T* X, int64_t ldx, const T* Y, int64_t ldy, int m, int n, double* partial, double* out,
           hipStream_t s) {
  hipLaunchKernelGGL((sumsq_kernel<T>), dim3(kNormBlocks), dim3(256), 0, s, X, ldx, Y, ldy, m, n, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partial, kNormBlocks, out);
}

// ---------------------------------------------------------------------------
// Elementwise: Y = a X + b Y ; C = (C + C^H)/2 ; A += s I
// ---------------------------------------------------------------------------
template <typename T>
__global__ void axpby_kernel(int m, int n, double a, const T* X, int64_t ldx, double b, T* Y, int64_t ldy) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  Y[r + c * ldy] = s_add(s_scale(X[r + c * ldx], a), s_scale(Y[r + c * ldy], b));
}
template <typename T>
void axpby(int m, int n, double a, const T* X, int64_t ldx, double b, T* Y, int64_t ldy, hipStream_t s) {
  const int64_t tot = (int64_t)m * n;
  if (tot <= 0) return;
  hipLaunchKernelGGL((axpby_kernel<T>), dim3((unsigned)cdiv64(tot, 256)), dim3(256), 0, s, m, n, a, X, ldx, b, Y, ldy);
}

template <typename T>
__global__ void herm_part_kernel(T* C, int64_t ldc, int n) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  if (r > c) return;
  const T x = C[r + c * ldc], y = C[c + r * ldc];
  const T v = s_scale(s_add(x, s_conj(y)), 0.5);
  C[r + c * ldc] = v;
  C[c + r * ldc] = s_conj(v);
}
template <typename T> void herm_part(T* C, int64_t ldc, int n, hipStream_t s) {
  const int64_t tot = (int64_t)n * n;
  if (tot <= 0) return;
  hipLaunchKernelGGL((herm_part_kernel<T>), dim3((unsigned)cdiv64(tot, 256)), dim3(256), 0, s, C, ldc, n);
}

template <typename T>
__global__ void add_diag_kernel(T* A, int64_t lda, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i + i * lda] = s_add(A[i + i * lda], s_from_real<T>(v));
}
template <typename T> void add_diag(T* A, int64_t lda, int n, double v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((add_diag_kernel<T>), dim3(cdiv(n, 256)), dim3(256), 0, s, A, lda, n, v);
}

template <typename T>
__global__ void zero_kernel(T* A, int64_t lda, int m, int n) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  A[r + c * lda] = s_zero(T());
}
template <typename T> void fill_zero(T* A, int64_t lda, int m, int n, hipStream_t s) {
  const int64_t tot = (int64_t)m * n;
  if (tot <= 0) return;
  hipLaunchKernelGGL((zero_kernel<T>), dim3((unsigned)cdiv64(tot, 256)), dim3(256), 0, s, A, lda, m, n);
}

#define MAK_INST_BLAS(T)                                                                                   \
  template void potrf_upper<T>(T*, int64_t, int, T*, int64_t, T*, int*, hipStream_t);                    \
  template void copy_mat<T>(int, int, const T*, int64_t, T*, int64_t, hipStream_t);                       \
  template void triangular_right_inplace<T>(T*, int64_t, int, int, const T*, int64_t, double, hipStream_t); \
  template void sumsq<T>(const T*, int64_t, const T*, int64_t, int, int, double*, double*, hipStream_t);   \
  template void axpby<T>(int, int, double, const T*, int64_t, double, T*, int64_t, hipStream_t);         \
  template void herm_part<T>(T*, int64_t, int, hipStream_t);                                              \
  template void add_diag<T>(T*, int64_t, int, double, hipStream_t);                                       \
  template void fill_zero<T>(T*, int64_t, int, int, hipStream_t);
MAK_INST_BLAS(double)
MAK_INST_BLAS(zd)

}  // namespace mak
