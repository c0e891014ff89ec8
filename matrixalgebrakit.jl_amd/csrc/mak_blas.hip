// Dense building blocks for the polar iterations (FP64 / ComplexF64):
//   * potrf_upper: blocked right-looking Cholesky Z = W^H W (W upper) that also
//     returns W^{-1}; the 64x64 diagonal blocks are factored in LDS by one
//     workgroup (one barrier per column), panels and trailing updates are MFMA
//     GEMMs;
//   * Frobenius norms (deterministic two-pass reductions);
//   * elementwise helpers (identity fill, axpby, Hermitian part).
#include <algorithm>

#include "mak_common.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

constexpr int DB = 64;  // diagonal block of the blocked Cholesky

template <typename T> void fill_zero(T* A, int64_t lda, int m, int n, hipStream_t s);

// Factor the kb x kb diagonal block Z (upper part read) into W (upper, written
// over Z's upper part; strict lower part set to zero) and write W^{-1} (upper)
// to Y.  info[0] receives j+1 (plus the block offset) when the block is not
// positive definite; the factorisation then stops.
template <typename T>
__global__ void __launch_bounds__(256) potrf_diag_kernel(T* Z, int64_t ldz, int kb, T* Y, int64_t ldy,
                                                         int* info, int offset) {
  __shared__ T S[DB][DB + 1];
  __shared__ T Wl[DB][DB + 1];
  __shared__ int bad;
  const int tid = threadIdx.x;
  for (int i = tid; i < DB * DB; i += 256) {
    const int r = i % DB, c = i / DB;
    S[r][c] = (r <= c && c < kb) ? Z[r + c * ldz] : s_zero(T());
    Wl[r][c] = s_zero(T());
  }
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    const double d2 = s_re(S[j][j]);
    if (!(d2 > 0.0)) {  // uniform: every thread reads the same value
      if (tid == 0) {
        bad = 1;
        if (info && info[0] == 0) info[0] = offset + j + 1;
      }
      break;
    }
    const double d = sqrt(d2), rd = 1.0 / d, rd2 = 1.0 / d2;
    for (int c = j + tid; c < kb; c += 256) Wl[j][c] = (c == j) ? s_from_real<T>(d) : s_scale(S[j][c], rd);
    // trailing update of the upper part: rows j < r <= c
    for (int i = tid; i < DB * DB; i += 256) {
      const int r = i % DB, c = i / DB;
      if (r > j && r <= c && c < kb)
        S[r][c] = s_sub(S[r][c], s_scale(s_mul(s_conj(S[j][r]), S[j][c]), rd2));
    }
    __syncthreads();
  }
  __syncthreads();
  if (bad) return;
  // Y = W^{-1}: init diag(1/W_rr), then for jj = kb-1..0, rows r < jj:
  // Y[r][:] -= (W[r][jj] / W[r][r]) Y[jj][:]
  for (int i = tid; i < DB * DB; i += 256) {
    const int r = i % DB, c = i / DB;
    S[r][c] = (r == c && r < kb) ? s_div(s_one(T()), Wl[r][r]) : s_zero(T());
  }
  __syncthreads();
  for (int jj = kb - 1; jj > 0; --jj) {
    for (int i = tid; i < DB * DB; i += 256) {
      const int r = i % DB, c = i / DB;
      if (r < jj && c >= jj && c < kb) S[r][c] = s_sub(S[r][c], s_mul(s_div(Wl[r][jj], Wl[r][r]), S[jj][c]));
    }
    __syncthreads();
  }
  for (int i = tid; i < DB * DB; i += 256) {
    const int r = i % DB, c = i / DB;
    if (r < kb && c < kb) {
      Z[r + c * ldz] = (r <= c) ? Wl[r][c] : s_zero(T());
      Y[r + c * ldy] = (r <= c) ? S[r][c] : s_zero(T());
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

// Z (n x n, upper part) -> W (upper, in place, strict lower zeroed) with
// Z = W^H W, and Winv = W^{-1} (upper, n x n, fully overwritten).  tmp holds
// DB x n scratch.  dinfo (device) receives j > 0 when the leading j x j minor
// is not positive definite.
template <typename T>
void potrf_upper(T* Z, int64_t ldz, int n, T* Winv, int64_t ldw, T* tmp, int* dinfo, hipStream_t s) {
  hipMemsetAsync(dinfo, 0, sizeof(int), s);
  fill_zero<T>(Winv, ldw, n, n, s);
  for (int k = 0; k < n; k += DB) {
    const int kb = std::min(DB, n - k);
    T* Zkk = Z + k + (int64_t)k * ldz;
    T* Ykk = Winv + k + (int64_t)k * ldw;
    hipLaunchKernelGGL((potrf_diag_kernel<T>), dim3(1), dim3(256), 0, s, Zkk, ldz, kb, Ykk, ldw, dinfo, k);
    const int rest = n - k - kb;
    if (rest > 0) {
      // W[k, k+kb:] = W_kk^{-H} Z[k, k+kb:]  (through tmp)
      T* Zrow = Z + k + (int64_t)(k + kb) * ldz;
      gemm<T>(OP_C, OP_N, gargs<T>(kb, rest, kb, Ykk, ldw, Zrow, ldz, tmp, DB, 1.0, 0.0), s);
      copy_mat<T>(kb, rest, tmp, DB, Zrow, ldz, s);
      // Z[k+kb:, k+kb:] -= W[k, k+kb:]^H W[k, k+kb:]
      T* Z22 = Z + (k + kb) + (int64_t)(k + kb) * ldz;
      gemm<T>(OP_C, OP_N, gargs<T>(rest, rest, kb, Zrow, ldz, Zrow, ldz, Z22, ldz, -1.0, 1.0), s);
    }
  }
  // off-diagonal blocks of W^{-1}, block column by block column:
  //   Y[0:k, k:k+kb] = -(Y[0:k, 0:k] W[0:k, k:k+kb]) Y_kk
  for (int k = DB; k < n; k += DB) {
    const int kb = std::min(DB, n - k);
    T* Yc = Winv + (int64_t)k * ldw;
    gemm<T>(OP_N, OP_N, gargs<T>(k, kb, k, Winv, ldw, Z + (int64_t)k * ldz, ldz, Yc, ldw, 1.0, 0.0), s);
    triangular_right_inplace<T>(Yc, ldw, k, kb, Winv + k + (int64_t)k * ldw, ldw, -1.0, s);
  }
  hipLaunchKernelGGL((zero_lower_kernel<T>), dim3((unsigned)cdiv64((int64_t)n * n, 256)), dim3(256), 0, s, Z, ldz, n);
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
