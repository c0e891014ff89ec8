// Dense building blocks for the polar iterations (FP64 / ComplexF64):
//   * potrf_upper: Cholesky Z = W^H W (W upper) that also returns W^{-1}:
//     the persistent tile kernel (mak_chol.hip), inside a blocked loop of MFMA
//     GEMMs for large n;
//   * Frobenius norms (deterministic two-pass reductions);
//   * elementwise helpers (identity fill, axpby, Hermitian part).
#include <algorithm>
#include <mutex>

#include "mak_common.hpp"
#include "mak_dense32.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

constexpr int DB = 64;  // row block of the in-place triangular multiply

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

// After the blocked loop: every 32 x 32 block strictly below the diagonal
// that lies outside the diagonal blocks of the loop holds W^H there; move it
// (conjugate transposed) to the mirrored upper block and zero it.  Blocks
// inside a diagonal block were written by the tile kernel.  One workgroup per
// lower block pair, the transpose staged through LDS.
template <typename T>
__global__ void __launch_bounds__(256) potrf_finish_kernel(T* Z, int64_t ldz, int n, int nb) {
  __shared__ T tile[32][33];
  const int bc = blockIdx.x, br = blockIdx.y;  // lower block (rows br, cols bc), br > bc
  if (br <= bc || (br * 32) / nb == (bc * 32) / nb) return;
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

// Diagonal block width of the blocked loop (n > kCholTileMaxN).
constexpr int kCholBlock = 1024;

// The blocked loop's look-ahead stream, one per device (default priority,
// created on first use, kept until the process ends), and the two events
// that fork work onto it and join it back.  A call holds the device's mutex
// while it enqueues (the events are reused call after call); the work itself
// runs on the caller's stream and this one.  nullptr: no side stream (the
// loop then runs on the caller's stream alone).
struct PotrfSide {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  std::mutex mu;
};
constexpr int kPotrfMaxDevices = 64;
PotrfSide g_potrf_side[kPotrfMaxDevices];
std::once_flag g_potrf_side_once[kPotrfMaxDevices];

static PotrfSide* potrf_side(hipStream_t s) {
  int dev = -1;
  if (s != nullptr) {
    if (hipStreamGetDevice(s, &dev) != hipSuccess) return nullptr;
  } else if (hipGetDevice(&dev) != hipSuccess) {
    return nullptr;
  }
  if (dev < 0 || dev >= kPotrfMaxDevices) return nullptr;
  PotrfSide* ps = &g_potrf_side[dev];
  std::call_once(g_potrf_side_once[dev], [dev, ps] {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return;
    hipStream_t t = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipStreamCreateWithFlags(&t, hipStreamNonBlocking) == hipSuccess &&
        hipEventCreateWithFlags(&a, hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&b, hipEventDisableTiming) == hipSuccess) {
      ps->s = t;
      ps->fork = a;
      ps->join = b;
    }
    if (cur != dev) (void)hipSetDevice(cur);
  });
  return ps->s != nullptr ? ps : nullptr;
}

size_t potrf_tmp_elems(int n) { return (size_t)64 * std::max(n, 1) + (size_t)n * n / 4; }

// Z (n x n, upper part) -> W (upper, in place, strict lower zeroed) with
// Z = W^H W, and Winv = W^{-1} (upper, n x n, fully overwritten).  tmp holds
// potrf_tmp_elems(n) elements.  dinfo (device) receives j > 0 when the leading
// j x j minor is not positive definite.
//
// n <= kCholTileMaxN: one launch of the persistent tile kernel.  Larger n:
// diagonal blocks of kCholBlock columns, each factored by the tile kernel;
// with Wkk, Xkk = W_kk, W_kk^{-1}:
//   panel     W_k,r = Xkk^H Z_k,r  (written conjugate-transposed into the free
//             lower block Z_r,k: the product cannot run in place)
//   trailing  Z_r,r -= W_k,r^H W_k,r   (upper triangle only)
//   inverse   X_0:k,k = -(X_0:k,0:k W_0:k,k) Xkk   (left-looking; W_0:k,k
//             read from the lower blocks)
// and potrf_finish_kernel moves the W blocks to the upper part at the end.
template <typename T>
void potrf_upper(T* Z, int64_t ldz, int n, T* Winv, int64_t ldw, T* tmp, int* dinfo, hipStream_t s) {
  hipMemsetAsync(dinfo, 0, sizeof(int), s);
  if (n <= 0) return;
  // the tile kernel's flags live at the end of tmp
  const size_t tmpn = potrf_tmp_elems(n);
  const size_t fl_elems = (chol_tile_flag_ints(std::min(n, kCholTileMaxN)) * sizeof(int) + sizeof(T) - 1) / sizeof(T);
  int* flags = reinterpret_cast<int*>(tmp + (tmpn - fl_elems));
  if (n <= kCholTileMaxN) {
    chol_tile_factor<T>(Z, ldz, n, Winv, ldw, flags, dinfo, 0, s);
    return;
  }
  hipLaunchKernelGGL((zero_lower_kernel<T>), dim3((unsigned)cdiv64((int64_t)n * n, 256)), dim3(256), 0, s, Winv, ldw,
                     n);
  // Look-ahead: once block k's panel has updated the next diagonal block, the
  // tile kernel's chain for block k + 1 (latency-bound, on half the CUs) runs
  // beside the rest of block k's trailing update and its inverse update on
  // the side stream; the caller's stream waits for them before block k + 1's
  // panel (which reads the updated columns) and before the finish.
  PotrfSide* side = potrf_side(s);
  std::unique_lock<std::mutex> lock;
  if (side != nullptr) lock = std::unique_lock<std::mutex>(side->mu);
  const hipStream_t s2 = side != nullptr ? side->s : s;
  bool pending = false;  // side work enqueued since the last join
  auto join = [&] {
    if (pending) {
      hipEventRecord(side->join, s2);
      hipStreamWaitEvent(s, side->join, 0);
      pending = false;
    }
  };
  // X_0:k,k = -(X_0:k,0:k W_0:k,k) Xkk on stream t (W_0:k,k^H in the lower blocks Z_k,0:k)
  auto inverse_update = [&](int k0, int kb, hipStream_t t) {
    const T* Xkk = Winv + k0 + (int64_t)k0 * ldw;
    const T* Zk0 = Z + k0;
    GemmArgs g = gargs<T>(k0, kb, k0, Winv, ldw, Zk0, ldz, tmp, k0, 1.0, 0.0);
    g.tri = TRI_A_UPPER;
    gemm<T>(OP_N, OP_C, g, t);
    g = gargs<T>(k0, kb, kb, tmp, k0, Xkk, ldw, Winv + (int64_t)k0 * ldw, ldw, -1.0, 0.0);
    g.tri = TRI_B_UPPER;
    gemm<T>(OP_N, OP_N, g, t);
  };
  for (int k0 = 0; k0 < n; k0 += kCholBlock) {
    const int kb = std::min(kCholBlock, n - k0);
    const int rest = n - k0 - kb;
    T* Zkk = Z + k0 + (int64_t)k0 * ldz;
    T* Xkk = Winv + k0 + (int64_t)k0 * ldw;
    chol_tile_factor<T>(Zkk, ldz, kb, Xkk, ldw, flags, dinfo, k0, s);
    join();  // block k - 1's trailing columns and inverse
    if (rest > 0) {
      T* Zkr = Z + k0 + (int64_t)(k0 + kb) * ldz;
      T* Zrk = Z + (k0 + kb) + (int64_t)k0 * ldz;
      T* Zrr = Z + (k0 + kb) + (int64_t)(k0 + kb) * ldz;
      // Z_r,k = W_k,r^H = Z_k,r^H Xkk
      GemmArgs g = gargs<T>(rest, kb, kb, Zkr, ldz, Xkk, ldw, Zrk, ldz, 1.0, 0.0);
      g.tri = TRI_B_UPPER;
      gemm<T>(OP_C, OP_N, g, s);
      // Z_r,r -= Z_r,k Z_r,k^H (upper triangle): the next diagonal block here,
      // the columns beyond it on the side stream
      const int nb1 = std::min(kb, rest), rest2 = rest - nb1;
      g = gargs<T>(nb1, nb1, kb, Zrk, ldz, Zrk, ldz, Zrr, ldz, -1.0, 1.0);
      g.upperC = 1;
      gemm<T>(OP_N, OP_C, g, s);
      if (side != nullptr) {
        hipEventRecord(side->fork, s);
        hipStreamWaitEvent(s2, side->fork, 0);
        pending = true;
      }
      if (rest2 > 0) {
        // rows of the next block, columns beyond it
        g = gargs<T>(nb1, rest2, kb, Zrk, ldz, Zrk + nb1, ldz, Zrr + (int64_t)nb1 * ldz, ldz, -1.0, 1.0);
        gemm<T>(OP_N, OP_C, g, s2);
        // the square beyond it (upper triangle)
        g = gargs<T>(rest2, rest2, kb, Zrk + nb1, ldz, Zrk + nb1, ldz, Zrr + nb1 + (int64_t)nb1 * ldz, ldz, -1.0,
                     1.0);
        g.upperC = 1;
        gemm<T>(OP_N, OP_C, g, s2);
      }
      if (k0 > 0) inverse_update(k0, kb, s2);
    } else if (k0 > 0) {
      inverse_update(k0, kb, s);
    }
  }
  join();
  const int nb32 = cdiv(n, 32);
  hipLaunchKernelGGL((potrf_finish_kernel<T>), dim3(nb32, nb32), dim3(256), 0, s, Z, ldz, n, kCholBlock);
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
