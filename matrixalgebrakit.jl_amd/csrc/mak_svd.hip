// SVD through the polar decomposition and the polar pushforwards.
//
// svd_compact! / svd_vals! with SVDViaPolar (src/interface/decompositions.jl:
// 173-183; the reference reaches it through gesvdp!, src/implementations/
// svd.jl:161,184-206, and fixes the gauge with gaugefix!(svd_compact!, U, Vh),
// src/common/gauge.jl:72-80):
//   m >= n:  A = W P (QDWH, FP64), P = Z diag(lambda) Z^H (eigh, ascending),
//            S = lambda reversed, V = Z reversed, U = W V, Vh = V^H;
//   m <  n:  A = P Wh, P = Z diag(lambda) Z^H, U = Z reversed, Vh = U^H Wh;
// then the entry of largest magnitude (the first such) of every column of U
// is made real and positive (U[:, j] *= conj(s_j), Vh[j, :] *= s_j).
//
// left_polar_pushforward! / right_polar_pushforward! (src/pushforwards/
// polar.jl:1-25): with P = Z diag(lambda) Z^H the Sylvester equation
// P K + K P = dP - dP^H is solved exactly, K = Z [(Z^H C Z)_ij /
// (lambda_i + lambda_j)] Z^H, and P^{-1} = Z diag(1/lambda) Z^H; the rest
// are MFMA GEMMs:
//   dP = W^H dA, dW = dA P^{-1} - W (W^H dA P^{-1}) + W K, dP -= K P.
#include <algorithm>
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

inline size_t al256(size_t b) { return ((b + 255) / 256) * 256; }

template <typename T> GemmArgs gargs(int M, int N, int K, const T* A, int64_t lda, const T* B, int64_t ldb, T* C,
                                     int64_t ldc, double alpha, double beta) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha_re = alpha; g.beta_re = beta;
  g.ksplit = 1; g.kchunk = K; g.batch = 1;
  return g;
}

// V[:, j] = Z[:, k-1-j] (k x k), S[j] = max(lambda[k-1-j], 0) (S may be null).
template <typename T>
__global__ void reverse_eig_kernel(int k, const T* Z, int64_t ldz, const double* lam, T* V, int64_t ldv, double* S) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)k * k) return;
  const int r = (int)(idx % k), c = (int)(idx / k);
  V[r + (int64_t)c * ldv] = Z[r + (int64_t)(k - 1 - c) * ldz];
  if (r == 0 && S) S[c] = fmax(lam[k - 1 - c], 0.0);
}

// One workgroup per column j of U (mu x k): s = sign of the first entry of
// largest magnitude; U[:, j] *= conj(s), Vh[j, :] *= s (Vh k x nv).
template <typename T>
__global__ void __launch_bounds__(256) svd_gauge_kernel(T* U, int64_t ldu, int mu, T* Vh, int64_t ldvh, int nv) {
  __shared__ double bv[256];
  __shared__ int bi[256];
  __shared__ T sgn;
  const int j = blockIdx.x;
  T* u = U + (int64_t)j * ldu;
  double best = -1.0;
  int bidx = 0;
  for (int i = threadIdx.x; i < mu; i += 256) {
    const double a = s_abs2(u[i]);
    if (a > best) { best = a; bidx = i; }  // strictly larger: the first index wins within a thread
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = bidx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double o = bv[threadIdx.x + w];
      const int oi = bi[threadIdx.x + w];
      if (o > bv[threadIdx.x] || (o == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
        bv[threadIdx.x] = o;
        bi[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const T x = u[bi[0]];
    const double a = sqrt(s_abs2(x));
    sgn = a > 0.0 ? s_scale(x, 1.0 / a) : s_zero(T());
  }
  __syncthreads();
  const T s = sgn;
  for (int i = threadIdx.x; i < mu; i += 256) u[i] = s_mul(u[i], s_conj(s));
  if (Vh)
    for (int c = threadIdx.x; c < nv; c += 256) Vh[j + (int64_t)c * ldvh] = s_mul(s, Vh[j + (int64_t)c * ldvh]);
}

// C = X - X^H (n x n)
template <typename T> __global__ void antiherm_kernel(int n, const T* X, int64_t ldx, T* C, int64_t ldc) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  C[r + (int64_t)c * ldc] = s_sub(X[r + (int64_t)c * ldx], s_conj(X[c + (int64_t)r * ldx]));
}

// X_ij /= (lambda_i + lambda_j)
template <typename T> __global__ void sylv_divide_kernel(int n, T* X, int64_t ldx, const double* lam) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  X[r + (int64_t)c * ldx] = s_scale(X[r + (int64_t)c * ldx], 1.0 / (lam[r] + lam[c]));
}

// X[:, j] *= 1 / lambda_j (rows m)
template <typename T> __global__ void scale_cols_inv_kernel(int m, int n, T* X, int64_t ldx, const double* lam) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  X[r + (int64_t)c * ldx] = s_scale(X[r + (int64_t)c * ldx], 1.0 / lam[c]);
}

// B (n x m) = A^H (A m x n)
template <typename T>
__global__ void __launch_bounds__(256) adj_kernel(int m, int n, const T* A, int64_t lda, T* B, int64_t ldb) {
  __shared__ T tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int q = ty; q < 32; q += 8) {
    const int r = r0 + tx, c = c0 + q;
    if (r < m && c < n) tile[q][tx] = A[r + (int64_t)c * lda];
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int br = c0 + tx, bc = r0 + q;
    if (br < n && bc < m) B[br + (int64_t)bc * ldb] = s_conj(tile[tx][q]);
  }
}
template <typename T> void adjoint_to(int m, int n, const T* A, int64_t lda, T* B, int64_t ldb, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL((adj_kernel<T>), dim3(cdiv(m, 32), cdiv(n, 32)), dim3(256), 0, s, m, n, A, lda, B, ldb);
}

inline unsigned blocks_for(int64_t n) { return (unsigned)cdiv64(std::max<int64_t>(n, 1), 256); }

struct SvdLayout {
  size_t offW, offP, offV, offD, offEig, offPol, bytes;
};
template <typename T> SvdLayout svd_layout(int m, int n) {
  const int k = std::min(m, n);
  SvdLayout L{};
  size_t off = 0;
  auto take = [&](size_t b) { size_t o = off; off += al256(b); return o; };
  L.offW = take((size_t)m * n * sizeof(T));
  L.offP = take((size_t)k * k * sizeof(T));
  L.offV = take((size_t)k * k * sizeof(T));
  L.offD = take((size_t)k * sizeof(double));
  L.offEig = take(eigh_workspace_bytes(k, 1));
  L.offPol = take(polar_workspace_bytes<T>(m, n));
  L.bytes = off;
  return L;
}

struct PushLayout {
  size_t offZ, offD, offC, offT, offK, offM, offEig, offWt, offAt, offDWt, offDPt, bytes;
};
// left form: W big x k, P k x k; right = 1 adds the conjugate transposes of
// Wh, dA, dWh (big x k) and dP (k x k)
template <typename T> PushLayout push_layout(int big, int k, int right) {
  PushLayout L{};
  size_t off = 0;
  auto take = [&](size_t b) { size_t o = off; off += al256(b); return o; };
  const size_t kk = (size_t)k * k * sizeof(T), bk = (size_t)big * k * sizeof(T);
  L.offZ = take(kk);
  L.offD = take((size_t)k * sizeof(double));
  L.offC = take(kk);
  L.offT = take(kk);
  L.offK = take(kk);
  L.offM = take(bk);
  L.offEig = take(eigh_workspace_bytes(k, 1));
  if (right) {
    L.offWt = take(bk);
    L.offAt = take(bk);
    L.offDWt = take(bk);
    L.offDPt = take(kk);
  }
  L.bytes = off;
  return L;
}

// The left pushforward on big x k operands (see the top of the file).
template <typename T>
int push_left_core(hipStream_t s, int big, int k, const T* W, int64_t ldw, const T* P, int64_t ldp, const T* dA,
                   int64_t ldda, T* dW, int64_t lddw, T* dP, int64_t lddp, char* ws, const PushLayout& L) {
  T* Z = reinterpret_cast<T*>(ws + L.offZ);
  double* lam = reinterpret_cast<double*>(ws + L.offD);
  T* Cm = reinterpret_cast<T*>(ws + L.offC);
  T* Tm = reinterpret_cast<T*>(ws + L.offT);
  T* Km = reinterpret_cast<T*>(ws + L.offK);
  T* Mm = reinterpret_cast<T*>(ws + L.offM);
  // dP = W^H dA (k x k)
  gemm<T>(OP_C, OP_N, gargs<T>(k, k, big, W, ldw, dA, ldda, dP, lddp, 1.0, 0.0), s);
  // P = Z diag(lam) Z^H
  copy_mat<T>(k, k, P, ldp, Z, k, s);
  const int est = eigh_batched<T>(s, k, Z, k, 0, lam, 0, 1, ws + L.offEig);
  if (est != 0) return est;
  // P^{-1} and the Sylvester solve need positive eigenvalues
  double lmin = 0.0;
  if (hipMemcpyAsync(&lmin, lam, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return 1;
  if (!(lmin > 0.0)) return 2;
  const unsigned bkk = blocks_for((int64_t)k * k);
  // K = Z [(Z^H C Z) ./ (lam_i + lam_j)] Z^H, C = dP - dP^H
  hipLaunchKernelGGL((antiherm_kernel<T>), dim3(bkk), dim3(256), 0, s, k, (const T*)dP, lddp, Cm, k);
  gemm<T>(OP_N, OP_N, gargs<T>(k, k, k, Cm, k, Z, k, Tm, k, 1.0, 0.0), s);     // C Z
  gemm<T>(OP_C, OP_N, gargs<T>(k, k, k, Z, k, Tm, k, Cm, k, 1.0, 0.0), s);     // Z^H C Z
  hipLaunchKernelGGL((sylv_divide_kernel<T>), dim3(bkk), dim3(256), 0, s, k, Cm, k, (const double*)lam);
  gemm<T>(OP_N, OP_N, gargs<T>(k, k, k, Z, k, Cm, k, Tm, k, 1.0, 0.0), s);     // Z X
  gemm<T>(OP_N, OP_C, gargs<T>(k, k, k, Tm, k, Z, k, Km, k, 1.0, 0.0), s);     // K = Z X Z^H
  // dW = dA P^{-1} = ((dA Z) diag(1/lam)) Z^H
  gemm<T>(OP_N, OP_N, gargs<T>(big, k, k, dA, ldda, Z, k, Mm, big, 1.0, 0.0), s);
  hipLaunchKernelGGL((scale_cols_inv_kernel<T>), dim3(blocks_for((int64_t)big * k)), dim3(256), 0, s, big, k, Mm,
                     (int64_t)big, (const double*)lam);
  gemm<T>(OP_N, OP_C, gargs<T>(big, k, k, Mm, big, Z, k, dW, lddw, 1.0, 0.0), s);
  // dW -= W (W^H dW);  dW += W K
  gemm<T>(OP_C, OP_N, gargs<T>(k, k, big, W, ldw, dW, lddw, Tm, k, 1.0, 0.0), s);
  gemm<T>(OP_N, OP_N, gargs<T>(big, k, k, W, ldw, Tm, k, dW, lddw, -1.0, 1.0), s);
  gemm<T>(OP_N, OP_N, gargs<T>(big, k, k, W, ldw, Km, k, dW, lddw, 1.0, 1.0), s);
  // dP -= K P
  gemm<T>(OP_N, OP_N, gargs<T>(k, k, k, Km, k, P, ldp, dP, lddp, -1.0, 1.0), s);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace

template <typename T> size_t svd_polar_workspace_bytes(int m, int n) {
  return svd_layout<T>(std::max(m, 0), std::max(n, 0)).bytes;
}

// svd_compact! (U, Vh non-null) or svd_vals! (both null) via the polar
// decomposition.  A is only read.  S: min(m, n) doubles (device).
template <typename T>
int svd_polar_driver(hipStream_t s, int m, int n, const T* A, int64_t lda, double* S, T* U, int64_t ldu, T* Vh,
                     int64_t ldvh, int maxiter, void* work, size_t work_bytes, mak_polar_info* info) {
  if (m < 0) return -2;
  if (n < 0) return -3;
  if ((int64_t)m * n > 0 && A == nullptr) return -4;
  if (lda < std::max(1, m)) return -5;
  const int k = std::min(m, n);
  if (k > 0 && S == nullptr) return -6;
  if ((U == nullptr) != (Vh == nullptr)) return U ? -9 : -7;
  if (U && ldu < std::max(1, m)) return -8;
  if (Vh && ldvh < std::max(1, k)) return -10;
  const size_t need = svd_polar_workspace_bytes<T>(m, n);
  if (work == nullptr || work_bytes < need) return -13;
  if (k == 0) {
    if (U) set_identity<T>(U, ldu, m, k, s);
    if (Vh) set_identity<T>(Vh, ldvh, k, n, s);
    return hipGetLastError() == hipSuccess ? 0 : 1;
  }
  if (!eigh_available()) return 3;
  char* ws = reinterpret_cast<char*>(work);
  const SvdLayout L = svd_layout<T>(m, n);
  T* W = reinterpret_cast<T*>(ws + L.offW);
  T* P = reinterpret_cast<T*>(ws + L.offP);
  T* V = reinterpret_cast<T*>(ws + L.offV);
  double* D = reinterpret_cast<double*>(ws + L.offD);
  const bool right = m < n;
  int st = right ? polar_driver<T>(s, 1, m, n, A, lda, W, m, P, k, 64, maxiter, 0.0, ws + L.offPol,
                                   need - L.offPol, info)
                 : polar_driver<T>(s, 0, m, n, A, lda, W, m, P, k, 64, maxiter, 0.0, ws + L.offPol,
                                   need - L.offPol, info);
  if (st != 0 && st != 6) return st < 0 ? 1 : st;
  const int est = eigh_batched<T>(s, k, P, k, 0, D, 0, 1, ws + L.offEig);
  if (est != 0) return est;
  hipLaunchKernelGGL((reverse_eig_kernel<T>), dim3(blocks_for((int64_t)k * k)), dim3(256), 0, s, k, P, k,
                     (const double*)D, V, k, S);
  if (U) {
    if (!right) {
      gemm<T>(OP_N, OP_N, gargs<T>(m, k, k, W, m, V, k, U, ldu, 1.0, 0.0), s);   // U = W V
      adjoint_to<T>(k, k, V, k, Vh, ldvh, s);                                      // Vh = V^H
    } else {
      copy_mat<T>(k, k, V, k, U, ldu, s);                                          // U = V
      gemm<T>(OP_C, OP_N, gargs<T>(k, n, k, V, k, W, m, Vh, ldvh, 1.0, 0.0), s);  // Vh = V^H Wh
    }
    hipLaunchKernelGGL((svd_gauge_kernel<T>), dim3(k), dim3(256), 0, s, U, ldu, m, Vh, ldvh, n);
  }
  if (hipGetLastError() != hipSuccess) return 1;
  return st;
}

template <typename T> size_t polar_pushforward_workspace_bytes(int m, int n, int right) {
  const int big = std::max(std::max(m, n), 0), k = std::max(std::min(m, n), 0);
  return push_layout<T>(big, k, right).bytes;
}

// left_polar_pushforward!(dA, A, (W, P), (dW, dP)) for m >= n: W, dA, dW
// m x n, P, dP n x n.  right = 1: right_polar_pushforward!(dA, A, (P, Wh),
// (dP, dWh)) for m <= n (src/pushforwards/polar.jl:14-25): Wh, dA, dWh m x n,
// P, dP m x m; it equals the left form on the conjugate transposes (A^H =
// Wh^H P), with dWh and dP its outputs conjugate-transposed.  dW / dWh and
// dP are overwritten.
template <typename T>
int polar_pushforward_driver(hipStream_t s, int right, int m, int n, const T* W, int64_t ldw, const T* P, int64_t ldp,
                             const T* dA, int64_t ldda, T* dW, int64_t lddw, T* dP, int64_t lddp, void* work,
                             size_t work_bytes) {
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (!right && m < n) return -3;
  if (right && n < m) return -4;
  const int big = std::max(m, n), k = std::min(m, n);
  if (k > 0 && (W == nullptr || P == nullptr || dA == nullptr || dW == nullptr || dP == nullptr)) return -5;
  if (ldw < std::max(1, m) || ldp < std::max(1, k) || ldda < std::max(1, m) || lddw < std::max(1, m) ||
      lddp < std::max(1, k))
    return -6;
  const size_t need = polar_pushforward_workspace_bytes<T>(m, n, right);
  if (work == nullptr || work_bytes < need) return -15;
  if (k == 0) return 0;
  if (!eigh_available()) return 3;
  char* ws = reinterpret_cast<char*>(work);
  const PushLayout L = push_layout<T>(big, k, right);
  if (!right) return push_left_core<T>(s, big, k, W, ldw, P, ldp, dA, ldda, dW, lddw, dP, lddp, ws, L);
  T* Wt = reinterpret_cast<T*>(ws + L.offWt);
  T* At = reinterpret_cast<T*>(ws + L.offAt);
  T* dWt = reinterpret_cast<T*>(ws + L.offDWt);
  T* dPt = reinterpret_cast<T*>(ws + L.offDPt);
  adjoint_to<T>(m, n, W, ldw, Wt, big, s);    // Wh^H (n x m)
  adjoint_to<T>(m, n, dA, ldda, At, big, s);  // dA^H
  const int st = push_left_core<T>(s, big, k, Wt, big, P, ldp, At, big, dWt, big, dPt, k, ws, L);
  if (st != 0) return st;
  adjoint_to<T>(big, k, dWt, big, dW, lddw, s);  // dWh = (dW')^H
  adjoint_to<T>(k, k, dPt, k, dP, lddp, s);      // dP = (dP')^H
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

template size_t svd_polar_workspace_bytes<double>(int, int);
template size_t svd_polar_workspace_bytes<zd>(int, int);
template size_t polar_pushforward_workspace_bytes<double>(int, int, int);
template size_t polar_pushforward_workspace_bytes<zd>(int, int, int);

}  // namespace mak

extern "C" {

int mak_svd_polar_workspace_size(int dtype, int m, int n, size_t* bytes) {
  if (bytes == nullptr) return -4;
  if (m < 0) return -2;
  if (n < 0) return -3;
  if (dtype == MAK_F64) *bytes = mak::svd_polar_workspace_bytes<double>(m, n);
  else if (dtype == MAK_C64) *bytes = mak::svd_polar_workspace_bytes<mak::zd>(m, n);
  else return -1;
  return 0;
}

int mak_svd_polar_d(void* s, int m, int n, const double* A, int lda, double* S, double* U, int ldu, double* Vh,
                    int ldvh, int maxiter, void* work, size_t work_bytes, mak_polar_info* info) {
  return mak::svd_polar_driver<double>((hipStream_t)s, m, n, A, lda, S, U, ldu, Vh, ldvh, maxiter, work, work_bytes,
                                       info);
}
int mak_svd_polar_z(void* s, int m, int n, const void* A, int lda, double* S, void* U, int ldu, void* Vh, int ldvh,
                    int maxiter, void* work, size_t work_bytes, mak_polar_info* info) {
  return mak::svd_polar_driver<mak::zd>((hipStream_t)s, m, n, (const mak::zd*)A, lda, S, (mak::zd*)U, ldu,
                                        (mak::zd*)Vh, ldvh, maxiter, work, work_bytes, info);
}

int mak_polar_pushforward_workspace_size(int dtype, int right, int m, int n, size_t* bytes) {
  if (bytes == nullptr) return -5;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (dtype == MAK_F64) *bytes = mak::polar_pushforward_workspace_bytes<double>(m, n, right);
  else if (dtype == MAK_C64) *bytes = mak::polar_pushforward_workspace_bytes<mak::zd>(m, n, right);
  else return -1;
  return 0;
}

int mak_left_polar_pushforward_d(void* s, int m, int n, const double* W, int ldw, const double* P, int ldp,
                                 const double* dA, int ldda, double* dW, int lddw, double* dP, int lddp, void* work,
                                 size_t work_bytes) {
  return mak::polar_pushforward_driver<double>((hipStream_t)s, 0, m, n, W, ldw, P, ldp, dA, ldda, dW, lddw, dP, lddp,
                                               work, work_bytes);
}
int mak_left_polar_pushforward_z(void* s, int m, int n, const void* W, int ldw, const void* P, int ldp,
                                 const void* dA, int ldda, void* dW, int lddw, void* dP, int lddp, void* work,
                                 size_t work_bytes) {
  return mak::polar_pushforward_driver<mak::zd>((hipStream_t)s, 0, m, n, (const mak::zd*)W, ldw, (const mak::zd*)P,
                                                ldp, (const mak::zd*)dA, ldda, (mak::zd*)dW, lddw, (mak::zd*)dP, lddp,
                                                work, work_bytes);
}
int mak_right_polar_pushforward_d(void* s, int m, int n, const double* Wh, int ldwh, const double* P, int ldp,
                                  const double* dA, int ldda, double* dWh, int lddwh, double* dP, int lddp,
                                  void* work, size_t work_bytes) {
  return mak::polar_pushforward_driver<double>((hipStream_t)s, 1, m, n, Wh, ldwh, P, ldp, dA, ldda, dWh, lddwh, dP,
                                               lddp, work, work_bytes);
}
int mak_right_polar_pushforward_z(void* s, int m, int n, const void* Wh, int ldwh, const void* P, int ldp,
                                  const void* dA, int ldda, void* dWh, int lddwh, void* dP, int lddp, void* work,
                                  size_t work_bytes) {
  return mak::polar_pushforward_driver<mak::zd>((hipStream_t)s, 1, m, n, (const mak::zd*)Wh, ldwh,
                                                (const mak::zd*)P, ldp, (const mak::zd*)dA, ldda, (mak::zd*)dWh,
                                                lddwh, (mak::zd*)dP, lddp, work, work_bytes);
}

}  // extern "C"
