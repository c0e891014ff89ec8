// Polar decomposition by QDWH (Nakatsukasa, Bai, Gygi, "Optimizing Halley's
// iteration for computing the matrix polar decomposition", SIAM J. Matrix
// Anal. Appl. 31, 2010; Nakatsukasa and Higham, SIAM J. Sci. Comput. 35, 2013)
// behind left_polar! / right_polar! (MatrixAlgebraKit
// src/implementations/polar.jl:59-125).
//
// X_0 = A / ||A||_F, l_0 a lower bound of sigma_min(X_0).  With the dynamic
// weights (a, b, c) of l_k:
//   c > 100 : X+ = (b/c) X + (a - b/c)/sqrt(c) Q1 Q2^H,  [sqrt(c) X; I] = [Q1; Q2] R
//   c <= 100: X+ = (b/c) X + (a - b/c) X (I + c X^H X)^{-1}
// The QR of the stacked matrix is a CholeskyQR2 (two Gram + Cholesky +
// triangular-solve passes): the stacked matrix has condition number at most
// sqrt(1 + c), so for c <= kMaxCholQrC CholeskyQR2 is as accurate as
// Householder QR and runs entirely on MFMA GEMMs.  l_0 defaults to 1e-10 (the
// iteration converges for any l_0; a smaller true sigma_min only costs
// iterations).  P = herm(W^H A) as in the reference's Newton path
// (src/implementations/polar.jl:108).
#include <algorithm>
#include <cmath>
#include <vector>

#include "mak_common.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

namespace {

constexpr double kMaxCholQrC = 1e14;  // sqrt(1 + c) <= 1e7: (1e7)^2 u << 1
// Reduced-precision bulk (precision 32 / 16, real data):
//   c > kMixedCholMaxC      QR-based steps as in FP64 (the iterate is ill
//                           conditioned; its first CholeskyQR pass stays FP64)
//   100 < c <= kMixedCholMaxC  Cholesky-based steps in FP64: their error
//                           ~ c u_64 is far below the FP32 level of the result
//   c <= 100                the polynomial phase takes over (below)
constexpr double kMixedCholMaxC = 1e5;
constexpr double kPolyTarget = 1e-4;  // polynomial phase ends once 1 - l <= this
constexpr int kMaxPoly = 16;
constexpr int kMaxNs = 6;             // FP64-Gram Newton-Schulz steps, at most
constexpr double kNsLast = 1e-6;      // ||I - X^H X||_F below this: one more step, then stop
constexpr double kNsDone = 1e-14;     // ... below this: no step needed

inline size_t al256(size_t b) { return ((b + 255) / 256) * 256; }

struct PolarLayout {
  size_t offX1, offX2, offQ1, offG, offY1, offY2, offTm, offTmp, offPart, offScal, offInfo, offLo, bytes;
};

template <typename T> PolarLayout polar_layout(int m, int n) {
  PolarLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += al256(bytes); return o; };
  const size_t mn = (size_t)m * n * sizeof(T), nn = (size_t)n * n * sizeof(T);
  L.offX1 = take(mn);
  L.offX2 = take(mn);
  L.offQ1 = take(mn);
  L.offG = take(nn);
  L.offY1 = take(nn);
  L.offY2 = take(nn);
  L.offTm = take(nn);  // n x n scratch of the closed-form second CholeskyQR pass
  L.offTmp = take(((size_t)64 * std::max(n, 1) + (size_t)n * n / 4) * sizeof(T));  // Cholesky scratch
  L.offPart = take(1024 * sizeof(double));
  L.offScal = take(16 * sizeof(double));
  L.offInfo = take(16 * sizeof(int));
  // rounded operand copies of the FP32 / BF16 products (real data only): the
  // largest is (m x n)(n x n)
  L.offLo = take(ScalarTraits<T>::is_complex
                     ? 0
                     : std::max(gemm_prec_workspace_bytes(m, n, n), gemm_prec_workspace_bytes(n, n, m)));
  L.bytes = off;
  return L;
}

void qdwh_params(double l, double& a, double& b, double& c) {
  const double l2 = l * l;
  const double g = std::cbrt(4.0 * (1.0 - l2) / (l2 * l2));
  a = std::sqrt(1.0 + g) + 0.5 * std::sqrt(8.0 - 4.0 * g + 8.0 * (2.0 - l2) / (l2 * std::sqrt(1.0 + g)));
  b = (a - 1.0) * (a - 1.0) / 4.0;
  c = a + b - 1.0;
}

// tri / upperC: structure hints (GemmArgs): the triangular inverse factors
// have exact zeros below the diagonal, and the Cholesky factorisation reads
// only the upper triangle of a Gram matrix.
template <typename T> GemmArgs ga(int M, int N, int K, const T* A, int64_t lda, const T* B, int64_t ldb, T* C,
                                  int64_t ldc, double alpha, double beta, int tri = 0, int upperC = 0) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha_re = alpha; g.beta_re = beta;
  g.ksplit = 1; g.kchunk = K; g.batch = 1;
  g.tri = tri; g.upperC = upperC;
  return g;
}

template <typename T>
__global__ void conj_transpose_scale_kernel(int m, int n, double a, const T* A, int64_t lda, T* B, int64_t ldb) {
  // B (n x m) = a A^H, A m x n
  __shared__ T tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // bx over rows of A, by over cols of A
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int r = bx + tx, c = by + k;
    tile[k][tx] = (r < m && c < n) ? A[r + (int64_t)c * lda] : s_zero(T());
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int r = by + tx, c = bx + k;  // B row = A col, B col = A row
    if (r < n && c < m) B[r + (int64_t)c * ldb] = s_scale(s_conj(tile[tx][k]), a);
  }
}
template <typename T>
void conj_transpose_scale(int m, int n, double a, const T* A, int64_t lda, T* B, int64_t ldb, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  dim3 grid(cdiv(m, 32), cdiv(n, 32));
  hipLaunchKernelGGL((conj_transpose_scale_kernel<T>), grid, dim3(256), 0, s, m, n, a, A, lda, B, ldb);
}

// G := a I + b Gu, Gu the Hermitian matrix whose upper triangle G holds (the
// strict lower part of G is not read).  One workgroup per 32 x 32 block pair
// (rows bx, cols by) with bx <= by: it reads only its upper block and writes
// that block and its mirror, so no two workgroups touch the same element.
template <typename T>
__global__ void __launch_bounds__(256) shift_herm_kernel(T* G, int64_t ldg, int n, double a, double b) {
  __shared__ T tile[32][33];
  const int bx = blockIdx.x, by = blockIdx.y;
  if (bx > by) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int r = bx * 32 + tx, c = by * 32 + k;
    if (r < n && c < n && (bx < by || tx <= k)) tile[tx][k] = G[r + (int64_t)c * ldg];
  }
  __syncthreads();
  if (bx == by) {  // complete the diagonal block from its upper triangle
    for (int k = ty; k < 32; k += 8)
      if (tx > k) tile[tx][k] = s_conj(tile[k][tx]);
    __syncthreads();
  }
  for (int k = ty; k < 32; k += 8) {
    const int r = bx * 32 + tx, c = by * 32 + k;
    if (r < n && c < n) {
      T v = s_scale(tile[tx][k], b);
      if (r == c) v = s_add(v, s_scale(s_one(T()), a));
      G[r + (int64_t)c * ldg] = v;
    }
    const int r2 = by * 32 + tx, c2 = bx * 32 + k;  // mirror: (r2, c2) = conj of (c2, r2)
    if (bx < by && r2 < n && c2 < n) G[r2 + (int64_t)c2 * ldg] = s_scale(s_conj(tile[k][tx]), b);
  }
}

template <typename T> void shift_herm(T* G, int64_t ldg, int n, double a, double b, hipStream_t s) {
  hipLaunchKernelGGL((shift_herm_kernel<T>), dim3(cdiv(n, 32), cdiv(n, 32)), dim3(256), 0, s, G, ldg, n, a, b);
}

// The odd cubic p(x) = alpha x - beta x^3 closest to 1 in the max norm on
// [l, 1]: it equioscillates at l, at its interior maximum x* = sqrt(s / 3)
// and at 1 (s = 1 + l + l^2, p(l) = p(1) gives alpha = beta s), then is
// scaled so that its maximum on [l, 1] is 1.  Returns min p / max p, the
// lower bound of the next iterate's singular values.  l = 1 gives the
// Newton-Schulz step (alpha, beta) = (3/2, 1/2).
double poly_cubic(double l, double& alpha, double& beta) {
  const double s = 1.0 + l + l * l;
  const double xs = std::sqrt(s / 3.0);
  const double b = 2.0 / ((2.0 / 3.0) * s * xs + s - 1.0);
  const double e = 1.0 - b * (s - 1.0);  // 1 - p(l) = 1 - p(1) = p(x*) - 1
  alpha = b * s / (1.0 + e);
  beta = b / (1.0 + e);
  return (1.0 - e) / (1.0 + e);
}

// Second CholeskyQR pass with G = I + E close to the identity: the Cholesky
// factor W = I + U and its inverse follow from E to within O(||E||^3):
// U1 = up(E), U = up(E - U1^H U1), W^{-1} = I - U + U^2, where up() keeps the
// strict upper part and half the (real) diagonal -- from (I + U)^H (I + U) =
// I + E.  Only the upper triangle of G (and of Sub) is read.
template <typename T>
__global__ void near_identity_sumsq_kernel(const T* G, int64_t ldg, int n, double* partial) {
  __shared__ double red[4];
  const int64_t tot = (int64_t)n * n;
  double acc = 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * 256) {
    const int r = (int)(idx % n), c = (int)(idx / n);
    if (r > c) continue;
    T v = G[r + (int64_t)c * ldg];
    if (r == c) v = s_sub(v, s_one(T()));
    acc += (r < c ? 2.0 : 1.0) * s_abs2(v);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ void sum_double_kernel(const double* partial, int nb, double* out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) acc += partial[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = red[0] + red[1] + red[2] + red[3];
}
// U = up(G - I - Sub) (Sub may be null), zero below the diagonal.
template <typename T>
__global__ void up_part_kernel(const T* G, int64_t ldg, const T* Sub, int64_t lds, int n, T* U, int64_t ldu) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  T u = s_zero(T());
  if (r <= c) {
    T v = G[r + (int64_t)c * ldg];
    if (Sub) v = s_sub(v, Sub[r + (int64_t)c * lds]);
    if (r == c) u = s_scale(s_add(s_sub(v, s_one(T())), s_conj(s_sub(v, s_one(T())))), 0.25);
    else u = v;
  }
  U[r + (int64_t)c * ldu] = u;
}
// Y = I - U + U2 (upper), zero below the diagonal; Y may alias U.
template <typename T>
__global__ void inv_near_identity_kernel(T* Y, int64_t ldy, const T* U2, int64_t ldu2, int n) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  T v = s_zero(T());
  if (r <= c) {
    v = s_add(s_neg(Y[r + (int64_t)c * ldy]), U2[r + (int64_t)c * ldu2]);
    if (r == c) v = s_add(v, s_one(T()));
  }
  Y[r + (int64_t)c * ldy] = v;
}

template <typename T> double read_scalar(const double* d, hipStream_t s) {
  double h = 0.0;
  hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, s);
  hipStreamSynchronize(s);
  return h;
}

// The iteration on an m x n (m >= n) matrix held in Xw (ld ldx), already
// scaled to ||X||_2 <= 1.  On return Xw holds the unitary polar factor.
//
// precision 32 / 16 (real data): the QDWH steps run while c > 100 (see
// kMixedCholMaxC), then a polynomial phase on FP32 MFMA replaces the
// Cholesky-based steps: with l the lower bound of sigma_min(X),
//   X+ = X (alpha I - beta X^H X),  (alpha, beta) = poly_cubic(l),
// formed as X + X ((alpha - 1) I - beta G) with G = X^H X (FP32 Gram, upper
// triangle) and the correction product in FP32 (BF16 for precision 16 once
// 1 - l <= 0.1, where the correction is small), the sum in FP64; until
// 1 - l <= kPolyTarget.  FP64 Newton-Schulz steps X+ = X + X (I - X^H X) / 2
// (FP64 Gram, FP32 correction product) then restore FP64 orthogonality.
template <typename T>
int qdwh_iterate(hipStream_t s, int m, int n, T* Xw, int64_t ldx, int precision, int maxiter, double l0,
                 char* ws, const PolarLayout& L, mak_polar_info* info) {
  T* X2 = reinterpret_cast<T*>(ws + L.offX2);
  T* Q1 = reinterpret_cast<T*>(ws + L.offQ1);
  T* G = reinterpret_cast<T*>(ws + L.offG);
  T* Y1 = reinterpret_cast<T*>(ws + L.offY1);
  T* Y2 = reinterpret_cast<T*>(ws + L.offY2);
  T* Tm = reinterpret_cast<T*>(ws + L.offTm);
  T* tmp = reinterpret_cast<T*>(ws + L.offTmp);
  double* part = reinterpret_cast<double*>(ws + L.offPart);
  double* scal = reinterpret_cast<double*>(ws + L.offScal);
  int* dinfo = reinterpret_cast<int*>(ws + L.offInfo);
  void* lo = ws + L.offLo;
  const int64_t ldm = m, ldn = n;
  const int prec = (precision == 32 || precision == 16) ? precision : 64;
  const bool mixed = prec != 64 && !ScalarTraits<T>::is_complex;
  // the second CholeskyQR pass of a QR-based step uses FP32 products in both
  // reduced precisions (the iterate is still ill conditioned there)
  const int prec_qr = mixed ? 32 : 64;

  // three m x n buffers rotate roles: the current iterate and two spares
  T* buf[3] = {Xw, X2, Q1};
  const int64_t ldb[3] = {ldx, ldm, ldm};
  int cur = 0;
  double l = l0;
  const double u = 1.1102230246251565e-16;
  const double tol = std::cbrt(5.0 * u);
  int it = 0, nqr = 0, nch = 0;
  double diff = 1.0;
  int status = 0;
  bool converged = false;
  for (it = 0; it < maxiter; ++it) {
    double a, b, c;
    if (1.0 - l < 1e-15) { a = 3.0; b = 1.0; c = 3.0; }  // Halley's limit (l = 1)
    else qdwh_params(l, a, b, c);
    // keep the stacked matrix within CholeskyQR2's range: a larger l only costs
    // iterations (QDWH converges for any l in (0, 1])
    while (c > kMaxCholQrC) {
      l *= 2.0;
      qdwh_params(l, a, b, c);
    }
    if (mixed && c <= 100.0) break;  // the polynomial phase takes over
    const bool qr_step = c > (mixed ? kMixedCholMaxC : 100.0);
    const int s1 = (cur + 1) % 3, s2 = (cur + 2) % 3;
    T* Xc = buf[cur];
    const int64_t ldc = ldb[cur];
    // G = I + c X^H X  (FP64: it feeds the Cholesky factorisation)
    set_identity<T>(G, ldn, n, n, s);
    gemm<T>(OP_C, OP_N, ga<T>(n, n, m, Xc, ldc, Xc, ldc, G, ldn, c, 1.0, 0, 1), s);
    potrf_upper<T>(G, ldn, n, Y1, ldn, tmp, dinfo, s);
    int nxt;
    if (qr_step) {
      // pass 1: Q1 = sqrt(c) X W1^{-1} (buf[s2]), Q2 = W1^{-1} (Y1).  FP64 in
      // every precision: it feeds the second Gram, which is then I + E with
      // ||E|| ~ (1 + c) u and factors in closed form (below)
      gemm<T>(OP_N, OP_N, ga<T>(m, n, n, Xc, ldc, Y1, ldn, buf[s2], ldb[s2], std::sqrt(c), 0.0, TRI_B_UPPER), s);
      // pass 2: G = Q2^H Q2 + Q1^H Q1 = W2^H W2
      gemm<T>(OP_C, OP_N, ga<T>(n, n, n, Y1, ldn, Y1, ldn, G, ldn, 1.0, 0.0, TRI_A_LOWER | TRI_B_UPPER, 1), s);
      gemm<T>(OP_C, OP_N, ga<T>(n, n, m, buf[s2], ldb[s2], buf[s2], ldb[s2], G, ldn, 1.0, 1.0, 0, 1), s);
      // W2^{-1} (Y2): in closed form when ||G - I||_F <= 1e-6 (||G - I|| is about
      // (1 + c) u after the first pass), else by the Cholesky factorisation
      hipLaunchKernelGGL((near_identity_sumsq_kernel<T>), dim3(256), dim3(256), 0, s, G, ldn, n, part);
      hipLaunchKernelGGL(sum_double_kernel, dim3(1), dim3(256), 0, s, part, 256, scal + 2);
      const double e2 = read_scalar<T>(scal + 2, s);
      if (e2 <= 1e-12) {
        const unsigned nb = (unsigned)cdiv64((int64_t)n * n, 256);
        hipLaunchKernelGGL((up_part_kernel<T>), dim3(nb), dim3(256), 0, s, G, ldn, (const T*)nullptr, ldn, n, Y2, ldn);
        gemm<T>(OP_C, OP_N, ga<T>(n, n, n, Y2, ldn, Y2, ldn, Tm, ldn, 1.0, 0.0, TRI_A_LOWER | TRI_B_UPPER, 1), s);
        hipLaunchKernelGGL((up_part_kernel<T>), dim3(nb), dim3(256), 0, s, G, ldn, (const T*)Tm, ldn, n, Y2, ldn);
        gemm<T>(OP_N, OP_N, ga<T>(n, n, n, Y2, ldn, Y2, ldn, Tm, ldn, 1.0, 0.0, TRI_A_UPPER | TRI_B_UPPER), s);
        hipLaunchKernelGGL((inv_near_identity_kernel<T>), dim3(nb), dim3(256), 0, s, Y2, ldn, (const T*)Tm, ldn, n);
      } else {
        potrf_upper<T>(G, ldn, n, Y2, ldn, tmp, dinfo, s);
      }
      // Q1' = Q1 W2^{-1} (buf[s1]), Q2' = Q2 W2^{-1} (G)
      gemm_prec<T>(prec_qr, OP_N, OP_N, ga<T>(m, n, n, buf[s2], ldb[s2], Y2, ldn, buf[s1], ldb[s1], 1.0, 0.0, TRI_B_UPPER), lo, s);
      gemm<T>(OP_N, OP_N, ga<T>(n, n, n, Y1, ldn, Y2, ldn, G, ldn, 1.0, 0.0, TRI_A_UPPER | TRI_B_UPPER), s);
      // X+ = (b/c) X + (a - b/c)/sqrt(c) Q1' Q2'^H  (buf[s2])
      copy_mat<T>(m, n, Xc, ldc, buf[s2], ldb[s2], s);
      gemm_prec<T>(prec_qr, OP_N, OP_C,
                   ga<T>(m, n, n, buf[s1], ldb[s1], G, ldn, buf[s2], ldb[s2], (a - b / c) / std::sqrt(c), b / c,
                         TRI_B_LOWER), lo, s);
      nxt = s2;
      ++nqr;
    } else {
      // Z^{-1} = W^{-1} W^{-H}:  Tm = X W^{-1} (buf[s2]);
      // X+ = (b/c) X + (a - b/c) Tm W^{-H}  (buf[s1]); two triangular products
      // instead of W^{-1} W^{-H} followed by a full product (FP64: a
      // reduced-precision run only gets here with c > 100)
      gemm<T>(OP_N, OP_N, ga<T>(m, n, n, Xc, ldc, Y1, ldn, buf[s2], ldb[s2], 1.0, 0.0, TRI_B_UPPER), s);
      copy_mat<T>(m, n, Xc, ldc, buf[s1], ldb[s1], s);
      gemm<T>(OP_N, OP_C, ga<T>(m, n, n, buf[s2], ldb[s2], Y1, ldn, buf[s1], ldb[s1], a - b / c, b / c, TRI_B_LOWER),
              s);
      nxt = s1;
      ++nch;
    }
    // convergence: ||X+ - X||_F / ||X+||_F, plus the Cholesky status
    sumsq<T>(buf[nxt], ldb[nxt], Xc, ldc, m, n, part, scal, s);
    sumsq<T>(buf[nxt], ldb[nxt], nullptr, 0, m, n, part + 512, scal + 1, s);
    double hs[2] = {0, 0};
    int pinfo = 0;
    hipMemcpyAsync(hs, scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(&pinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, s);
    hipStreamSynchronize(s);
    cur = nxt;
    if (pinfo != 0) { status = 2; ++it; break; }
    diff = std::sqrt(hs[0]) / std::max(std::sqrt(hs[1]), 1e-300);
    l = l * (a + b * l * l) / (1.0 + c * l * l);
    if (1.0 - l < 10.0 * u && diff <= (prec == 64 ? tol : std::sqrt(tol))) { ++it; converged = true; break; }
    if (it + 1 == maxiter) status = 6;  // not converged within maxiter (a warning)
  }
  int npoly = 0, nns = 0;
  if (mixed && status == 0) {
    // polynomial phase (no host synchronisation: the coefficients follow from l)
    while (!converged && 1.0 - l > kPolyTarget && npoly < kMaxPoly) {
      double al, be;
      const double lnext = poly_cubic(l, al, be);
      const int nx = (cur + 1) % 3;
      gemm_prec<T>(32, OP_C, OP_N, ga<T>(n, n, m, buf[cur], ldb[cur], buf[cur], ldb[cur], G, ldn, 1.0, 0.0, 0, 1), lo,
                   s);
      shift_herm<T>(G, ldn, n, al - 1.0, -be, s);
      copy_mat<T>(m, n, buf[cur], ldb[cur], buf[nx], ldb[nx], s);
      const int pp = (prec == 16 && 1.0 - l <= 0.1) ? 16 : 32;
      gemm_prec<T>(pp, OP_N, OP_N, ga<T>(m, n, n, buf[cur], ldb[cur], G, ldn, buf[nx], ldb[nx], 1.0, 1.0), lo, s);
      cur = nx;
      l = lnext;
      ++npoly;
    }
    // FP64 Newton-Schulz refinement X <- X (3I - X^H X) / 2 = X + X E / 2,
    // E = I - X^H X from an FP64 Gram; E is small, so its FP32 product costs
    // about |E| 6e-8 absolute, and the sum with X is formed in FP64.  Each
    // step reads ||E||_F: the step after ||E||_F <= kNsLast is the last
    // (the error then falls to about 0.75 ||E||^2).
    for (int k = 0; k < kMaxNs; ++k) {
      const int nx = (cur + 1) % 3;
      gemm<T>(OP_C, OP_N, ga<T>(n, n, m, buf[cur], ldb[cur], buf[cur], ldb[cur], G, ldn, 1.0, 0.0, 0, 1), s);
      hipLaunchKernelGGL((near_identity_sumsq_kernel<T>), dim3(256), dim3(256), 0, s, G, ldn, n, part);
      hipLaunchKernelGGL(sum_double_kernel, dim3(1), dim3(256), 0, s, part, 256, scal + 4);
      const double e = std::sqrt(read_scalar<T>(scal + 4, s));
      diff = e;
      if (!std::isfinite(e)) { status = 5; break; }
      if (e <= kNsDone) break;
      shift_herm<T>(G, ldn, n, 1.0, -1.0, s);
      copy_mat<T>(m, n, buf[cur], ldb[cur], buf[nx], ldb[nx], s);
      gemm_prec<T>(32, OP_N, OP_N, ga<T>(m, n, n, buf[cur], ldb[cur], G, ldn, buf[nx], ldb[nx], 0.5, 1.0), lo, s);
      cur = nx;
      ++nns;
      if (e <= kNsLast) break;
      if (k + 1 == kMaxNs) status = 6;
    }
  }
  if (cur != 0) copy_mat<T>(m, n, buf[cur], ldb[cur], Xw, ldx, s);
  if (info) {
    info->iterations = it;
    info->qr_steps = nqr;
    info->chol_steps = nch;
    info->poly_steps = npoly;
    info->ns_steps = nns;
    info->last_diff = diff;
    info->l_final = l;
  }
  return status;
}

// B = pseudo-random entries in [-1, 1) (a fixed hash of the position): the
// direction that completes a partial isometry (below).
__device__ inline double hash_unit(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return (double)x * (2.0 / 4294967296.0) - 1.0;
}
template <typename T> __global__ void hash_fill_kernel(int m, int n, T* B, int64_t ldb);
template <> __global__ void hash_fill_kernel<double>(int m, int n, double* B, int64_t ldb) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  B[r + (int64_t)c * ldb] = hash_unit((uint32_t)(r * 2654435761U) ^ (uint32_t)(c * 40503U + 0x9e3779b9U));
}
template <> __global__ void hash_fill_kernel<zd>(int m, int n, zd* B, int64_t ldb) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  const uint32_t h = (uint32_t)(r * 2654435761U) ^ (uint32_t)(c * 40503U + 0x9e3779b9U);
  B[r + (int64_t)c * ldb] = mkz(hash_unit(h), hash_unit(h ^ 0x5bd1e995U));
}

// The polar factor of the m x n (m >= n) iterate X (already scaled to
// ||X||_2 <= 1), in place.  QDWH maps zero singular values to zero, so for a
// rank-deficient input it returns a partial isometry W0 = U_r V_r^H (+ the
// images of the null directions); the reference's SVD route returns an
// isometry (src/implementations/polar.jl:59-70).  When ||W0||_F^2 < n (1 -
// 1e-8), the iteration is run once more on
//   Z = W0 + (I - W0 W0^H) B (I - W0^H W0),   B pseudo-random:
// the added term lives in range(U_r)^perp x null(A) and is generically of
// full rank there, so the polar factor of Z keeps W0 on range(A^H) and maps
// null(A) isometrically into range(U_r)^perp -- an isometry W with A = W P.
// *completed receives 1 when that second run happened.
template <typename T>
int polar_isometry(hipStream_t s, int m, int n, T* X, int64_t ldx, int precision, int maxiter, double l0,
                   char* ws, const PolarLayout& L, mak_polar_info* info, int* completed) {
  *completed = 0;
  int st = qdwh_iterate<T>(s, m, n, X, ldx, precision, maxiter, l0, ws, L, info);
  if (st != 0 && st != 6) return st;
  double* part = reinterpret_cast<double*>(ws + L.offPart);
  double* scal = reinterpret_cast<double*>(ws + L.offScal);
  sumsq<T>(X, ldx, nullptr, 0, m, n, part, scal + 3, s);
  const double nrm2 = read_scalar<T>(scal + 3, s);
  if (!(nrm2 < n * (1.0 - 1e-8))) return st;
  mak_polar_info first{};
  if (info) first = *info;
  T* B = reinterpret_cast<T*>(ws + L.offX2);
  T* C1 = reinterpret_cast<T*>(ws + L.offQ1);
  T* Gn = reinterpret_cast<T*>(ws + L.offG);
  T* Gx = reinterpret_cast<T*>(ws + L.offY1);
  const int64_t ldm = m, ldn = n;
  hipLaunchKernelGGL((hash_fill_kernel<T>), dim3((unsigned)cdiv64((int64_t)m * n, 256)), dim3(256), 0, s, m, n, B,
                     ldm);
  // C1 = (I - X X^H) B
  gemm<T>(OP_C, OP_N, ga<T>(n, n, m, X, ldx, B, ldm, Gn, ldn, 1.0, 0.0), s);
  copy_mat<T>(m, n, B, ldm, C1, ldm, s);
  gemm<T>(OP_N, OP_N, ga<T>(m, n, n, X, ldx, Gn, ldn, C1, ldm, -1.0, 1.0), s);
  // B = C1 (I - X^H X);  X += B;  X /= ||X||_F
  gemm<T>(OP_C, OP_N, ga<T>(n, n, m, X, ldx, X, ldx, Gx, ldn, 1.0, 0.0), s);
  copy_mat<T>(m, n, C1, ldm, B, ldm, s);
  gemm<T>(OP_N, OP_N, ga<T>(m, n, n, C1, ldm, Gx, ldn, B, ldm, -1.0, 1.0), s);
  axpby<T>(m, n, 1.0, B, ldm, 1.0, X, ldx, s);
  sumsq<T>(X, ldx, nullptr, 0, m, n, part, scal + 3, s);
  const double z = std::sqrt(read_scalar<T>(scal + 3, s));
  if (!(z > 0.0) || !std::isfinite(z)) return 5;
  axpby<T>(m, n, 0.0, X, ldx, 1.0 / z, X, ldx, s);
  *completed = 1;
  st = qdwh_iterate<T>(s, m, n, X, ldx, precision, maxiter, l0, ws, L, info);
  if (info) {
    info->iterations += first.iterations;
    info->qr_steps += first.qr_steps;
    info->chol_steps += first.chol_steps;
    info->ns_steps += first.ns_steps;
    info->poly_steps += first.poly_steps;
  }
  return st;
}

// m x n identity (ones on the main diagonal).
template <typename T> void eye_rect(T* X, int64_t ldx, int m, int n, hipStream_t s) { set_identity<T>(X, ldx, m, n, s); }

}  // namespace

template <typename T> size_t polar_workspace_bytes(int m, int n) {
  const int mm = std::max(m, n), nn = std::min(m, n);
  return polar_layout<T>(mm, nn).bytes;
}

template <typename T>
int polar_driver(hipStream_t s, int right, int m, int n, const T* A, int64_t lda, T* W, int64_t ldw, T* P,
                 int64_t ldp, int precision, int maxiter, double l0, void* work, size_t work_bytes,
                 mak_polar_info* info) {
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (!right && m < n) return -3;  // left_polar! needs m >= n (src/implementations/polar.jl:9-10)
  if (right && n < m) return -4;   // right_polar! needs n >= m (:21-22)
  if (lda < std::max(1, m)) return -6;
  if (ldw < std::max(1, m)) return -8;
  const int k = std::min(m, n), big = std::max(m, n);
  if (P && ldp < std::max(1, k)) return -10;
  if (precision != 64 && precision != 32 && precision != 16) return -11;
  if (ScalarTraits<T>::is_complex && precision != 64) return -11;
  if (maxiter <= 0) maxiter = 20;
  if (!(l0 > 0.0 && l0 < 1.0)) l0 = 1e-10;
  const size_t need = polar_workspace_bytes<T>(m, n);
  if (work == nullptr || work_bytes < need) return -14;
  if (k == 0) return 0;
  PolarLayout L = polar_layout<T>(big, k);
  char* ws = reinterpret_cast<char*>(work);
  double* part = reinterpret_cast<double*>(ws + L.offPart);
  double* scal = reinterpret_cast<double*>(ws + L.offScal);
  // alpha = ||A||_F
  sumsq<T>(A, lda, nullptr, 0, m, n, part, scal, s);
  const double alpha = std::sqrt(read_scalar<T>(scal, s));
  if (info) { *info = mak_polar_info{}; info->alpha = alpha; info->l0 = l0; }
  if (!std::isfinite(alpha)) return 5;
  if (alpha == 0.0) {
    // A = 0: every isometry is a polar factor; the SVD route gives U = I,
    // V = I, so W = I (m x n) and P = 0
    eye_rect<T>(W, ldw, m, n, s);
    if (P) fill_zero<T>(P, ldp, k, k, s);
    return hipGetLastError() == hipSuccess ? 0 : 1;
  }
  int st;
  int completed = 0;
  // P in FP32 at reduced precision: its rounding (~1e-7 relative) is below
  // the polar factor's own error there
  const int pprec = precision == 64 ? 64 : 32;
  if (!right) {
    // X0 = A / alpha, iterated in W
    copy_mat<T>(m, n, A, lda, W, ldw, s);
    axpby<T>(m, n, 0.0, A, lda, 1.0 / alpha, W, ldw, s);
    st = polar_isometry<T>(s, m, n, W, ldw, precision, maxiter, l0, ws, L, info, &completed);
    if ((st == 0 || st == 6) && P) {
      // P = herm(W^H A)
      gemm_prec<T>(pprec, OP_C, OP_N, ga<T>(n, n, m, W, ldw, A, lda, P, ldp, 1.0, 0.0), ws + L.offLo, s);
      herm_part<T>(P, ldp, n, s);
    }
  } else {
    // right_polar!(A) via the left polar of A^H (n x m): A^H = W' P'
    T* Xh = reinterpret_cast<T*>(ws + L.offX1);
    conj_transpose_scale<T>(m, n, 1.0 / alpha, A, lda, Xh, n, s);
    st = polar_isometry<T>(s, n, m, Xh, n, precision, maxiter, l0, ws, L, info, &completed);
    if (st == 0 || st == 6) {
      conj_transpose_scale<T>(n, m, 1.0, Xh, n, W, ldw, s);   // Wh = (W')^H
      if (P) {
        // P = herm(A Wh^H)  (src/implementations/polar.jl:122)
        gemm_prec<T>(pprec, OP_N, OP_C, ga<T>(m, m, n, A, lda, W, ldw, P, ldp, 1.0, 0.0), ws + L.offLo, s);
        herm_part<T>(P, ldp, m, s);
      }
    }
  }
  if (info) info->completed = completed;
  if (hipGetLastError() != hipSuccess) return 1;
  return st;
}

template size_t polar_workspace_bytes<double>(int, int);
template size_t polar_workspace_bytes<zd>(int, int);
template int polar_driver<double>(hipStream_t, int, int, int, const double*, int64_t, double*, int64_t, double*,
                                  int64_t, int, int, double, void*, size_t, mak_polar_info*);
template int polar_driver<zd>(hipStream_t, int, int, int, const zd*, int64_t, zd*, int64_t, zd*, int64_t, int, int,
                              double, void*, size_t, mak_polar_info*);

}  // namespace mak

extern "C" {

int mak_polar_workspace_size(int dtype, int m, int n, size_t* bytes) {
  if (bytes == nullptr) return -4;
  if (m < 0) return -2;
  if (n < 0) return -3;
  if (dtype == MAK_F64) *bytes = mak::polar_workspace_bytes<double>(m, n);
  else if (dtype == MAK_C64) *bytes = mak::polar_workspace_bytes<mak::zd>(m, n);
  else return -1;
  return 0;
}

int mak_left_polar_d(void* s, int m, int n, const double* A, int lda, double* W, int ldw, double* P, int ldp,
                     int precision, int maxiter, double l0, void* work, size_t work_bytes, mak_polar_info* info) {
  return mak::polar_driver<double>((hipStream_t)s, 0, m, n, A, lda, W, ldw, P, ldp, precision, maxiter, l0, work,
                                   work_bytes, info);
}
int mak_left_polar_z(void* s, int m, int n, const void* A, int lda, void* W, int ldw, void* P, int ldp,
                     int precision, int maxiter, double l0, void* work, size_t work_bytes, mak_polar_info* info) {
  return mak::polar_driver<mak::zd>((hipStream_t)s, 0, m, n, (const mak::zd*)A, lda, (mak::zd*)W, ldw,
                                    (mak::zd*)P, ldp, precision, maxiter, l0, work, work_bytes, info);
}
int mak_right_polar_d(void* s, int m, int n, const double* A, int lda, double* P, int ldp, double* Wh, int ldw,
                      int precision, int maxiter, double l0, void* work, size_t work_bytes, mak_polar_info* info) {
  return mak::polar_driver<double>((hipStream_t)s, 1, m, n, A, lda, Wh, ldw, P, ldp, precision, maxiter, l0, work,
                                   work_bytes, info);
}
int mak_right_polar_z(void* s, int m, int n, const void* A, int lda, void* P, int ldp, void* Wh, int ldw,
                      int precision, int maxiter, double l0, void* work, size_t work_bytes, mak_polar_info* info) {
  return mak::polar_driver<mak::zd>((hipStream_t)s, 1, m, n, (const mak::zd*)A, lda, (mak::zd*)Wh, ldw,
                                    (mak::zd*)P, ldp, precision, maxiter, l0, work, work_bytes, info);
}

}  // extern "C"
