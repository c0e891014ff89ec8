// Tall panel factorisation by CholeskyQR2 with Householder reconstruction,
// gated on the device by a conditioning check, with the Householder TSQR
// path as the fallback (the fallback's launches are always enqueued and exit
// at once unless the gate opens; no host synchronisation).
//
//   pass 1: G1 = X^H X, G1 = R1^H R1            (gram_chol_kernel)
//           Q1 = X R1^{-1}                       (MFMA GEMM)
//   pass 2: G2 = Q1^H Q1, G2 = R2^H R2, accepted only if ||G2 - I||_F <= 0.1
//           (then kappa(Q1) <= 1.1 and the second pass is orthogonal to
//           working precision; Yamamoto, Nakatsukasa, Yanagisawa, Fukaya,
//           ETNA 44, 2015), R = R2 R1, Q = Q1 R2^{-1}
//   reconstruction (Ballard, Demmel, Grigori, Jacquelin, Nguyen, Solomonik,
//   IPDPS 2014): Q - [D; 0] = V U, LU without pivoting with
//   d_j = -sign(pivot_j); T = -U D^H V1^{-H}, the panel's R = D R and
//   V = [V1; Q_rest U^{-1}] -- the compact WY form LAPACK's geqrt produces.
// A panel whose first Gram is not numerically positive definite or whose
// second Gram fails the check (kappa(X) beyond ~1e7) is factored by the TSQR
// path instead, so rank-deficient inputs keep Householder's stability.
#include <algorithm>

#include "mak_common.hpp"
#include "mak_dense32.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

namespace {

constexpr int NB = 32;

// G = X^H X over the row blocks of this workgroup -> partial[blockIdx.x];
// the last workgroup to finish sums the partials in a fixed order, factors G
// and writes the factors.  pass 1: always runs, sets *flag = (G not PD).
// pass 2: runs only if *flag == 0; sets *flag when G is not PD or far from I;
// otherwise writes Rout = W2 * R1 and Rinv = W2^{-1}.
template <typename T>
__global__ void __launch_bounds__(256) gram_chol_kernel(const T* X, int64_t ldx, int H, int nbc, T* partial,
                                                        int* counter, int pass, const T* R1, T* Rout, T* Rinv,
                                                        int* flag) {
  __shared__ T Xs[64][NB + 1];
  __shared__ T G[NB][NB + 1];
  __shared__ T Yl[NB][NB + 1];
  __shared__ T Rl[NB][NB + 1];
  __shared__ int last_s, bad_s;
  __shared__ double dev_s[4];
  if (pass == 2 && *flag != 0) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int ti = w & 1, tj = w >> 1;
  constexpr bool CPLX = ScalarTraits<T>::is_complex;
  v4d cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  for (int r0 = blockIdx.x * 64; r0 < H; r0 += gridDim.x * 64) {
    __syncthreads();
    for (int e = tid; e < 64 * NB; e += 256) {
      const int r = e % 64, c = e / 64;
      Xs[r][c] = (r0 + r < H && c < nbc) ? X[(r0 + r) + (int64_t)c * ldx] : s_zero(T());
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 64; kk += 4) {
      const T a = Xs[kk + lk][ti * 16 + lr];  // (X^H)(i, k) = conj(X(k, i))
      const T b = Xs[kk + lk][tj * 16 + lr];
      if constexpr (CPLX) {
        cr = mfma_f64(s_re(a), s_re(b), cr);
        cr = mfma_f64(s_im(a), s_im(b), cr);
        ci = mfma_f64(s_re(a), s_im(b), ci);
        ci = mfma_f64(-s_im(a), s_re(b), ci);
      } else {
        cr = mfma_f64(a, b, cr);
      }
    }
  }
  T* P = partial + (int64_t)blockIdx.x * NB * NB;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = ti * 16 + mfma_f64_row(lane, r), j = tj * 16 + lr;
    if constexpr (CPLX) P[i + j * NB] = mkz(cr[r], ci[r]);
    else P[i + j * NB] = cr[r];
  }
  // last workgroup: reduce, factor
  __threadfence();
  __syncthreads();
  if (tid == 0) last_s = (atomicAdd(counter, 1) == (int)gridDim.x - 1);
  __syncthreads();
  if (!last_s) return;
  __threadfence();
  double dev = 0.0;
  for (int e = tid; e < NB * NB; e += 256) {
    const int i = e % NB, j = e / NB;
    // eight independent partial sums keep eight loads in flight
    T a8[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a8[u] = s_zero(T());
    const int nb = (int)gridDim.x;
    int b = 0;
    for (; b + 8 <= nb; b += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u) a8[u] = s_add(a8[u], partial[(int64_t)(b + u) * NB * NB + e]);
    }
    for (; b < nb; ++b) a8[0] = s_add(a8[0], partial[(int64_t)b * NB * NB + e]);
    T acc = s_add(s_add(s_add(a8[0], a8[1]), s_add(a8[2], a8[3])), s_add(s_add(a8[4], a8[5]), s_add(a8[6], a8[7])));
    if (i >= nbc || j >= nbc) acc = (i == j) ? s_one(T()) : s_zero(T());
    G[i][j] = acc;
    const T dlt = s_sub(acc, (i == j) ? s_one(T()) : s_zero(T()));
    dev += s_abs2(dlt);
  }
  dev = wave_sum(dev);
  if (lane == 0) dev_s[w] = dev;
  if (tid == 0) *counter = 0;  // ready for the next use
  __syncthreads();
  if (w == 0) {
    const int b = wave_chol_inv32<T>(G, Yl);
    if (lane == 0) bad_s = b;
  }
  __syncthreads();
  const double devF2 = dev_s[0] + dev_s[1] + dev_s[2] + dev_s[3];
  const bool fail = bad_s != 0 || (pass == 2 && !(devF2 <= 0.01));
  if (pass == 2) {
    if (fail) {
      if (tid == 0) *flag = 1;
      return;
    }
    for (int e = tid; e < NB * NB; e += 256) Rl[e % NB][e / NB] = R1[e];
    __syncthreads();
    mm32<T>(Rl, G, false, Rl, false, 1.0, false);  // R = W2 R1 (W2 in G)
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      Rout[e] = Rl[i][j];
      Rinv[e] = Yl[i][j];
    }
  } else {
    if (tid == 0) *flag = fail ? 1 : 0;
    if (fail) return;
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      Rout[e] = G[i][j];
      Rinv[e] = Yl[i][j];
    }
  }
}

// In-wave LU of a 32 x 32 block without pivoting, with the sign choice
// d_j = -sign(M(j, j)) subtracted from each pivot first (|pivot| >= 1 for the
// top of an orthonormal Q).  Lane c keeps column c: U(r, c) for r <= c,
// L(r, c) (unit lower) for r > c; other lanes' entries come from lane
// broadcasts.  dv receives d.
template <typename T, int J>
__device__ __forceinline__ void lu32_step(T (&s)[32], int c, T* dv) {
  const T piv = rdl(s[J], J);
  const T d = s_neg(sign_safe(piv));
  const T p = s_sub(piv, d);
  if (c == 0) dv[J] = d;
  const T inv = s_div(s_one(T()), p);
  const T mjc = s[J];
  s[J] = sel(c == J, p, mjc);
#pragma unroll
  for (int r = J + 1; r < 32; ++r) {
    const T lrj = s_mul(rdl(s[r], J), inv);  // L(r, J), from lane J before its scaling
    s[r] = sel(c == J, lrj, sel(c > J, s_sub(s[r], s_mul(lrj, mjc)), s[r]));
  }
}
template <typename T, int... Js>
__device__ __forceinline__ void lu32_steps(T (&s)[32], int c, T* dv, std::integer_sequence<int, Js...>) {
  (lu32_step<T, Js>(s, c, dv), ...);
}

// Householder reconstruction of a CholeskyQR2 panel (see the file comment).
// Runs only if *flag == 0.  Q: H x nbc (ld ldq) explicit Q; R: 32 x 32.
// Writes T (32 x 32) to Tout, U^{-1} (32 x 32) to Uinv, and the top nbc rows
// of the panel Ap: D R on and above the diagonal, V1 below.
template <typename T>
__global__ void __launch_bounds__(256) cholqr_recon_kernel(const T* Q, int64_t ldq, int nbc, const T* R, T* Tout,
                                                           T* Uinv, T* Ap, int64_t lda, const int* flag) {
  __shared__ T Ls[NB][NB + 1];   // L (unit lower), U on and above
  __shared__ T Li[NB][NB + 1];   // L^{-1}
  __shared__ T Ms[NB][NB + 1];   // U D^H, then T
  __shared__ T Ui[NB][NB + 1];   // U^{-1}
  __shared__ T dv[NB];
  if (*flag != 0) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int c = lane & 31;
  if (w == 0) {
    T s[32];
#pragma unroll
    for (int r = 0; r < 32; ++r) {
      const T q = Q[(int64_t)min(r, nbc - 1) + (int64_t)min(c, nbc - 1) * ldq];
      s[r] = sel(r < nbc && c < nbc, q, sel(r == c, s_one(T()), s_zero(T())));
    }
    lu32_steps<T>(s, c, dv, std::make_integer_sequence<int, 32>{});
#pragma unroll
    for (int r = 0; r < 32; ++r)
      if (lane < 32) Ls[r][c] = s[r];
    // U^{-1} (upper): rows from the bottom, U(r, k) of lane k
    T y[32];
#pragma unroll
    for (int r = 0; r < 32; ++r) y[r] = s_zero(T());
    trinv32_rows<T>(y, s, c, std::make_integer_sequence<int, 32>{});
#pragma unroll
    for (int r = 0; r < 32; ++r)
      if (lane < 32) Ui[r][c] = y[r];
  }
  __syncthreads();
  if (w == 1) {
    // L^{-1} (unit lower), column c: z(c) = 1, z(r) = -sum_{c<=k<r} L(r, k) z(k)
    T z[32];
#pragma unroll
    for (int r = 0; r < 32; ++r) {
      T acc = s_zero(T());
#pragma unroll
      for (int k = 0; k < r; ++k) acc = s_fma(Ls[r][k], z[k], acc);   // z(k) = 0 for k < c
      z[r] = sel(r == c, s_one(T()), sel(r > c, s_neg(acc), s_zero(T())));
    }
#pragma unroll
    for (int r = 0; r < 32; ++r)
      if (lane < 32) Li[r][c] = z[r];
  }
  for (int e = tid; e < NB * NB; e += 256) {
    const int r = e % NB, cc = e / NB;
    Ms[r][cc] = (r <= cc) ? s_mul(Ls[r][cc], s_conj(dv[cc])) : s_zero(T());
  }
  __syncthreads();
  mm32<T>(Ms, Ms, false, Li, true, -1.0, false);  // T = -(U D^H) L^{-H}
  for (int e = tid; e < NB * NB; e += 256) {
    const int r = e % NB, cc = e / NB;
    const bool in = r < nbc && cc < nbc;
    Tout[e] = (in && r <= cc) ? Ms[r][cc] : s_zero(T());
    Uinv[e] = (in && r <= cc) ? Ui[r][cc] : s_zero(T());
    if (in) Ap[r + (int64_t)cc * lda] = (r <= cc) ? s_mul(dv[r], R[r + cc * NB]) : Ls[r][cc];
  }
}

template <typename T> GemmArgs gg(int M, int N, int K, const T* A, int64_t lda, const T* B, int64_t ldb, T* C,
                                  int64_t ldc, const int* gate) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha_re = 1.0; g.beta_re = 0.0;
  g.ksplit = 1; g.kchunk = K; g.batch = 1;
  g.gate = gate; g.gate_want = 0;
  return g;
}

}  // namespace

__global__ void zero_ints_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
void zero_ints(int* p, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(zero_ints_kernel, dim3(cdiv(n, 64)), dim3(64), 0, s, p, n);
}

size_t cholqr_workspace_elems(int m) {
  // Q1, Q2 (m x 32), 256 partial Grams, R1, R1inv, R, R2inv, Uinv
  return (size_t)2 * std::max(m, 1) * NB + (size_t)256 * NB * NB + (size_t)5 * NB * NB;
}

template <typename T>
void cholqr_panel(T* Ap, int64_t lda, int H, int nbc, T* Tpan, T* cw, int* flag, int* counter, hipStream_t s) {
  T* Q1 = cw;
  T* Q2 = Q1 + (size_t)H * NB;
  T* part = Q2 + (size_t)H * NB;
  T* R1 = part + (size_t)256 * NB * NB;
  T* R1i = R1 + NB * NB;
  T* R = R1i + NB * NB;
  T* R2i = R + NB * NB;
  T* Ui = R2i + NB * NB;
  const int nblk = std::max(1, std::min(64, cdiv(H, 256)));  // 64 partials at most: a short final sum
  hipLaunchKernelGGL((gram_chol_kernel<T>), dim3(nblk), dim3(256), 0, s, Ap, lda, H, nbc, part, counter, 1,
                     (const T*)nullptr, R1, R1i, flag);
  gemm<T>(OP_N, OP_N, gg<T>(H, nbc, nbc, Ap, lda, R1i, NB, Q1, H, flag), s);
  hipLaunchKernelGGL((gram_chol_kernel<T>), dim3(nblk), dim3(256), 0, s, Q1, (int64_t)H, H, nbc, part, counter, 2,
                     (const T*)R1, R, R2i, flag);
  gemm<T>(OP_N, OP_N, gg<T>(H, nbc, nbc, Q1, H, R2i, NB, Q2, H, flag), s);
  hipLaunchKernelGGL((cholqr_recon_kernel<T>), dim3(1), dim3(256), 0, s, (const T*)Q2, (int64_t)H, nbc,
                     (const T*)R, Tpan, Ui, Ap, lda, (const int*)flag);
  if (H > nbc) gemm<T>(OP_N, OP_N, gg<T>(H - nbc, nbc, nbc, Q2 + nbc, H, Ui, NB, Ap + nbc, lda, flag), s);
}

template void cholqr_panel<double>(double*, int64_t, int, int, double*, double*, int*, int*, hipStream_t);
template void cholqr_panel<zd>(zd*, int64_t, int, int, zd*, zd*, int*, int*, hipStream_t);

}  // namespace mak
