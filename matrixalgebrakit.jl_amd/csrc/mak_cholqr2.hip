// Whole-matrix CholeskyQR2 for qr_compact!(A, (Q, R); positive = true), m >= n
// (the result is the unique QR factorisation with a positive real diagonal of
// R, the one MatrixAlgebraKit's qr_householder! + gaugefix! returns,
// src/implementations/qr.jl:132-188, src/common/gauge.jl:16-25):
//
//   G1 = A^H A                       upper triangle, split-K MFMA product
//   G1 = R1^H R1, X1 = R1^{-1}       persistent tile Cholesky (mak_chol.hip)
//   Q1 = A X1                        triangular MFMA product tasks inside the
//                                    Cholesky launch, as X1 fills
//   G2 = Q1^H Q1 = I + E             upper triangle, split-K MFMA product
//   W2 = I + U with (I + U)^H (I + U) = I + E, in closed form:
//        U1 = up(E), U = up(E - U1^H U1), W2^{-1} = I - U + U^2
//        (up(): strict upper part and half the real diagonal; the error is
//        O(||E||^3); when ||E||_F <= 1e-9 the second-order terms are below
//        1e-18 and U = U1, W2^{-1} = I - U are taken directly)
//   Q = Q1 W2^{-1},  R = W2 R1
//
// The call is accepted on the device when the Cholesky factorisation
// succeeded and ||E||_F <= 1e-6: then kappa(Q1) <= 1 + 1e-6 and Q is
// orthonormal to working precision (Yamamoto, Nakatsukasa, Yanagisawa,
// Fukaya, ETNA 44, 2015).  A is only read, so the caller falls back to the
// Householder path for a rejected input (kappa(A) beyond about 1e5).
// Every large product is an FP64 MFMA GEMM; the only sequential part is the
// n x n Cholesky factorisation.
#include <algorithm>
#include <cmath>

#include "mak_common.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

namespace {

inline size_t al256(size_t b) { return ((b + 255) / 256) * 256; }

constexpr int kRedWgs = 1024;  // workgroups of the second pass's sweep (and its ||E|| partials)

// Largest ||E||_F for which the final Q = Q1 + Q1 D (D = W2^{-1} - I, upper
// triangular, ||D||_F ~ ||E||_F) takes the product Q1 D from FP32 MFMA with
// FP32 accumulation (complex data through its real embedding, where each
// entry of the product is a sum of 2n real terms): its rounding error is at
// most about
// n u_32 ||D||_F <= 1e3 * 6e-8 * 1e-10 = 6e-15 per row even in the worst case
// of the accumulation (typically sqrt(n) u_32 ||E|| ~ 2e-16), against the
// FP64 product's own 1e-16.  Gaussian inputs have ||E||_F ~ 1e-13.
constexpr double kLoCorrMaxE = 1e-10;
// ... and on BF16 MFMA (FP32 accumulation) when ||E||_F <= kBf16CorrMaxE: the
// correction Q1 D has ||D||_F ~ ||E||_F / 2, and rounding both operands to
// BF16 (relative 2^-8) errs by at most 2^-8 ||Q1||_F ||D||_F ~ 2e-3 sqrt(n)
// ||E||_F <= 1e-16 sqrt(n): no more than Q's own FP64 rounding (u sqrt(n)),
// and in practice (errors of random sign) some thirty times less.
constexpr double kBf16CorrMaxE = 5e-14;

struct CQ2Layout {
  size_t offQ1, offG, offE, offX, offU, offW, offT, offSl, offFlags, offPart, offSt, offXc, bytes;
  int ks;    // split-K slices of the Gram products
  int ldxc;  // leading dimension of the Cholesky's copy of X1
};

// Split-K slices of an upper-triangle n x n Gram over m rows on 128 x 128
// tiles: two workgroups per CU (about 512; one per CU leaves the MFMA waiting
// on loads), slices of at least 256 rows.  At n = 1024, m = 8192: 36 tiles x
// 14 slices, 190 us (45 TF/s on the upper-triangle count) against 300 us for
// 64 x 64 tiles or one workgroup per CU.
constexpr int kGramTile = 128;
int gram_slices(int m, int n) {
  const int t = cdiv(n, kGramTile);
  const int upper = t * (t + 1) / 2;
  int ks = (512 + upper / 2) / std::max(upper, 1);
  ks = std::min(ks, std::max(1, m / 256));
  return std::max(1, std::min(ks, 32));
}

template <typename T> CQ2Layout cq2_layout(int m, int n) {
  CQ2Layout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += al256(bytes); return o; };
  const size_t nn = (size_t)n * n * sizeof(T);
  L.ks = gram_slices(m, n);
  L.offQ1 = take((size_t)m * n * sizeof(T));
  L.offG = take(nn);
  L.offE = take(nn);
  L.offX = take(nn);
  L.offU = take(nn);
  L.offW = take(nn);
  L.offT = take(nn);
  // the Gram slices; later the packed operands of the BF16 / FP32 correction
  L.offSl = take(std::max(nn * L.ks, ScalarTraits<T>::is_complex ? gemm_lo_complex_workspace_bytes(m, n, n)
                                                                    : gemm_prec_workspace_bytes(m, n, n)));
  L.offFlags = take(chol_tile_flag_ints(n) * sizeof(int));
  L.offPart = take(kRedWgs * sizeof(double));
  L.offSt = take(64 * sizeof(int));
  L.ldxc = (n + 15) / 16 * 16;
  L.offXc = take((size_t)L.ldxc * n * sizeof(T));
  L.bytes = off;
  return L;
}

template <typename T> GemmArgs gq(int M, int N, int K, const T* A, int64_t lda, const T* B, int64_t ldb, T* C,
                                  int64_t ldc, double alpha, double beta, int tri, const int* gate = nullptr,
                                  int want = 1) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha_re = alpha; g.beta_re = beta;
  g.ksplit = 1; g.kchunk = K; g.batch = 1;
  g.tri = tri;
  g.gate = gate; g.gate_want = want;
  return g;
}

// G (n x n, upper triangle; the strict lower part is left unspecified) =
// X^H X for X m x n: split-K slices then their sum.
// G[upper] = sum of the ks slices (n x n each, upper triangles only).
template <typename T>
__global__ void __launch_bounds__(256) sum_upper_slices_kernel(const T* P, int ks, int n, T* G) {
  const int64_t tot = (int64_t)n * (n + 1) / 2;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) {
    // e -> (r, c), r <= c, column by column of the upper triangle
    int c = (int)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
    while ((int64_t)c * (c + 1) / 2 > e) --c;
    while ((int64_t)(c + 1) * (c + 2) / 2 <= e) ++c;
    const int r = (int)(e - (int64_t)c * (c + 1) / 2);
    const int64_t idx = r + (int64_t)c * n;
    T acc[4] = {s_zero(T()), s_zero(T()), s_zero(T()), s_zero(T())};
    int z = 0;
    for (; z + 4 <= ks; z += 4)
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] = s_add(acc[u], P[idx + (int64_t)(z + u) * n * n]);
    for (; z < ks; ++z) acc[0] = s_add(acc[0], P[idx + (int64_t)z * n * n]);
    G[idx] = s_add(s_add(acc[0], acc[1]), s_add(acc[2], acc[3]));
  }
}

// The split-K slices of the upper Gram X^H X; returns their count.
template <typename T> int gram_upper_slices(const T* X, int64_t ldx, int m, int n, T* slices, int ks, hipStream_t s) {
  GemmArgs g = gq<T>(n, n, m, X, ldx, X, ldx, slices, n, 1.0, 0.0, 0);
  g.upperC = 1;
  if (!ScalarTraits<T>::is_complex) g.tile = kGramTile;
  g.kchunk = cdiv(cdiv(m, ks), 16) * 16;
  g.ksplit = cdiv(m, g.kchunk);
  g.splitStride = (int64_t)n * n;
  gemm<T>(OP_C, OP_N, g, s);
  return g.ksplit;
}

template <typename T>
void gram_upper(const T* X, int64_t ldx, int m, int n, T* G, T* slices, int ks, hipStream_t s) {
  const int nsl = gram_upper_slices<T>(X, ldx, m, n, slices, ks, s);
  const int64_t tot = (int64_t)n * (n + 1) / 2;
  hipLaunchKernelGGL((sum_upper_slices_kernel<T>), dim3((unsigned)std::min<int64_t>(cdiv64(tot, 256), 2048)), dim3(256),
                     0, s, slices, nsl, n, G);
}

// The second pass's first kernel, in one sweep over the n x n elements:
// G2 = I + E on and above the diagonal from the ks slices (into E), U1 =
// up(E) into U (zero below the diagonal: later products read whole diagonal
// tiles), and this workgroup's part of ||E||_F^2 (off-diagonal entries
// counted twice) into part.
template <typename T>
__global__ void __launch_bounds__(256) cq2_sum_up_kernel(const T* P, int ks, int n, T* E, T* U, double* part) {
  __shared__ double red[4];
  const int64_t tot = (int64_t)n * n;
  double dev = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) {
    const int r = (int)(e % n), c = (int)(e / n);
    T u = s_zero(T());
    if (r <= c) {
      T acc[4] = {s_zero(T()), s_zero(T()), s_zero(T()), s_zero(T())};
      int z = 0;
      for (; z + 4 <= ks; z += 4)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = s_add(acc[q], P[e + (int64_t)(z + q) * tot]);
      for (; z < ks; ++z) acc[0] = s_add(acc[0], P[e + (int64_t)z * tot]);
      const T g = s_add(s_add(acc[0], acc[1]), s_add(acc[2], acc[3]));
      E[e] = g;
      if (r == c) {
        const double d = s_re(g) - 1.0;
        u = s_from_real<T>(0.5 * d);
        dev += s_abs2(s_sub(g, s_one(T())));
      } else {
        u = g;
        dev += 2.0 * s_abs2(g);
      }
    }
    U[e] = u;
  }
  dev = wave_sum(dev);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dev;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// R (n x n, ld ldr) = sum of the ks slices (n x n each) on and above the
// diagonal, zero below; only when *gate == 1.
constexpr int kRSlices = 4;
template <typename T>
__global__ void __launch_bounds__(256) sum_r_slices_kernel(const T* P, int ks, int n, T* R, int64_t ldr,
                                                           const int* gate) {
  if (*gate != 1) return;
  const int64_t tot = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) {
    const int r = (int)(e % n), c = (int)(e / n);
    T acc = s_zero(T());
    if (r <= c)
      for (int z = 0; z < ks; ++z) acc = s_add(acc, P[e + (int64_t)z * tot]);
    R[r + (int64_t)c * ldr] = acc;
  }
}

// st[0] = 1 (accept) when the Cholesky succeeded (info == 0) and ||E||_F <=
// 1e-6; st[1] = 1 when the second-order terms are needed (||E||_F > 1e-9);
// of the accepted calls, exactly one of st[7], st[3], st[6] is 1: Q = Q1 +
// Q1 (W2^{-1} - I) takes its correction from BF16 MFMA (st[7]: lo_ok and
// ||E||_F <= kBf16CorrMaxE) or from FP32 MFMA (st[3]: lo_ok and ||E||_F <=
// kLoCorrMaxE), else Q = Q1 W2^{-1} runs on FP64 MFMA (st[6]);
// st[4..5] = ||E||_F as a double.
__global__ void cq2_decide_kernel(const double* part, int np, const int* cinfo, int lo_ok, int* st) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) acc += part[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double e = sqrt(red[0] + red[1] + red[2] + red[3]);
    const bool ok = (*cinfo == 0) && (e <= 1e-6);  // false for NaN
    st[0] = ok ? 1 : 0;
    st[1] = (ok && e > 1e-9) ? 1 : 0;
    st[7] = (ok && lo_ok && e <= kBf16CorrMaxE) ? 1 : 0;
    st[3] = (ok && lo_ok && st[7] == 0 && e <= kLoCorrMaxE) ? 1 : 0;
    st[6] = (ok && st[3] == 0 && st[7] == 0) ? 1 : 0;
    double* ed = reinterpret_cast<double*>(st + 4);
    ed[0] = e;
  }
}

// U = up(G - I - Sub) (Sub null: not subtracted), zero below the diagonal;
// when gate is given the kernel runs only if *gate == 1.
template <typename T>
__global__ void cq2_up_kernel(const T* G, const T* Sub, int n, T* U, const int* gate) {
  if (gate != nullptr && *gate != 1) return;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  T u = s_zero(T());
  if (r <= c) {
    T v = G[idx];
    if (Sub) v = s_sub(v, Sub[idx]);
    if (r == c) u = s_from_real<T>(0.5 * (s_re(v) - 1.0));
    else u = v;
  }
  U[idx] = u;
}

// W = I + U and Wi = I - U (+ U2 when st[1]), zero below the diagonal; D
// (optional) = Wi - I.
template <typename T>
__global__ void cq2_w_kernel(const T* U, const T* U2, const int* st, int n, T* W, T* Wi, T* D) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx % n), c = (int)(idx / n);
  T w = s_zero(T()), wi = s_zero(T());
  if (r <= c) {
    const T u = U[idx];
    w = u;
    wi = s_neg(u);
    if (st[1] == 1) wi = s_add(wi, U2[idx]);
    if (D != nullptr) D[idx] = wi;
    if (r == c) {
      w = s_add(w, s_one(T()));
      wi = s_add(wi, s_one(T()));
    }
  } else if (D != nullptr) {
    D[idx] = s_zero(T());
  }
  W[idx] = w;
  Wi[idx] = wi;
}

}  // namespace

template <typename T> size_t cholqr2_workspace_bytes(int m, int n) {
  if (!cholqr2_eligible(m, n)) return 0;
  return cq2_layout<T>(m, n).bytes;
}

bool cholqr2_eligible(int m, int n) { return m >= n && n >= 1 && n <= kCholTileMaxN; }

// Enqueues the whole-matrix CholeskyQR2 on s; Q (m x n) and R (n x n or null)
// receive the factors when the device-side check accepts (status word at
// cholqr2_status(ws): [0] = 1 accepted).  ws: cholqr2_workspace_bytes(m, n).
template <typename T>
void cholqr2_enqueue(hipStream_t s, int m, int n, const T* A, int64_t lda, T* Q, int64_t ldq, T* R, int64_t ldr,
                     char* ws) {
  const CQ2Layout L = cq2_layout<T>(m, n);
  T* Q1 = reinterpret_cast<T*>(ws + L.offQ1);
  T* G = reinterpret_cast<T*>(ws + L.offG);   // G1, then R1
  T* E = reinterpret_cast<T*>(ws + L.offE);   // G2 = I + E
  T* X = reinterpret_cast<T*>(ws + L.offX);   // X1 = R1^{-1}, then W2^{-1}
  T* U = reinterpret_cast<T*>(ws + L.offU);
  T* W = reinterpret_cast<T*>(ws + L.offW);
  T* Tm = reinterpret_cast<T*>(ws + L.offT);
  T* Sl = reinterpret_cast<T*>(ws + L.offSl);
  int* flags = reinterpret_cast<int*>(ws + L.offFlags);
  double* part = reinterpret_cast<double*>(ws + L.offPart);
  int* st = reinterpret_cast<int*>(ws + L.offSt);
  const int64_t nn = (int64_t)n * n;
  const unsigned eg = (unsigned)cdiv64(nn, 256);
  zero_ints(st, 8, s);  // st[2]: the Cholesky's info
  // pass 1: Q1 = A X1 inside the Cholesky launch, column block by column
  // block as X1 becomes final
  gram_upper<T>(A, lda, m, n, G, Sl, L.ks, s);
  {
    const CholRightProduct rp{A, lda, Q1, m, m, ws + L.offXc, L.ldxc};
    chol_tile_factor<T>(G, n, n, X, n, flags, st + 2, 0, s, 0, nullptr, &rp);
  }
  // pass 2: G2 = I + E (slices summed with U1 = up(E) and ||E||_F in one
  // sweep), the decision, and W2, W2^{-1} in closed form
  const int ks2 = gram_upper_slices<T>(Q1, m, m, n, Sl, L.ks, s);
  hipLaunchKernelGGL((cq2_sum_up_kernel<T>), dim3(kRedWgs), dim3(256), 0, s, Sl, ks2, n, E, U, part);
  hipLaunchKernelGGL(cq2_decide_kernel, dim3(1), dim3(256), 0, s, part, kRedWgs, st + 2, 1, st);
  // second order (st[1] == 1): U = up(E - U1^H U1), U2 = U U
  {
    GemmArgs g = gq<T>(n, n, n, U, n, U, n, Tm, n, 1.0, 0.0, TRI_A_LOWER | TRI_B_UPPER, st + 1, 1);
    g.upperC = 1;
    gemm<T>(OP_C, OP_N, g, s);
  }
  hipLaunchKernelGGL((cq2_up_kernel<T>), dim3(eg), dim3(256), 0, s, E, (const T*)Tm, n, U, (const int*)(st + 1));
  gemm<T>(OP_N, OP_N, gq<T>(n, n, n, U, n, U, n, Tm, n, 1.0, 0.0, TRI_A_UPPER | TRI_B_UPPER, st + 1, 1), s);
  // D = W2^{-1} - I into E, which is no longer read
  T* D = E;
  hipLaunchKernelGGL((cq2_w_kernel<T>), dim3(eg), dim3(256), 0, s, U, Tm, st, n, W, X, D);
  // Q = Q1 W2^{-1};  R = W2 R1 (R1 is upper with a zero strict lower part).
  // ||E|| small: Q = Q1 + Q1 D with the correction on BF16 (st[7]) or FP32
  // (st[3]) MFMA (complex data through its real embedding; the packed
  // operands go where the Gram slices were); else Q1 W2^{-1} on FP64
  {
    GemmArgs g = gq<T>(m, n, n, Q1, m, D, n, Q, ldq, 1.0, 1.0, TRI_B_UPPER);
    g.Cin = Q1;
    g.ldcin = m;
    gemm_lo_reduced(g, OP_N, OP_N, st + 7, st + 3, Sl, s, ScalarTraits<T>::is_complex);
    gemm<T>(OP_N, OP_N, gq<T>(m, n, n, Q1, m, X, n, Q, ldq, 1.0, 0.0, TRI_B_UPPER, st + 6, 1), s);
  }
  if (R != nullptr) {
    // complex data: R = W2 R1 = R1 + U R1 with the correction on BF16 (st[7])
    // or FP32 (st[3]) MFMA, else (st[6]) on FP64.  U R1 (upper x upper) is
    // exactly upper triangular and R1's strict lower part is zero, so R's is
    // too.  (Real data: the split-K FP64 product below is as fast at n ~ 1e3
    // as the packing and launches of the low-precision one.)
    const int* r64 = st;
    if constexpr (ScalarTraits<T>::is_complex) {
      GemmArgs g = gq<T>(n, n, n, U, n, G, n, R, ldr, 1.0, 1.0, TRI_B_UPPER);
      g.Cin = G;
      g.ldcin = n;
      gemm_lo_reduced(g, OP_N, OP_N, st + 7, st + 3, Sl, s, true);
      r64 = st + 6;
    }
    if (!ScalarTraits<T>::is_complex && n >= 512 && L.ks >= kRSlices) {
      // upper x upper on 32 x 32 tiles with K split kRSlices ways (each
      // workgroup a short MFMA chain; partials in the Gram slices' space),
      // then the gated sum, zero below the diagonal
      GemmArgs g = gq<T>(n, n, n, W, n, G, n, Sl, n, 1.0, 0.0, TRI_A_UPPER | TRI_B_UPPER, r64, 1);
      g.tile = 32;
      g.kchunk = cdiv(cdiv(n, kRSlices), 16) * 16;
      g.ksplit = cdiv(n, g.kchunk);
      g.splitStride = nn;
      gemm<T>(OP_N, OP_N, g, s);
      hipLaunchKernelGGL((sum_r_slices_kernel<T>), dim3((unsigned)std::min<int64_t>(cdiv64(nn, 256), 2048)), dim3(256),
                         0, s, Sl, g.ksplit, n, R, ldr, r64);
    } else {
      gemm<T>(OP_N, OP_N, gq<T>(n, n, n, W, n, G, n, R, ldr, 1.0, 0.0, TRI_A_UPPER | TRI_B_UPPER, r64, 1), s);
    }
  }
}

template <typename T> int* cholqr2_status(char* ws, int m, int n) {
  return reinterpret_cast<int*>(ws + cq2_layout<T>(m, n).offSt);
}
template <typename T> T* cholqr2_scratch(char* ws, int m, int n) {
  return reinterpret_cast<T*>(ws + cq2_layout<T>(m, n).offQ1);
}

template size_t cholqr2_workspace_bytes<double>(int, int);
template size_t cholqr2_workspace_bytes<zd>(int, int);
template void cholqr2_enqueue<double>(hipStream_t, int, int, const double*, int64_t, double*, int64_t, double*,
                                      int64_t, char*);
template void cholqr2_enqueue<zd>(hipStream_t, int, int, const zd*, int64_t, zd*, int64_t, zd*, int64_t, char*);
template int* cholqr2_status<double>(char*, int, int);
template int* cholqr2_status<zd>(char*, int, int);
template double* cholqr2_scratch<double>(char*, int, int);
template zd* cholqr2_scratch<zd>(char*, int, int);

}  // namespace mak
