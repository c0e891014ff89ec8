// Dense 32 x 32 building blocks in LDS for one workgroup / one wave (gfx950).
#pragma once

#include <utility>

#include "mak_common.hpp"

namespace mak {

// ---------------------------------------------------------------------------
// 32x32x32 products on FP64 MFMA for one 256-thread workgroup:
//   C = alpha op(A) op(B) (+ C when acc), op = identity or conjugate transpose.
// Wave w computes the 16x16 block (w % 2, w / 2).  The result is written after
// a workgroup barrier, so C may alias A or B.  All 256 threads must call it.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline void mm32(T (*C)[33], const T (*A)[33], bool cA, const T (*B)[33], bool cB, double alpha,
                            bool acc) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int bi = w & 1, bj = w >> 1;
  v4d cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < 32; kk += 4) {
    const int k = kk + (lane >> 4);
    const int i = bi * 16 + (lane & 15);
    const int j = bj * 16 + (lane & 15);
    const T a = cA ? s_conj(A[k][i]) : A[i][k];
    const T b = cB ? s_conj(B[j][k]) : B[k][j];
    // transposed block: lane l holds row (l % 16) of the output
    cr = mfma_f64(s_re(b), s_re(a), cr);
    if constexpr (ScalarTraits<T>::is_complex) {
      cr = mfma_f64(-s_im(b), s_im(a), cr);
      ci = mfma_f64(s_im(b), s_re(a), ci);
      ci = mfma_f64(s_re(b), s_im(a), ci);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = bi * 16 + (lane & 15);
    const int col = bj * 16 + mfma_f64_row(lane, r);
    T v;
    if constexpr (ScalarTraits<T>::is_complex) v = mkz(alpha * cr[r], alpha * ci[r]);
    else v = alpha * cr[r];
    C[row][col] = acc ? s_add(C[row][col], v) : v;
  }
  __syncthreads();
}

// C = alpha (op(A0) op(B0) + op(A1) op(B1)) (+ C when acc): two products
// summed in the MFMA accumulators, one write (after a barrier, so C may alias
// any operand).  A0 = nullptr drops the first product.
template <typename T>
__device__ inline void mm32_2(T (*C)[33], const T (*A0)[33], const T (*B0)[33], const T (*A1)[33],
                              const T (*B1)[33], bool cA, bool cB, double alpha, bool acc) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int bi = w & 1, bj = w >> 1;
  v4d cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  auto product = [&](const T (*A)[33], const T (*B)[33]) {
#pragma unroll
    for (int kk = 0; kk < 32; kk += 4) {
      const int k = kk + (lane >> 4);
      const int i = bi * 16 + (lane & 15);
      const int j = bj * 16 + (lane & 15);
      const T a = cA ? s_conj(A[k][i]) : A[i][k];
      const T b = cB ? s_conj(B[j][k]) : B[k][j];
      cr = mfma_f64(s_re(b), s_re(a), cr);
      if constexpr (ScalarTraits<T>::is_complex) {
        cr = mfma_f64(-s_im(b), s_im(a), cr);
        ci = mfma_f64(s_im(b), s_re(a), ci);
        ci = mfma_f64(s_re(b), s_im(a), ci);
      }
    }
  };
  if (A0 != nullptr) product(A0, B0);
  product(A1, B1);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = bi * 16 + (lane & 15);
    const int col = bj * 16 + mfma_f64_row(lane, r);
    T v;
    if constexpr (ScalarTraits<T>::is_complex) v = mkz(alpha * cr[r], alpha * ci[r]);
    else v = alpha * cr[r];
    C[row][col] = acc ? s_add(C[row][col], v) : v;
  }
  __syncthreads();
}

// Value of v in lane `src` (uniform over the wave) broadcast to the wave through a scalar register.
__device__ inline double rdl(double v, int src) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ inline zd rdl(zd v, int src) { return mkz(rdl(v.x, src), rdl(v.y, src)); }

// Broadcasts issued together before their arithmetic: four real or two
// complex values (eight scalar registers).
template <typename T> constexpr int kBc = ScalarTraits<T>::is_complex ? 2 : 4;

// Makes `v` appear to depend on `dep` through an empty asm (no instruction is
// emitted): broadcasts of `v` cannot then be scheduled before `dep` is
// computed.  Without it the compiler hoists all broadcasts of a factorisation
// step (or of a whole triangular inverse) into scalar registers at once and
// spills them to VGPR lanes, which costs more than the arithmetic.
__device__ __forceinline__ void order_after(double& v, double dep) { asm volatile("" : "+v"(v) : "v"(dep)); }
// The lane's column index through an empty asm, so that the per-step lane
// masks (c > j, c == j) are formed in their step instead of all being hoisted
// into scalar registers up front.
__device__ __forceinline__ int step_lane(int c) {
  asm volatile("" : "+v"(c));
  return c;
}
__device__ __forceinline__ void order_after(zd& v, const zd& dep) {
  asm volatile("" : "+v"(v.x), "+v"(v.y) : "v"(dep.x));
}
// An index (zero) that appears to depend on `dep`: a load through it cannot
// be issued before `dep` is computed.
__device__ __forceinline__ void index_after(int& z, double dep) { asm volatile("" : "+v"(z) : "v"(dep)); }
__device__ __forceinline__ void index_after(int& z, const zd& dep) {
  asm volatile("" : "+v"(z) : "v"(dep.x), "v"(dep.y));
}

// Cholesky with inverse of a 32 x 32 Hermitian positive definite block, by
// ONE wave (no workgroup barrier, no LDS traffic inside the factorisation):
//   S (LDS, 32 x 33) holds the upper triangle on entry (rows/columns beyond the
//   active size padded with the identity by the caller); on exit S = W (upper,
//   zero below) with S_in = W^H W, and Y = W^{-1} (upper, zero below).
//   Returns 0, or j + 1 for the first pivot j that is not positive (uniform
//   over the wave; S and Y are then unspecified).
// Lane c (c = lane % 32; lanes 32..63 mirror lanes 0..31) keeps column c of W
// and of Y in registers; row entries of other columns come from lane
// broadcasts (v_readlane), four at a time.  Cholesky step j: W(j, c) =
// S(j, c) / d_j, then S(r, c) -= conj(W(j, r)) W(j, c) for r > j (entries
// below the diagonal only receive the mirrored update and are ignored).
// `dinv` of lane j receives 1 / W(j, j).
template <typename T, int N, int J>
__device__ __forceinline__ void cholN_step(T (&s)[N], int c0, int& bad, T& dinv) {
  const int c = step_lane(c0);
  const double d2 = s_re(rdl(s[J], J));
  if (!bad && !(d2 > 0.0)) bad = J + 1;
  const double rd = (d2 > 0.0) ? rsqrt_nr(d2) : 0.0;
  const double d = d2 * rd;
  dinv = sel(c == J, s_from_real<T>(rd), dinv);
  const T w = sel(c > J, s_scale(s[J], rd), sel(c == J, s_from_real<T>(d), s_zero(T())));
  s[J] = w;
  T src = w;
#pragma unroll
  for (int r0 = J + 1; r0 < N; r0 += kBc<T>) {
    T b[kBc<T>];
#pragma unroll
    for (int u = 0; u < kBc<T>; ++u) b[u] = (r0 + u < N) ? rdl(src, r0 + u) : s_zero(T());
#pragma unroll
    for (int u = 0; u < kBc<T>; ++u)
      if (r0 + u < N) s[r0 + u] = s_sub(s[r0 + u], s_mul(s_conj(b[u]), w));
    order_after(src, s[r0]);
  }
}
template <typename T, int N, int... Js>
__device__ __forceinline__ void cholN_steps(T (&s)[N], int c, int& bad, T& dinv, std::integer_sequence<int, Js...>) {
  (cholN_step<T, N, Js>(s, c, bad, dinv), ...);
}

// Inverse Y of an upper triangular N x N block U held by lanes as above
// (lane c: s[r] = U(r, c)), `dinv` of lane r = 1 / U(r, r).  Rows from the
// bottom: Y(r, c) = (delta_rc - sum_{k > r} U(r, k) Y(k, c)) / U(r, r).
template <typename T, int N, int I>
__device__ __forceinline__ void trinvN_row(T (&y)[N], const T (&s)[N], const T& dinv, int c0) {
  constexpr int r = N - 1 - I;
  const int c = step_lane(c0);
  T acc0 = s_zero(T()), acc1 = s_zero(T());
  T src = s[r];
  if constexpr (r < N - 1) order_after(src, y[r < N - 1 ? r + 1 : r]);
#pragma unroll
  for (int k0 = r + 1; k0 < N; k0 += kBc<T>) {
    T b[kBc<T>];
#pragma unroll
    for (int u = 0; u < kBc<T>; ++u) b[u] = (k0 + u < N) ? rdl(src, k0 + u) : s_zero(T());   // U(r, k)
#pragma unroll
    for (int u = 0; u < kBc<T>; ++u) {
      if (k0 + u < N) {
        if (u & 1) acc1 = s_fma(b[u], y[k0 + u], acc1);
        else acc0 = s_fma(b[u], y[k0 + u], acc0);
      }
    }
    order_after(src, acc0);
  }
  const T inv = rdl(dinv, r);
  y[r] = sel(r < c, s_neg(s_mul(s_add(acc0, acc1), inv)), sel(r == c, inv, s_zero(T())));
}
template <typename T, int N, int... Is>
__device__ __forceinline__ void trinvN_rows(T (&y)[N], const T (&s)[N], const T& dinv, int c,
                                            std::integer_sequence<int, Is...>) {
  (trinvN_row<T, N, Is>(y, s, dinv, c), ...);
}
template <typename T, int... Is>
__device__ __forceinline__ void trinv32_rows(T (&y)[32], const T (&s)[32], const T& dinv, int c,
                                             std::integer_sequence<int, Is...> seq) {
  trinvN_rows<T, 32>(y, s, dinv, c, seq);
}

// Inverse Z of a unit lower triangular N x N block L (lane c: s[r] = L(r, c)
// for r > c), rows from the top: Z(r, c) = delta_rc - sum_{c <= k < r} L(r, k) Z(k, c).
template <typename T, int N, int R>
__device__ __forceinline__ void trilinvN_row(T (&z)[N], const T (&s)[N], int c0) {
  const int c = step_lane(c0);
  T acc0 = s_zero(T()), acc1 = s_zero(T());
  T src = s[R];
  if constexpr (R > 0) order_after(src, z[R > 0 ? R - 1 : 0]);
#pragma unroll
  for (int k0 = 0; k0 < R; k0 += kBc<T>) {
    T b[kBc<T>];
#pragma unroll
    for (int u = 0; u < kBc<T>; ++u) b[u] = (k0 + u < R) ? rdl(src, k0 + u) : s_zero(T());   // L(R, k)
#pragma unroll
    for (int u = 0; u < kBc<T>; ++u) {
      if (k0 + u < R) {
        if (u & 1) acc1 = s_fma(b[u], z[k0 + u], acc1);
        else acc0 = s_fma(b[u], z[k0 + u], acc0);
      }
    }
    order_after(src, acc0);
  }
  z[R] = sel(R == c, s_one(T()), sel(R > c, s_neg(s_add(acc0, acc1)), s_zero(T())));
}
template <typename T, int N, int... Rs>
__device__ __forceinline__ void trilinvN_rows(T (&z)[N], const T (&s)[N], int c, std::integer_sequence<int, Rs...>) {
  (trilinvN_row<T, N, Rs>(z, s, c), ...);
}

// 16 x 16 x 16 product by one wave on FP64 MFMA, operands and result in LDS
// (row stride 33): D = alpha op(A) B (+ D when acc), op(A) = A^H when CA.
// Every operand is read before D is written, so D may alias A or B.
// MFMA f64 16x16x4 layout: A operand lane l = (row l % 16, k l / 16), B
// operand lane l = (col l % 16, k l / 16), result register r of lane l =
// (row 4 r + l / 16, col l % 16).
template <typename T, bool CA>
__device__ __forceinline__ void wave_mm16(T (*D)[33], const T (*A)[33], const T (*B)[33], double alpha, bool acc) {
  const int l = lane_id();
  const int lr = l & 15, lk = l >> 4;
  v4d cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  T a[4], b[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = 4 * q + lk;
    a[q] = CA ? s_conj(A[k][lr]) : A[lr][k];
    b[q] = B[k][lr];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if constexpr (ScalarTraits<T>::is_complex) {
      cr = mfma_f64(s_re(a[q]), s_re(b[q]), cr);
      cr = mfma_f64(-s_im(a[q]), s_im(b[q]), cr);
      ci = mfma_f64(s_re(a[q]), s_im(b[q]), ci);
      ci = mfma_f64(s_im(a[q]), s_re(b[q]), ci);
    } else {
      cr = mfma_f64(a[q], b[q], cr);
    }
  }
  T old[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) old[r] = acc ? D[4 * r + lk][lr] : s_zero(T());
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    T v;
    if constexpr (ScalarTraits<T>::is_complex) v = mkz(alpha * cr[r], alpha * ci[r]);
    else v = alpha * cr[r];
    D[4 * r + lk][lr] = s_add(old[r], v);
  }
}

// Cholesky with inverse of a 16 x 16 diagonal block (rows / columns r0..r0+15
// of S and Y) by one wave, lane c (c = lane % 16) holding column c; W and
// W^{-1} (upper, zero below) replace the block of S and fill the block of Y.
template <typename T>
__device__ __forceinline__ int wave_chol_inv16(T (*S)[33], T (*Y)[33], int r0) {
  const int l = lane_id();
  const int c = l & 15;
  const bool writer = l < 16;
  T s[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const T v = S[r0 + r][r0 + c];
    s[r] = sel(r <= c, v, s_zero(T()));
  }
  int bad = 0;
  T dinv = s_zero(T());
  cholN_steps<T, 16>(s, c, bad, dinv, std::make_integer_sequence<int, 16>{});
#pragma unroll
  for (int r = 0; r < 16; ++r)
    if (writer) S[r0 + r][r0 + c] = (r <= c) ? s[r] : s_zero(T());
  T y[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) y[r] = s_zero(T());
  trinvN_rows<T, 16>(y, s, dinv, c, std::make_integer_sequence<int, 16>{});
#pragma unroll
  for (int r = 0; r < 16; ++r)
    if (writer) Y[r0 + r][r0 + c] = y[r];
  return bad;
}

// Cholesky with inverse of a 32 x 32 Hermitian positive definite block, by
// ONE wave (no workgroup barrier):
//   S (LDS, 32 x 33) holds the upper triangle on entry (rows/columns beyond the
//   active size padded with the identity by the caller); on exit S = W (upper,
//   zero below) with S_in = W^H W, and Y = W^{-1} (upper, zero below).
//   Returns 0, or j + 1 for the first pivot j that is not positive (uniform
//   over the wave; S and Y are then unspecified).
// 2 x 2 blocks of 16: W11, Y11 in-wave; W12 = Y11^H S12; S22 -= W12^H W12;
// W22, Y22 in-wave; Y12 = -Y11 (W12 Y22) -- the off-diagonal products on MFMA.
template <typename T>
__device__ __forceinline__ int wave_chol_inv32(T (*S)[33], T (*Y)[33]) {
  auto blk = [](T (*M)[33], int r0, int c0) { return reinterpret_cast<T (*)[33]>(&M[r0][c0]); };
  int bad = wave_chol_inv16<T>(S, Y, 0);
  wave_mm16<T, true>(blk(S, 0, 16), blk(Y, 0, 0), blk(S, 0, 16), 1.0, false);     // W12 = Y11^H S12
  wave_mm16<T, true>(blk(S, 16, 16), blk(S, 0, 16), blk(S, 0, 16), -1.0, true);   // S22 -= W12^H W12
  const int b2 = wave_chol_inv16<T>(S, Y, 16);
  if (bad == 0 && b2 != 0) bad = 16 + b2;
  wave_mm16<T, false>(blk(Y, 0, 16), blk(S, 0, 16), blk(Y, 16, 16), 1.0, false);  // P = W12 Y22
  wave_mm16<T, false>(blk(Y, 0, 16), blk(Y, 0, 0), blk(Y, 0, 16), -1.0, false);   // Y12 = -Y11 P
  const int l = lane_id();
  for (int e = l; e < 256; e += 64) {
    const int i = e & 15, j = e >> 4;
    S[16 + i][j] = s_zero(T());
    Y[16 + i][j] = s_zero(T());
  }
  wave_lds_sync();
  return bad;
}

// Cholesky factorisation with inverse of a 32 x 32 Hermitian positive definite
// tile by ONE wave, eliminating [G | I] in a single pass:
//   S (LDS, 32 x 33): upper triangle of G on entry (strict lower part not
//   read); on exit R (upper, zero below) with G = R^H R.
//   Y (LDS, 32 x 33): on exit X = R^{-1} (upper, zero below).
//   row: LDS scratch of 64 elements, 16-byte aligned (two pivot-row buffers).
// Lane c < 32 keeps column c of G, lane 32 + c column c of the identity (that
// becomes M = R^{-H}, lower).  Step k, with the pivot d = G'(k, k) read from
// lane k and the raw pivot row G'(k, r) in LDS:
//   x[r] -= conj(G'(k, r)) x[k] / d  (r > k),   x[k] <- x[k] / sqrt(d)
// is the same instruction stream for both halves: the Cholesky update
// G'(r, c) -= conj(R(k, r)) R(k, c) and the forward elimination of the
// identity Z(r, c) -= conj(R(k, r)) M(k, c).
// Software pipelined: the update of row k + 1 and its publication as the next
// pivot row (the dependency chain: one LDS round trip, the pivot reciprocal
// and two FMAs per column) are issued before the bulk of the step's updates,
// which then overlap the next round trip; the pivot rows alternate between
// two buffers so the bulk reads of step k never race the write of step k + 1.
// Returns 0, or k + 1 for the first pivot k that is not positive (S and Y are
// then unspecified).
template <typename T>
__device__ __forceinline__ int wave_chol_aug32(T (*S)[33], T (*Y)[33], T* row) {
  const int l = lane_id();
  const int c = l & 31;
  const bool lo = l < 32;
  T x[32];
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    const T v = S[r][c];
    x[r] = lo ? (r <= c ? v : s_zero(T())) : (r == c ? s_one(T()) : s_zero(T()));
  }
  // `row` is read at wave-uniform addresses; an index offset the compiler
  // cannot see is zero keeps those values in vector registers (as uniform
  // values they would be moved to scalar registers and hoisted, and spill).
  int zoff = 0;
  asm volatile("" : "+v"(zoff));
  if (lo) row[c] = x[0];
  int bad = 0;
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    T* cur = row + (k & 1) * 32;
    T* nxt = row + ((k + 1) & 1) * 32;
    wave_lds_sync();  // the pivot row of step k is visible
    // the whole pivot row in one burst of 16-byte reads (issued together, so
    // their latency is paid once per step; read one at a time, as the
    // scheduler otherwise does, they cost most of the step)
    T rv[32];
    {
      using V2 = typename ScalarTraits<T>::vec2;
      const V2* cv = reinterpret_cast<const V2*>(cur + zoff);
#pragma unroll
      for (int r = (k + 1) & ~1; r < 32; r += 2) {
        const V2 v = cv[r / 2];
        rv[r] = ScalarTraits<T>::lo(v);
        rv[r + 1] = ScalarTraits<T>::hi(v);
      }
    }
    const T raw = x[k];
    const double d = s_re(rdl(raw, k));
    bad = bad ? bad : (d > 0.0 ? 0 : k + 1);
    const double dd = (d > 0.0) ? d : 1.0;
    const double inv = recip_nr(dd);
    const T w = s_scale(raw, inv);
    if (k < 31) {
      // the chain: row k + 1, published as the next pivot row
      x[k + 1] = s_sub(x[k + 1], s_mul(s_conj(rv[k + 1]), w));
      if (lo) nxt[c] = x[k + 1];
    }
#pragma unroll
    for (int r = k + 2; r < 32; ++r) x[r] = s_sub(x[r], s_mul(s_conj(rv[r]), w));
    // one step of overlap at most: step k + 1's chain may run ahead of this
    // bulk, but row k + 3 (the chain of step k + 2) waits for all of it
#pragma unroll
    for (int r = k + 4; r < 32; ++r) order_after(x[k + 3], x[r]);
    x[k] = s_scale(raw, rsqrt_nr(dd));
  }
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    if (lo) S[r][c] = (r <= c) ? x[r] : s_zero(T());
    else Y[c][r] = (r >= c) ? s_conj(x[r]) : s_zero(T());
  }
  wave_lds_sync();
  return bad;
}

}  // namespace mak
