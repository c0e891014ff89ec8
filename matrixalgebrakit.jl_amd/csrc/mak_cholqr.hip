// Tall panel factorisation by CholeskyQR2 with Householder reconstruction,
// gated on the device by a conditioning check, with the Householder TSQR
// path as the fallback (the fallback's launches are always enqueued and exit
// at once unless the gate opens; no host synchronisation).
//
//   pass 1: G1 = X^H X, G1 = R1^H R1            (gram_chol_kernel)
//   pass 2: Q1 = X R1^{-1}, G2 = Q1^H Q1, G2 = R2^H R2, accepted only if
//           ||G2 - I||_F <= 0.1 (then kappa(Q1) <= 1.1 and the second pass is
//           orthogonal to working precision; Yamamoto, Nakatsukasa,
//           Yanagisawa, Fukaya, ETNA 44, 2015), R = R2 R1, Q = Q1 R2^{-1}
//   reconstruction (Ballard, Demmel, Grigori, Jacquelin, Nguyen, Solomonik,
//   IPDPS 2014): Q - [D; 0] = V U, LU without pivoting with
//   d_j = -sign(pivot_j); T = -U D^H V1^{-H}, the panel's R = D R and
//   V = [V1; Q_rest U^{-1}] -- the compact WY form LAPACK's geqrt produces.
//   Pass 2 and the reconstruction run in one launch (cholqr2_finish_kernel):
//   only the top 32 rows of Q are formed there, and V_rest = Q1_rest M with
//   M = R2^{-1} U^{-1} is one GEMM afterwards.
// A panel whose first Gram is not numerically positive definite or whose
// second Gram fails the check (kappa(X) beyond ~1e7) is factored by the TSQR
// path instead, so rank-deficient inputs keep Householder's stability.
//
// Carried first Gram (consecutive panels, host-checked mode): the finish
// kernel of panel k also forms Gw = Xn^H Xn of the next panel's columns over
// the same rows j0.. (before panel k's trailing update).  The update applies
// the unitary H_k^H to those rows, which leaves Xn^H Xn unchanged, so the next
// panel's first Gram (rows j0 + 32..) is Gw - Rt^H Rt, Rt the 32 rows of its
// columns that become R (formed and factored at the start of the next finish
// kernel: no pass over the panel and no launch of its own).  The
// difference loses accuracy when Xn lies mostly in the span of panel k; R1
// then only preconditions worse, and the second pass's check still decides
// whether the panel is accepted.
#include <algorithm>

#include "mak_common.hpp"
#include "mak_dense32.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

namespace {

constexpr int NB = 32;
constexpr int kMaxGramBlocks = 32;  // partial Grams summed by the last workgroup

// rows of X staged per tile: 16 elements per thread in both cases
template <typename T> struct CqTile;
template <> struct CqTile<double> { static constexpr int TR = 128; };
template <> struct CqTile<zd> { static constexpr int TR = 64; };

// How gram_tiles forms Y from the tile X of each row block:
//   kPlain     Y = X;
//   kTimesB    Y = X B (B: 32 x 32 in LDS), Y also written to Yout (H x ycols,
//              ld ldy; by default H x 32, ld H);
//   kProduct   as kTimesB, without the Gram (only Yout is formed).
enum GramMode { kPlain = 0, kTimesB = 1, kProduct = 2 };

// G += Y^H Y over the row tiles tfirst, tfirst + tstep, ... of X (H x nbc,
// columns nbc..31 read as zero; by default tfirst = blockIdx.x, tstep =
// gridDim.x), with Y formed as MODE says.
// The next tile's loads are issued before the current tile is used.
// Wave w accumulates the 16 x 16 block (w % 2, w / 2) of G in (cr, ci).
// With Xn (not null), the same row tiles of Xn (H x nbcn, ld ldx) are staged
// in Xns alongside and Gn += Xn^H Xn accumulates in (nr, ni).
struct NoPre {
  __device__ void operator()() const {}
};

// pre() runs once, after the first tile's loads are issued and before the
// tile is staged (it may use the LDS of Xs and Bs and must end with a
// workgroup barrier).
template <typename T, int MODE, typename Pre = NoPre>
__device__ inline void gram_tiles(const T* X, int64_t ldx, int H, int nbc, T (*Xs)[NB + 1], const T (*Bs)[NB + 1],
                                  T* Yout, v4d& cr, v4d& ci, const T* Xn = nullptr, int nbcn = 0,
                                  T (*Xns)[NB + 1] = nullptr, v4d* nr = nullptr, v4d* ni = nullptr,
                                  Pre pre = Pre(), int tfirst = -1, int tstep = 0, int64_t ldy = 0,
                                  int ycols = NB) {
  constexpr int TR = CqTile<T>::TR;
  constexpr int PER = TR * NB / 256;
  constexpr bool CPLX = ScalarTraits<T>::is_complex;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int ti = w & 1, tj = w >> 1;
  const int ntile = cdiv(H, TR);
  const bool next = Xn != nullptr;
  T cur[PER], curn[PER];
  auto load = [&](int t) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + 256 * q;
      const int r = e % TR, c = e / TR;
      const int row = t * TR + r;
      cur[q] = (row < H && c < nbc) ? X[row + (int64_t)c * ldx] : s_zero(T());
    }
    if (next) {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int e = tid + 256 * q;
        const int r = e % TR, c = e / TR;
        const int row = t * TR + r;
        curn[q] = (row < H && c < nbcn) ? Xn[row + (int64_t)c * ldx] : s_zero(T());
      }
    }
  };
  const int step = tstep > 0 ? tstep : (int)gridDim.x;
  const int64_t ldyo = ldy > 0 ? ldy : (int64_t)H;
  int t = tfirst >= 0 ? tfirst : (int)blockIdx.x;
  if (t < ntile) load(t);
  pre();
  for (; t < ntile; t += step) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + 256 * q;
      Xs[e % TR][e / TR] = cur[q];
    }
    if (next) {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int e = tid + 256 * q;
        Xns[e % TR][e / TR] = curn[q];
      }
    }
    if (t + step < ntile) load(t + step);
    __syncthreads();
    if (next) {
#pragma unroll 4
      for (int kk = 0; kk < TR; kk += 4) {
        const T a = Xns[kk + lk][ti * 16 + lr];
        const T b = Xns[kk + lk][tj * 16 + lr];
        if constexpr (CPLX) {
          *nr = mfma_f64(s_re(a), s_re(b), *nr);
          *nr = mfma_f64(s_im(a), s_im(b), *nr);
          *ni = mfma_f64(s_re(a), s_im(b), *ni);
          *ni = mfma_f64(-s_im(a), s_re(b), *ni);
        } else {
          *nr = mfma_f64(a, b, *nr);
        }
      }
    }
    if constexpr (MODE == kTimesB || MODE == kProduct) {
      // Y = X B as transposed 16 x 16 blocks (Y^T = B^T X^T): lane l, register
      // r holds Y(i0 + l % 16, j0 + 4 r + l / 16), so stores are row-contiguous.
      // Wave w owns rows [w TR / 4, (w + 1) TR / 4).
      constexpr int RB = TR / 64;
      v4d yr[RB][2], yi[RB][2];
#pragma unroll
      for (int ib = 0; ib < RB; ++ib)
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
          yr[ib][jb] = v4d{0.0, 0.0, 0.0, 0.0};
          yi[ib][jb] = v4d{0.0, 0.0, 0.0, 0.0};
        }
#pragma unroll
      for (int kk = 0; kk < NB; kk += 4) {
        const int k = kk + lk;
        T a[2];
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) a[jb] = Bs[k][jb * 16 + lr];
#pragma unroll
        for (int ib = 0; ib < RB; ++ib) {
          const T b = Xs[w * (TR / 4) + ib * 16 + lr][k];
#pragma unroll
          for (int jb = 0; jb < 2; ++jb) {
            if constexpr (CPLX) {
              yr[ib][jb] = mfma_f64(s_re(a[jb]), s_re(b), yr[ib][jb]);
              yr[ib][jb] = mfma_f64(-s_im(a[jb]), s_im(b), yr[ib][jb]);
              yi[ib][jb] = mfma_f64(s_im(a[jb]), s_re(b), yi[ib][jb]);
              yi[ib][jb] = mfma_f64(s_re(a[jb]), s_im(b), yi[ib][jb]);
            } else {
              yr[ib][jb] = mfma_f64(a[jb], b, yr[ib][jb]);
            }
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int ib = 0; ib < RB; ++ib) {
        const int i = w * (TR / 4) + ib * 16 + lr;
        const int row = t * TR + i;
#pragma unroll
        for (int jb = 0; jb < 2; ++jb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int j = jb * 16 + 4 * r + lk;
            T v;
            if constexpr (CPLX) v = mkz(yr[ib][jb][r], yi[ib][jb][r]);
            else v = yr[ib][jb][r];
            if constexpr (MODE == kTimesB) Xs[i][j] = v;
            if (row < H && j < ycols) Yout[row + (int64_t)j * ldyo] = v;
          }
      }
      __syncthreads();
    }
    if constexpr (MODE == kProduct) continue;
#pragma unroll 4
    for (int kk = 0; kk < TR; kk += 4) {
      const T a = Xs[kk + lk][ti * 16 + lr];  // (Y^H)(i, k) = conj(Y(k, i))
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
}

// Stores this workgroup's partial Gram (32 x 32 at partial + blockIdx.x * 1024).
template <typename T>
__device__ inline void store_partial(T* partial, const v4d& cr, const v4d& ci) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ti = w & 1, tj = w >> 1;
  T* P = partial + (int64_t)blockIdx.x * NB * NB;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = ti * 16 + mfma_f64_row(lane, r), j = tj * 16 + (lane & 15);
    if constexpr (ScalarTraits<T>::is_complex) P[i + j * NB] = mkz(cr[r], ci[r]);
    else P[i + j * NB] = cr[r];
  }
}

// Stores this workgroup's partial Gram (32 x 32 at partial + blockIdx.x * 1024)
// and takes a ticket; returns true (uniformly) only in the last workgroup to
// finish, after an acquire fence, so every partial is visible to it.
template <typename T>
__device__ inline bool partial_is_last(T* partial, const v4d& cr, const v4d& ci, int* counter, int* last_s) {
  const int tid = threadIdx.x;
  store_partial<T>(partial, cr, ci);
  __threadfence();
  __syncthreads();
  if (tid == 0) *last_s = (atomicAdd(counter, 1) == (int)gridDim.x - 1);
  __syncthreads();
  if (!*last_s) return false;
  __threadfence();
  return true;
}

// G = sum over the nb partials in workgroup order (rows / columns >= nbc:
// identity).  Thread t owns elements 4t .. 4t + 3 and reads them with 16-byte
// loads, 32 of them in flight per batch of partials.
// Resets *counter for the next use; returns ||G - I||_F^2.
template <typename T>
__device__ inline double reduce_partials(const T* partial, int nb, int nbc, T (*G)[NB + 1], int* counter,
                                         double* dev_s) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  constexpr bool CPLX = ScalarTraits<T>::is_complex;
  constexpr int LD = CPLX ? 4 : 2;     // 16-byte loads per partial and thread
  constexpr int BATCH = 32 / LD;       // partials per batch
  constexpr int64_t PSTR = (int64_t)NB * NB * sizeof(T) / 16;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const d2* base = reinterpret_cast<const d2*>(partial) + tid * LD;
  d2 acc[LD];
#pragma unroll
  for (int q = 0; q < LD; ++q) acc[q] = d2{0.0, 0.0};
  int b = 0;
  for (; b + BATCH <= nb; b += BATCH) {
    d2 v[BATCH][LD];
#pragma unroll
    for (int p = 0; p < BATCH; ++p)
#pragma unroll
      for (int q = 0; q < LD; ++q) v[p][q] = base[(b + p) * PSTR + q];
#pragma unroll
    for (int p = 0; p < BATCH; ++p)
#pragma unroll
      for (int q = 0; q < LD; ++q) acc[q] += v[p][q];
  }
  for (; b < nb; ++b)
#pragma unroll
    for (int q = 0; q < LD; ++q) acc[q] += base[b * PSTR + q];
  double dev = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = 4 * tid + u;
    const int i = e % NB, j = e / NB;
    T a;
    if constexpr (CPLX) a = mkz(acc[u].x, acc[u].y);
    else a = (u & 1) ? acc[u >> 1].y : acc[u >> 1].x;
    const T id = (i == j) ? s_one(T()) : s_zero(T());
    const T g = (i >= nbc || j >= nbc) ? id : a;
    G[i][j] = g;
    dev += s_abs2(s_sub(g, id));
  }
  dev = wave_sum(dev);
  if (lane == 0) dev_s[w] = dev;
  if (tid == 0) *counter = 0;  // ready for the next use
  __syncthreads();
  return dev_s[0] + dev_s[1] + dev_s[2] + dev_s[3];
}

// status[0]: panel rejected (the flag), status[1]: ticket counter,
// status[2]: first column of the first rejected panel (sticky mode).
//
// Pass 1: G1 = X^H X over nblk workgroups; the last one factors G1 = W^H W and
// writes R1 = W, R1inv = W^{-1} (32 x 32, ld 32).  When G1 is not numerically
// positive definite the flag is set (with status[2] = j0 if it was clear);
// otherwise it is cleared, unless `sticky` (then a set flag stays set).
template <typename T>
__global__ void __launch_bounds__(256) gram_chol_kernel(const T* X, int64_t ldx, int H, int nbc, T* partial,
                                                        int* status, int j0, int sticky, T* R1, T* R1inv) {
  int* counter = status + 1;
  constexpr int TR = CqTile<T>::TR;
  __shared__ T Xs[TR][NB + 1];
  __shared__ T Yl[NB][NB + 1];
  __shared__ int last_s, bad_s;
  __shared__ double dev_s[4];
  const int tid = threadIdx.x;
  v4d cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  gram_tiles<T, kPlain>(X, ldx, H, nbc, Xs, nullptr, nullptr, cr, ci);
  if (!partial_is_last<T>(partial, cr, ci, counter, &last_s)) return;
  T (*G)[NB + 1] = Xs;
  reduce_partials<T>(partial, (int)gridDim.x, nbc, G, counter, dev_s);
  if ((tid >> 6) == 0) {
    const int b = wave_chol_inv32<T>(G, Yl);
    if (tid == 0) bad_s = b;
  }
  __syncthreads();
  if (tid == 0) {
    const int prev = status[0];
    if (bad_s != 0 && prev == 0) status[2] = j0;
    status[0] = (bad_s != 0 || (sticky && prev != 0)) ? 1 : 0;
  }
  if (bad_s != 0) return;
  for (int e = tid; e < NB * NB; e += 256) {
    R1[e] = G[e % NB][e / NB];
    R1inv[e] = Yl[e % NB][e / NB];
  }
}

// In-wave LU of an N x N block without pivoting, with the sign choice
// d_j = -sign(M(j, j)) subtracted from each pivot first (|pivot| >= 1 for the
// top of an orthonormal Q).  Lane c keeps column c: U(r, c) for r <= c and,
// below the diagonal, the eliminated entry M(r, c) (scaled into L(r, c) =
// M(r, c) / U(c, c) by luN_scale_l afterwards); other lanes' entries come
// from lane broadcasts.  Step j: lanes c > j update S(r, c) -= M(r, j) /
// U(j, j) * S(j, c) for r > j (lanes c <= j multiply by zero, which leaves
// their finite entries exact).  dv receives d (when dv is not null), `pinv`
// of lane j receives 1 / U(j, j).
template <typename T, int N, int J>
__device__ __forceinline__ void luN_step(T (&s)[N], int c0, T* dv, T& pinv) {
  const int c = step_lane(c0);
  const T piv = rdl(s[J], J);
  const T d = s_neg(sign_safe(piv));
  const T p = s_sub(piv, d);
  if (dv != nullptr && c == 0) dv[J] = d;
  const T inv = s_recip(p);
  pinv = sel(c == J, inv, pinv);
  const T mjc = sel(c > J, s[J], s_zero(T()));  // U(J, c) for the lanes that update
  s[J] = sel(c == J, p, s[J]);
  const T f = s_neg(s_mul(inv, mjc));            // -U(J, c) / U(J, J)
  T dep = p;
#pragma unroll
  for (int r0 = J + 1; r0 < N; r0 += kBc<T>) {
    T b[kBc<T>];
#pragma unroll
    for (int u = 0; u < kBc<T>; ++u) {
      if (r0 + u < N) order_after(s[r0 + u], dep);
      b[u] = (r0 + u < N) ? rdl(s[r0 + u], J) : s_zero(T());   // M(r, J)
    }
#pragma unroll
    for (int u = 0; u < kBc<T>; ++u)
      if (r0 + u < N) s[r0 + u] = s_fma(b[u], f, s[r0 + u]);
    dep = s[r0];
  }
}
template <typename T, int N, int... Js>
__device__ __forceinline__ void luN_steps(T (&s)[N], int c, T* dv, T& pinv, std::integer_sequence<int, Js...>) {
  (luN_step<T, N, Js>(s, c, dv, pinv), ...);
}
// L(r, c) = M(r, c) / U(c, c) below the diagonal (lane c holds 1 / U(c, c) in pinv).
template <typename T, int N>
__device__ __forceinline__ void luN_scale_l(T (&s)[N], int c, const T& pinv) {
#pragma unroll
  for (int r = 0; r < N; ++r) s[r] = sel(r > c, s_mul(s[r], pinv), s[r]);
}

// LU of the 16 x 16 diagonal block (r0, r0) of In by the calling wave (lanes
// c = lane % 16 hold column c): L \ U and d go to Out and dv (when dv is not
// null), U^{-1} to Ui (when not null), L^{-1} to Li (when not null), all at
// the same block position.  In and Out must differ.
template <typename T>
__device__ __forceinline__ void wave_lu16(const T (*In)[33], T (*Out)[33], int r0, T* dv, T (*Ui)[33],
                                          T (*Li)[33]) {
  const int l = lane_id();
  const int c = l & 15;
  const bool writer = l < 16;
  T s[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) s[r] = In[r0 + r][r0 + c];
  T pinv = s_zero(T());
  luN_steps<T, 16>(s, c, dv != nullptr ? dv + r0 : (T*)nullptr, pinv, std::make_integer_sequence<int, 16>{});
  luN_scale_l<T, 16>(s, c, pinv);
  if (dv != nullptr) {
#pragma unroll
    for (int r = 0; r < 16; ++r)
      if (writer) Out[r0 + r][r0 + c] = s[r];
  }
  T y[16];
  if (Ui != nullptr) {
#pragma unroll
    for (int r = 0; r < 16; ++r) y[r] = s_zero(T());
    trinvN_rows<T, 16>(y, s, pinv, c, std::make_integer_sequence<int, 16>{});
#pragma unroll
    for (int r = 0; r < 16; ++r)
      if (writer) Ui[r0 + r][r0 + c] = y[r];
  }
  if (Li != nullptr) {
    trilinvN_rows<T, 16>(y, s, c, std::make_integer_sequence<int, 16>{});
#pragma unroll
    for (int r = 0; r < 16; ++r)
      if (writer) Li[r0 + r][r0 + c] = y[r];
  }
}

// LU of the 32 x 32 block Qt (padded on entry; its lower-right block is used
// as scratch) with the sign choice d, as 2 x 2 blocks of 16 by waves 0 and 1
// of the workgroup (every thread must call it): Ls = L \ U, dv = d,
// Ui = U^{-1}, Li = L^{-1}.  Both waves factor each diagonal block; wave 0
// writes L \ U and U^{-1}, wave 1 L^{-1}; the off-diagonal blocks are MFMA
// products: U12 = L11^{-1} A12, L21 = A21 U11^{-1}, A22 - L21 U12,
// (U^{-1})12 = -U11^{-1} U12 U22^{-1}, (L^{-1})21 = -L22^{-1} L21 L11^{-1}.
template <typename T>
__device__ __forceinline__ void lu32_recon(T (*Qt)[33], T (*Ls)[33], T* dv, T (*Ui)[33], T (*Li)[33]) {
  auto blk = [](T (*M)[33], int r0, int c0) { return reinterpret_cast<T (*)[33]>(&M[r0][c0]); };
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  for (int e = tid; e < NB * NB; e += 256) Ls[e % NB][e / NB] = Qt[e % NB][e / NB];
  __syncthreads();
  if (w == 0) wave_lu16<T>(Qt, Ls, 0, dv, Ui, nullptr);
  else if (w == 1) wave_lu16<T>(Qt, Ls, 0, nullptr, nullptr, Li);
  __syncthreads();
  if (w == 0) wave_mm16<T, false>(blk(Ls, 0, 16), blk(Li, 0, 0), blk(Ls, 0, 16), 1.0, false);        // U12
  else if (w == 1) wave_mm16<T, false>(blk(Ls, 16, 0), blk(Ls, 16, 0), blk(Ui, 0, 0), 1.0, false);  // L21
  __syncthreads();
  if (w == 0) wave_mm16<T, false>(blk(Qt, 16, 16), blk(Ls, 16, 0), blk(Ls, 0, 16), -1.0, true);  // A22 - L21 U12
  __syncthreads();
  if (w == 0) wave_lu16<T>(Qt, Ls, 16, dv, Ui, nullptr);
  else if (w == 1) wave_lu16<T>(Qt, Ls, 16, nullptr, nullptr, Li);
  __syncthreads();
  const int l = tid & 63;
  if (w == 0) {
    wave_mm16<T, false>(blk(Ui, 0, 16), blk(Ls, 0, 16), blk(Ui, 16, 16), 1.0, false);   // U12 U22^{-1}
    wave_mm16<T, false>(blk(Ui, 0, 16), blk(Ui, 0, 0), blk(Ui, 0, 16), -1.0, false);    // -U11^{-1} (..)
    for (int e = l; e < 256; e += 64) Ui[16 + (e & 15)][e >> 4] = s_zero(T());
  } else if (w == 1) {
    wave_mm16<T, false>(blk(Li, 16, 0), blk(Ls, 16, 0), blk(Li, 0, 0), 1.0, false);    // L21 L11^{-1}
    wave_mm16<T, false>(blk(Li, 16, 0), blk(Li, 16, 16), blk(Li, 16, 0), -1.0, false);  // -L22^{-1} (..)
    for (int e = l; e < 256; e += 64) Li[e & 15][16 + (e >> 4)] = s_zero(T());
  }
  __syncthreads();
}

// Pass 2 and the reconstruction (runs only if the flag is clear after pass 1):
// every workgroup forms its tiles of Q1 = X R1^{-1} (to Q1, H x 32, ld H) and
// their partial Gram; the last one factors G2 = W2^H W2 and checks it (sets
// the flag and status[2] = j0 and returns when it fails), then forms R = W2 R1, the top of
// Q = Q1 W2^{-1}, its LU with signs D, T = -U D^H L^{-H}, M = W2^{-1} U^{-1}
// (to Mout, 32 x 32) and the top nbc rows of the panel Ap (D R on and above
// the diagonal, L below).  T goes to Tout (32 x 32, zero outside nbc x nbc).
// With Xn (not null) it also forms Gw = Xn^H Xn (H x nbcn, ld ldx; 32 x 32, zero
// outside nbcn x nbcn) as one 32 x 32 partial per workgroup in partialN, summed
// into Gw here when reduce_next is set (otherwise the next trailing update's
// split-K product sums them on the side, GemmArgs::rPart).
// With carried, pass 1 runs here instead of in a kernel of its own: every
// workgroup forms G1 = Gw - Rt^H Rt (Gw as the previous panel's call left it,
// Rt the 32 rows above X) and factors it, redundantly and identically, while
// its first tile's loads are in flight; workgroup 0 writes R1 for the last
// workgroup (published by the ticket).  A G1 that is not positive definite
// sets the flag (and status[2] = j0) and every workgroup returns.
template <typename T>
__global__ void __launch_bounds__(256) cholqr2_finish_kernel(const T* X, int64_t ldx, int H, int nbc, T* R1,
                                                             const T* R1inv, T* Q1, T* partial, int* status,
                                                             int j0, T* Tout, T* Mout, T* Ap, int64_t lda,
                                                             const T* Xn, int nbcn, T* partialN, T* Gw,
                                                             int carried, int reduce_next) {
  constexpr int TR = CqTile<T>::TR;
  constexpr int XM = TR / NB;  // 32-row matrices spanned by the staged tile
  // the staged tiles of X and Xn, R1^{-1}; at least the 8 matrices of the end
  constexpr int NPOOL = (2 * XM + 1 > 8) ? 2 * XM + 1 : 8;
  __shared__ T pool[NPOOL][NB][NB + 1];
  __shared__ T dv[NB];
  __shared__ int last_s, bad_s;
  __shared__ double dev_s[4];
  if (status[0] != 0) return;
  int* counter = status + 1;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int c = lane & 31;
  {
    T (*Bs)[NB + 1] = pool[XM];
    if (!carried)
      for (int e = tid; e < NB * NB; e += 256) Bs[e % NB][e / NB] = R1inv[e];
    auto chol1 = [&]() {
      if (!carried) return;
      T (*G1)[NB + 1] = pool[0];
      T (*Rs)[NB + 1] = pool[1];
      for (int e = tid; e < NB * NB; e += 256) {
        const int i = e % NB, j = e / NB;
        G1[i][j] = Gw[e];
        Rs[i][j] = (j < nbc) ? X[i - NB + (int64_t)j * ldx] : s_zero(T());
      }
      __syncthreads();
      mm32<T>(G1, Rs, true, Rs, false, -1.0, true);  // G1 = Gw - Rt^H Rt
      for (int e = tid; e < NB * NB; e += 256) {    // identity outside nbc x nbc
        const int i = e % NB, j = e / NB;
        if (i >= nbc || j >= nbc) G1[i][j] = (i == j) ? s_one(T()) : s_zero(T());
      }
      __syncthreads();
      if (w == 0) {
        const int b = wave_chol_inv32<T>(G1, Bs);
        if (lane == 0) bad_s = b;
      }
      __syncthreads();
      if (bad_s == 0 && blockIdx.x == 0)
        for (int e = tid; e < NB * NB; e += 256) R1[e] = G1[e % NB][e / NB];
      __syncthreads();
    };
    v4d cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
    v4d nr = {0.0, 0.0, 0.0, 0.0}, ni = {0.0, 0.0, 0.0, 0.0};
    // (gram_tiles begins each tile with a workgroup barrier)
    gram_tiles<T, kTimesB>(X, ldx, H, nbc, pool[0], Bs, Q1, cr, ci, Xn, nbcn, pool[XM + 1], &nr, &ni, chol1);
    if (carried && bad_s != 0) {
      if (tid == 0 && blockIdx.x == 0) {
        status[2] = j0;
        status[0] = 1;
      }
      return;
    }
    if (Xn != nullptr) store_partial<T>(partialN, nr, ni);
    if (!partial_is_last<T>(partial, cr, ci, counter, &last_s)) return;
  }
  if (Xn != nullptr && reduce_next) {  // Gw = sum of the partials (zero outside nbcn x nbcn)
    T (*Gn)[NB + 1] = pool[1];
    reduce_partials<T>(partialN, (int)gridDim.x, NB, Gn, counter, dev_s);
    for (int e = tid; e < NB * NB; e += 256) Gw[e] = Gn[e % NB][e / NB];
    __syncthreads();
  }
  T (*G)[NB + 1] = pool[0];   // G2, then W2
  T (*Yl)[NB + 1] = pool[1];  // W2^{-1}
  T (*Rl)[NB + 1] = pool[2];  // R1, then R = W2 R1
  T (*Qt)[NB + 1] = pool[3];  // top of Q1, then of Q; later M
  T (*Ls)[NB + 1] = pool[4];  // L (unit lower), U on and above
  T (*Li)[NB + 1] = pool[5];  // L^{-1}
  T (*Ms)[NB + 1] = pool[6];  // U D^H, then T
  T (*Ui)[NB + 1] = pool[7];  // U^{-1}
  const double devF2 = reduce_partials<T>(partial, (int)gridDim.x, nbc, G, counter, dev_s);
  // G2 = I + E.  For ||E||_F <= 1e-6 (the usual case: ||E|| ~ kappa(X)^2 u
  // after pass 1) its Cholesky factor and inverse follow from E to within
  // O(||E||^3) <= 1e-18:  U1 = up(E), U = up(E - U1^H U1), W2 = I + U,
  // W2^{-1} = I - U + U^2, where up() keeps the strict upper part and half
  // the diagonal (from (I + U)^H (I + U) = I + E).  Otherwise one wave factors G2.
  const bool near_identity = devF2 <= 1e-12;
  if (!near_identity && w == 0) {
    const int b = wave_chol_inv32<T>(G, Yl);
    if (lane == 0) bad_s = b;
  }
  if (near_identity && tid == 0) bad_s = 0;
  for (int e = tid; e < NB * NB; e += 256) {
    const int i = e % NB, j = e / NB;
    Rl[i][j] = R1[e];
    Qt[i][j] = Q1[i + (int64_t)j * H];  // H >= 64: rows 0..31 exist
  }
  if (near_identity) {
    T (*Tmp)[NB + 1] = pool[6];
    auto up = [&](int i, int j, T v) {  // v = E(i, j); the diagonal of a Hermitian matrix is real
      return (i < j) ? v : ((i == j) ? s_scale(s_add(v, s_conj(v)), 0.25) : s_zero(T()));
    };
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      Yl[i][j] = up(i, j, s_sub(G[i][j], (i == j) ? s_one(T()) : s_zero(T())));  // U1
    }
    __syncthreads();
    mm32<T>(Tmp, Yl, true, Yl, false, 1.0, false);  // U1^H U1
    __syncthreads();
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      const T u = up(i, j, s_sub(s_sub(G[i][j], (i == j) ? s_one(T()) : s_zero(T())), Tmp[i][j]));
      Yl[i][j] = u;                                                    // U
      G[i][j] = (i == j) ? s_add(s_one(T()), u) : u;                   // W2 = I + U
    }
    __syncthreads();
    mm32<T>(Tmp, Yl, false, Yl, false, 1.0, false);  // U^2
    __syncthreads();
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      const T v = s_add(s_neg(Yl[i][j]), Tmp[i][j]);
      Yl[i][j] = (i == j) ? s_add(s_one(T()), v) : v;                  // W2^{-1} = I - U + U^2
    }
  }
  __syncthreads();
  if (bad_s != 0 || !(devF2 <= 0.01)) {
    if (tid == 0) {
      status[2] = j0;
      status[0] = 1;
    }
    return;
  }
  mm32<T>(Rl, G, false, Rl, false, 1.0, false);   // R = W2 R1
  mm32<T>(Qt, Qt, false, Yl, false, 1.0, false);  // top of Q = Q1 W2^{-1}
  for (int e = tid; e < NB * NB; e += 256) {  // pad the top of Q with the identity outside nbc x nbc
    const int r = e % NB, cc = e / NB;
    Qt[r][cc] = sel(r < nbc && cc < nbc, Qt[r][cc], sel(r == cc, s_one(T()), s_zero(T())));
  }
  lu32_recon<T>(Qt, Ls, dv, Ui, Li);
  for (int e = tid; e < NB * NB; e += 256) {
    const int r = e % NB, cc = e / NB;
    Ms[r][cc] = (r <= cc) ? s_mul(Ls[r][cc], s_conj(dv[cc])) : s_zero(T());
  }
  __syncthreads();
  mm32<T>(Ms, Ms, false, Li, true, -1.0, false);  // T = -(U D^H) L^{-H}
  mm32<T>(Qt, Yl, false, Ui, false, 1.0, false);  // M = W2^{-1} U^{-1}
  for (int e = tid; e < NB * NB; e += 256) {
    const int r = e % NB, cc = e / NB;
    const bool in = r < nbc && cc < nbc;
    Tout[e] = (in && r <= cc) ? Ms[r][cc] : s_zero(T());
    Mout[e] = in ? Qt[r][cc] : s_zero(T());
    if (in) Ap[r + (int64_t)cc * lda] = (r <= cc) ? s_mul(dv[r], Rl[r][cc]) : Ls[r][cc];
  }
}

// Wave w's 16 x 16 block (w % 2, w / 2) of a Gram accumulated by gram_tiles
// into G (32 x 32 in LDS; identity outside nbc x nbc).  Returns nothing; the
// caller synchronises.
template <typename T>
__device__ inline void acc_to_lds(T (*G)[NB + 1], const v4d& cr, const v4d& ci, int nbc) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ti = w & 1, tj = w >> 1;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = ti * 16 + mfma_f64_row(lane, r), j = tj * 16 + (lane & 15);
    T v;
    if constexpr (ScalarTraits<T>::is_complex) v = mkz(cr[r], ci[r]);
    else v = cr[r];
    G[i][j] = (i < nbc && j < nbc) ? v : ((i == j) ? s_one(T()) : s_zero(T()));
  }
}

// One panel per workgroup from a task list (the panels of a batch of
// independent problems): the same CholeskyQR2 with Householder
// reconstruction as cholqr_panel, every pass in one workgroup:
//   G1 = X^H X, G1 = R1^H R1 (one wave), Q1 = X R1^{-1} to the task's Q1
//   scratch with G2 = Q1^H Q1 alongside, W2 (closed form or one wave), the
//   check, R = W2 R1, the LU of the top of Q with signs D, T, M, and
//   V_rest = Q1_rest M into the panel.
// X is only read until the check has passed: a rejected panel (status 1) is
// untouched, and the Householder launches gated on its status factor it.
template <typename T>
__global__ void __launch_bounds__(256) cholqr2_panel_tasks_kernel(const CholQrTask* tasks) {
  constexpr int TR = CqTile<T>::TR;
  constexpr int XM = TR / NB;
  constexpr int NPOOL = (XM + 4 > 8) ? XM + 4 : 8;
  __shared__ T pool[NPOOL][NB][NB + 1];
  __shared__ T dv[NB];
  __shared__ int bad_s;
  __shared__ double dev_s[4];
  const CholQrTask tk = tasks[blockIdx.x];
  T* X = reinterpret_cast<T*>(tk.X);
  T* Q1 = reinterpret_cast<T*>(tk.Q1);
  const int64_t ldx = tk.ldx, ldq = tk.ldq;
  const int H = tk.H, nbc = tk.nbc;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  T (*Bs)[NB + 1] = pool[XM];      // R1^{-1}
  T (*R1)[NB + 1] = pool[XM + 1];  // G1, then R1, then R = W2 R1
  // pass 1
  {
    v4d cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
    gram_tiles<T, kPlain>(X, ldx, H, nbc, pool[0], nullptr, nullptr, cr, ci, nullptr, 0, nullptr, nullptr, nullptr,
                          NoPre(), 0, 1);
    acc_to_lds<T>(R1, cr, ci, nbc);
  }
  __syncthreads();
  if (w == 0) {
    const int b = wave_chol_inv32<T>(R1, Bs);
    if (lane == 0) bad_s = b;
  }
  __syncthreads();
  if (bad_s != 0) {
    if (tid == 0) *tk.status = 1;
    return;
  }
  // pass 2: Q1 = X R1^{-1} (nbc columns) and G2 = Q1^H Q1
  T (*G)[NB + 1] = pool[XM + 2];  // G2, then W2
  {
    v4d cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
    gram_tiles<T, kTimesB>(X, ldx, H, nbc, pool[0], Bs, Q1, cr, ci, nullptr, 0, nullptr, nullptr, nullptr, NoPre(),
                           0, 1, ldq, nbc);
    acc_to_lds<T>(G, cr, ci, nbc);
  }
  __syncthreads();
  double dev = 0.0;
  for (int e = tid; e < NB * NB; e += 256) {
    const int i = e % NB, j = e / NB;
    dev += s_abs2(s_sub(G[i][j], (i == j) ? s_one(T()) : s_zero(T())));
  }
  dev = wave_sum(dev);
  if (lane == 0) dev_s[w] = dev;
  __syncthreads();
  const double devF2 = dev_s[0] + dev_s[1] + dev_s[2] + dev_s[3];
  // The reconstruction's matrices.  Pool slots 0 .. XM - 1 stage the row
  // tiles again in the last pass, so they hold only what is dead by then;
  // M (in Qt) and L (in Ls) outlive it, in R1^{-1}'s slot and the last one.
  T (*Yl)[NB + 1] = pool[0];                          // W2^{-1}
  T (*Li)[NB + 1] = pool[1];                          // L^{-1}
  T (*Ms)[NB + 1] = pool[XM >= 4 ? 2 : XM + 3];       // closed-form scratch, U D^H, then T
  T (*Ui)[NB + 1] = pool[XM >= 4 ? 3 : XM + 4];       // U^{-1}
  T (*Qt)[NB + 1] = pool[XM];                         // top of Q1, then of Q; then M
  T (*Ls)[NB + 1] = pool[NPOOL - 1];                  // L (unit lower), U on and above
  static_assert(NPOOL - 1 > XM + 2 && (XM >= 4 || XM + 4 < NPOOL - 1), "distinct slots");
  // G2 = I + E: closed form for ||E||_F <= 1e-6 (as in the whole-panel
  // kernel), else one wave factors G2
  const bool near_identity = devF2 <= 1e-12;
  if (!near_identity && w == 0) {
    const int b = wave_chol_inv32<T>(G, Yl);
    if (lane == 0) bad_s = b;
  }
  for (int e = tid; e < NB * NB; e += 256) {
    const int i = e % NB, j = e / NB;
    Qt[i][j] = (j < nbc) ? Q1[i + (int64_t)j * ldq] : s_zero(T());  // H >= 64: rows 0..31 exist
  }
  if (near_identity) {
    auto up = [&](int i, int j, T v) {
      return (i < j) ? v : ((i == j) ? s_scale(s_add(v, s_conj(v)), 0.25) : s_zero(T()));
    };
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      Yl[i][j] = up(i, j, s_sub(G[i][j], (i == j) ? s_one(T()) : s_zero(T())));  // U1
    }
    __syncthreads();
    mm32<T>(Ms, Yl, true, Yl, false, 1.0, false);  // U1^H U1
    __syncthreads();
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      const T u = up(i, j, s_sub(s_sub(G[i][j], (i == j) ? s_one(T()) : s_zero(T())), Ms[i][j]));
      Yl[i][j] = u;                                   // U
      G[i][j] = (i == j) ? s_add(s_one(T()), u) : u;  // W2 = I + U
    }
    __syncthreads();
    mm32<T>(Ms, Yl, false, Yl, false, 1.0, false);  // U^2
    __syncthreads();
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      const T v = s_add(s_neg(Yl[i][j]), Ms[i][j]);
      Yl[i][j] = (i == j) ? s_add(s_one(T()), v) : v;  // W2^{-1} = I - U + U^2
    }
  }
  __syncthreads();
  if ((!near_identity && bad_s != 0) || !(devF2 <= 0.01)) {
    if (tid == 0) *tk.status = 1;
    return;
  }
  mm32<T>(R1, G, false, R1, false, 1.0, false);   // R = W2 R1
  mm32<T>(Qt, Qt, false, Yl, false, 1.0, false);  // top of Q = Q1 W2^{-1}
  for (int e = tid; e < NB * NB; e += 256) {      // identity outside nbc x nbc
    const int r = e % NB, cc = e / NB;
    Qt[r][cc] = sel(r < nbc && cc < nbc, Qt[r][cc], sel(r == cc, s_one(T()), s_zero(T())));
  }
  lu32_recon<T>(Qt, Ls, dv, Ui, Li);
  for (int e = tid; e < NB * NB; e += 256) {
    const int r = e % NB, cc = e / NB;
    Ms[r][cc] = (r <= cc) ? s_mul(Ls[r][cc], s_conj(dv[cc])) : s_zero(T());
  }
  __syncthreads();
  mm32<T>(Ms, Ms, false, Li, true, -1.0, false);  // T = -(U D^H) L^{-H}
  mm32<T>(Qt, Yl, false, Ui, false, 1.0, false);  // M = W2^{-1} U^{-1}
  T* Tout = reinterpret_cast<T*>(tk.Tout);
  for (int e = tid; e < NB * NB; e += 256) {
    const int r = e % NB, cc = e / NB;
    const bool in = r < nbc && cc < nbc;
    Tout[e] = (in && r <= cc) ? Ms[r][cc] : s_zero(T());
    // padding outside nbc x nbc keeps M's product with the zero-padded tiles exact
    if (!in) Qt[r][cc] = s_zero(T());
  }
  __syncthreads();
  // V_rest = Q1_rest M (rows nbc .. H - 1) into the panel, then its top rows:
  // D R on and above the diagonal, L below (the top rows are read by nothing
  // after this point)
  if (H > nbc) {
    v4d cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
    gram_tiles<T, kProduct>(Q1 + nbc, ldq, H - nbc, nbc, pool[0], Qt, X + nbc, cr, ci, nullptr, 0, nullptr,
                            nullptr, nullptr, NoPre(), 0, 1, ldx, nbc);
  }
  for (int e = tid; e < NB * NB; e += 256) {
    const int r = e % NB, cc = e / NB;
    if (r < nbc && cc < nbc) X[r + (int64_t)cc * ldx] = (r <= cc) ? s_mul(dv[r], R1[r][cc]) : Ls[r][cc];
  }
  if (tid == 0) *tk.status = 0;
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

int cholqr_gram_blocks(int H) { return std::max(1, std::min(kMaxGramBlocks, cdiv(H, 256))); }

size_t cholqr_workspace_elems(int m) {
  // the carried Gram and its partials, the partial Grams, R1, R1^{-1}, M,
  // Q1 (m x 32)
  return (size_t)std::max(m, 1) * NB + (size_t)kMaxGramBlocks * NB * NB + (size_t)3 * NB * NB +
         (size_t)kMaxGramBlocks * NB * NB + (size_t)NB * NB;
}

template <typename T>
void cholqr_panel(T* Ap, int64_t lda, int H, int nbc, T* Tpan, T* cw, int* status, int j0, int sticky,
                  int carried, int nbc_next, hipStream_t s) {
  // fixed offsets first (the carried Gram must sit at the same address for
  // the next panel, whose H is smaller), Q1 last
  T* Gw = cw;
  T* partN = Gw + NB * NB;
  T* part = partN + (size_t)kMaxGramBlocks * NB * NB;
  T* R1 = part + (size_t)kMaxGramBlocks * NB * NB;
  T* R1i = R1 + NB * NB;
  T* M = R1i + NB * NB;
  T* Q1 = M + NB * NB;
  const int nblk = cholqr_gram_blocks(H);
  if (carried == 0)
    hipLaunchKernelGGL((gram_chol_kernel<T>), dim3(nblk), dim3(256), 0, s, (const T*)Ap, lda, H, nbc, part,
                       status, j0, sticky, R1, R1i);
  const T* Xn = nbc_next > 0 ? (const T*)(Ap + (int64_t)nbc * lda) : nullptr;
  hipLaunchKernelGGL((cholqr2_finish_kernel<T>), dim3(nblk), dim3(256), 0, s, (const T*)Ap, lda, H, nbc, R1,
                     (const T*)R1i, Q1, part, status, j0, Tpan, M, Ap, lda, Xn, nbc_next, partN, Gw,
                     carried == 1 ? 1 : 0, ScalarTraits<T>::is_complex ? 1 : 0);
  if (H > nbc) gemm<T>(OP_N, OP_N, gg<T>(H - nbc, nbc, nbc, Q1 + nbc, H, M, NB, Ap + nbc, lda, status), s);
}

template <typename T> void launch_cholqr2_panel_tasks(const CholQrTask* tasks, int ntasks, hipStream_t s) {
  if (ntasks > 0) hipLaunchKernelGGL((cholqr2_panel_tasks_kernel<T>), dim3(ntasks), dim3(256), 0, s, tasks);
}
template void launch_cholqr2_panel_tasks<double>(const CholQrTask*, int, hipStream_t);
template void launch_cholqr2_panel_tasks<zd>(const CholQrTask*, int, hipStream_t);

template <typename T> void cholqr_carry_buffers(T* cw, T** Gw, T** partN, T** R1, T** R1inv) {
  *Gw = cw;
  *partN = cw + NB * NB;
  T* part = cw + NB * NB + (size_t)kMaxGramBlocks * NB * NB;
  *R1 = part + (size_t)kMaxGramBlocks * NB * NB;
  *R1inv = *R1 + NB * NB;
}
template void cholqr_carry_buffers<double>(double*, double**, double**, double**, double**);
template void cholqr_carry_buffers<zd>(zd*, zd**, zd**, zd**, zd**);

template void cholqr_panel<double>(double*, int64_t, int, int, double*, double*, int*, int, int, int, int,
                                   hipStream_t);
template void cholqr_panel<zd>(zd*, int64_t, int, int, zd*, zd*, int*, int, int, int, int, hipStream_t);

}  // namespace mak
