// FP64 MFMA GEMM for gfx950 (v_mfma_f64_16x16x4_f64), real and complex.
//
// C = alpha op(A) op(B) + beta C, column-major, op in {N, T, C}.  One
// workgroup (4 waves) computes a BM x BN tile; each wave owns a
// (BM/WM) x (BN/WN) sub-tile of 16x16 MFMA blocks.  Tiles of op(A) and op(B)
// are staged through LDS as [k][row] / [k][col] so a fragment read is 16
// consecutive doubles per k.  Complex products use four real MFMAs on split
// real/imaginary LDS tiles (there is no complex MFMA).
//
// Split-K (ksplit > 1) writes per-slice partial products; the WY update uses
// it for the tall inner products V^H C whose output is only 32 rows.
#include <algorithm>

#include "mak_common.hpp"
#include "mak_dense32.hpp"
#include "mak_kernels.hpp"

namespace mak {

template <int OPA>
__device__ inline double load_a(const double* A, int64_t lda, int gi, int gk, int unitLower) {
  // op(A)(gi, gk); stored index (r, c)
  const int r = (OPA == OP_N) ? gi : gk;
  const int c = (OPA == OP_N) ? gk : gi;
  if (unitLower && r <= c) return r == c ? 1.0 : 0.0;
  return A[r + (int64_t)c * lda];
}
template <int OPA>
__device__ inline zd load_a(const zd* A, int64_t lda, int gi, int gk, int unitLower) {
  const int r = (OPA == OP_N) ? gi : gk;
  const int c = (OPA == OP_N) ? gk : gi;
  if (unitLower && r <= c) return r == c ? mkz(1.0, 0.0) : mkz(0.0, 0.0);
  zd v = A[r + (int64_t)c * lda];
  if (OPA == OP_C) v.y = -v.y;
  return v;
}
template <int OPB>
__device__ inline double load_b(const double* B, int64_t ldb, int gk, int gj) {
  return (OPB == OP_N) ? B[gk + (int64_t)gj * ldb] : B[gj + (int64_t)gk * ldb];
}
template <int OPB>
__device__ inline zd load_b(const zd* B, int64_t ldb, int gk, int gj) {
  zd v = (OPB == OP_N) ? B[gk + (int64_t)gj * ldb] : B[gj + (int64_t)gk * ldb];
  if (OPB == OP_C) v.y = -v.y;
  return v;
}

// Tile order: linear workgroup ids are dealt round-robin to the 8 XCDs; regroup
// them so every XCD walks a contiguous range of tiles (shared A / B panels stay
// in that XCD's L2).  Under structure hints the work per tile varies with its
// row or column, and contiguous ranges would leave one XCD with all the long
// tiles; those launches keep the round-robin deal (every XCD gets every
// column), and the index whose K range grows (TRI_A_LOWER: rows, TRI_B_UPPER:
// columns; upperC: columns) is reversed so that the longest tiles go first.
//
// upperC launches use a 1-D grid over the tiles on or above the diagonal only
// (times the z slabs; see upper_tiles): with a 2-D grid every tile of tile
// row r would land on XCD r % 8, and the upper triangle's first rows hold
// most of the tiles.  Consecutive ids walk the triangle column by column from
// the last column (the longest K ranges under triangular hints come first).
__device__ inline void tile_coords(int& tm, int& tn, int tri, int upperC) {
  const int gx = gridDim.x, gy = gridDim.y;
  const int total = gx * gy;
  int L = blockIdx.x + blockIdx.y * gx;
  if ((total & 7) == 0 && tri == 0 && upperC == 0) L = (L & 7) * (total >> 3) + (L >> 3);
  tm = L % gx;
  tn = L / gx;
  if (tri & TRI_A_LOWER) tm = gx - 1 - tm;
  if ((tri & TRI_B_UPPER) || upperC) tn = gy - 1 - tn;
}

// Tiles (tm <= tn) of a gx x gy tile grid with square tiles.
__host__ __device__ inline int upper_tiles(int gx, int gy) {
  int t = 0;
  for (int j = 0; j < gy; ++j) t += (j + 1 < gx ? j + 1 : gx);
  return t;
}
// Workgroup -> (tile row, tile column, z slab) of a compact upperC launch.
__device__ inline void upper_tile_coords(int gx, int gy, int& tm, int& tn, int& bz) {
  const int nup = upper_tiles(gx, gy);
  int L = blockIdx.x;
  bz = L / nup;
  L %= nup;
  int j = gy - 1;
  for (;; --j) {
    const int c = (j + 1 < gx ? j + 1 : gx);
    if (L < c || j == 0) break;
    L -= c;
  }
  tn = j;
  tm = L;
}

// Clip [kbeg, kend) to the terms a tile (rows m0.., cols n0..) can have
// nonzero under the triangular-operand hints (GemmArgs::tri).
__device__ inline void tri_k_range(int tri, int m0, int bm, int n0, int bn, int& kbeg, int& kend) {
  if (tri & TRI_A_LOWER) kend = min(kend, m0 + bm);
  if (tri & TRI_B_UPPER) kend = min(kend, n0 + bn);
  if (tri & TRI_A_UPPER) kbeg = max(kbeg, m0);
  if (tri & TRI_B_LOWER) kbeg = max(kbeg, n0);
}

// The carried first Gram (GemmArgs::cGw) in the workgroup of tile (0, 0),
// after its C stores: Rt and G1 in the LDS of the operand tiles (each at least
// 32 x 33 doubles), W^{-1} over Rt once G1 is formed.
__device__ inline void carry_chol(const GemmArgs& g, double* lds_a, double* lds_b) {
  __syncthreads();  // this workgroup's C stores are complete and visible to it
  const int* status = g.cStatus;
  if (status[0] != 0) return;
  const int tid = threadIdx.x;
  const int nbc = g.cNbc;
  double (*Rt)[33] = reinterpret_cast<double (*)[33]>(lds_a);
  double (*G)[33] = reinterpret_cast<double (*)[33]>(lds_b);
  const double* C = reinterpret_cast<const double*>(g.C);
  const double* Gw = reinterpret_cast<const double*>(g.cGw);
  for (int e = tid; e < 1024; e += 256) {
    const int i = e % 32, j = e / 32;
    Rt[i][j] = (j < nbc) ? C[i + (int64_t)j * g.ldc] : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < 1024; e += 256) {
    const int i = e % 32, j = e / 32;
    double v = Gw[e];
#pragma unroll 8
    for (int k = 0; k < 32; ++k) v = fma(-Rt[k][i], Rt[k][j], v);
    G[i][j] = (i < nbc && j < nbc) ? v : (i == j ? 1.0 : 0.0);
  }
  __shared__ int bad_s;
  __syncthreads();
  if ((tid >> 6) == 0) {
    const int b = wave_chol_inv32<double>(G, Rt);
    if (tid == 0) bad_s = b;
  }
  __syncthreads();
  if (bad_s != 0) {
    if (tid == 0) {
      g.cStatus[2] = g.cJ0;
      g.cStatus[0] = 1;
    }
    return;
  }
  double* R1 = reinterpret_cast<double*>(g.cR1);
  double* R1i = reinterpret_cast<double*>(g.cR1i);
  for (int e = tid; e < 1024; e += 256) {
    R1[e] = G[e % 32][e / 32];
    R1i[e] = Rt[e % 32][e / 32];
  }
}

// Real FP64 GEMM.  Each thread owns fixed load coordinates; the next K-tile is
// fetched into registers while the MFMAs of the current tile run.
// SIDE: the QR trailing update's side jobs -- with BM >= 64 the carried
// Cholesky in the workgroup of tile (0, 0) (GemmArgs::cGw), with BM == 32 the
// side reduction in an extra z slice (GemmArgs::rPart).
template <int BM, int BN, int BK, int WM, int WN, int OPA, int OPB, bool SIDE = false>
__global__ void __launch_bounds__(256, 2) gemm_f64_kernel(GemmArgs g) {
  constexpr int WTM = BM / WM, WTN = BN / WN, TM = WTM / 16, TN = WTN / 16;
  constexpr int NA = BM * BK / 256, NB_ = BN * BK / 256;
  static_assert(WM * WN == 4, "4 waves");
  static_assert(NA * 256 == BM * BK && NB_ * 256 == BN * BK, "tile / thread count");
  __shared__ double As[BK][BM + 2];
  __shared__ double Bs[BK][BN + 2];
  if constexpr (SIDE && BM == 32) {
    // the side reduction (GemmArgs::rPart): the first workgroup of the extra z slice
    if (blockIdx.z == (unsigned)(g.batch * g.ksplit)) {
      if (blockIdx.x == 0 && blockIdx.y == 0) {
        const double* P = reinterpret_cast<const double*>(g.rPart);
        double* out = reinterpret_cast<double*>(g.rOut);
        for (int e = threadIdx.x; e < 1024; e += 256) {
          double acc[4] = {0.0, 0.0, 0.0, 0.0};
          int q = 0;
          for (; q + 4 <= g.rN; q += 4)
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[u] += P[(int64_t)(q + u) * 1024 + e];
          for (; q < g.rN; ++q) acc[0] += P[(int64_t)q * 1024 + e];
          out[e] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
        }
      }
      return;
    }
  }
  if (gate_closed(g.gate, g.gate_want)) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int wm = w % WM, wn = w / WM;
  int tmi, tni, bz;
  if (BM == BN && g.upperC) {
    upper_tile_coords(cdiv(g.M, BM), cdiv(g.N, BN), tmi, tni, bz);
  } else {
    tile_coords(tmi, tni, g.tri, g.upperC);
    bz = blockIdx.z;
  }
  const int bat = bz / g.ksplit;
  const int ks = bz % g.ksplit;
  const int m0 = tmi * BM, n0 = tni * BN;
  if (g.upperC && m0 >= n0 + BN) return;
  int kbeg = ks * g.kchunk;
  int kend = min(g.K, kbeg + g.kchunk);
  tri_k_range(g.tri, m0, BM, n0, BN, kbeg, kend);
  const double* A = reinterpret_cast<const double*>(g.A) + (int64_t)bat * g.strideA;
  const double* B = reinterpret_cast<const double*>(g.B) + (int64_t)bat * g.strideB;
  double* C = reinterpret_cast<double*>(g.C) + (int64_t)bat * g.strideC + (int64_t)ks * g.splitStride;
  const int64_t lda = g.lda, ldb = g.ldb;

  // load coordinates: A element u of this thread is (ia[u], ka[u]) of the tile
  // OP_N: i fixed, k = k_a0 + u * (256 / BM); OP_T/C: k fixed, i = i_a0 + u * (256 / BK)
  const int ia0 = (OPA == OP_N) ? (tid % BM) : (tid / BK);
  const int ka0 = (OPA == OP_N) ? (tid / BM) : (tid % BK);
  constexpr int dA = (OPA == OP_N) ? (256 / BM) : (256 / BK);
  const int jb0 = (OPB == OP_N) ? (tid / BK) : (tid % BN);
  const int kb0 = (OPB == OP_N) ? (tid % BK) : (tid / BN);
  constexpr int dB = (OPB == OP_N) ? (256 / BK) : (256 / BN);

  double pa[NA], pb[NB_];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int i = (OPA == OP_N) ? ia0 : ia0 + u * dA;
      const int k = (OPA == OP_N) ? ka0 + u * dA : ka0;
      const int gi = m0 + i, gk = k0 + k;
      pa[u] = (gi < g.M && gk < kend) ? load_a<OPA>(A, lda, gi, gk, g.unitLowerA) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < NB_; ++u) {
      const int j = (OPB == OP_N) ? jb0 + u * dB : jb0;
      const int k = (OPB == OP_N) ? kb0 : kb0 + u * dB;
      const int gj = n0 + j, gk = k0 + k;
      pb[u] = (gj < g.N && gk < kend) ? load_b<OPB>(B, ldb, gk, gj) : 0.0;
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int i = (OPA == OP_N) ? ia0 : ia0 + u * dA;
      const int k = (OPA == OP_N) ? ka0 + u * dA : ka0;
      As[k][i] = pa[u];
    }
#pragma unroll
    for (int u = 0; u < NB_; ++u) {
      const int j = (OPB == OP_N) ? jb0 + u * dB : jb0;
      const int k = (OPB == OP_N) ? kb0 : kb0 + u * dB;
      Bs[k][j] = pb[u];
    }
  };

  v4d acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  if (kbeg < kend) fetch(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    stash();
    __syncthreads();
    if (k0 + BK < kend) fetch(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = As[kk + (lane >> 4)][wm * WTM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = Bs[kk + (lane >> 4)][wn * WTN + j * 16 + (lane & 15)];
      // operands swapped: the MFMA computes the transposed block, so lane l
      // holds rows (l % 16) of C and stores are 128-byte column segments
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = mfma_f64(bf[j], af[i], acc[i][j]);
    }
    __syncthreads();
  }
  const double alpha = g.alpha_re, beta = g.beta_re;
  const int rowb = m0 + wm * WTM + (lane & 15);
  const int colb = n0 + wn * WTN;
  if (beta != 0.0) {
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      double cv[TN][4];
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rowb + i * 16, col = colb + j * 16 + mfma_f64_row(lane, r);
          cv[j][r] = (row < g.M && col < g.N) ? C[row + (int64_t)col * g.ldc] : 0.0;
        }
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rowb + i * 16, col = colb + j * 16 + mfma_f64_row(lane, r);
          if (row < g.M && col < g.N) C[row + (int64_t)col * g.ldc] = fma(beta, cv[j][r], alpha * acc[i][j][r]);
        }
    }
  } else {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rowb + i * 16, col = colb + j * 16 + mfma_f64_row(lane, r);
          if (row < g.M && col < g.N) C[row + (int64_t)col * g.ldc] = alpha * acc[i][j][r];
        }
  }
  if constexpr (SIDE && BM >= 64) {
    static_assert(BK * (BM + 2) >= 32 * 33 && BK * (BN + 2) >= 32 * 33, "carry needs 32 x 33 of LDS per tile");
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && g.cGw != nullptr)
      carry_chol(g, &As[0][0], &Bs[0][0]);
  }
}

template <int BM, int BN, int BK, int WM, int WN, int OPA, int OPB>
__global__ void __launch_bounds__(256) gemm_z64_kernel(GemmArgs g) {
  constexpr int WTM = BM / WM, WTN = BN / WN, TM = WTM / 16, TN = WTN / 16;
  static_assert(WM * WN == 4, "4 waves");
  __shared__ double Ar[BK][BM + 2], Ai[BK][BM + 2];
  __shared__ double Br[BK][BN + 2], Bi[BK][BN + 2];
  if (gate_closed(g.gate, g.gate_want)) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int wm = w % WM, wn = w / WM;
  int tmi, tni, bz;
  if (BM == BN && g.upperC) {
    upper_tile_coords(cdiv(g.M, BM), cdiv(g.N, BN), tmi, tni, bz);
  } else {
    tile_coords(tmi, tni, g.tri, g.upperC);
    bz = blockIdx.z;
  }
  const int bat = bz / g.ksplit;
  const int ks = bz % g.ksplit;
  const int m0 = tmi * BM, n0 = tni * BN;
  if (g.upperC && m0 >= n0 + BN) return;
  int kbeg = ks * g.kchunk;
  int kend = min(g.K, kbeg + g.kchunk);
  tri_k_range(g.tri, m0, BM, n0, BN, kbeg, kend);
  const zd* A = reinterpret_cast<const zd*>(g.A) + (int64_t)bat * g.strideA;
  const zd* B = reinterpret_cast<const zd*>(g.B) + (int64_t)bat * g.strideB;
  zd* C = reinterpret_cast<zd*>(g.C) + (int64_t)bat * g.strideC + (int64_t)ks * g.splitStride;

  // Three real products per complex multiply-add (Gauss):
  //   p1 = sum ar br,  p2 = sum ai bi,  p3 = sum (ar + ai)(br + bi),
  //   re = p1 - p2,  im = p3 - p1 - p2
  // (three MFMAs instead of four; the error stays within a small multiple
  // of the usual componentwise bound sum |a| |b|).
  v4d p1[TM][TN], p2[TM][TN], p3[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      p1[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
      p2[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
      p3[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  // the next slab is loaded into registers while the current one is used
  constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
  static_assert(NA * 256 == BM * BK && NB * 256 == BN * BK, "whole slabs per thread");
  zd pa[NA], pb[NB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int e = tid + 256 * u;
      int i, k;
      if (OPA == OP_N) { i = e % BM; k = e / BM; }
      else { k = e % BK; i = e / BK; }
      const int gi = m0 + i, gk = k0 + k;
      pa[u] = (gi < g.M && gk < kend) ? load_a<OPA>(A, g.lda, gi, gk, g.unitLowerA) : mkz(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int e = tid + 256 * u;
      int j, k;
      if (OPB == OP_N) { k = e % BK; j = e / BK; }
      else { j = e % BN; k = e / BN; }
      const int gj = n0 + j, gk = k0 + k;
      pb[u] = (gj < g.N && gk < kend) ? load_b<OPB>(B, g.ldb, gk, gj) : mkz(0.0, 0.0);
    }
  };
  if (kbeg < kend) fetch(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int e = tid + 256 * u;
      int i, k;
      if (OPA == OP_N) { i = e % BM; k = e / BM; }
      else { k = e % BK; i = e / BK; }
      Ar[k][i] = pa[u].x;
      Ai[k][i] = pa[u].y;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int e = tid + 256 * u;
      int j, k;
      if (OPB == OP_N) { k = e % BK; j = e / BK; }
      else { j = e % BN; k = e / BN; }
      Br[k][j] = pb[u].x;
      Bi[k][j] = pb[u].y;
    }
    __syncthreads();
    if (k0 + BK < kend) fetch(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double ar[TM], ai[TM], as[TM], br[TN], bi[TN], bs[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        ar[i] = Ar[kk + (lane >> 4)][wm * WTM + i * 16 + (lane & 15)];
        ai[i] = Ai[kk + (lane >> 4)][wm * WTM + i * 16 + (lane & 15)];
        as[i] = ar[i] + ai[i];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        br[j] = Br[kk + (lane >> 4)][wn * WTN + j * 16 + (lane & 15)];
        bi[j] = Bi[kk + (lane >> 4)][wn * WTN + j * 16 + (lane & 15)];
        bs[j] = br[j] + bi[j];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          p1[i][j] = mfma_f64(br[j], ar[i], p1[i][j]);
          p2[i][j] = mfma_f64(bi[j], ai[i], p2[i][j]);
          p3[i][j] = mfma_f64(bs[j], as[i], p3[i][j]);
        }
    }
    __syncthreads();
  }
  const zd alpha = mkz(g.alpha_re, g.alpha_im), beta = mkz(g.beta_re, g.beta_im);
  const bool use_beta = (g.beta_re != 0.0 || g.beta_im != 0.0);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * WTM + i * 16 + (lane & 15);
        const int col = n0 + wn * WTN + j * 16 + mfma_f64_row(lane, r);
        if (row < g.M && col < g.N) {
          const double re = p1[i][j][r] - p2[i][j][r], im = p3[i][j][r] - p1[i][j][r] - p2[i][j][r];
          zd v = s_mul(alpha, mkz(re, im));
          zd* cp = C + row + (int64_t)col * g.ldc;
          if (use_beta) v = s_fma(beta, *cp, v);
          *cp = v;
        }
      }
}

// ---------------------------------------------------------------------------
// Dispatch.  Tile shapes: a square 128x128 tile for general products, a 32-row
// tile for the skinny V^H C products (M <= 32) of the WY updates.
// ---------------------------------------------------------------------------
template <int BM, int BN, int BK, int WM, int WN, int OPA, int OPB, typename T>
static void launch_gemm_tile(const GemmArgs& g, hipStream_t s) {
  dim3 grid(cdiv(g.M, BM), cdiv(g.N, BN), g.batch * g.ksplit);
  if constexpr (BM == BN) {
    // upperC: compact 1-D grid over the upper-triangle tiles (square tiles;
    // the kernels decode it with upper_tile_coords)
    if (g.upperC) grid = dim3((unsigned)(upper_tiles(cdiv(g.M, BM), cdiv(g.N, BN)) * g.batch * g.ksplit), 1, 1);
  }
  if constexpr (ScalarTraits<T>::is_complex)
    hipLaunchKernelGGL((gemm_z64_kernel<BM, BN, BK, WM, WN, OPA, OPB>), grid, dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, BK, WM, WN, OPA, OPB>), grid, dim3(256), 0, s, g);
}
template <int BM, int BN>
static void launch_gemm_carry(const GemmArgs& g, hipStream_t s) {
  dim3 grid(cdiv(g.M, BM), cdiv(g.N, BN), 1);
  hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, 16, 2, 2, OP_N, OP_N, true>), grid, dim3(256), 0, s, g);
}
// skinny split-K product with the side reduction (one extra z slice)
template <int OPA, int OPB>
static void launch_gemm_sidered(const GemmArgs& g, hipStream_t s) {
  dim3 grid(cdiv(g.M, 32), cdiv(g.N, 64), g.batch * g.ksplit + 1);
  hipLaunchKernelGGL((gemm_f64_kernel<32, 64, 16, 2, 2, OPA, OPB, true>), grid, dim3(256), 0, s, g);
}

// The largest tile that still gives the launch enough workgroups to cover the
// chip (256 CUs): 128x128 (real) / 64x64 while that yields >= 192 / 128
// workgroups, then 32x32 -- a 512^3 product on 128x128 tiles runs 16
// workgroups and is ~10x slower than on 32x32 tiles.
template <typename T, int OPA, int OPB>
static void gemm_shape(const GemmArgs& g, hipStream_t s) {
  constexpr bool cplx = ScalarTraits<T>::is_complex;
  const int64_t slabs = (int64_t)std::max(g.batch, 1) * std::max(g.ksplit, 1);
  auto wgs = [&](int bm, int bn) { return (int64_t)cdiv(g.M, bm) * cdiv(g.N, bn) * slabs; };
  // tiles that do work: upperC drops those strictly below the diagonal
  auto busy = [&](int b) {
    if (!g.upperC) return wgs(b, b);
    const int64_t tm = cdiv(g.M, b), tn = cdiv(g.N, b);
    int64_t t = 0;
    for (int64_t j = 0; j < tn; ++j) t += std::min<int64_t>(tm, (j * b + b + b - 1) / b);
    return t * slabs;
  };
  const bool hinted = g.tri != 0 || g.upperC != 0;
  if constexpr (!cplx && OPA == OP_N && OPB == OP_N) {
    if (g.cGw != nullptr) {  // QR trailing update carrying the next panel's first Gram (K = 32)
      launch_gemm_carry<64, 64>(g, s);
      return;
    }
  }
  // complex products never take 128x128 tiles (their 256 accumulators per
  // lane do not fit beside the operands), so that kernel is not instantiated
  if constexpr (!cplx) {
    if (g.tile != 0) {
      if (g.tile >= 128) launch_gemm_tile<128, 128, 16, 2, 2, OPA, OPB, T>(g, s);
      else if (g.tile >= 64) launch_gemm_tile<64, 64, 16, 2, 2, OPA, OPB, T>(g, s);
      else launch_gemm_tile<32, 32, 16, 2, 2, OPA, OPB, T>(g, s);
      return;
    }
  }
  if (g.M <= 32) {
    if constexpr (!cplx && OPA == OP_T && OPB == OP_N) {
      if (g.rPart != nullptr) {
        launch_gemm_sidered<OPA, OPB>(g, s);
        return;
      }
    }
    launch_gemm_tile<32, 64, 16, 2, 2, OPA, OPB, T>(g, s);
  } else if (!cplx && hinted && busy(128) < 1024 && busy(64) >= 512) {
    // hinted products spread uneven work over the tiles: more, smaller tiles
    // balance the CUs (and avoid a second round of a few 128x128 tiles)
    launch_gemm_tile<64, 64, 16, 2, 2, OPA, OPB, T>(g, s);
  } else if (!cplx && g.K <= 128 && wgs(64, 64) >= 512) {
    // short-K updates stream C: smaller tiles keep more loads in flight
    launch_gemm_tile<64, 64, 16, 2, 2, OPA, OPB, T>(g, s);
  } else if (!cplx && wgs(128, 128) >= 192) {
    if constexpr (!cplx) launch_gemm_tile<128, 128, 16, 2, 2, OPA, OPB, T>(g, s);
  } else if (wgs(64, 64) >= 128) {
    launch_gemm_tile<64, 64, 16, 2, 2, OPA, OPB, T>(g, s);
  } else {
    launch_gemm_tile<32, 32, 16, 2, 2, OPA, OPB, T>(g, s);
  }
}

template <typename T> void gemm(int opA, int opB, GemmArgs g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  if (g.batch <= 0) g.batch = 1;
  if (g.ksplit <= 0) g.ksplit = 1;
  if (g.kchunk <= 0) g.kchunk = g.K;
  if (g.K <= 0) {
    // C = beta C: run with an empty K range (the kernel writes alpha*0 + beta*C)
    g.K = 0;
    g.kchunk = 1;
  }
  constexpr bool cplx = ScalarTraits<T>::is_complex;
  const int b = (opB == OP_N) ? 0 : ((opB == OP_T || !cplx) ? 1 : 2);
  const int a = (opA == OP_N) ? 0 : ((opA == OP_T || !cplx) ? 1 : 2);
  switch (a * 3 + b) {
    case 0: gemm_shape<T, OP_N, OP_N>(g, s); break;
    case 1: gemm_shape<T, OP_N, OP_T>(g, s); break;
    case 2: gemm_shape<T, OP_N, OP_C>(g, s); break;
    case 3: gemm_shape<T, OP_T, OP_N>(g, s); break;
    case 4: gemm_shape<T, OP_T, OP_T>(g, s); break;
    case 5: gemm_shape<T, OP_T, OP_C>(g, s); break;
    case 6: gemm_shape<T, OP_C, OP_N>(g, s); break;
    case 7: gemm_shape<T, OP_C, OP_T>(g, s); break;
    default: gemm_shape<T, OP_C, OP_C>(g, s); break;
  }
}
template void gemm<double>(int, int, GemmArgs, hipStream_t);
template void gemm<zd>(int, int, GemmArgs, hipStream_t);

void splitk_plan(int N, int K, int bn, int& ksplit, int& kchunk) {
  if (N <= 0 || K <= 0) {  // empty products: one (empty) slice
    ksplit = 1;
    kchunk = std::max(K, 1);
    return;
  }
  const int tiles = cdiv(N, bn);
  int ks = cdiv(kSplitTargetWgs, tiles);
  ks = std::min(ks, std::max(1, K / 256));
  ks = std::max(ks, 1);
  kchunk = cdiv(cdiv(K, ks), 16) * 16;
  ksplit = cdiv(K, kchunk);
}

// ---------------------------------------------------------------------------
// WY reduction: W2 = op(T) * sum_z Wp[z]
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) wy_reduce_kernel(const T* __restrict__ Wp, int ksplit, int64_t slice,
                                                        const T* __restrict__ Tm, int transT, int nbc,
                                                        int ncols, T* __restrict__ W2, const int* gate) {
  // 8 columns per workgroup; thread (r, cc) sums the ksplit partials of one
  // output with eight independent accumulators (the loads of eight slices in
  // flight at once).
  __shared__ T S[32][9];
  __shared__ T Tl[32][33];
  if (gate_closed(gate, 0)) return;
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * 8;
  for (int i = tid; i < 32 * 32; i += 256) {
    const int r = i % 32, c = i / 32;
    T v = s_zero(T());
    if (r < nbc && c < nbc) v = transT ? s_conj(Tm[c + r * 32]) : Tm[r + c * 32];
    Tl[r][c] = v;
  }
  const int r = tid % 32, cc = tid / 32;
  const int c = c0 + cc;
  {
    T acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = s_zero(T());
    if (r < nbc && c < ncols) {
      const T* p = Wp + r + (int64_t)c * nbc;
      int z = 0;
      for (; z + 8 <= ksplit; z += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] = s_add(acc[u], p[(int64_t)(z + u) * slice]);
      }
      for (; z < ksplit; ++z) acc[z & 7] = s_add(acc[z & 7], p[(int64_t)z * slice]);
    }
    S[r][cc] = s_add(s_add(s_add(acc[0], acc[1]), s_add(acc[2], acc[3])),
                     s_add(s_add(acc[4], acc[5]), s_add(acc[6], acc[7])));
  }
  __syncthreads();
  if (r < nbc && c < ncols) {
    T v = s_zero(T());
#pragma unroll 8
    for (int k = 0; k < 32; ++k) v = s_fma(Tl[r][k], S[k][cc], v);
    W2[r + (int64_t)c * nbc] = v;
  }
}

template <typename T>
void wy_reduce(const void* Wp, int ksplit, int64_t slice, const void* Tm, int transT, int nbc, int ncols,
               void* W2, hipStream_t s, const int* gate) {
  if (ncols <= 0) return;
  hipLaunchKernelGGL((wy_reduce_kernel<T>), dim3(cdiv(ncols, 8)), dim3(256), 0, s,
                     reinterpret_cast<const T*>(Wp), ksplit, slice, reinterpret_cast<const T*>(Tm), transT,
                     nbc, ncols, reinterpret_cast<T*>(W2), gate);
}
template void wy_reduce<double>(const void*, int, int64_t, const void*, int, int, int, void*, hipStream_t,
                                const int*);
template void wy_reduce<zd>(const void*, int, int64_t, const void*, int, int, int, void*, hipStream_t, const int*);

// ---------------------------------------------------------------------------
// Block reflectors from several 32-column panels (LAPACK xLARFT's recursion
// applied blockwise): for panels 0..np-1 with compact WY factors (V_p, T_p),
//   H_0 H_1 ... H_{np-1} = I - V T V^H,  T = [[T_a, -T_a S_ab T_b], [0, T_b]]
// built one panel at a time, with S = V^H V.
// ---------------------------------------------------------------------------

// Vc (rows x kg, ld ldvc) = the unit lower trapezoidal V stored in A below the
// diagonal (ones on the diagonal, zeros above).
template <typename T>
__global__ void copy_unit_lower_kernel(const T* A, int64_t lda, int rows, int kg, T* Vc, int64_t ldvc) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)rows * kg) return;
  const int r = (int)(idx % rows), c = (int)(idx / rows);
  T v = s_zero(T());
  if (r > c) v = A[r + (int64_t)c * lda];
  else if (r == c) v = s_one(T());
  Vc[r + (int64_t)c * ldvc] = v;
}
template <typename T> void copy_unit_lower(const T* A, int64_t lda, int rows, int kg, T* Vc, int64_t ldvc,
                                           hipStream_t s) {
  const int64_t tot = (int64_t)rows * kg;
  if (tot <= 0) return;
  hipLaunchKernelGGL((copy_unit_lower_kernel<T>), dim3((unsigned)cdiv64(tot, 256)), dim3(256), 0, s, A, lda, rows,
                     kg, Vc, ldvc);
}
template void copy_unit_lower<double>(const double*, int64_t, int, int, double*, int64_t, hipStream_t);
template void copy_unit_lower<zd>(const zd*, int64_t, int, int, zd*, int64_t, hipStream_t);

// out (M x N, ld ldo) = sum over the ks slices P + z * slice (M x N, ld M), in
// slice order.
template <typename T>
__global__ void sum_slices_kernel(const T* __restrict__ P, int ks, int64_t slice, int M, int N, T* __restrict__ out,
                                  int64_t ldo) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)M * N) return;
  const int i = (int)(idx % M), j = (int)(idx / M);
  const T* p = P + idx;
  T acc[4] = {s_zero(T()), s_zero(T()), s_zero(T()), s_zero(T())};
  int z = 0;
  for (; z + 4 <= ks; z += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = s_add(acc[u], p[(int64_t)(z + u) * slice]);
  }
  for (; z < ks; ++z) acc[0] = s_add(acc[0], p[(int64_t)z * slice]);
  out[i + (int64_t)j * ldo] = s_add(s_add(acc[0], acc[1]), s_add(acc[2], acc[3]));
}
template <typename T> void sum_slices(const T* P, int ks, int64_t slice, int M, int N, T* out, int64_t ldo,
                                      hipStream_t s) {
  const int64_t tot = (int64_t)M * N;
  if (tot <= 0) return;
  hipLaunchKernelGGL((sum_slices_kernel<T>), dim3((unsigned)cdiv64(tot, 256)), dim3(256), 0, s, P, ks, slice, M, N,
                     out, ldo);
}
template void sum_slices<double>(const double*, int, int64_t, int, int, double*, int64_t, hipStream_t);
template void sum_slices<zd>(const zd*, int, int64_t, int, int, zd*, int64_t, hipStream_t);

// T (kg x kg, ld kg, upper) from S = V^H V (kg x kg, ld kg) and the panels'
// T_p (32 x 32 each, consecutive, zero outside their nbc x nbc block), one
// workgroup: Y_c = S(c, b) T_b, T(a, b) = -sum_{a <= c < b} T(a, c) Y_c, block
// by block on FP64 MFMA.  Float64 keeps every block of T in LDS; ComplexF64
// (twice the bytes) re-reads the finished blocks from Tout.
template <typename T>
__global__ void __launch_bounds__(256) wy_merge_kernel(const T* S, const T* Tp, int np, int kg, T* Tout, int npan,
                                                       int minmn) {
  constexpr bool LDS_T = !ScalarTraits<T>::is_complex;
  if (npan > 0) {  // one group of four panels per workgroup (see wy_merge_groups)
    const int g = blockIdx.x;
    np = min(4, npan - 4 * g);
    kg = min(minmn - 128 * g, 128);
    S += (int64_t)g * 128 * 128;
    Tp += (int64_t)g * 4 * 1024;
    Tout += (int64_t)g * 128 * 128;
  }
  __shared__ T Tt[LDS_T ? 10 : 1][32][33];  // block (a, c), a <= c, at c (c + 1) / 2 + a
  __shared__ T Y[3][32][33];
  __shared__ T Z[3][32][33];
  __shared__ T X[32][33];
  __shared__ T Acc[32][33];
  const int tid = threadIdx.x;
  auto load = [&](T (*D)[33], const T* src, int64_t ld, int r0, int c0) {
    for (int e = tid; e < 1024; e += 256) {
      const int i = e % 32, j = e / 32;
      const int r = r0 + i, c = c0 + j;
      D[i][j] = (r < kg && c < kg) ? src[r + (int64_t)c * ld] : s_zero(T());
    }
  };
  auto store = [&](const T (*D)[33], int r0, int c0) {
    for (int e = tid; e < 1024; e += 256) {
      const int i = e % 32, j = e / 32;
      const int r = r0 + i, c = c0 + j;
      if (r < kg && c < kg) Tout[r + (int64_t)c * kg] = D[i][j];
    }
  };
  auto tix = [](int a, int c) { return c * (c + 1) / 2 + a; };
  // diagonal blocks T_b, zeros below them
  for (int b = 0; b < np; ++b) {
    if constexpr (LDS_T) {
      for (int e = tid; e < 1024; e += 256) Tt[tix(b, b)][e % 32][e / 32] = Tp[(int64_t)b * 1024 + e];
    } else {
      for (int e = tid; e < 1024; e += 256) {
        const int r = 32 * b + e % 32, c = 32 * b + e / 32;
        if (r < kg && c < kg) Tout[r + (int64_t)c * kg] = Tp[(int64_t)b * 1024 + e];
      }
    }
    for (int a = b + 1; a < np; ++a)
      for (int e = tid; e < 1024; e += 256) {
        const int r = 32 * a + e % 32, c = 32 * b + e / 32;
        if (r < kg && c < kg) Tout[r + (int64_t)c * kg] = s_zero(T());
      }
  }
  for (int b = 1; b < np; ++b) {
    __syncthreads();
    for (int c = 0; c < b; ++c) load(Z[c], S, kg, 32 * c, 32 * b);
    if constexpr (!LDS_T)
      for (int e = tid; e < 1024; e += 256) X[e % 32][e / 32] = Tp[(int64_t)b * 1024 + e];
    __syncthreads();
    for (int c = 0; c < b; ++c) {
      if constexpr (LDS_T) mm32<T>(Y[c], Z[c], false, Tt[tix(b, b)], false, 1.0, false);  // Y_c = S(c, b) T_b
      else mm32<T>(Y[c], Z[c], false, X, false, 1.0, false);
    }
    for (int a = 0; a < b; ++a) {
      T (*dst)[33] = LDS_T ? Tt[tix(a, b)] : Acc;
      for (int c = a; c < b; ++c) {
        if constexpr (LDS_T) {
          mm32<T>(dst, Tt[tix(a, c)], false, Y[c], false, -1.0, c > a);  // T(a, b) = -sum T(a, c) Y_c
        } else {
          __syncthreads();
          load(Z[0], Tout, kg, 32 * a, 32 * c);
          __syncthreads();
          mm32<T>(dst, Z[0], false, Y[c], false, -1.0, c > a);
        }
      }
      if constexpr (!LDS_T) {
        store(Acc, 32 * a, 32 * b);
        __syncthreads();
      }
    }
  }
  if constexpr (LDS_T) {
    __syncthreads();
    for (int b = 0; b < np; ++b)
      for (int a = 0; a <= b; ++a) store(Tt[tix(a, b)], 32 * a, 32 * b);
  }
}
template <typename T> void wy_merge(const T* S, const T* Tp, int np, int kg, T* Tout, hipStream_t s) {
  if (np <= 0 || np > 4 || kg <= 0) return;
  hipLaunchKernelGGL((wy_merge_kernel<T>), dim3(1), dim3(256), 0, s, S, Tp, np, kg, Tout, 0, 0);
}
template <typename T> void wy_merge_groups(const T* S, const T* Tp, int npan, int minmn, T* Tout, hipStream_t s) {
  const int ngr = cdiv(npan, 4);
  if (ngr <= 0) return;
  hipLaunchKernelGGL((wy_merge_kernel<T>), dim3(ngr), dim3(256), 0, s, S, Tp, 0, 0, Tout, npan, minmn);
}
template void wy_merge<double>(const double*, const double*, int, int, double*, hipStream_t);
template void wy_merge<zd>(const zd*, const zd*, int, int, zd*, hipStream_t);
template void wy_merge_groups<double>(const double*, const double*, int, int, double*, hipStream_t);
template void wy_merge_groups<zd>(const zd*, const zd*, int, int, zd*, hipStream_t);

// ---------------------------------------------------------------------------
// Elementwise helpers
// ---------------------------------------------------------------------------
template <typename T>
__global__ void set_identity_kernel(T* Q, int64_t ldq, int m, int n) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  Q[r + c * ldq] = (r == c) ? s_one(T()) : s_zero(T());
}
template <typename T> void set_identity(void* Q, int64_t ldq, int m, int n, hipStream_t s) {
  const int64_t tot = (int64_t)m * n;
  if (tot <= 0) return;
  hipLaunchKernelGGL((set_identity_kernel<T>), dim3((unsigned)cdiv64(tot, 256)), dim3(256), 0, s,
                     reinterpret_cast<T*>(Q), ldq, m, n);
}
template void set_identity<double>(void*, int64_t, int, int, hipStream_t);
template void set_identity<zd>(void*, int64_t, int, int, hipStream_t);

template <typename T>
__global__ void qr_extract_r_kernel(const T* A, int64_t lda, T* R, int64_t ldr, int rrows, int n, int minmn,
                                    int positive) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)rrows * n) return;
  const int r = (int)(idx % rrows), c = (int)(idx / rrows);
  T v = s_zero(T());
  if (r < minmn && r <= c) {
    v = A[r + c * lda];
    // positive gauge: row r times conj(sign(R'_rr)); the diagonal is |R'_rr|, exactly real
    if (positive) v = (r == c) ? s_from_real<T>(s_absv(v)) : s_mul(s_conj(sign_safe(A[r + r * lda])), v);
  }
  R[r + c * ldr] = v;
}
template <typename T>
void qr_extract_r(const void* A, int64_t lda, void* R, int64_t ldr, int rrows, int n, int minmn, int positive,
                  hipStream_t s) {
  const int64_t tot = (int64_t)rrows * n;
  if (tot <= 0) return;
  hipLaunchKernelGGL((qr_extract_r_kernel<T>), dim3((unsigned)cdiv64(tot, 256)), dim3(256), 0, s,
                     reinterpret_cast<const T*>(A), lda, reinterpret_cast<T*>(R), ldr, rrows, n, minmn, positive);
}
template void qr_extract_r<double>(const void*, int64_t, void*, int64_t, int, int, int, int, hipStream_t);
template void qr_extract_r<zd>(const void*, int64_t, void*, int64_t, int, int, int, int, hipStream_t);

template <typename T>
__global__ void qr_gauge_q_kernel(const T* A, int64_t lda, T* Q, int64_t ldq, int m, int minmn) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * minmn) return;
  const int r = (int)(idx % m), c = (int)(idx / m);
  Q[r + c * ldq] = s_mul(Q[r + c * ldq], sign_safe(A[c + c * lda]));
}
template <typename T>
void qr_gauge_q(const void* A, int64_t lda, void* Q, int64_t ldq, int m, int minmn, hipStream_t s) {
  const int64_t tot = (int64_t)m * minmn;
  if (tot <= 0) return;
  hipLaunchKernelGGL((qr_gauge_q_kernel<T>), dim3((unsigned)cdiv64(tot, 256)), dim3(256), 0, s,
                     reinterpret_cast<const T*>(A), lda, reinterpret_cast<T*>(Q), ldq, m, minmn);
}
template void qr_gauge_q<double>(const void*, int64_t, void*, int64_t, int, int, hipStream_t);
template void qr_gauge_q<zd>(const void*, int64_t, void*, int64_t, int, int, hipStream_t);

}  // namespace mak
