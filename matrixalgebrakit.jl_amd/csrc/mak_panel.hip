// Householder panel factorisation for gfx950.
//
// The panel (rows x NB columns, NB <= 32) of a blocked Householder QR is
// factored by a TSQR tree whose nodes run `geqr2_node_kernel`: one workgroup
// holds its block in registers (a row per thread and RPT rows per thread at
// most), and every column costs ONE fused workgroup reduction of NB values,
// computed with wavefront shuffles and exchanged across waves through LDS:
//   red[c <  j] = sum_{i>j} conj(a_ic) a_ij   (column j of the T factor)
//   red[c == j] = sum_{i>j} |a_ij|^2           (reflector norm)
//   red[c >  j] = sum_{i>j} conj(a_ij) a_ic    (reflector dot products)
// Reflectors follow LAPACK's xLARFG convention (beta = -sign(Re alpha) ||x||,
// tau = (beta - alpha) / beta, v = x / (alpha - beta), v_j = 1), the T factor
// follows xLARFT (forward, columnwise), so H_1 ... H_k = I - V T V^H exactly as
// in LAPACK's geqrt (MatrixAlgebraKit src/yalapack.jl:303-333).
//
// When the panel is taller than one node, the tree's explicit Q is formed top
// down (`tsqr_root_finish_kernel`, `tsqr_expand_kernel`) and the Householder
// vectors are reconstructed from it (Ballard, Demmel, Grigori, Jacquelin,
// Nguyen, Solomonik, "Reconstructing Householder vectors from tall-skinny QR",
// 2014): Q - [D; 0] = V U is an LU factorisation without pivoting when each
// d_j = -sign(pivot_j); then T = -U D^H V1^{-H} and R = D R_tree.
#include <utility>

#include "mak_common.hpp"
#include "mak_dense32.hpp"
#include "mak_kernels.hpp"

namespace mak {

// ---------------------------------------------------------------------------
// Transposed wave reduction: each lane holds NB partial sums; afterwards lane l
// holds the full wave sum of column col_of_lane<NB>(l) (every column appears
// on 64 / NB lanes).  Costs NB - 1 + log2(64 / NB) shuffles instead of
// 6 NB for NB independent butterflies.
// ---------------------------------------------------------------------------
template <int NB> __device__ inline int col_of_lane(int l) {
  // halving steps use masks 32, 16, ..., 64 / NB; lane bit for mask M selects
  // the upper half of the current range.
  int c = 0;
  int half = NB / 2;
#pragma unroll
  for (int m = 32; m >= 64 / NB; m >>= 1) {
    if (l & m) c += half;
    half >>= 1;
  }
  return c;
}

// Halving step across lane bit 5 (mask 32) / bit 4 (mask 16) with the gfx950
// lane-swap instructions: after v_permlane32_swap(x, y) the lower half-wave
// holds (x_l, x_{l+32}) and the upper half-wave (y_{l-32}, y_l), so x + y is
// the pair sum of column k (lower lanes) and column k + half (upper lanes).
__device__ inline double swap_add32(double x, double y) {
  const uint64_t xb = __builtin_bit_cast(uint64_t, x), yb = __builtin_bit_cast(uint64_t, y);
  auto lo = __builtin_amdgcn_permlane32_swap((unsigned)xb, (unsigned)yb, false, false);
  auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(xb >> 32), (unsigned)(yb >> 32), false, false);
  const double xs = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
  const double ys = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
  return xs + ys;
}
__device__ inline double swap_add16(double x, double y) {
  const uint64_t xb = __builtin_bit_cast(uint64_t, x), yb = __builtin_bit_cast(uint64_t, y);
  auto lo = __builtin_amdgcn_permlane16_swap((unsigned)xb, (unsigned)yb, false, false);
  auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(xb >> 32), (unsigned)(yb >> 32), false, false);
  const double xs = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
  const double ys = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
  return xs + ys;
}
__device__ inline zd swap_add32(zd x, zd y) { return mkz(swap_add32(x.x, y.x), swap_add32(x.y, y.y)); }
__device__ inline zd swap_add16(zd x, zd y) { return mkz(swap_add16(x.x, y.x), swap_add16(x.y, y.y)); }

template <typename T, int NB> __device__ inline T wave_reduce_transposed(T (&p)[NB]) {
  const int l = lane_id();
  static_assert(NB == 32, "reduction levels are laid out for 32 slots");
  // masks 32 and 16: lane-swap instructions, no selects
#pragma unroll
  for (int k = 0; k < 16; ++k) p[k] = swap_add32(p[k], p[k + 16]);
#pragma unroll
  for (int k = 0; k < 8; ++k) p[k] = swap_add16(p[k], p[k + 8]);
  // masks 8, 4, 2: shuffles with value selects
#pragma unroll
  for (int len = 8, m = 8; len > 1; len >>= 1, m >>= 1) {
    const bool up = (l & m) != 0;
#pragma unroll
    for (int k = 0; k < len / 2; ++k) {
      const T lo = p[k];
      const T hi = p[k + len / 2];
      const T keep = sel(up, hi, lo);
      const T send = sel(up, lo, hi);
      p[k] = s_add(keep, shfl_xor_s(send, m));
    }
  }
  // mask 1: the two lanes that share a column
  return s_add(p[0], shfl_xor_s(p[0], 1));
}

// Rows of node `i` out of `nodes` over `H` stacked rows: fixed chunks when
// chunk > 0, a balanced split otherwise.
__device__ __host__ inline void node_rows(int i, int nodes, int H, int chunk, int& r0, int& r1) {
  if (chunk > 0) {
    r0 = i * chunk;
    r1 = min(H, r0 + chunk);
  } else {
    r0 = (int)(((int64_t)i * H) / nodes);
    r1 = (int)(((int64_t)(i + 1) * H) / nodes);
  }
}


// Shared state of geqr2_node_kernel.
template <typename T, int NB, int NW> struct NodeLds {
  T red_w[2][NW][NB];
  T rowj[2][NB];
  T gsh[NW][NB];
  T Ysh[NB][NB + 1];   // Ysh[j][c] = V_c^H v_j for c < j
  T tau_s[NB];
};

// One column step j of the node factorisation, with the current column held
// in register slot OFF (slot i holds column (i + j - OFF) mod NB).
template <typename T, int NB, int RPT, int NT, int OFF>
__device__ inline void node_step(T (&a)[RPT][NB], NodeLds<T, NB, NT / WAVE>& S, int j, int h, int nbc) {
  constexpr int NW = NT / WAVE;
  const int tid = threadIdx.x;
  const int wid = tid / WAVE;
  const int l = lane_id();
  const int par = j & 1;
  // partial sums over rows r > j of conj(a_rj) a_r(slot i)
  T p[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) p[i] = s_zero(T());
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + NT * q;
    const T x = (r > j && r < h) ? a[q][OFF] : s_zero(T());
#pragma unroll
    for (int i = 0; i < NB; ++i) p[i] = s_cfma(x, a[q][i], p[i]);
  }
  const T tot = wave_reduce_transposed<T, NB>(p);
  if ((l & (64 / NB - 1)) == 0) S.red_w[par][wid][col_of_lane<NB>(l)] = tot;
  __syncthreads();

  // every wave: lane i < NB finishes slot i
  T ri = s_zero(T());
  double nu2 = 0.0;
  const int is = l & (NB - 1);
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    ri = s_add(ri, S.red_w[par][w][is]);
    nu2 += s_re(S.red_w[par][w][OFF]);
  }
  const T alpha = S.rowj[par][OFF];
  const T rji = S.rowj[par][is];
  T tau, beta, scale;
  const double bn = sqrt(s_abs2(alpha) + nu2);
  if ((nu2 == 0.0 && s_im(alpha) == 0.0) || !(bn > 0.0)) {
    // nothing to annihilate -- or a column whose squares underflow (entries
    // below ~1e-154): no reflection, the column is left as it is
    tau = s_zero(T());
    beta = alpha;
    scale = s_zero(T());
  } else {
    const double br = s_re(alpha) >= 0.0 ? -bn : bn;
    beta = s_from_real<T>(br);
    tau = s_scale(s_sub(beta, alpha), 1.0 / br);
    scale = s_div(s_one(T()), s_sub(alpha, beta));
  }
  {
    int c = is + j - OFF;  // column held by slot is
    if (c >= NB) c -= NB;
    if (c < 0) c += NB;
    T g = s_zero(T());
    if (c > j && c < nbc) g = s_mul(s_conj(tau), s_cfma(scale, ri, rji));   // conj(tau) v^H a_c
    if (l < NB) S.gsh[wid][is] = g;
    if (wid == 0 && l < NB && c < j) S.Ysh[j][c] = s_fma(scale, s_conj(ri), s_conj(rji));  // V_c^H v_j
    if (tid == 0) S.tau_s[j] = tau;
  }
  wave_lds_sync();
  T g[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) g[i] = (i == OFF) ? s_zero(T()) : S.gsh[wid][i];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + NT * q;
    const bool below = (r > j && r < h);
    const bool onrow = (r == j);
    const T v = s_mul(scale, a[q][OFF]);
    const T coef = below ? v : (onrow ? s_one(T()) : s_zero(T()));
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i != OFF) a[q][i] = s_sub(a[q][i], s_mul(coef, g[i]));
    a[q][OFF] = below ? v : (onrow ? beta : a[q][OFF]);
  }
}

template <typename T, int NB, int RPT, int NT>
__device__ inline void publish_row(const T (&a)[RPT][NB], NodeLds<T, NB, NT / WAVE>& S, int row, int kmax) {
  // the owner of `row` (thread row, register row 0) publishes it for step `row`
  if (threadIdx.x == row && row < kmax) {
#pragma unroll
    for (int i = 0; i < NB; ++i) S.rowj[row & 1][i] = a[0][i];
  }
}

template <typename T, int NB, int RPT, int NT>
__global__ void __launch_bounds__(NT) geqr2_node_kernel(Geqr2NodeArgs args) {
  if (gate_closed(args.gate, args.gate_want)) return;
  // Register layout: thread t holds rows t + NT*q (q < RPT).  Rows are kept
  // ROTATED in registers: steps go in pairs with the current column in slot 0
  // then slot 1, and the row is rotated by two slots after each pair, so no
  // register is indexed by the runtime column.
  //
  // One workgroup barrier per column: the cross-wave partials and row j are
  // double-buffered by step parity, and every wave finishes the reduction for
  // itself (lane i owns slot i), broadcasting the update coefficients through
  // a wave-private LDS row.  T is formed after the loop from the stored
  // products Y(c, j) = V_c^H v_j (LAPACK xLARFT recurrence).
  constexpr int NW = NT / WAVE;
  __shared__ NodeLds<T, NB, NW> S;

  const int node = blockIdx.x;
  int r0, r1;
  node_rows(node, args.nodes, args.H, args.chunk, r0, r1);
  const int h = r1 - r0;
  const int nbc = args.nbc;
  T* X = reinterpret_cast<T*>(args.X) + r0;
  const int64_t ldx = args.ldx;
  const int tid = threadIdx.x;
  const int wid = tid / WAVE;
  const int l = lane_id();

  T a[RPT][NB];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + NT * q;
#pragma unroll
    for (int c = 0; c < NB; ++c)
      a[q][c] = (r < h && c < nbc) ? X[r + c * ldx] : s_zero(T());
  }
  for (int i = tid; i < NB * (NB + 1); i += NT) (&S.Ysh[0][0])[i] = s_zero(T());
  if (tid < NB) S.tau_s[tid] = s_zero(T());
  const int kmax = min(h, nbc);
  publish_row<T, NB, RPT, NT>(a, S, 0, kmax);

  for (int j = 0; j < kmax; j += 2) {
    node_step<T, NB, RPT, NT, 0>(a, S, j, h, nbc);
    publish_row<T, NB, RPT, NT>(a, S, j + 1, kmax);
    if (j + 1 < kmax) node_step<T, NB, RPT, NT, 1>(a, S, j + 1, h, nbc);
    // rotate by two slots: slot i <- slot i+2
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const T f0 = a[q][0], f1 = a[q][1];
#pragma unroll
      for (int i = 0; i < NB - 2; ++i) a[q][i] = a[q][i + 2];
      a[q][NB - 2] = f0;
      a[q][NB - 1] = f1;
    }
    publish_row<T, NB, RPT, NT>(a, S, j + 2, kmax);
  }
  __syncthreads();

  // slot i now holds column (i + krot) mod NB
  const int krot = 2 * ((kmax + 1) / 2);
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + NT * q;
    if (r < h) {
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        int c = (i + krot) % NB;
        if (c < nbc) X[r + c * ldx] = a[q][i];
      }
    }
  }
  if (args.Tout && wid == 0) {
    // T(0:j, j) = -tau_j T(0:j, 0:j) Y(0:j, j); T(j, j) = tau_j.  Lane r keeps
    // row r of T in registers.
    const int r = l;
    T Treg[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) Treg[k] = s_zero(T());
    for (int j = 0; j < kmax; ++j) {
      T acc[4] = {s_zero(T()), s_zero(T()), s_zero(T()), s_zero(T())};
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        const T y = S.Ysh[j][k];
        const T t = s_mul(Treg[k], y);
        acc[k & 3] = s_add(acc[k & 3], (k >= r && k < j) ? t : s_zero(T()));
      }
      const T tj = S.tau_s[j];
      const T val = (r < j) ? s_neg(s_mul(tj, s_add(s_add(acc[0], acc[1]), s_add(acc[2], acc[3]))))
                            : ((r == j) ? tj : s_zero(T()));
#pragma unroll
      for (int k = 0; k < NB; ++k) Treg[k] = (k == j) ? val : Treg[k];
    }
    T* To = reinterpret_cast<T*>(args.Tout) + (int64_t)node * args.tstride;
    if (r < NB) {
#pragma unroll
      for (int c = 0; c < NB; ++c) To[r + c * NB] = (r <= c) ? Treg[c] : s_zero(T());
    }
  }
  if (args.Rnext) {
    // R of this node: rows [node*nbc, node*nbc + nbc) of the next level
    T* Rn = reinterpret_cast<T*>(args.Rnext) + (int64_t)node * nbc;
    const int r = tid;
    if (r < nbc) {
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int c = (i + krot) % NB;
        if (c < nbc) Rn[r + (int64_t)c * args.ldr] = (r <= c && r < h) ? a[0][i] : s_zero(T());
      }
    }
  }
  if (args.V1out) {
    T* V1 = reinterpret_cast<T*>(args.V1out) + (int64_t)node * NB * NB;
    const int r = tid;
    if (r < NB) {
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int c = (i + krot) % NB;
        V1[r + c * NB] = (r > c && r < h && c < nbc) ? a[0][i] : (r == c ? s_one(T()) : s_zero(T()));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Node factorisation with a 2-D lane layout (the production node kernel).
//
// A wave holds WR = 4 KR rows of the 32-column block: lane l owns columns
// cl = l & 15 and cl + 16 of rows 4k + (l >> 4), k < KR.  Per column j:
//   * the 4 lanes holding column j copy its rows below j into wave-private LDS,
//     every lane reads back the KR entries of its own rows;
//   * each lane forms conj(x)^T a for its two columns over its KR rows, the
//     four row groups are summed with two lane-swap steps (permlane32/16), so
//     lane l ends with column (l & 15) + 16 (l >> 5) of the wave's sum;
//   * the NW wave sums meet in LDS behind ONE workgroup barrier per column;
//   * reflector (xLARFG convention), update coefficients and the xLARFT
//     products are formed redundantly per lane; the rank-1 update is 2 FMAs
//     per row and lane.
// The column loop runs two steps per iteration (step parity is a template
// argument); registers are only ever selected by value, never indexed.
// ---------------------------------------------------------------------------
template <typename T, int NW, int WR> struct Node2Lds {
  T red[2][32][NW];  // wave partial sums per column, double-buffered by step parity
  T rowj[2][32];     // row j of the block, published at the end of step j - 1
  T xs[NW][WR];      // wave-private copy of the current column (rows of the wave)
  T Ysh[32][33];     // Ysh[j][c] = V_c^H v_j, c < j
  T tau_s[32];
  T scale_s[32];     // v_j = scale_j x_j below the diagonal (applied after the loop)
  T beta_s[32];      // R(j, j)
};

template <typename T, int NW, int KR, int PAR>
__device__ __forceinline__ void node2_step(T (&a)[KR][2], Node2Lds<T, NW, 4 * KR>& S, const int J, int kmax,
                                           int nbc) {
  // J is uniform (the column loop counter, J & 1 == PAR); J < 32 <= WR, so
  // rows <= J all live in registers k < 8 of wave 0.  Finished columns keep
  // their unscaled x below the diagonal until the loop ends.
  constexpr int WR = 4 * KR;
  const int l = lane_id();
  const int w = threadIdx.x / WAVE;
  const int g = l >> 4;
  const int cl = l & 15;
  const int d = w * WR + g - J;     // row of register k minus J: d + 4 k
  const bool hi = J >= 16;           // register half holding column J (uniform)

  // 1. the column-J lanes publish the column below row J
  if (cl == (J & 15)) {
    if (w == 0) {
#pragma unroll
      for (int k = 0; k < KR; ++k)
        S.xs[0][g * KR + k] = sel(k >= 8 || d + 4 * k > 0, sel(hi, a[k][1], a[k][0]), s_zero(T()));
    } else {
#pragma unroll
      for (int k = 0; k < KR; ++k) S.xs[w][g * KR + k] = sel(hi, a[k][1], a[k][0]);
    }
  }
  wave_lds_sync();
  // With eight waves (or complex data) a lane has too few registers for a
  // second copy: x is not kept but read from LDS where it is used (this
  // wave's copy, unchanged until its next step), so the column data alone
  // stays in registers.
  constexpr bool kXLds = NW * KR > 64 || ScalarTraits<T>::is_complex;
  T x[kXLds ? 1 : KR];
  auto xk = [&](int k) -> T {
    if constexpr (kXLds) return S.xs[w][g * KR + k];
    else return x[k];
  };
  if constexpr (!kXLds) {
#pragma unroll
    for (int k = 0; k < KR; ++k) x[k] = S.xs[w][g * KR + k];
  }

  // 2. partial products conj(x)^T a over this lane's rows
  T p0a = s_zero(T()), p0b = s_zero(T()), p1a = s_zero(T()), p1b = s_zero(T());
#pragma unroll
  for (int k = 0; k < KR; k += 2) {
    const T x0 = xk(k);
    p0a = s_cfma(x0, a[k][0], p0a);
    p1a = s_cfma(x0, a[k][1], p1a);
    if (k + 1 < KR) {
      const T x1 = xk(k + 1);
      p0b = s_cfma(x1, a[k + 1][0], p0b);
      p1b = s_cfma(x1, a[k + 1][1], p1b);
    }
  }
  // 3. sum the four row groups: lanes < 32 -> column cl, lanes >= 32 -> cl + 16
  T sw = swap_add32(s_add(p0a, p0b), s_add(p1a, p1b));
  sw = swap_add16(sw, sw);
  if ((l & 16) == 0) S.red[PAR][cl + ((l >> 5) << 4)][w] = sw;
  __syncthreads();

  // 4. workgroup totals of this lane's columns and of column J
  T t0 = s_zero(T()), t1 = s_zero(T());
  double nu2 = 0.0;
#pragma unroll
  for (int v = 0; v < NW; ++v) {
    t0 = s_add(t0, S.red[PAR][cl][v]);
    t1 = s_add(t1, S.red[PAR][cl + 16][v]);
    nu2 += s_re(S.red[PAR][J][v]);
  }
  const T alpha = S.rowj[PAR][J];
  const T r0 = S.rowj[PAR][cl];
  const T r1 = S.rowj[PAR][cl + 16];
  T tau, beta, scale;
  const double bn = sqrt(s_abs2(alpha) + nu2);
  if ((nu2 == 0.0 && s_im(alpha) == 0.0) || !(bn > 0.0)) {
    // nothing to annihilate -- or a column whose squares underflow (entries
    // below ~1e-154): no reflection, the column is left as it is
    tau = s_zero(T());
    beta = alpha;
    scale = s_zero(T());
  } else {
    const double br = s_re(alpha) >= 0.0 ? -bn : bn;
    beta = s_from_real<T>(br);
    tau = s_scale(s_sub(beta, alpha), 1.0 / br);
    scale = s_div(s_one(T()), s_sub(alpha, beta));
  }
  // update coefficients conj(tau) v^H a_c for J < c < nbc
  const T ctau = s_conj(tau);
  const T g0 = sel(cl > J && cl < nbc, s_mul(ctau, s_cfma(scale, t0, r0)), s_zero(T()));
  const T g1 = sel(cl + 16 > J && cl + 16 < nbc, s_mul(ctau, s_cfma(scale, t1, r1)), s_zero(T()));
  if (w == 0 && l < 16) {
    // V_c^H v_J for c < J: column c still holds x_c (v_c = scale_c x_c below c)
    if (cl < J) S.Ysh[J][cl] = s_mul(s_conj(S.scale_s[cl]), s_fma(scale, s_conj(t0), s_conj(r0)));
    if (cl + 16 < J) S.Ysh[J][cl + 16] = s_mul(s_conj(S.scale_s[cl + 16]), s_fma(scale, s_conj(t1), s_conj(r1)));
    if (l == 0) {
      S.tau_s[J] = tau;
      S.scale_s[J] = scale;
      S.beta_s[J] = beta;
    }
  }

  // 5. rank-1 update a_rc -= v_r g_c with v_r = scale x_r below row J and
  //    v_J = 1 (x was zeroed on rows <= J); g = 0 on columns <= J.
  if (w == 0) {
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const T coef = (k < 8) ? sel(d + 4 * k == 0, s_one(T()), s_mul(scale, xk(k))) : s_mul(scale, xk(k));
      a[k][0] = s_sub(a[k][0], s_mul(coef, g0));
      a[k][1] = s_sub(a[k][1], s_mul(coef, g1));
    }
  } else {
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const T v = s_mul(scale, xk(k));
      a[k][0] = s_sub(a[k][0], s_mul(v, g0));
      a[k][1] = s_sub(a[k][1], s_mul(v, g1));
    }
  }
  // 6. publish row J + 1 (register k1 = (J + 1) / 4 of row group (J + 1) & 3, wave 0)
  const int J1 = J + 1;
  if (J1 < kmax && w == 0 && g == (J1 & 3)) {
    const int k1 = J1 >> 2;
    T q0 = a[0][0], q1 = a[0][1];
#pragma unroll
    for (int k = 1; k < 8; ++k) {
      q0 = sel(k == k1, a[k][0], q0);
      q1 = sel(k == k1, a[k][1], q1);
    }
    S.rowj[PAR ^ 1][cl] = q0;
    S.rowj[PAR ^ 1][cl + 16] = q1;
  }
}

template <typename T, int NW, int KR>
__device__ __forceinline__ void node2_all_steps(T (&a)[KR][2], Node2Lds<T, NW, 4 * KR>& S, int kmax, int nbc) {
#pragma unroll 1
  for (int j = 0; j < kmax; j += 2) {
    node2_step<T, NW, KR, 0>(a, S, j, kmax, nbc);
    if (j + 1 < kmax) node2_step<T, NW, KR, 1>(a, S, j + 1, kmax, nbc);
  }
}

template <typename T, int J>
__device__ __forceinline__ void node2_tcol(T (&Treg)[32], const T (*Ysh)[33], const T* tau_s, int r, int kmax) {
  if (J >= kmax) return;
  T acc0 = s_zero(T()), acc1 = s_zero(T());
#pragma unroll
  for (int k = 0; k < J; ++k) {
    if (k & 1) acc1 = s_fma(Treg[k], Ysh[J][k], acc1);
    else acc0 = s_fma(Treg[k], Ysh[J][k], acc0);
  }
  const T tj = tau_s[J];
  Treg[J] = sel(r < J, s_neg(s_mul(tj, s_add(acc0, acc1))), sel(r == J, tj, s_zero(T())));
}
template <typename T, int... Js>
__device__ __forceinline__ void node2_tcols(T (&Treg)[32], const T (*Ysh)[33], const T* tau_s, int r, int kmax,
                                            std::integer_sequence<int, Js...>) {
  (node2_tcol<T, Js>(Treg, Ysh, tau_s, r, kmax), ...);
}

// Factor one node: X (h x nbc, ld ldx) in place; T (32 x 32, ld 32) to Tout,
// R (nbc x nbc) to Rn (ld ldr) and the top 32 x 32 of V to V1 when non-null.
template <typename T, int NW, int KR>
__device__ __forceinline__ void node2_body(Node2Lds<T, NW, 4 * KR>& S, T* X, int64_t ldx, int h, int nbc,
                                           T* Tout, T* Rn, int64_t ldr, T* V1) {
  constexpr int WR = 4 * KR;
  const int tid = threadIdx.x;
  const int l = lane_id();
  const int w = tid / WAVE;
  const int g = l >> 4;
  const int cl = l & 15;
  const int rbase = w * WR + g;

  T a[KR][2];
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int r = rbase + 4 * k;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int c = cl + 16 * hh;
      a[k][hh] = (r < h && c < nbc) ? X[r + c * ldx] : s_zero(T());
    }
  }
  for (int i = tid; i < 32 * 33; i += NW * WAVE) (&S.Ysh[0][0])[i] = s_zero(T());
  if (tid < 32) {
    S.tau_s[tid] = s_zero(T());
    S.scale_s[tid] = s_zero(T());
    S.beta_s[tid] = s_zero(T());
  }
  const int kmax = min(h, nbc);
  // row 0 is register 0 of row group 0 in wave 0
  if (w == 0 && g == 0) {
    S.rowj[0][cl] = a[0][0];
    S.rowj[0][cl + 16] = a[0][1];
  }
  __syncthreads();
  node2_all_steps<T, NW, KR>(a, S, kmax, nbc);
  __syncthreads();
  // finished columns: v_c = scale_c x_c below the diagonal, R(c, c) = beta_c
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    const int c = cl + 16 * hh;
    if (c < kmax) {
      const T sc = S.scale_s[c];
      const T bt = S.beta_s[c];
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        const int r = rbase + 4 * k;
        a[k][hh] = sel(r > c, s_mul(sc, a[k][hh]), sel(r == c, bt, a[k][hh]));
      }
    }
  }

#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int r = rbase + 4 * k;
    if (r < h) {
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int c = cl + 16 * hh;
        if (c < nbc) X[r + c * ldx] = a[k][hh];
      }
    }
  }
  if (w == 0) {
    // rows < 32 live in wave 0 (WR >= 32): registers k < 8
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int r = g + 4 * k;
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int c = cl + 16 * hh;
        if (Rn && r < nbc && c < nbc) Rn[r + (int64_t)c * ldr] = (r <= c && r < h) ? a[k][hh] : s_zero(T());
        if (V1) V1[r + c * 32] = (r > c && r < h && c < nbc) ? a[k][hh] : (r == c ? s_one(T()) : s_zero(T()));
      }
    }
  }
  if (Tout && w == 0) {
    // T(0:j, j) = -tau_j T(0:j, 0:j) Y(0:j, j); T(j, j) = tau_j (xLARFT).  Lane r
    // keeps row r of T; T(r, k) = 0 for k < r, so the sums need no masks.
    T Treg[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) Treg[c] = s_zero(T());
    node2_tcols<T>(Treg, S.Ysh, S.tau_s, l, kmax, std::make_integer_sequence<int, 32>{});
    if (l < 32) {
#pragma unroll
      for (int c = 0; c < 32; ++c) Tout[l + c * 32] = Treg[c];
    }
  }
}

template <typename T, int NW, int KR>
__global__ void __launch_bounds__(NW * WAVE) geqr2_node2_kernel(Geqr2NodeArgs args) {
  __shared__ Node2Lds<T, NW, 4 * KR> S;
  if (gate_closed(args.gate, args.gate_want)) return;
  const int node = blockIdx.x;
  int r0, r1;
  node_rows(node, args.nodes, args.H, args.chunk, r0, r1);
  T* Tout = args.Tout ? reinterpret_cast<T*>(args.Tout) + (int64_t)node * args.tstride : nullptr;
  T* Rn = args.Rnext ? reinterpret_cast<T*>(args.Rnext) + (int64_t)node * args.nbc : nullptr;
  T* V1 = args.V1out ? reinterpret_cast<T*>(args.V1out) + (int64_t)node * 32 * 32 : nullptr;
  node2_body<T, NW, KR>(S, reinterpret_cast<T*>(args.X) + r0, args.ldx, r1 - r0, args.nbc, Tout, Rn, args.ldr, V1);
}

// One node per workgroup from a task list (independent problems of a batch).
template <typename T, int NW, int KR>
__global__ void __launch_bounds__(NW * WAVE) geqr2_node2_tasks_kernel(const NodeTask* tasks) {
  __shared__ Node2Lds<T, NW, 4 * KR> S;
  const NodeTask t = tasks[blockIdx.x];
  if (t.gate != nullptr && *t.gate != 1) return;  // its panel was factored by CholeskyQR2
  node2_body<T, NW, KR>(S, reinterpret_cast<T*>(t.X), t.ldx, t.H, t.nbc, reinterpret_cast<T*>(t.Tout),
                        reinterpret_cast<T*>(t.Rn), t.ldr, reinterpret_cast<T*>(t.V1));
}

// Element (r, c) of a unit-lower-trapezoidal V stored strictly below the
// diagonal of `V` (the diagonal and above hold other data).
template <typename T>
__device__ inline T vload(const T* V, int64_t ldv, int r, int c) {
  if (r > c) return V[r + c * ldv];
  return r == c ? s_one(T()) : s_zero(T());
}

// One workgroup: finishes the tree after its top node was factored.
//   Mtop = T_top V_top1^H (stored for the expansion of the level below),
//   Qt   = top nbc rows of the tree's explicit Q (walked down the first node of
//          every lower level),
//   LU of Qt - D with d_j = -sign(pivot_j)  ->  U (stored as U^{-1}), L1, D,
//   T = -U D^H L1^{-H},  R = D R_top,  panel top rows <- [R ; L1].
// Every operand is staged in LDS first; triangular inverses run as one
// rank-1 update per step over all (row, col) pairs.
template <typename T, int NB, int NT>
__device__ __forceinline__ void tsqr_root_finish_body(const TsqrRootArgs& args) {
  if (gate_closed(args.gate, args.gate_want)) return;
  __shared__ T Vt[NB][NB + 1];   // top node's V1 (unit lower) ; R_top above
  __shared__ T Rt[NB][NB + 1];
  __shared__ T M[NB][NB + 1];
  __shared__ T Qt[NB][NB + 1];
  __shared__ T W[NB][NB + 1];
  __shared__ T Tn[NB][NB + 1];
  __shared__ T Li[NB][NB + 1];
  __shared__ T dvec[NB];
  const int tid = threadIdx.x;
  const int nbc = args.nbc;
  const T* Vg = reinterpret_cast<const T*>(args.Xtop);
  const int64_t ldt = args.ldxtop;
  const T* Tt = reinterpret_cast<const T*>(args.Ttop);

  for (int i = tid; i < NB * NB; i += NT) {
    const int r = i % NB, c = i / NB;
    const bool in = r < nbc && c < nbc;
    Vt[r][c] = in ? vload(Vg, ldt, r, c) : s_zero(T());
    Rt[r][c] = (in && r <= c) ? Vg[r + c * ldt] : s_zero(T());
    Tn[r][c] = in ? Tt[r + c * NB] : s_zero(T());
  }
  __syncthreads();
  // M = Tn Vt^H ; Qt = I - Vt M
  mm32<T>(M, Tn, false, Vt, true, 1.0, false);
  {
    T* Mo = reinterpret_cast<T*>(args.Mtop);
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      Mo[r + c * NB] = M[r][c];
      Qt[r][c] = (r == c) ? s_one(T()) : s_zero(T());
    }
  }
  __syncthreads();
  mm32<T>(Qt, Vt, false, M, false, -1.0, true);
  // walk down the first node of each lower level: Qt <- Qt - V1 (Tn (V1^H Qt))
  for (int lev = args.nchain - 1; lev >= 0; --lev) {
    const T* V1g = reinterpret_cast<const T*>(args.chainV1[lev]);
    const int64_t ld1 = args.chainLd[lev];
    const T* Tc = reinterpret_cast<const T*>(args.chainT[lev]);
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      const bool in = r < nbc && c < nbc;
      Vt[r][c] = in ? vload(V1g, ld1, r, c) : s_zero(T());
      Tn[r][c] = in ? Tc[r + c * NB] : s_zero(T());
    }
    __syncthreads();
    mm32<T>(W, Vt, true, Qt, false, 1.0, false);
    mm32<T>(M, Tn, false, W, false, 1.0, false);
    mm32<T>(Qt, Vt, false, M, false, -1.0, true);
  }
  // LU of Qt - D without pivoting, d_j = -sign(pivot_j)
  for (int j = 0; j < nbc; ++j) {
    const T piv = Qt[j][j];
    const T d = s_neg(sign_safe(piv));
    const T pj = s_sub(piv, d);
    __syncthreads();
    if (tid == 0) {
      dvec[j] = d;
      Qt[j][j] = pj;
    }
    for (int r = j + 1 + tid; r < nbc; r += NT) Qt[r][j] = s_div(Qt[r][j], pj);
    __syncthreads();
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      if (r > j && c > j && r < nbc && c < nbc) Qt[r][c] = s_sub(Qt[r][c], s_mul(Qt[r][j], Qt[j][c]));
    }
    __syncthreads();
  }
  // Li = L1^{-1} (unit lower) and W = U^{-1} (upper), one rank-1 step per j
  for (int i = tid; i < NB * NB; i += NT) {
    const int r = i % NB, c = i / NB;
    Li[r][c] = (r == c) ? s_one(T()) : s_zero(T());
    W[r][c] = (r == c) ? (r < nbc ? s_div(s_one(T()), Qt[r][r]) : s_one(T())) : s_zero(T());
  }
  __syncthreads();
  for (int j = 0; j < nbc; ++j) {
    // forward: rows r > j of Li:  Li[r][:] -= L[r][j] Li[j][:]
    // backward (column nbc-1-j of U): rows r < jj of W: W[r][:] -= U[r][jj]/U[r][r] W[jj][:]
    const int jj = nbc - 1 - j;
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      if (r > j && r < nbc && c <= j) Li[r][c] = s_sub(Li[r][c], s_mul(Qt[r][j], Li[j][c]));
    }
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      if (r < jj && c >= jj && c < nbc)
        W[r][c] = s_sub(W[r][c], s_mul(s_div(Qt[r][jj], Qt[r][r]), W[jj][c]));
    }
    __syncthreads();
  }
  // T = -(U D^H) Li^H : UD in M (upper), result in Tn
  for (int i = tid; i < NB * NB; i += NT) {
    const int r = i % NB, c = i / NB;
    M[r][c] = (r <= c && c < nbc) ? s_mul(Qt[r][c], s_conj(dvec[c])) : s_zero(T());
  }
  __syncthreads();
  mm32<T>(Tn, M, false, Li, true, -1.0, false);
  {
    T* To = reinterpret_cast<T*>(args.Tout);
    T* Uo = reinterpret_cast<T*>(args.Uout);
    T* Ap = reinterpret_cast<T*>(args.Apanel);
    const int64_t lda = args.lda;
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      To[r + c * NB] = (r <= c && c < nbc) ? Tn[r][c] : s_zero(T());
      Uo[r + c * NB] = (r <= c && c < nbc) ? W[r][c] : s_zero(T());
      if (r < nbc && c < nbc) {
        if (r <= c) Ap[r + c * lda] = s_mul(dvec[r], Rt[r][c]);
        else Ap[r + c * lda] = Qt[r][c];
      }
    }
  }
}

// Expansion of one tree level: every node forms its explicit Q rows from its
// parent's Q block, Q = [Qb; 0] - V M with M = T (V1^H Qb); at the leaves the
// rows become Householder vectors V = Q U^{-1} (rows at or below nbc of the
// panel).  Both products run on FP64 MFMA: a workgroup owns 64 rows of one
// node, each wave 16 rows x 32 columns (two 16x16 blocks, K = 32).

template <typename T, int NB, int NT>
__global__ void __launch_bounds__(NT) tsqr_root_finish_kernel(TsqrRootArgs args) {
  tsqr_root_finish_body<T, NB, NT>(args);
}
template <typename T, int NB, int NT>
__global__ void __launch_bounds__(NT) tsqr_root_finish_tasks_kernel(const TsqrRootArgs* tasks) {
  tsqr_root_finish_body<T, NB, NT>(tasks[blockIdx.x]);
}

template <typename T> struct ExpandLds;
template <> struct ExpandLds<double> {
  double Ar[32][64 + 2];   // A operand, k-major: Ar[k][row]
  double Mr[32][32 + 2];   // B operand: M (first product) / Uinv (second)
  double Ur[32][32 + 2];
};
template <> struct ExpandLds<zd> {
  double Ar[32][64 + 2], Ai[32][64 + 2];
  double Mr[32][32 + 2], Mi[32][32 + 2];
  double Ur[32][32 + 2], Ui[32][32 + 2];
};

// acc[j] (16 rows x 16 cols block j) = A[16 rows of wave w] * B[:, 16 j ..]
template <typename T>
__device__ inline void expand_mma(const ExpandLds<T>& L, bool useU, int w, int lane, v4d (&cr)[2], v4d (&ci)[2]) {
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    cr[j] = v4d{0.0, 0.0, 0.0, 0.0};
    ci[j] = v4d{0.0, 0.0, 0.0, 0.0};
  }
#pragma unroll
  for (int kk = 0; kk < 32; kk += 4) {
    const int k = kk + (lane >> 4);
    const int row = w * 16 + (lane & 15);
    const double ar = L.Ar[k][row];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = j * 16 + (lane & 15);
      const double br = useU ? L.Ur[k][col] : L.Mr[k][col];
      cr[j] = mfma_f64(ar, br, cr[j]);
      if constexpr (ScalarTraits<T>::is_complex) {
        const double ai = L.Ai[k][row];
        const double bi = useU ? L.Ui[k][col] : L.Mi[k][col];
        cr[j] = mfma_f64(-ai, bi, cr[j]);
        ci[j] = mfma_f64(ar, bi, ci[j]);
        ci[j] = mfma_f64(ai, br, ci[j]);
      }
    }
  }
}

template <typename T> __device__ inline void lds_put(ExpandLds<T>& L, int which, int k, int c, T v);
template <> __device__ inline void lds_put<double>(ExpandLds<double>& L, int which, int k, int c, double v) {
  if (which == 0) L.Ar[k][c] = v; else if (which == 1) L.Mr[k][c] = v; else L.Ur[k][c] = v;
}
template <> __device__ inline void lds_put<zd>(ExpandLds<zd>& L, int which, int k, int c, zd v) {
  if (which == 0) { L.Ar[k][c] = v.x; L.Ai[k][c] = v.y; }
  else if (which == 1) { L.Mr[k][c] = v.x; L.Mi[k][c] = v.y; }
  else { L.Ur[k][c] = v.x; L.Ui[k][c] = v.y; }
}
template <typename T> __device__ inline T mk_from(double re, double im);
template <> __device__ inline double mk_from<double>(double re, double) { return re; }
template <> __device__ inline zd mk_from<zd>(double re, double im) { return mkz(re, im); }

template <typename T, int NB, int NT>
__device__ __forceinline__ void tsqr_expand_body(const TsqrExpandArgs& args, const int node, const int tile_y) {
  static_assert(NB == 32 && NT == 256, "expand layout");
  __shared__ T Tn[NB][NB + 1];
  __shared__ T V1[NB][NB + 1];
  __shared__ T Qb[NB][NB + 1];
  __shared__ T W[NB][NB + 1];
  __shared__ T Qb2[NB][NB + 1];
  __shared__ ExpandLds<T> L;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int nbc = args.nbc;
  if (node >= args.nodes) return;  // a task-list launch covers the most nodes of any task
  int r0, r1;
  node_rows(node, args.nodes, args.H, args.chunk, r0, r1);
  const int hnode = r1 - r0;
  const int tile0 = tile_y * 64;  // first local row of this workgroup
  if (tile0 >= hnode) return;         // uniform per workgroup
  if (gate_closed(args.gate, args.gate_want)) return;

  const T* Tsrc = reinterpret_cast<const T*>(args.Tnodes) + (int64_t)node * NB * NB;
  const T* V1src = reinterpret_cast<const T*>(args.V1base) + (int64_t)node * args.v1stride;
  for (int i = tid; i < NB * NB; i += NT) {
    const int r = i % NB, c = i / NB;
    const bool in = r < nbc && c < nbc;
    Tn[r][c] = in ? Tsrc[r + c * NB] : s_zero(T());
    V1[r][c] = in ? vload(V1src, (int64_t)args.v1ld, r, c) : s_zero(T());
    if (args.Umat)
      lds_put<T>(L, 2, r, c, in ? reinterpret_cast<const T*>(args.Umat)[r + c * NB] : s_zero(T()));
    if (args.Mtop) {
      W[r][c] = in ? reinterpret_cast<const T*>(args.Mtop)[r + c * NB] : s_zero(T());
    } else {
      const T* Qp = reinterpret_cast<const T*>(args.Qparent) + (int64_t)node * nbc;
      Qb[r][c] = in ? Qp[r + (int64_t)c * args.ldqp] : s_zero(T());
    }
  }
  __syncthreads();
  if (args.Mtop) {
    // the parent is the top node: its Q block for this child is
    // E - V_top[node*nbc + r, :] Mtop, with E the identity for node 0
    const T* Vt = reinterpret_cast<const T*>(args.Xtop);
    const int64_t ldt = args.ldxtop;
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      const int g = node * nbc + r;
      Tn[r][c] = s_zero(T());  // Tn is reloaded below; reuse it as staging
      V1[r][c] = (r < nbc && c < nbc) ? vload(Vt, ldt, g, c) : s_zero(T());
    }
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      const int g = node * nbc + r;
      Qb[r][c] = (r < nbc && c < nbc && g == c) ? s_one(T()) : s_zero(T());
    }
    __syncthreads();
    mm32<T>(Qb, V1, false, W, false, -1.0, true);
    const T* V1src2 = reinterpret_cast<const T*>(args.V1base) + (int64_t)node * args.v1stride;
    const T* Tsrc2 = reinterpret_cast<const T*>(args.Tnodes) + (int64_t)node * NB * NB;
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      const bool in = r < nbc && c < nbc;
      Tn[r][c] = in ? Tsrc2[r + c * NB] : s_zero(T());
      V1[r][c] = in ? vload(V1src2, (int64_t)args.v1ld, r, c) : s_zero(T());
    }
    __syncthreads();
  }
  // W = V1^H Qb ; M = Tn W -> B operand of the first product
  mm32<T>(W, V1, true, Qb, false, 1.0, false);
  mm32<T>(Qb2, Tn, false, W, false, 1.0, false);
  for (int i = tid; i < NB * NB; i += NT) lds_put<T>(L, 1, i % NB, i / NB, Qb2[i % NB][i / NB]);
  // A operand: 64 rows of V (unit lower in node-local rows), k-major
  T* X = reinterpret_cast<T*>(args.X);
  const int64_t ldx = args.ldx;
  for (int i = tid; i < 64 * NB; i += NT) {
    const int rr = i % 64, k = i / 64;
    const int lrow = tile0 + rr;
    T v = s_zero(T());
    if (lrow < hnode && k < nbc)
      v = (lrow > k) ? X[(r0 + lrow) + k * ldx] : (lrow == k ? s_one(T()) : s_zero(T()));
    lds_put<T>(L, 0, k, rr, v);
  }
  __syncthreads();
  v4d cr[2], ci[2];
  expand_mma<T>(L, false, w, lane, cr, ci);
  // q = Qb_part - V M
  T q[2][4];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = w * 16 + mfma_f64_row(lane, r);
      const int col = j * 16 + (lane & 15);
      const int lrow = tile0 + rr;
      const T base = (lrow < nbc && col < nbc) ? Qb[lrow][col] : s_zero(T());
      q[j][r] = s_sub(base, mk_from<T>(cr[j][r], ci[j][r]));
    }
  if (!args.Umat) {
    T* Qo = reinterpret_cast<T*>(args.Qout);
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = w * 16 + mfma_f64_row(lane, r);
        const int col = j * 16 + (lane & 15);
        const int lrow = tile0 + rr;
        if (lrow < hnode && col < nbc) Qo[(r0 + lrow) + col * (int64_t)args.ldqo] = q[j][r];
      }
    return;
  }
  // leaves: V = Q Uinv.  Every wave only touches its own 16 rows of L.A.
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = w * 16 + mfma_f64_row(lane, r);
      const int col = j * 16 + (lane & 15);
      lds_put<T>(L, 0, col, rr, q[j][r]);
    }
  __syncthreads();
  expand_mma<T>(L, true, w, lane, cr, ci);
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = w * 16 + mfma_f64_row(lane, r);
      const int col = j * 16 + (lane & 15);
      const int lrow = tile0 + rr;
      const int grow = r0 + lrow;
      if (lrow < hnode && col < nbc && grow >= nbc) X[grow + col * ldx] = mk_from<T>(cr[j][r], ci[j][r]);
    }
}


template <typename T, int NB, int NT>
__global__ void __launch_bounds__(NT) tsqr_expand_kernel(TsqrExpandArgs args) {
  tsqr_expand_body<T, NB, NT>(args, blockIdx.x, blockIdx.y);
}
template <typename T, int NB, int NT>
__global__ void __launch_bounds__(NT) tsqr_expand_tasks_kernel(const TsqrExpandArgs* tasks) {
  tsqr_expand_body<T, NB, NT>(tasks[blockIdx.z], blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
// node kernel configuration: waves per node, rows per lane (4 * KR rows per wave)
template <typename T> struct NodeCfg;
template <> struct NodeCfg<double> { static constexpr int NW = 4, KR = 32; };
template <> struct NodeCfg<zd> { static constexpr int NW = 4, KR = 16; };

// Geqr2NodeArgs::variant: 1: per-thread-row kernel, 2: 2-D lane layout
// (NodeCfg), 3: 2-D lane layout with another wave count; 0: the
// measured-fastest per element type (real: 1, 78 us per 512 x 32 node;
// complex: 2, 137 us per 256 x 32 node, on MI355X)

template <typename T> int geqr2_node_capacity() {
  return ScalarTraits<T>::is_complex ? 256 : 512;
}
template int geqr2_node_capacity<double>();
template int geqr2_node_capacity<zd>();

template <typename T>
void launch_geqr2_node(const Geqr2NodeArgs& a, int leaf_rows_max, hipStream_t s) {
  // the register layout holds geqr2_node_capacity<T>() rows per workgroup
  static_assert(NodeCfg<T>::NW * 4 * NodeCfg<T>::KR == (ScalarTraits<T>::is_complex ? 256 : 512),
                "node capacity");
  const int variant = a.variant != 0 ? a.variant : (ScalarTraits<T>::is_complex ? 2 : 1);
  if (variant == 2) {
    constexpr int NW = NodeCfg<T>::NW, KR = NodeCfg<T>::KR;
    hipLaunchKernelGGL((geqr2_node2_kernel<T, NW, KR>), dim3(a.nodes), dim3(NW * WAVE), 0, s, a);
  } else if (variant == 3) {
    // alternative shape of the same kernel (same capacity), for tuning
    constexpr int NW = ScalarTraits<T>::is_complex ? 8 : 8, KR = ScalarTraits<T>::is_complex ? 8 : 16;
    hipLaunchKernelGGL((geqr2_node2_kernel<T, NW, KR>), dim3(a.nodes), dim3(NW * WAVE), 0, s, a);
  } else if constexpr (ScalarTraits<T>::is_complex) {
    hipLaunchKernelGGL((geqr2_node_kernel<T, 32, 1, 256>), dim3(a.nodes), dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL((geqr2_node_kernel<T, 32, 2, 256>), dim3(a.nodes), dim3(256), 0, s, a);
  }
  (void)leaf_rows_max;
}

// Rows of the node size classes of the task-list launches (0: none).  A
// ComplexF64 node of more than 256 rows does not fit a lane's registers (it
// spills to scratch), so complex panels beyond 256 rows go to TSQR leaves.
template <typename T> int node_class_rows(int cls) {
  if (cls == 0) return 256;
  if (ScalarTraits<T>::is_complex) return 0;
  return cls == 1 ? 512 : 1024;
}

template <typename T> void launch_geqr2_node_tasks(const NodeTask* tasks, int ntasks, int cls, hipStream_t s) {
  if (ntasks <= 0) return;
  if constexpr (ScalarTraits<T>::is_complex) {
    (void)cls;  // one class: 256 rows
    hipLaunchKernelGGL((geqr2_node2_tasks_kernel<T, 4, 16>), dim3(ntasks), dim3(256), 0, s, tasks);
  } else {
    if (cls == 0) hipLaunchKernelGGL((geqr2_node2_tasks_kernel<T, 4, 16>), dim3(ntasks), dim3(256), 0, s, tasks);
    else if (cls == 1) hipLaunchKernelGGL((geqr2_node2_tasks_kernel<T, 4, 32>), dim3(ntasks), dim3(256), 0, s, tasks);
    else hipLaunchKernelGGL((geqr2_node2_tasks_kernel<T, 8, 32>), dim3(ntasks), dim3(512), 0, s, tasks);
  }
}
template <typename T> void launch_tsqr_root_finish_tasks(const TsqrRootArgs* tasks, int ntasks, hipStream_t s) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL((tsqr_root_finish_tasks_kernel<T, 32, 256>), dim3(ntasks), dim3(256), 0, s, tasks);
}
template <typename T>
void launch_tsqr_expand_tasks(const TsqrExpandArgs* tasks, int ntasks, int nodes, int max_node_rows, hipStream_t s) {
  if (ntasks <= 0) return;
  dim3 grid(nodes, cdiv(max_node_rows, 64), ntasks);
  hipLaunchKernelGGL((tsqr_expand_tasks_kernel<T, 32, 256>), grid, dim3(256), 0, s, tasks);
}
template void launch_tsqr_root_finish_tasks<double>(const TsqrRootArgs*, int, hipStream_t);
template void launch_tsqr_root_finish_tasks<zd>(const TsqrRootArgs*, int, hipStream_t);
template void launch_tsqr_expand_tasks<double>(const TsqrExpandArgs*, int, int, int, hipStream_t);
template void launch_tsqr_expand_tasks<zd>(const TsqrExpandArgs*, int, int, int, hipStream_t);
template int node_class_rows<double>(int);
template int node_class_rows<zd>(int);
template void launch_geqr2_node_tasks<double>(const NodeTask*, int, int, hipStream_t);
template void launch_geqr2_node_tasks<zd>(const NodeTask*, int, int, hipStream_t);

template <typename T> void launch_tsqr_root_finish(const TsqrRootArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((tsqr_root_finish_kernel<T, 32, 256>), dim3(1), dim3(256), 0, s, a);
}

template <typename T> void launch_tsqr_expand(const TsqrExpandArgs& a, int max_node_rows, hipStream_t s) {
  dim3 grid(a.nodes, cdiv(max_node_rows, 64));
  hipLaunchKernelGGL((tsqr_expand_kernel<T, 32, 256>), grid, dim3(256), 0, s, a);
}

template void launch_geqr2_node<double>(const Geqr2NodeArgs&, int, hipStream_t);
template void launch_geqr2_node<zd>(const Geqr2NodeArgs&, int, hipStream_t);
template void launch_tsqr_root_finish<double>(const TsqrRootArgs&, hipStream_t);
template void launch_tsqr_root_finish<zd>(const TsqrRootArgs&, hipStream_t);
template void launch_tsqr_expand<double>(const TsqrExpandArgs&, int, hipStream_t);
template void launch_tsqr_expand<zd>(const TsqrExpandArgs&, int, hipStream_t);

}  // namespace mak
