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
#include "mak_common.hpp"
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

template <typename T, int NB> __device__ inline T wave_reduce_transposed(T (&p)[NB]) {
  const int l = lane_id();
  // halving steps
#pragma unroll
  for (int len = NB, m = 32; len > 1; len >>= 1, m >>= 1) {
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
  T v = p[0];
  // remaining full-reduction steps over the lanes that share a column
#pragma unroll
  for (int m = 32 / NB; m >= 1; m >>= 1) v = s_add(v, shfl_xor_s(v, m));
  return v;
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

template <typename T, int NB, int RPT, int NT>
__global__ void __launch_bounds__(NT) geqr2_node_kernel(Geqr2NodeArgs args) {
  // Register layout: thread t holds rows t + NT*q (q < RPT).  The row is kept
  // ROTATED: at step j, slot i holds column (i + j) mod NB, so the current
  // column is always slot 0 and no register is indexed by the runtime j.
  constexpr int NW = NT / WAVE;
  __shared__ T red_w[NW][NB];
  __shared__ T rowj[NB];   // row j, rotated order
  __shared__ T gvec[NB];   // conj(tau) * v^H a_c, rotated order (0 unless trailing)
  __shared__ T yvec[NB];   // V_c^H v_j, natural order
  __shared__ T scal[3];    // scale, tau, beta
  __shared__ T Tl[NB][NB + 1];

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
  for (int i = tid; i < NB * (NB + 1); i += NT) (&Tl[0][0])[i] = s_zero(T());
  __syncthreads();

  const int kmax = min(h, nbc);
  for (int j = 0; j < kmax; ++j) {
    // partial sums t_i = sum_{rows r > j} conj(a_rj) a_r(slot i)
    T p[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) p[i] = s_zero(T());
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + NT * q;
      const T x = (r > j && r < h) ? a[q][0] : s_zero(T());
#pragma unroll
      for (int i = 0; i < NB; ++i) p[i] = s_cfma(x, a[q][i], p[i]);
    }
    T tot = wave_reduce_transposed<T, NB>(p);
    if ((l & (64 / NB - 1)) == 0) red_w[wid][col_of_lane<NB>(l)] = tot;
    if (tid == j) {
#pragma unroll
      for (int i = 0; i < NB; ++i) rowj[i] = a[0][i];
    }
    __syncthreads();
    if (tid < NB) {
      const int i = tid;             // slot
      int c = i + j;                 // column held by slot i
      if (c >= NB) c -= NB;
      T ri = s_zero(T());
      double nu2 = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        ri = s_add(ri, red_w[w][i]);
        nu2 += s_re(red_w[w][0]);
      }
      const T alpha = rowj[0];
      T tau, beta, scale;
      if (nu2 == 0.0 && s_im(alpha) == 0.0) {
        tau = s_zero(T());
        beta = alpha;
        scale = s_zero(T());
      } else {
        const double bn = sqrt(s_abs2(alpha) + nu2);
        const double br = s_re(alpha) >= 0.0 ? -bn : bn;
        beta = s_from_real<T>(br);
        tau = s_scale(s_sub(beta, alpha), 1.0 / br);
        scale = s_div(s_one(T()), s_sub(alpha, beta));
      }
      T g = s_zero(T());
      if (c > j && c < nbc) {
        // w_c = a_jc + conj(scale) red ; g = conj(tau) w_c
        g = s_mul(s_conj(tau), s_cfma(scale, ri, rowj[i]));
      } else if (c < j) {
        // y_c = conj(a_jc) + scale conj(red)
        yvec[c] = s_fma(scale, s_conj(ri), s_conj(rowj[i]));
      }
      gvec[i] = g;
      if (i == 0) {
        scal[0] = scale;
        scal[1] = tau;
        scal[2] = beta;
      }
    }
    __syncthreads();
    {
      const T scale = scal[0];
      const T tau = scal[1];
      const T beta = scal[2];
      T g[NB];
#pragma unroll
      for (int i = 1; i < NB; ++i) g[i] = gvec[i];
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int r = tid + NT * q;
        const bool below = (r > j && r < h);
        const bool onrow = (r == j);
        const T v = s_mul(scale, a[q][0]);
        const T coef = below ? v : (onrow ? s_one(T()) : s_zero(T()));
#pragma unroll
        for (int i = 1; i < NB; ++i) a[q][i] = s_sub(a[q][i], s_mul(coef, g[i]));
        a[q][0] = below ? v : (onrow ? beta : a[q][0]);
      }
      // T(0:j, j) = -tau T(0:j, 0:j) y ; T(j, j) = tau
      if (tid < j) {
        T s = s_zero(T());
        for (int k = tid; k < j; ++k) s = s_fma(Tl[tid][k], yvec[k], s);
        Tl[tid][j] = s_neg(s_mul(tau, s));
      } else if (tid == j) {
        Tl[j][j] = tau;
      }
      // rotate: slot i <- slot i+1
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const T first = a[q][0];
#pragma unroll
        for (int i = 0; i < NB - 1; ++i) a[q][i] = a[q][i + 1];
        a[q][NB - 1] = first;
      }
    }
    __syncthreads();
  }

  // slot i now holds column (i + kmax) mod NB
  // write back the factored block (V strictly below, R on and above)
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + NT * q;
    if (r < h) {
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        int c = i + kmax;
        if (c >= NB) c -= NB;
        if (c < nbc) X[r + c * ldx] = a[q][i];
      }
    }
  }
  if (args.Tout) {
    T* To = reinterpret_cast<T*>(args.Tout) + (int64_t)node * args.tstride;
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      To[r + c * NB] = (r <= c) ? Tl[r][c] : s_zero(T());
    }
  }
  if (args.Rnext) {
    // R of this node: rows [node*nbc, node*nbc + nbc) of the next level
    T* Rn = reinterpret_cast<T*>(args.Rnext) + (int64_t)node * nbc;
    const int r = tid;
    if (r < nbc) {
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        int c = i + kmax;
        if (c >= NB) c -= NB;
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
        int c = i + kmax;
        if (c >= NB) c -= NB;
        V1[r + c * NB] = (r > c && r < h && c < nbc) ? a[0][i] : (r == c ? s_one(T()) : s_zero(T()));
      }
    }
  }
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
template <typename T, int NB, int NT>
__global__ void __launch_bounds__(NT) tsqr_root_finish_kernel(TsqrRootArgs args) {
  __shared__ T M[NB][NB + 1];
  __shared__ T Qt[NB][NB + 1];
  __shared__ T W[NB][NB + 1];
  __shared__ T Tn[NB][NB + 1];
  __shared__ T Li[NB][NB + 1];
  __shared__ T dvec[NB];
  const int tid = threadIdx.x;
  const int nbc = args.nbc;
  const T* Vt = reinterpret_cast<const T*>(args.Xtop);
  const int64_t ldt = args.ldxtop;

  {
    const T* Tt = reinterpret_cast<const T*>(args.Ttop);
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      Tn[r][c] = (r < nbc && c < nbc) ? Tt[r + c * NB] : s_zero(T());
    }
    __syncthreads();
    T* Mo = reinterpret_cast<T*>(args.Mtop);
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      T s = s_zero(T());
      if (r < nbc && c < nbc) {
        for (int k = r; k < nbc; ++k) s = s_fma(Tn[r][k], s_conj(vload(Vt, ldt, c, k)), s);
      }
      M[r][c] = s;
      Mo[r + c * NB] = s;
    }
    __syncthreads();
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      T s = (r == c) ? s_one(T()) : s_zero(T());
      if (r < nbc && c < nbc) {
        for (int k = 0; k <= r; ++k) s = s_sub(s, s_mul(vload(Vt, ldt, r, k), M[k][c]));
      }
      Qt[r][c] = s;
    }
    __syncthreads();
  }
  // walk down the first node of each lower level: Qt <- Qt - V1 (Tn (V1^H Qt))
  for (int lev = args.nchain - 1; lev >= 0; --lev) {
    const T* V1 = reinterpret_cast<const T*>(args.chainV1[lev]);
    const int64_t ld1 = args.chainLd[lev];
    const T* Tc = reinterpret_cast<const T*>(args.chainT[lev]);
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      Tn[r][c] = (r < nbc && c < nbc) ? Tc[r + c * NB] : s_zero(T());
      T s = s_zero(T());
      if (r < nbc && c < nbc)
        for (int k = r; k < nbc; ++k) s = s_cfma(vload(V1, ld1, k, r), Qt[k][c], s);
      W[r][c] = s;
    }
    __syncthreads();
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      T s = s_zero(T());
      if (r < nbc && c < nbc)
        for (int k = r; k < nbc; ++k) s = s_fma(Tn[r][k], W[k][c], s);
      M[r][c] = s;
    }
    __syncthreads();
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      T s = Qt[r][c];
      if (r < nbc && c < nbc)
        for (int k = 0; k <= r; ++k) s = s_sub(s, s_mul(vload(V1, ld1, r, k), M[k][c]));
      W[r][c] = s;
    }
    __syncthreads();
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      Qt[r][c] = W[r][c];
    }
    __syncthreads();
  }
  // LU of Qt - D without pivoting, d_j = -sign(pivot_j)
  for (int j = 0; j < nbc; ++j) {
    if (tid == 0) {
      const T piv = Qt[j][j];
      const T d = s_neg(sign_safe(piv));
      dvec[j] = d;
      Qt[j][j] = s_sub(piv, d);
    }
    __syncthreads();
    for (int r = j + 1 + tid; r < nbc; r += NT) Qt[r][j] = s_div(Qt[r][j], Qt[j][j]);
    __syncthreads();
    for (int i = tid; i < nbc * nbc; i += NT) {
      const int r = i % nbc, c = i / nbc;
      if (r > j && c > j) Qt[r][c] = s_sub(Qt[r][c], s_mul(Qt[r][j], Qt[j][c]));
    }
    __syncthreads();
  }
  // Li = L1^{-1} (unit lower), one column per thread
  if (tid < NB) {
    const int c = tid;
    for (int r = 0; r < NB; ++r) Li[r][c] = s_zero(T());
    if (c < nbc) {
      Li[c][c] = s_one(T());
      for (int r = c + 1; r < nbc; ++r) {
        T s = s_zero(T());
        for (int k = c; k < r; ++k) s = s_fma(Qt[r][k], Li[k][c], s);
        Li[r][c] = s_neg(s);
      }
    }
  }
  // W = U^{-1} (upper), one column per thread, back substitution
  if (tid < NB) {
    const int c = tid;
    for (int r = 0; r < NB; ++r) W[r][c] = s_zero(T());
    if (c < nbc) {
      W[c][c] = s_div(s_one(T()), Qt[c][c]);
      for (int r = c - 1; r >= 0; --r) {
        T s = s_zero(T());
        for (int k = r + 1; k <= c; ++k) s = s_fma(Qt[r][k], W[k][c], s);
        W[r][c] = s_neg(s_div(s, Qt[r][r]));
      }
    }
  }
  __syncthreads();
  {
    T* To = reinterpret_cast<T*>(args.Tout);
    T* Uo = reinterpret_cast<T*>(args.Uout);
    T* Ap = reinterpret_cast<T*>(args.Apanel);
    const int64_t lda = args.lda;
    for (int i = tid; i < NB * NB; i += NT) {
      const int r = i % NB, c = i / NB;
      T t = s_zero(T());
      if (r <= c && c < nbc) {
        for (int k = r; k <= c; ++k)
          t = s_fma(s_mul(Qt[r][k], s_conj(dvec[k])), s_conj(Li[c][k]), t);
        t = s_neg(t);
      }
      To[r + c * NB] = t;
      Uo[r + c * NB] = (r <= c && c < nbc) ? W[r][c] : s_zero(T());
      if (r < nbc && c < nbc) {
        if (r <= c) Ap[r + c * lda] = s_mul(dvec[r], Vt[r + c * ldt]);
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
__global__ void __launch_bounds__(NT) tsqr_expand_kernel(TsqrExpandArgs args) {
  static_assert(NB == 32 && NT == 256, "expand layout");
  __shared__ T Tn[NB][NB + 1];
  __shared__ T V1[NB][NB + 1];
  __shared__ T Qb[NB][NB + 1];
  __shared__ T W[NB][NB + 1];
  __shared__ ExpandLds<T> L;
  const int node = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int nbc = args.nbc;
  int r0, r1;
  node_rows(node, args.nodes, args.H, args.chunk, r0, r1);
  const int hnode = r1 - r0;
  const int tile0 = blockIdx.y * 64;  // first local row of this workgroup
  if (tile0 >= hnode) return;         // uniform per workgroup

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
      T s = s_zero(T());
      if (r < nbc && c < nbc) {
        const int g = node * nbc + r;
        s = (g == c) ? s_one(T()) : s_zero(T());
        const int kend = min(nbc, g + 1);
        for (int k = 0; k < kend; ++k) s = s_sub(s, s_mul(vload(Vt, ldt, g, k), W[k][c]));
      }
      Qb[r][c] = s;
    }
    __syncthreads();
  }
  // W = V1^H Qb
  for (int i = tid; i < NB * NB; i += NT) {
    const int r = i % NB, c = i / NB;
    T s = s_zero(T());
    for (int k = r; k < nbc; ++k) s = s_cfma(V1[k][r], Qb[k][c], s);
    W[r][c] = s;
  }
  __syncthreads();
  // M = Tn W  -> B operand of the first product
  for (int i = tid; i < NB * NB; i += NT) {
    const int r = i % NB, c = i / NB;
    T s = s_zero(T());
    for (int k = r; k < nbc; ++k) s = s_fma(Tn[r][k], W[k][c], s);
    lds_put<T>(L, 1, r, c, s);
  }
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

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <typename T> int geqr2_node_capacity() {
  return ScalarTraits<T>::is_complex ? 256 : 1024;
}
template int geqr2_node_capacity<double>();
template int geqr2_node_capacity<zd>();

template <typename T>
void launch_geqr2_node(const Geqr2NodeArgs& a, int leaf_rows_max, hipStream_t s) {
  // the register layout holds geqr2_node_capacity<T>() rows per workgroup
  if constexpr (ScalarTraits<T>::is_complex) {
    hipLaunchKernelGGL((geqr2_node_kernel<T, 32, 1, 256>), dim3(a.nodes), dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL((geqr2_node_kernel<T, 32, 2, 512>), dim3(a.nodes), dim3(512), 0, s, a);
  }
  (void)leaf_rows_max;
}

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
