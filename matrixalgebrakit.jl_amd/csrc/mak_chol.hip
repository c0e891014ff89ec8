// Cholesky factorisation with inverse, G = R^H R -> R (upper) and X = R^{-1},
// for FP64 / ComplexF64 Hermitian positive definite matrices.
//
// Small and medium n (n <= kCholTileMaxN): ONE persistent launch.  The
// matrix is cut into 32 x 32 tiles and [G | I] is eliminated right-looking:
//   D(k)        R_kk, X_kk = chol_inv(G_kk)                  (one wave)
//   P(k, j)     R_kj = X_kk^H G_kj                           (j > k)
//   U(k, i, j)  G_ij -= R_ki^H R_kj                          (k < i <= j)
//   U'(k, j, l) Z_jl -= X_jk R_kl                            (j <= k < l)
//   P'(k, j)    X_jk = Z_jk X_kk                             (j < k)
// (X R = I column block by column block: X_jl = -(sum_{k=j}^{l-1} X_jk R_kl) X_ll.)
// Every task is a 32 x 32 x 32 product on FP64 MFMA except D, whose in-wave
// elimination carries the identity along (the inverse costs no extra
// dependency chain).  Workgroups take tasks from an atomic counter in a fixed
// topological order with a one-step look-ahead (the next diagonal block's D
// and P tasks come before the bulk of the current step's updates), and wait
// for the per-tile version counters the task needs:
//   gst[i][j]  G/R tile (i <= j): number of U updates applied; i + 1 = R final
//   zst[j][l]  X tile (j <= l): number of U' updates applied; l - j + 1 = final
// Every dependency of a task precedes it in the order, and a workgroup only
// waits after it has taken its task, so the waits always end; each wait is
// still bounded (kSpinLimit polls), and a timeout or a non-positive pivot sets
// an abort flag that every later wait and task honours, so the grid drains.
//
// Larger n: a blocked loop over kCholBlock-column diagonal blocks, each
// factored by the tile kernel, with MFMA GEMMs for the panel, the trailing
// update and the inverse's column blocks.
//
// Replaces the potrf / trtri pair of the Cholesky-based steps (the reference
// has none on this path: its polar goes through the SVD,
// src/implementations/polar.jl:59-95; the QR's CholeskyQR2 route and the
// QDWH polar iteration are this backend's own).
#include <algorithm>
#include <vector>

#include "mak_common.hpp"
#include "mak_dense32.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

namespace {

constexpr int TB = 32;                 // tile size
constexpr long kSpinLimit = 1L << 22;  // polls per wait before the abort flag is raised

// ---------------------------------------------------------------- task order
enum CholTaskType { CT_D = 0, CT_P = 1, CT_UQ = 2, CT_PQ = 3, CT_U = 4 };

struct CholTask {
  int type, k, i, j;
};

// Tasks of iteration k, by segment, in the order they are taken:
//   D(k)  P(k, k+1)                               (the panel workgroup's)
//   s1  P(k, j)            j = k+2 .. nt-1
//   s4a U(k-1, k+1, j)     j = k+1 .. nt-1        (row k+1 of step k-1)
//   --  U(k, k+1, k+1)                            (the panel workgroup's)
//   s6  U(k, k+1, j)       j = k+2 .. nt-1
//   s2  U'(k-1, j, k)      j < k
//   s3  P'(k, j)           j < k
//   s4b U(k-1, i, j)       k+2 <= i <= j <= nt-1
//   s5  U'(k-1, j, l)      j < k < l
// The inputs of the next panel steps (P(k, k+2), row k+1 of step k-1 and
// U(k, k+1, k+2)) come first; the bulk updates and the inverse follow.
// panel = 1 leaves out the panel workgroup's tasks (D(k), P(k, k+1) and
// U(k, k+1, k+1)): the order the other workgroups take theirs in.
constexpr int kCholSegs = 10;
__host__ __device__ inline void chol_iter_counts(int nt, int k, int c[kCholSegs], int panel = 0) {
  const int rest = nt - 2 - k > 0 ? nt - 2 - k : 0;  // columns k+2 .. nt-1
  c[0] = panel ? 0 : 1;                              // D(k)
  c[1] = (!panel && k <= nt - 2) ? 1 : 0;            // P(k, k+1)
  c[2] = rest;                                       // s1
  c[3] = (k >= 1 && k <= nt - 2) ? nt - 1 - k : 0;   // s4a
  c[4] = (!panel && k <= nt - 2) ? 1 : 0;            // U(k, k+1, k+1)
  c[5] = rest;                                       // s6
  c[6] = k;                                          // s2
  c[7] = k;                                          // s3
  c[8] = k >= 1 ? rest * (rest + 1) / 2 : 0;         // s4b
  c[9] = k * (nt - 1 - k);                           // s5
}

__host__ __device__ inline int chol_iter_total(int nt, int k, int panel = 0) {
  int c[kCholSegs];
  chol_iter_counts(nt, k, c, panel);
  int t = 0;
  for (int q = 0; q < kCholSegs; ++q) t += c[q];
  return t;
}

// u-th task of iteration k (0 <= u < chol_iter_total(nt, k, panel)).
__host__ __device__ inline CholTask chol_task_in_iter(int nt, int k, int u, int panel = 0) {
  int c[kCholSegs];
  chol_iter_counts(nt, k, c, panel);
  CholTask t{CT_D, k, k, k};
  int q = 0;
  while (q < kCholSegs - 1 && u >= c[q]) u -= c[q++];
  switch (q) {
    case 0: break;
    case 1: t.type = CT_P; t.j = k + 1; break;
    case 2: t.type = CT_P; t.j = k + 2 + u; break;
    case 3: t.type = CT_U; t.k = k - 1; t.i = k + 1; t.j = k + 1 + u; break;
    case 4: t.type = CT_U; t.i = k + 1; t.j = k + 1; break;
    case 5: t.type = CT_U; t.i = k + 1; t.j = k + 2 + u; break;
    case 6: t.type = CT_UQ; t.k = k - 1; t.i = u; t.j = k; break;
    case 7: t.type = CT_PQ; t.i = u; t.j = k; break;
    case 8: {  // pairs k+2 <= i <= j <= nt-1, row by row
      int i = k + 2, len = nt - i;
      while (u >= len) { u -= len; ++i; --len; }
      t.type = CT_U; t.k = k - 1; t.i = i; t.j = i + u;
      break;
    }
    default: {
      const int w = nt - 1 - k;
      t.type = CT_UQ; t.k = k - 1; t.i = u / w; t.j = k + 1 + u % w;
      break;
    }
  }
  return t;
}

struct CholTileArgs {
  void* G;        // in: upper triangle of the Hermitian matrix; out: R (strict lower part zeroed)
  int64_t ldg;
  void* X;        // out: R^{-1} (upper, strict lower part zeroed)
  int64_t ldx;
  int n, nt, ntasks;
  int* gst;       // nt * nt
  int* zst;       // nt * nt
  int* ctl;       // [0] task counter, [1] abort flag, [2] first non-positive pivot (1-based) or 0
  int* info;      // optional: receives ctl[2] + info_offset on failure (when *info == 0)
  int info_offset;
  long long* trace;  // optional (tests): wall-clock stamps of the panel's iterations, 8 per iteration
  int iter_off[kCholTileMaxNt + 1];  // first queue task of every iteration
};

// ---------------------------------------------------------------- sync helpers
__device__ inline int ld_flag(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Thread 0 polls until *p >= want (or the abort flag is set, or the spin limit
// is reached, which raises it).  The caller follows with inputs_ready().
__device__ inline void wait_ge(const int* p, int want, int* ctl) {
  if (threadIdx.x == 0) {
    long spins = 0;
    while (ld_flag(p) < want) {
      if (ld_flag(ctl + 1) != 0) break;
      if (++spins > kSpinLimit) {
        __hip_atomic_store(ctl + 1, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
}

// Memory protocol between workgroups.  Tiles that tasks exchange are moved
// with agent-scope relaxed atomic loads and stores (sc1: coherent across the
// XCDs' L2 caches location by location), so no whole-cache writeback or
// invalidate is needed:
//   producer: every wave waits for its stores to complete (vmcnt(0)), then a
//             workgroup barrier, then thread 0 stores the version counter;
//   consumer: thread 0 polls the counter, then a workgroup barrier, then the
//             tile loads.
// A load issued after the counter was seen reads the completed stores.
__device__ inline bool inputs_ready(int* ctl) {
  __shared__ int abort_s;
  if (threadIdx.x == 0) abort_s = ld_flag(ctl + 1);
  __syncthreads();
  return abort_s == 0;
}

__device__ inline void stores_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // s_waitcnt vmcnt(0)

__device__ inline void publish(int* p, int v, int* p2 = nullptr, int v2 = 0) {
  stores_done();
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (p2) __hip_atomic_store(p2, v2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__device__ inline double ld_coh(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline zd ld_coh(const zd* p) {
  const double* q = reinterpret_cast<const double*>(p);
  return mkz(ld_coh(q), ld_coh(q + 1));
}
__device__ inline void st_coh(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_coh(zd* p, zd v) {
  double* q = reinterpret_cast<double*>(p);
  st_coh(q, v.x);
  st_coh(q + 1, v.y);
}

// ---------------------------------------------------------------- tile I/O
// Fast path (real data, a tile entirely inside the matrix, even leading
// dimension, 16-byte aligned base): thread t moves rows 4 (t % 8) .. +3 of
// column t / 8 with two 16-byte coherent (sc1) accesses, so a wave touches
// four 256-byte column segments instead of 256 separate 8-byte locations.
// Loads of up to three tiles are issued in one asm block and waited for
// inside it (the outputs are only valid after the wait).
typedef double v2d __attribute__((ext_vector_type(2)));

template <typename T> __device__ inline bool tile_fast(const T* M, int64_t ld, int n, int ti, int tj) {
  if constexpr (ScalarTraits<T>::is_complex) {
    return false;
  } else {
    return (ti + 1) * TB <= n && (tj + 1) * TB <= n && (ld & 1) == 0 &&
           (reinterpret_cast<uintptr_t>(M) & 15) == 0;
  }
}
__device__ inline const double* tile_chunk(const double* M, int64_t ld, int ti, int tj) {
  const int t = threadIdx.x;
  return M + (int64_t)(tj * TB + (t >> 3)) * ld + ti * TB + 4 * (t & 7);
}
__device__ inline void put_chunk(double (*S)[33], v2d a, v2d b) {
  const int t = threadIdx.x;
  const int c = t >> 3, r = 4 * (t & 7);
  S[r][c] = a.x;
  S[r + 1][c] = a.y;
  S[r + 2][c] = b.x;
  S[r + 3][c] = b.y;
}
__device__ inline void load_fast1(double (*S)[33], const double* p) {
  v2d a, b;
  asm volatile("global_load_dwordx4 %0, %2, off sc1\n\tglobal_load_dwordx4 %1, %2, off offset:16 sc1\n\t"
               "s_waitcnt vmcnt(0)"
               : "=&v"(a), "=&v"(b) : "v"(p) : "memory");
  put_chunk(S, a, b);
}
__device__ inline void load_fast2(double (*S0)[33], const double* p0, double (*S1)[33], const double* p1) {
  v2d a, b, c, d;
  asm volatile("global_load_dwordx4 %0, %4, off sc1\n\tglobal_load_dwordx4 %1, %4, off offset:16 sc1\n\t"
               "global_load_dwordx4 %2, %5, off sc1\n\tglobal_load_dwordx4 %3, %5, off offset:16 sc1\n\t"
               "s_waitcnt vmcnt(0)"
               : "=&v"(a), "=&v"(b), "=&v"(c), "=&v"(d) : "v"(p0), "v"(p1) : "memory");
  put_chunk(S0, a, b);
  put_chunk(S1, c, d);
}
__device__ inline void load_fast3(double (*S0)[33], const double* p0, double (*S1)[33], const double* p1,
                                  double (*S2)[33], const double* p2) {
  v2d a, b, c, d, e, f;
  asm volatile("global_load_dwordx4 %0, %6, off sc1\n\tglobal_load_dwordx4 %1, %6, off offset:16 sc1\n\t"
               "global_load_dwordx4 %2, %7, off sc1\n\tglobal_load_dwordx4 %3, %7, off offset:16 sc1\n\t"
               "global_load_dwordx4 %4, %8, off sc1\n\tglobal_load_dwordx4 %5, %8, off offset:16 sc1\n\t"
               "s_waitcnt vmcnt(0)"
               : "=&v"(a), "=&v"(b), "=&v"(c), "=&v"(d), "=&v"(e), "=&v"(f) : "v"(p0), "v"(p1), "v"(p2)
               : "memory");
  put_chunk(S0, a, b);
  put_chunk(S1, c, d);
  put_chunk(S2, e, f);
}
__device__ inline void store_fast(const double (*S)[33], double* M, int64_t ld, int ti, int tj) {
  const int t = threadIdx.x;
  const int c = t >> 3, r = 4 * (t & 7);
  double* p = M + (int64_t)(tj * TB + c) * ld + ti * TB + r;
  v2d a, b;
  a.x = S[r][c]; a.y = S[r + 1][c];
  b.x = S[r + 2][c]; b.y = S[r + 3][c];
  asm volatile("global_store_dwordx4 %0, %1, off sc1" :: "v"(p), "v"(a) : "memory");
  asm volatile("global_store_dwordx4 %0, %1, off offset:16 sc1" :: "v"(p), "v"(b) : "memory");
}
__device__ inline void zero_fast(double* M, int64_t ld, int ti, int tj) {
  double* p = const_cast<double*>(tile_chunk(M, ld, ti, tj));
  v2d z = {0.0, 0.0};
  asm volatile("global_store_dwordx4 %0, %1, off sc1" :: "v"(p), "v"(z) : "memory");
  asm volatile("global_store_dwordx4 %0, %1, off offset:16 sc1" :: "v"(p), "v"(z) : "memory");
}

// Tile (ti, tj) of an n x n matrix into LDS; entries outside the matrix are
// the identity (diagonal tiles) or zero.
template <typename T>
__device__ inline void load_tile(T (*S)[33], const T* M, int64_t ld, int n, int ti, int tj) {
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    const int r = e % TB, c = e / TB;
    const int gr = ti * TB + r, gc = tj * TB + c;
    S[r][c] = (gr < n && gc < n) ? ld_coh(M + gr + (int64_t)gc * ld) : (gr == gc ? s_one(T()) : s_zero(T()));
  }
}
template <typename T>
__device__ inline void store_tile(const T (*S)[33], T* M, int64_t ld, int n, int ti, int tj) {
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    const int r = e % TB, c = e / TB;
    const int gr = ti * TB + r, gc = tj * TB + c;
    if (gr < n && gc < n) st_coh(M + gr + (int64_t)gc * ld, S[r][c]);
  }
}
template <typename T> __device__ inline void zero_tile(T* M, int64_t ld, int n, int ti, int tj) {
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    const int r = e % TB, c = e / TB;
    const int gr = ti * TB + r, gc = tj * TB + c;
    if (gr < n && gc < n) st_coh(M + gr + (int64_t)gc * ld, s_zero(T()));
  }
}

// Tile loads / stores with the fast path when every tile involved allows it.
struct TileSrc {
  const void* M;
  int64_t ld;
  int ti, tj;
};
template <typename T>
__device__ inline void load_tiles(int cnt, T (*S0)[33], TileSrc s0, T (*S1)[33], TileSrc s1, T (*S2)[33], TileSrc s2,
                                  int n) {
  const T* M0 = reinterpret_cast<const T*>(s0.M);
  const T* M1 = reinterpret_cast<const T*>(s1.M);
  const T* M2 = reinterpret_cast<const T*>(s2.M);
  bool fast = tile_fast<T>(M0, s0.ld, n, s0.ti, s0.tj);
  if (cnt > 1) fast = fast && tile_fast<T>(M1, s1.ld, n, s1.ti, s1.tj);
  if (cnt > 2) fast = fast && tile_fast<T>(M2, s2.ld, n, s2.ti, s2.tj);
  if constexpr (!ScalarTraits<T>::is_complex) {
    if (fast) {
      const double* p0 = tile_chunk(M0, s0.ld, s0.ti, s0.tj);
      if (cnt == 1) {
        load_fast1(S0, p0);
      } else if (cnt == 2) {
        load_fast2(S0, p0, S1, tile_chunk(M1, s1.ld, s1.ti, s1.tj));
      } else {
        load_fast3(S0, p0, S1, tile_chunk(M1, s1.ld, s1.ti, s1.tj), S2, tile_chunk(M2, s2.ld, s2.ti, s2.tj));
      }
      return;
    }
  }
  load_tile<T>(S0, M0, s0.ld, n, s0.ti, s0.tj);
  if (cnt > 1) load_tile<T>(S1, M1, s1.ld, n, s1.ti, s1.tj);
  if (cnt > 2) load_tile<T>(S2, M2, s2.ld, n, s2.ti, s2.tj);
}
template <typename T> __device__ inline void put_tile(const T (*S)[33], T* M, int64_t ld, int n, int ti, int tj) {
  if constexpr (!ScalarTraits<T>::is_complex) {
    if (tile_fast<T>(M, ld, n, ti, tj)) {
      store_fast(S, M, ld, ti, tj);
      return;
    }
  }
  store_tile<T>(S, M, ld, n, ti, tj);
}
template <typename T> __device__ inline void clear_tile(T* M, int64_t ld, int n, int ti, int tj) {
  if constexpr (!ScalarTraits<T>::is_complex) {
    if (tile_fast<T>(M, ld, n, ti, tj)) {
      zero_fast(M, ld, ti, tj);
      return;
    }
  }
  zero_tile<T>(M, ld, n, ti, tj);
}

// ---------------------------------------------------------------- the kernel
// Factor one diagonal tile in LDS (S: in G, out R; Xo: out R^{-1}) by wave 0.
// Returns the bad pivot (0 = fine), uniform over the workgroup.
template <typename T>
__device__ inline int factor_diag(T (*S)[33], T (*Xo)[33], T* rowbuf) {
  __shared__ int bad_s;
  if (threadIdx.x == 0) bad_s = 0;
  __syncthreads();
  if (threadIdx.x < 64) {
    // complex tiles: the 2 x 2-blocked in-wave factorisation (the single-pass
    // elimination needs more registers than a wave has)
    int b;
    if constexpr (ScalarTraits<T>::is_complex) b = wave_chol_inv32<T>(S, Xo);
    else b = wave_chol_aug32<T>(S, Xo, rowbuf);
    if (threadIdx.x == 0) bad_s = b;
  }
  __syncthreads();
  return bad_s;
}

template <typename T>
__device__ inline void report_bad(const CholTileArgs& a, int k, int bad) {
  if (threadIdx.x == 0) {
    atomicCAS(a.ctl + 2, 0, k * TB + bad);
    __hip_atomic_store(a.ctl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (a.info != nullptr) atomicCAS(a.info, 0, a.info_offset + k * TB + bad);
  }
}

// The panel workgroup: the diagonal chain.  Iteration k >= 1 takes
// G_{k-1,k} and G_kk after the updates of steps < k-1 and, with X_{k-1,k-1}
// still in LDS from iteration k-1, runs
//   P(k-1, k):    R_{k-1,k} = X_{k-1,k-1}^H G_{k-1,k}      (published at once)
//   U(k-1, k, k): G_kk -= R_{k-1,k}^H R_{k-1,k}            (in LDS)
//   D(k):         R_kk, X_kk = chol_inv(G_kk)
// so a step costs one load of two tiles and two publications, not three
// task hand-offs between workgroups.
template <typename T>
__device__ void chol_panel(const CholTileArgs& a, T (*tA)[33], T (*tB)[33], T (*tC)[33], T (*tX)[33],
                           T* rowbuf) {
  T* G = reinterpret_cast<T*>(a.G);
  T* X = reinterpret_cast<T*>(a.X);
  const int nt = a.nt;
  int* gs = a.gst;
  int* zs = a.zst;
  int* ctl = a.ctl;
  long long* tr = a.trace;
  auto stamp = [&](int k, int slot) {
    if (tr != nullptr && threadIdx.x == 0) tr[k * 8 + slot] = (long long)__builtin_amdgcn_s_memrealtime();
  };
  for (int k = 0; k < nt; ++k) {
    stamp(k, 0);
    if (k == 0) {
      if (!inputs_ready(ctl)) return;
      load_tiles<T>(1, tB, TileSrc{G, a.ldg, 0, 0}, tB, TileSrc{}, tB, TileSrc{}, a.n);
    } else {
      wait_ge(&gs[(k - 1) * nt + k], k - 1, ctl);
      wait_ge(&gs[k * nt + k], k - 1, ctl);
      if (!inputs_ready(ctl)) return;
      stamp(k, 1);
      load_tiles<T>(2, tA, TileSrc{G, a.ldg, k - 1, k}, tB, TileSrc{G, a.ldg, k, k}, tB, TileSrc{}, a.n);
      __syncthreads();
      stamp(k, 2);
      mm32<T>(tC, tX, true, tA, false, 1.0, false);     // R_{k-1,k}
      put_tile<T>(tC, G, a.ldg, a.n, k - 1, k);
      clear_tile<T>(G, a.ldg, a.n, k, k - 1);
      publish(&gs[(k - 1) * nt + k], k);
      stamp(k, 3);
      mm32<T>(tB, tC, true, tC, false, -1.0, true);    // G_kk after step k-1
    }
    __syncthreads();
    stamp(k, 4);
    const int bad = factor_diag<T>(tB, tX, rowbuf);    // tX <- X_kk (X_{k-1,k-1} is no longer needed)
    if (bad != 0) {
      report_bad<T>(a, k, bad);
      return;
    }
    stamp(k, 5);
    put_tile<T>(tB, G, a.ldg, a.n, k, k);
    put_tile<T>(tX, X, a.ldx, a.n, k, k);
    publish(&zs[k * nt + k], 1, &gs[k * nt + k], k + 1);
    stamp(k, 6);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) chol_tile_kernel(CholTileArgs a) {
  __shared__ T tA[TB][33], tB[TB][33], tC[TB][33], tX[TB][33];
  __shared__ __align__(16) T rowbuf[2 * TB];
  __shared__ int task_s;
  if (blockIdx.x == 0) {
    chol_panel<T>(a, tA, tB, tC, tX, rowbuf);
    return;
  }
  T* G = reinterpret_cast<T*>(a.G);
  T* X = reinterpret_cast<T*>(a.X);
  const int nt = a.nt;
  int* ctl = a.ctl;
  int iter = 0;
  for (;;) {
    if (threadIdx.x == 0) task_s = atomicAdd(ctl, 1);
    __syncthreads();
    const int t = task_s;
    if (t >= a.ntasks) break;
    while (a.iter_off[iter + 1] <= t) ++iter;  // tasks are taken in increasing order
    const CholTask tk = chol_task_in_iter(nt, iter, t - a.iter_off[iter], 1);
    const int k = tk.k, i = tk.i, j = tk.j;
    int* gs = a.gst;
    int* zs = a.zst;
    switch (tk.type) {
      case CT_P: {  // R_kj = X_kk^H G_kj; the mirrored lower tile of G is zeroed
        wait_ge(&gs[k * nt + k], k + 1, ctl);
        wait_ge(&gs[k * nt + j], k, ctl);
        if (!inputs_ready(ctl)) break;
        load_tiles<T>(2, tA, TileSrc{X, a.ldx, k, k}, tB, TileSrc{G, a.ldg, k, j}, tB, TileSrc{}, a.n);
        __syncthreads();
        mm32<T>(tC, tA, true, tB, false, 1.0, false);
        put_tile<T>(tC, G, a.ldg, a.n, k, j);
        clear_tile<T>(G, a.ldg, a.n, j, k);
        publish(&gs[k * nt + j], k + 1);
        break;
      }
      case CT_PQ: {  // X_ik = Z_ik X_kk (i < k = j); the mirrored lower tile of X is zeroed
        const int kk = j;
        wait_ge(&gs[kk * nt + kk], kk + 1, ctl);
        wait_ge(&zs[i * nt + kk], kk - i, ctl);
        if (!inputs_ready(ctl)) break;
        load_tiles<T>(2, tA, TileSrc{X, a.ldx, i, kk}, tB, TileSrc{X, a.ldx, kk, kk}, tB, TileSrc{}, a.n);
        __syncthreads();
        mm32<T>(tC, tA, false, tB, false, 1.0, false);
        put_tile<T>(tC, X, a.ldx, a.n, i, kk);
        clear_tile<T>(X, a.ldx, a.n, kk, i);
        publish(&zs[i * nt + kk], kk - i + 1);
        break;
      }
      case CT_U: {  // G_ij -= R_ki^H R_kj
        wait_ge(&gs[k * nt + i], k + 1, ctl);
        wait_ge(&gs[k * nt + j], k + 1, ctl);
        wait_ge(&gs[i * nt + j], k, ctl);
        if (!inputs_ready(ctl)) break;
        load_tiles<T>(3, tA, TileSrc{G, a.ldg, k, i}, tB, TileSrc{G, a.ldg, k, j}, tC, TileSrc{G, a.ldg, i, j}, a.n);
        __syncthreads();
        mm32<T>(tC, tA, true, tB, false, -1.0, true);
        put_tile<T>(tC, G, a.ldg, a.n, i, j);
        publish(&gs[i * nt + j], k + 1);
        break;
      }
      case CT_UQ: {  // Z_il -= X_ik R_kl (i <= k < l = j); the first update (k == i) overwrites
        const int l = j;
        wait_ge(&zs[i * nt + k], k - i + 1, ctl);
        wait_ge(&gs[k * nt + l], k + 1, ctl);
        if (k > i) wait_ge(&zs[i * nt + l], k - i, ctl);
        if (!inputs_ready(ctl)) break;
        load_tiles<T>(k > i ? 3 : 2, tA, TileSrc{X, a.ldx, i, k}, tB, TileSrc{G, a.ldg, k, l}, tC,
                      TileSrc{X, a.ldx, i, l}, a.n);
        __syncthreads();
        mm32<T>(tC, tA, false, tB, false, -1.0, k > i);
        put_tile<T>(tC, X, a.ldx, a.n, i, l);
        publish(&zs[i * nt + l], k - i + 1);
        break;
      }
      default:
        break;
    }
    __syncthreads();  // task_s and the tiles are reused by the next task
  }
}

template <typename T> __global__ void chol_zero_ints_kernel(int* p, int n) {
  for (int e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) p[e] = 0;
}

}  // namespace

size_t chol_tile_flag_ints(int n) {
  const int nt = std::max(1, cdiv(std::min(n, kCholTileMaxN), TB));
  return (size_t)2 * nt * nt + 64;
}

// One persistent launch for n <= kCholTileMaxN.  flags: chol_tile_flag_ints(n)
// ints of device memory (cleared here by a kernel, not a memset, so the call
// can be captured).  info (device, optional): first non-positive pivot + offset.
template <typename T>
int chol_tile_factor(T* G, int64_t ldg, int n, T* X, int64_t ldx, int* flags, int* info, int info_offset,
                     hipStream_t s, int grid, long long* trace) {
  if (n <= 0) return 0;
  if (n > kCholTileMaxN) return -3;
  CholTileArgs a{};
  a.G = G; a.ldg = ldg; a.X = X; a.ldx = ldx;
  a.n = n;
  a.nt = cdiv(n, TB);
  int off = 0;
  for (int k = 0; k < a.nt; ++k) {
    a.iter_off[k] = off;
    off += chol_iter_total(a.nt, k, 1);
  }
  a.iter_off[a.nt] = off;
  a.ntasks = off;
  a.gst = flags;
  a.zst = flags + a.nt * a.nt;
  a.ctl = flags + 2 * a.nt * a.nt;
  a.info = info;
  a.info_offset = info_offset;
  a.trace = trace;
  const int nflags = 2 * a.nt * a.nt + 64;
  hipLaunchKernelGGL((chol_zero_ints_kernel<T>), dim3(cdiv(nflags, 256)), dim3(256), 0, s, flags, nflags);
  // the panel workgroup plus at least one for the other tasks
  if (grid <= 0) grid = 256;
  grid = std::max(2, std::min(grid, a.ntasks + 1));
  hipLaunchKernelGGL((chol_tile_kernel<T>), dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

template int chol_tile_factor<double>(double*, int64_t, int, double*, int64_t, int*, int*, int, hipStream_t, int,
                                      long long*);
template int chol_tile_factor<zd>(zd*, int64_t, int, zd*, int64_t, int*, int*, int, hipStream_t, int, long long*);

// Host-side check of the task order (test hook): every task once, every
// dependency earlier in the order, and the other workgroups' queue equal to
// the order without the panel workgroup's tasks (D(k), P(k, k+1),
// U(k, k+1, k+1)), which that workgroup runs in the same relative order.
// Returns the task count, or -1.
int chol_tile_order_check(int nt) {
  if (nt <= 0 || nt > kCholTileMaxNt) return -1;
  {
    std::vector<CholTask> full, queue;
    for (int k = 0; k < nt; ++k) {
      for (int u = 0; u < chol_iter_total(nt, k); ++u) {
        const CholTask t = chol_task_in_iter(nt, k, u);
        const bool panel = t.type == CT_D || (t.type == CT_P && t.j == t.k + 1) ||
                           (t.type == CT_U && t.i == t.k + 1 && t.j == t.k + 1);
        if (!panel) full.push_back(t);
      }
      for (int u = 0; u < chol_iter_total(nt, k, 1); ++u) queue.push_back(chol_task_in_iter(nt, k, u, 1));
    }
    if (full.size() != queue.size()) return -1;
    for (size_t q = 0; q < full.size(); ++q)
      if (full[q].type != queue[q].type || full[q].k != queue[q].k || full[q].i != queue[q].i ||
          full[q].j != queue[q].j)
        return -1;
  }
  std::vector<int> gst(nt * nt, 0), zst(nt * nt, 0);
  int total = 0;
  for (int k = 0; k < nt; ++k) {
    const int cnt = chol_iter_total(nt, k);
    for (int u = 0; u < cnt; ++u) {
      const CholTask t = chol_task_in_iter(nt, k, u);
      const int kk = t.k, i = t.i, j = t.j;
      auto G = [&](int r, int c) -> int& { return gst[r * nt + c]; };
      auto Z = [&](int r, int c) -> int& { return zst[r * nt + c]; };
      switch (t.type) {
        case CT_D:
          if (G(kk, kk) != kk) return -1;
          G(kk, kk) = kk + 1; Z(kk, kk) = 1; break;
        case CT_P:
          if (G(kk, kk) != kk + 1 || G(kk, j) != kk || j <= kk) return -1;
          G(kk, j) = kk + 1; break;
        case CT_PQ:
          if (G(j, j) != j + 1 || Z(i, j) != j - i || i >= j) return -1;
          Z(i, j) = j - i + 1; break;
        case CT_U:
          if (G(kk, i) != kk + 1 || G(kk, j) != kk + 1 || G(i, j) != kk || !(kk < i && i <= j)) return -1;
          G(i, j) = kk + 1; break;
        case CT_UQ:
          if (Z(i, kk) != kk - i + 1 || G(kk, j) != kk + 1 || Z(i, j) != kk - i || !(i <= kk && kk < j)) return -1;
          Z(i, j) = kk - i + 1; break;
      }
      ++total;
    }
  }
  for (int r = 0; r < nt; ++r)
    for (int c = r; c < nt; ++c)
      if (gst[r * nt + c] != r + 1 || zst[r * nt + c] != c - r + 1) return -1;
  return total;
}

}  // namespace mak
