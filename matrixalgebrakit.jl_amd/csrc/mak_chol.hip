// Cholesky factorisation with inverse, G = R^H R -> R (upper) and X = R^{-1},
// for FP64 / ComplexF64 Hermitian positive definite matrices.
//
// Small and medium n (n <= kCholTileMaxN): ONE persistent launch.  The
// matrix is cut into 32 x 32 tiles and [G | I] is eliminated right-looking,
// one PAIR of tile columns (64 columns) per step.  Pair K has tiles
// c0 = 2K, c1 = 2K + 1; its 2 x 2 tile diagonal block is factored by the
// panel workgroup entirely in LDS:
//   D(c0); R01 = X00^H G01; G11 -= R01^H R01; D(c1); X01 = -X00 R01 X11
// (D: one-wave Cholesky with inverse of a 32 x 32 tile).  The other
// workgroups run 32 x 32 tasks whose products contract over the pair (K = 64
// on FP64 MFMA):
//   P2(K, j)      [R_c0j; R_c1j] = X_KK^H [G_c0j; G_c1j]        j >= 2K+4
//   U2(K, i, j)   G_ij -= R_c0i^H R_c0j + R_c1i^H R_c1j        2K+2 <= i <= j, j >= 2K+4
//   UQ2(K, j, l)  Z_jl -= X_jc0 R_c0l + X_jc1 R_c1l            j < 2K+2 <= l
//   PQ2(K, j)     [X_jc0 X_jc1] = [Z_jc0 Z_jc1] X_KK           j < 2K
// (X R = I column block by column block: X_{j,L} = -(sum_K X_{j,K} R_{K,L}) X_LL.)
// The panel itself forms P2(K, j) for the next pair's columns j = 2K+2, 2K+3
// and U2(K) of the next pair's diagonal block, which it keeps in LDS: the
// chain of pair steps runs inside one workgroup, which only waits for the
// workers' updates of earlier pairs (long done by then).  Workgroups take tasks from an atomic
// counter in a fixed topological order (the next pair's inputs first, the
// bulk last) and wait for the per-tile version counters the task needs:
//   gst[i][j]  G/R tile (i <= j): pair steps applied; i/2 + 1 = R final
//   zst[j][l]  X tile (j <= l): UQ2 updates applied; l/2 - j/2 + 1 = final
// Every dependency of a task precedes it in the order, and a workgroup only
// waits after it has taken its task, so the waits always end; each wait is
// still bounded (kSpinLimit polls), and a timeout or a non-positive pivot sets
// an abort flag that every later wait and task honours, so the grid drains.
// An odd tile count is padded with one virtual tile: its loads are the
// identity (diagonal) or zero, its stores are dropped.
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
constexpr int kYRows = 128;            // rows of B / Y per right-product task (real data)
constexpr int kYRowsZ = 64;            // complex data (the same accumulator count)
template <typename T> constexpr int kYRowsOf = ScalarTraits<T>::is_complex ? kYRowsZ : kYRows;
constexpr int kCholGridY = 256;        // default workgroups with a right product
constexpr long kSpinLimit = 1L << 22;  // polls per wait before the abort flag is raised

// ---------------------------------------------------------------- task order
enum CholTaskType { CT_P2 = 1, CT_U2 = 2, CT_UQ2 = 3, CT_PQ2 = 4, CT_Y = 5 };

struct CholTask {
  int type, k, i, j;  // k: the pair step
};

// Tasks of iteration K (T tiles, T even), by segment, in queue order:
//   A  P2(K, j)        j = 2K+4 .. T-1
//   B  U2(K, i, j)     i = 2K+2, 2K+3, j = 2K+4 .. T-1  (the next pair's rows)
//   C  PQ2(K, j)       j = 0 .. 2K-1
//   E  UQ2(K, j, l)    j = 0 .. 2K+1, l = 2K+2 .. T-1
//   F  U2(K, i, j)     2K+4 <= i <= j <= T-1          (the bulk)
//   H  Y(K, r)         r = 0 .. ny-1                  (the right product, below)
// The optional right product Y = B X (B m x n, real data) is computed in the
// same launch: task Y(K, r) forms rows [64 r, 64 r + 64) of the pair's 64
// columns of Y, Y[:, pair K] = B[:, 0 : 64K + 64] X[0 : 64K + 64, pair K],
// once that column block of X is final (the panel's pair K and the 2K tasks
// PQ2(K, j); a per-pair counter records them).  X fills column block by
// column block as the elimination proceeds, so the product runs in the time
// the workgroups would otherwise spend waiting for the chain.
constexpr int kCholSegs = 6;
__host__ __device__ inline void chol_iter_counts(int T, int K, int ny, int c[kCholSegs], int ysplit = 1 << 30) {
  const int r0 = 2 * K + 2;
  const int rem = T - r0 > 0 ? T - r0 : 0;
  const int b = rem >= 2 ? rem - 2 : 0;
  c[0] = b;       // P2 past the panel's look-ahead columns
  c[1] = 2 * b;   // U2 of the next pair's rows, right of its diagonal block
  c[2] = 2 * K;
  c[3] = r0 * rem;
  c[4] = b * (b + 1) / 2;
  c[5] = K >= ysplit ? 2 * ny : ny;
}

__host__ __device__ inline int chol_iter_total(int T, int K, int ny, int ysplit = 1 << 30) {
  int c[kCholSegs];
  chol_iter_counts(T, K, ny, c, ysplit);
  return c[0] + c[1] + c[2] + c[3] + c[4] + c[5];
}

// u-th task of iteration K (0 <= u < chol_iter_total(T, K, ny, ysplit)).
// Right-product task j: row block j % ny, K half j / ny (pairs K >= ysplit
// split their contraction in two halves).
__host__ __device__ inline CholTask chol_task_in_iter(int T, int K, int ny, int u, int ysplit = 1 << 30) {
  int c[kCholSegs];
  chol_iter_counts(T, K, ny, c, ysplit);
  const int r0 = 2 * K + 2, rem = T - r0;
  CholTask t{CT_P2, K, 0, 0};
  int q = 0;
  while (q < kCholSegs - 1 && u >= c[q]) u -= c[q++];
  switch (q) {
    case 0: t.type = CT_P2; t.j = r0 + 2 + u; break;
    case 1:
      t.type = CT_U2;
      t.i = u < c[0] ? r0 : r0 + 1;
      t.j = r0 + 2 + (u < c[0] ? u : u - c[0]);
      break;
    case 2: t.type = CT_PQ2; t.j = u; break;
    case 3: t.type = CT_UQ2; t.i = u / rem; t.j = r0 + u % rem; break;
    case 5: t.type = CT_Y; t.j = u; break;
    default: {  // pairs r0+2 <= i <= j <= T-1, row by row
      int i = r0 + 2, len = T - i;
      while (u >= len) { u -= len; ++i; --len; }
      t.type = CT_U2; t.i = i; t.j = i + u;
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
  int n, nt;      // nt: tiles, rounded up to even
  int ntasks;
  int* gst;       // nt * nt
  int* zst;       // nt * nt
  int* ctl;       // [0] task counter, [1] abort flag, [2] first non-positive pivot (1-based) or 0
  int* info;      // optional: receives ctl[2] + info_offset on failure (when *info == 0)
  int info_offset;
  long long* trace;  // optional (tests): wall-clock stamps of the panel's iterations, 8 per iteration:
                     // 3, 4, 5 before D(c0), before D(c1), after X01; 6 at the end; real
                     // data: 1, 2, 0 after the look-ahead's waits, loads and R products
                     // (complex: 0, 1, 2 at the start); 7 the end of the pair's last
                     // right-product task.  From entry 256 the start and end of every
                     // right-product task.
  // the right product Y = B X (ny = 0: none), elements of the factor's type
  const void* B;
  int64_t ldb;
  void* Y;
  int64_t ldy;
  void* Xc;       // final tiles of X, written once each (ldxc: a multiple of 16)
  int64_t ldxc;
  int m;          // rows of B and Y
  int ny;         // Y row blocks per pair step: ceil(m / kYRowsOf<T>)
  int ysplit;     // pairs K >= ysplit run two tasks per row block, one per half
                  // of the contraction, adding atomically into a zeroed Y
                  // (two addends on zero: the sum is independent of their order)
  int* xdone;     // per pair: tasks that have published final tiles of its X column block
  int iter_off[kCholTileMaxNt / 2 + 1];  // first queue task of every iteration
};

// ---------------------------------------------------------------- sync helpers
__device__ inline int ld_flag(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Memory protocol between workgroups.  Tiles that tasks exchange are moved
// with agent-scope relaxed atomic loads and stores (sc1: coherent across the
// XCDs' L2 caches location by location), so no whole-cache writeback or
// invalidate is needed:
//   producer: every wave waits for its stores to complete (vmcnt(0)), then a
//             workgroup barrier, then thread 0 stores the version counter;
//   consumer: wave 0 polls the counters, then a workgroup barrier, then the
//             tile loads.
// A load issued after the counter was seen reads the completed stores.
__device__ inline bool inputs_ready(int* ctl) {
  __shared__ int abort_s;
  if (threadIdx.x == 0) abort_s = ld_flag(ctl + 1);
  __syncthreads();
  return abort_s == 0;
}

// Thread 0 polls a task's counters in turn until *p[q] >= v[q] for every q
// (or the abort flag is set, or the spin limit is reached, which raises it),
// then a workgroup barrier, the consumer's side of the protocol above; false
// when the launch is aborting.  (Polling all counters at once from the lanes
// of a wave measured slower: it multiplies the polling traffic of the
// waiting workgroups.)
template <int N>
__device__ inline bool wait_all(int* const (&p)[N], const int (&v)[N], int* ctl) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) {
      long spins = 0;
      while (ld_flag(p[q]) < v[q]) {
        if (ld_flag(ctl + 1) != 0) break;
        if (++spins > kSpinLimit) {
          __hip_atomic_store(ctl + 1, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
  }
  return inputs_ready(ctl);
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

__device__ inline void load_fast4(double (*S0)[33], const double* p0, double (*S1)[33], const double* p1,
                                  double (*S2)[33], const double* p2, double (*S3)[33], const double* p3) {
  v2d a, b, c, d, e, f, g, h;
  asm volatile("global_load_dwordx4 %0, %8, off sc1\n\tglobal_load_dwordx4 %1, %8, off offset:16 sc1\n\t"
               "global_load_dwordx4 %2, %9, off sc1\n\tglobal_load_dwordx4 %3, %9, off offset:16 sc1\n\t"
               "global_load_dwordx4 %4, %10, off sc1\n\tglobal_load_dwordx4 %5, %10, off offset:16 sc1\n\t"
               "global_load_dwordx4 %6, %11, off sc1\n\tglobal_load_dwordx4 %7, %11, off offset:16 sc1\n\t"
               "s_waitcnt vmcnt(0)"
               : "=&v"(a), "=&v"(b), "=&v"(c), "=&v"(d), "=&v"(e), "=&v"(f), "=&v"(g), "=&v"(h)
               : "v"(p0), "v"(p1), "v"(p2), "v"(p3)
               : "memory");
  put_chunk(S0, a, b);
  put_chunk(S1, c, d);
  put_chunk(S2, e, f);
  put_chunk(S3, g, h);
}
__device__ inline void load_fast5(double (*S0)[33], const double* p0, double (*S1)[33], const double* p1,
                                  double (*S2)[33], const double* p2, double (*S3)[33], const double* p3,
                                  double (*S4)[33], const double* p4) {
  v2d a, b, c, d, e, f, g, h, x, y;
  asm volatile("global_load_dwordx4 %0, %10, off sc1\n\tglobal_load_dwordx4 %1, %10, off offset:16 sc1\n\t"
               "global_load_dwordx4 %2, %11, off sc1\n\tglobal_load_dwordx4 %3, %11, off offset:16 sc1\n\t"
               "global_load_dwordx4 %4, %12, off sc1\n\tglobal_load_dwordx4 %5, %12, off offset:16 sc1\n\t"
               "global_load_dwordx4 %6, %13, off sc1\n\tglobal_load_dwordx4 %7, %13, off offset:16 sc1\n\t"
               "global_load_dwordx4 %8, %14, off sc1\n\tglobal_load_dwordx4 %9, %14, off offset:16 sc1\n\t"
               "s_waitcnt vmcnt(0)"
               : "=&v"(a), "=&v"(b), "=&v"(c), "=&v"(d), "=&v"(e), "=&v"(f), "=&v"(g), "=&v"(h), "=&v"(x),
                 "=&v"(y)
               : "v"(p0), "v"(p1), "v"(p2), "v"(p3), "v"(p4)
               : "memory");
  put_chunk(S0, a, b);
  put_chunk(S1, c, d);
  put_chunk(S2, e, f);
  put_chunk(S3, g, h);
  put_chunk(S4, x, y);
}

// CNT (1 .. 5) tiles into LDS (S0 .. S4 from s0 .. s4, the first CNT used):
// one asm block of coherent 16-byte loads when every tile allows the fast
// path, else element loads.  CNT is a template parameter so that nothing is
// indexed at run time (no private-memory arrays).
template <int CNT, typename T>
__device__ inline void load_n(T (*S0)[33], TileSrc s0, T (*S1)[33], TileSrc s1, T (*S2)[33], TileSrc s2,
                              T (*S3)[33], TileSrc s3, T (*S4)[33], TileSrc s4, int n) {
  auto fast_of = [&](const TileSrc& q) {
    return tile_fast<T>(reinterpret_cast<const T*>(q.M), q.ld, n, q.ti, q.tj);
  };
  bool fast = fast_of(s0);
  if constexpr (CNT > 1) fast = fast && fast_of(s1);
  if constexpr (CNT > 2) fast = fast && fast_of(s2);
  if constexpr (CNT > 3) fast = fast && fast_of(s3);
  if constexpr (CNT > 4) fast = fast && fast_of(s4);
  if constexpr (!ScalarTraits<T>::is_complex) {
    if (fast) {
      auto chunk = [&](const TileSrc& q) {
        return tile_chunk(reinterpret_cast<const double*>(q.M), q.ld, q.ti, q.tj);
      };
      if constexpr (CNT == 1) load_fast1(S0, chunk(s0));
      else if constexpr (CNT == 2) load_fast2(S0, chunk(s0), S1, chunk(s1));
      else if constexpr (CNT == 3) load_fast3(S0, chunk(s0), S1, chunk(s1), S2, chunk(s2));
      else if constexpr (CNT == 4) load_fast4(S0, chunk(s0), S1, chunk(s1), S2, chunk(s2), S3, chunk(s3));
      else load_fast5(S0, chunk(s0), S1, chunk(s1), S2, chunk(s2), S3, chunk(s3), S4, chunk(s4));
      return;
    }
  }
  auto slow = [&](T (*S)[33], const TileSrc& q) {
    load_tile<T>(S, reinterpret_cast<const T*>(q.M), q.ld, n, q.ti, q.tj);
  };
  slow(S0, s0);
  if constexpr (CNT > 1) slow(S1, s1);
  if constexpr (CNT > 2) slow(S2, s2);
  if constexpr (CNT > 3) slow(S3, s3);
  if constexpr (CNT > 4) slow(S4, s4);
}

// Seven tiles into LDS in one batch of coherent loads (the panel's
// look-ahead); element loads when a tile does not allow the fast path.
__device__ inline void load_fast7(double (*S0)[33], const double* p0, double (*S1)[33], const double* p1,
                                  double (*S2)[33], const double* p2, double (*S3)[33], const double* p3,
                                  double (*S4)[33], const double* p4, double (*S5)[33], const double* p5,
                                  double (*S6)[33], const double* p6) {
  v2d a0, b0, a1, b1, a2, b2, a3, b3, a4, b4, a5, b5, a6, b6;
  asm volatile("global_load_dwordx4 %0, %14, off sc1\n\tglobal_load_dwordx4 %1, %14, off offset:16 sc1\n\t"
               "global_load_dwordx4 %2, %15, off sc1\n\tglobal_load_dwordx4 %3, %15, off offset:16 sc1\n\t"
               "global_load_dwordx4 %4, %16, off sc1\n\tglobal_load_dwordx4 %5, %16, off offset:16 sc1\n\t"
               "global_load_dwordx4 %6, %17, off sc1\n\tglobal_load_dwordx4 %7, %17, off offset:16 sc1\n\t"
               "global_load_dwordx4 %8, %18, off sc1\n\tglobal_load_dwordx4 %9, %18, off offset:16 sc1\n\t"
               "global_load_dwordx4 %10, %19, off sc1\n\tglobal_load_dwordx4 %11, %19, off offset:16 sc1\n\t"
               "global_load_dwordx4 %12, %20, off sc1\n\tglobal_load_dwordx4 %13, %20, off offset:16 sc1\n\t"
               "s_waitcnt vmcnt(0)"
               : "=&v"(a0), "=&v"(b0), "=&v"(a1), "=&v"(b1), "=&v"(a2), "=&v"(b2), "=&v"(a3), "=&v"(b3),
                 "=&v"(a4), "=&v"(b4), "=&v"(a5), "=&v"(b5), "=&v"(a6), "=&v"(b6)
               : "v"(p0), "v"(p1), "v"(p2), "v"(p3), "v"(p4), "v"(p5), "v"(p6)
               : "memory");
  put_chunk(S0, a0, b0);
  put_chunk(S1, a1, b1);
  put_chunk(S2, a2, b2);
  put_chunk(S3, a3, b3);
  put_chunk(S4, a4, b4);
  put_chunk(S5, a5, b5);
  put_chunk(S6, a6, b6);
}
template <typename T>
__device__ inline void load_7(T (*S0)[33], TileSrc s0, T (*S1)[33], TileSrc s1, T (*S2)[33], TileSrc s2,
                              T (*S3)[33], TileSrc s3, T (*S4)[33], TileSrc s4, T (*S5)[33], TileSrc s5,
                              T (*S6)[33], TileSrc s6, int n) {
  auto fast_of = [&](const TileSrc& q) {
    return tile_fast<T>(reinterpret_cast<const T*>(q.M), q.ld, n, q.ti, q.tj);
  };
  if constexpr (!ScalarTraits<T>::is_complex) {
    if (fast_of(s0) && fast_of(s1) && fast_of(s2) && fast_of(s3) && fast_of(s4) && fast_of(s5) && fast_of(s6)) {
      auto chunk = [&](const TileSrc& q) {
        return tile_chunk(reinterpret_cast<const double*>(q.M), q.ld, q.ti, q.tj);
      };
      load_fast7(S0, chunk(s0), S1, chunk(s1), S2, chunk(s2), S3, chunk(s3), S4, chunk(s4), S5, chunk(s5), S6,
                 chunk(s6));
      __syncthreads();
      return;
    }
  }
  auto slow = [&](T (*S)[33], const TileSrc& q) {
    load_tile<T>(S, reinterpret_cast<const T*>(q.M), q.ld, n, q.ti, q.tj);
  };
  slow(S0, s0);
  slow(S1, s1);
  slow(S2, s2);
  slow(S3, s3);
  slow(S4, s4);
  slow(S5, s5);
  slow(S6, s6);
  __syncthreads();
}

// LDS tiles of the panel workgroup: ten for real data, seven for complex
// (ten 16-byte tiles would not fit the 160 KB of LDS).
template <typename T> constexpr int kPanelTiles = ScalarTraits<T>::is_complex ? 7 : 10;

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
__device__ inline void report_bad(const CholTileArgs& a, int tile, int bad) {
  if (threadIdx.x == 0) {
    atomicCAS(a.ctl + 2, 0, tile * TB + bad);
    __hip_atomic_store(a.ctl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (a.info != nullptr) atomicCAS(a.info, 0, a.info_offset + tile * TB + bad);
  }
}

// Several version counters at once (after the stores of the task completed).
__device__ inline void publish_n(int cnt, int* const* p, const int* v) {
  stores_done();
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < cnt; ++q) __hip_atomic_store(p[q], v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Y(K, r) of the right product: rows [BM r, BM r + BM) x the pair's columns
// [64 K, 64 K + 64) of Y = B X, over k < 64 K + 64 (X is upper triangular);
// BM = 128 rows for real data, 64 for complex.  X is read from Xc, the copy
// of its final tiles (below): plain loads, so a column block is shared
// through each XCD's L2 by the pair's tasks.  Four waves, each a BM/2 x 32
// quarter: eight FP64 MFMA accumulators (real: six LDS reads feed eight
// MFMAs; complex: four real and imaginary accumulator pairs, four MFMAs per
// complex multiply-add); the slabs of B and X for the next two 16-deep steps
// are in flight (in registers) while the current one is used.  LDS strides
// keep every read at two passes (real: B's slab rows 144 doubles apart, X's
// slab stored transposed with rows 36 doubles apart; complex: 80 and 17
// elements, 16-byte lanes).
template <typename T> constexpr int kYLdA = kYRowsOf<T> + 16;
template <typename T> constexpr int kYLdX = ScalarTraits<T>::is_complex ? 17 : 16 + 20;
template <typename T> constexpr int kYLdsElems = 16 * kYLdA<T> + 64 * kYLdX<T>;
// half: 0 = the whole contraction, stored; 1 / 2 = its first / second half
// (split at a multiple of 16), added atomically.
template <typename T>
__device__ inline void y_task(const CholTileArgs& a, int K, int r, int half, T* lds) {
  constexpr bool CX = ScalarTraits<T>::is_complex;
  constexpr int BM = kYRowsOf<T>, BN = 64, BK = 16, LDA = kYLdA<T>, LDX = kYLdX<T>;
  constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
  constexpr int WM = BM / 2, MI = WM / 16;  // rows per wave, 16-row blocks per wave
  T (*As)[LDA] = reinterpret_cast<T (*)[LDA]>(lds);
  T (*Xs)[LDX] = reinterpret_cast<T (*)[LDX]>(lds + BK * LDA);  // Xs[j][k]
  const T* B = static_cast<const T*>(a.B);
  const T* Xc = static_cast<const T*>(a.Xc);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int m0 = r * BM, n0 = K * 64;
  const int kall = min(a.n, n0 + 64);
  const int kmid = (kall / 2 + 15) / 16 * 16;
  const int kbeg = half == 2 ? kmid : 0;
  const int kend = half == 1 ? kmid : kall;
  auto fetch = [&](int k0, T (&pa)[NA], T (&px)[NB]) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int gi = m0 + (tid & (BM - 1)), gk = k0 + (tid / BM) + (256 / BM) * u;
      pa[u] = (gi < a.m && gk < kend) ? B[gi + (int64_t)gk * a.ldb] : s_zero(T());
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int gk = k0 + (tid & 15), gj = n0 + (tid >> 4) + 16 * u;
      px[u] = (gj < a.n && gk < kend) ? Xc[gk + (int64_t)gj * a.ldxc] : s_zero(T());
    }
  };
  v4d acc[MI][2], acci[CX ? MI : 1][2];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
      if constexpr (CX) acci[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    }
  // one 16-deep step: the slabs in (pa, px) to LDS, the loads of step
  // k0 + 2 BK into the freed registers, the MFMAs of this step
  auto step = [&](int k0, T (&pa)[NA], T (&px)[NB]) {
#pragma unroll
    for (int u = 0; u < NA; ++u) As[(tid / BM) + (256 / BM) * u][tid & (BM - 1)] = pa[u];
#pragma unroll
    for (int u = 0; u < NB; ++u) Xs[(tid >> 4) + 16 * u][tid & 15] = px[u];
    __syncthreads();
    if (k0 + 2 * BK < kend) fetch(k0 + 2 * BK, pa, px);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[MI], bf[2];
#pragma unroll
      for (int i = 0; i < MI; ++i) af[i] = As[kk + (lane >> 4)][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = Xs[wn * 32 + j * 16 + (lane & 15)][kk + (lane >> 4)];
      // operands swapped: lane l holds rows l % 16 of the block
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if constexpr (CX) {
            acc[i][j] = mfma_f64(bf[j].x, af[i].x, acc[i][j]);
            acc[i][j] = mfma_f64(-bf[j].y, af[i].y, acc[i][j]);
            acci[i][j] = mfma_f64(bf[j].y, af[i].x, acci[i][j]);
            acci[i][j] = mfma_f64(bf[j].x, af[i].y, acci[i][j]);
          } else {
            acc[i][j] = mfma_f64(bf[j], af[i], acc[i][j]);
          }
        }
    }
    __syncthreads();
  };
  T pa0[NA], px0[NB], pa1[NA], px1[NB];
  if (kbeg < kend) fetch(kbeg, pa0, px0);
  if (kbeg + BK < kend) fetch(kbeg + BK, pa1, px1);
  for (int k0 = kbeg; k0 < kend; k0 += 2 * BK) {
    step(k0, pa0, px0);
    if (k0 + BK < kend) step(k0 + BK, pa1, px1);
  }
  T* const Yt = static_cast<T*>(a.Y);
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm * WM + i * 16 + (lane & 15);
        const int col = n0 + wn * 32 + j * 16 + mfma_f64_row(lane, q);
        if (row < a.m && col < a.n) {
          T* y = Yt + row + (int64_t)col * a.ldy;
          if constexpr (CX) {
            if (half == 0) {
              *y = mkz(acc[i][j][q], acci[i][j][q]);
            } else {
              atomicAdd(reinterpret_cast<double*>(y), acc[i][j][q]);
              atomicAdd(reinterpret_cast<double*>(y) + 1, acci[i][j][q]);
            }
          } else {
            if (half == 0) *y = acc[i][j][q];
            else atomicAdd(y, acc[i][j][q]);
          }
        }
      }
}

// The panel workgroup: the 2 x 2 tile diagonal block of every pair, in LDS.
// Iteration K runs D(c0), R01 = X00^H G01, G11 -= R01^H R01, D(c1),
// X01 = -X00 (R01 X11), and publishes R and X of the block.  Then it looks
// ahead to the next pair (rows c0, c1 against columns r0 = c0 + 2, r0 + 1):
// it forms and publishes R_c0,r0 .. R_c1,r0+1 (the work of P2(K, r0 .. r0+1))
// and applies pair K's update to the next diagonal block, which it keeps in
// LDS for iteration K + 1 (the work of U2(K) on that block), so the chain
// hands nothing to the workers and back between two pair steps.  The block
// it loads must carry the K earlier updates, which the workers apply (U2 of
// the earlier pairs).  Tiles: tA, tB, tC the block, tX, tE, tD its X, and
// the look-ahead's R tiles (tF .. tI for real data, where the next block
// lands in tA .. tC; complex data reuses the block's tiles and tF, and the
// roles rotate every iteration).
template <typename T>
__device__ void chol_panel(const CholTileArgs& a, T (*lds)[TB][33], T* rowbuf) {
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
  T(*tA)[33] = lds[0];
  T(*tB)[33] = lds[1];
  T(*tC)[33] = lds[2];
  T(*tD)[33] = lds[3];
  T(*tE)[33] = lds[4];
  T(*tX)[33] = lds[5];
  T(*tF)[33] = lds[6];
  // real data only (ten tiles): the look-ahead's R tiles are tF .. tI
  constexpr bool kCx = ScalarTraits<T>::is_complex;
  T(*tG)[33] = lds[kCx ? 6 : 7];
  T(*tH)[33] = lds[kCx ? 6 : 8];
  T(*tI)[33] = lds[kCx ? 6 : 9];
  for (int K = 0; 2 * K < nt; ++K) {
    const int c0 = 2 * K, c1 = 2 * K + 1, r0 = c0 + 2;
    if (kCx) stamp(K, 0);
    if (K == 0) {
      if (!inputs_ready(ctl)) return;
      load_n<3, T>(tA, TileSrc{G, a.ldg, c0, c0}, tB, TileSrc{G, a.ldg, c0, c1}, tC, TileSrc{G, a.ldg, c1, c1},
                   tC, TileSrc{}, tC, TileSrc{}, a.n);
      __syncthreads();
    }
    if (kCx) {
      stamp(K, 1);
      stamp(K, 2);
    }
    // h = 0: D(c0) on tA -> R00 (tA), X00 (tX); h = 1: R01 = X00^H G01 (tB),
    // G11 -= R01^H R01, D(c1) on tC -> R11 (tC), X11 (tD).  One loop body, so
    // the in-wave factorisation is instantiated once.
    for (int h = 0; h < 2; ++h) {
      if (h == 1) {
        mm32<T>(tB, tX, true, tB, false, 1.0, false);
        mm32<T>(tC, tB, true, tB, false, -1.0, true);
      }
      stamp(K, 3 + h);
      const int bad = factor_diag<T>(h ? tC : tA, h ? tD : tX, rowbuf);
      if (bad != 0) {
        report_bad<T>(a, c0 + h, bad);
        return;
      }
    }
    stamp(K, 5);
    mm32<T>(tE, tB, false, tD, false, 1.0, false);        // R01 X11
    mm32<T>(tE, tX, false, tE, false, -1.0, false);       // X01 = -X00 (R01 X11)
    if constexpr (!ScalarTraits<T>::is_complex) {
      // Real data, ten tiles: the pair's block first (R, X and the Xc copy,
      // one drain, its counters), so the workers' tasks of this pair start
      // at once; then the look-ahead's seven loads in one batch (G of rows
      // c0, c1 in columns r0, r0 + 1 into tF .. tI, the next block into
      // tA .. tC), its R tiles, their stores draining behind the update of
      // the next block, and their counters.
      put_tile<T>(tA, G, a.ldg, a.n, c0, c0);
      put_tile<T>(tB, G, a.ldg, a.n, c0, c1);
      put_tile<T>(tC, G, a.ldg, a.n, c1, c1);
      clear_tile<T>(G, a.ldg, a.n, c1, c0);
      put_tile<T>(tX, X, a.ldx, a.n, c0, c0);
      put_tile<T>(tE, X, a.ldx, a.n, c0, c1);
      put_tile<T>(tD, X, a.ldx, a.n, c1, c1);
      clear_tile<T>(X, a.ldx, a.n, c1, c0);
      if (a.ny > 0) {
        T* Xc = static_cast<T*>(a.Xc);
        put_tile<T>(tX, Xc, a.ldxc, a.n, c0, c0);
        put_tile<T>(tE, Xc, a.ldxc, a.n, c0, c1);
        put_tile<T>(tD, Xc, a.ldxc, a.n, c1, c1);
        clear_tile<T>(Xc, a.ldxc, a.n, c1, c0);
      }
      {
        int* fp[6] = {&zs[c0 * nt + c0], &zs[c0 * nt + c1], &zs[c1 * nt + c1],
                      &gs[c0 * nt + c0], &gs[c0 * nt + c1], &gs[c1 * nt + c1]};
        const int fv[6] = {1, 1, 1, K + 1, K + 1, K + 1};
        publish_n(6, fp, fv);
        if (a.ny > 0 && threadIdx.x == 0) atomicAdd(a.xdone + K, 1);
      }
      if (r0 < nt) {
        {
          int* const wp[7] = {&gs[c0 * nt + r0], &gs[c1 * nt + r0], &gs[c0 * nt + r0 + 1], &gs[c1 * nt + r0 + 1],
                              &gs[r0 * nt + r0], &gs[r0 * nt + r0 + 1], &gs[(r0 + 1) * nt + r0 + 1]};
          const int wv[7] = {K, K, K, K, K, K, K};
          if (!wait_all<7>(wp, wv, ctl)) return;
        }
        stamp(K, 1);
        load_7(tF, TileSrc{G, a.ldg, c0, r0}, tG, TileSrc{G, a.ldg, c1, r0}, tH, TileSrc{G, a.ldg, c0, r0 + 1},
               tI, TileSrc{G, a.ldg, c1, r0 + 1}, tA, TileSrc{G, a.ldg, r0, r0}, tB, TileSrc{G, a.ldg, r0, r0 + 1},
               tC, TileSrc{G, a.ldg, r0 + 1, r0 + 1}, a.n);
        stamp(K, 2);
        mm32_2<T>(tG, tE, tF, tD, tG, true, false, 1.0, false);  // R_c1,r0 = X01^H G_c0,r0 + X11^H G_c1,r0
        mm32<T>(tF, tX, true, tF, false, 1.0, false);            // R_c0,r0 = X00^H G_c0,r0
        mm32_2<T>(tI, tE, tH, tD, tI, true, false, 1.0, false);  // R_c1,r0+1
        mm32<T>(tH, tX, true, tH, false, 1.0, false);            // R_c0,r0+1
        stamp(K, 0);
        put_tile<T>(tF, G, a.ldg, a.n, c0, r0);
        put_tile<T>(tG, G, a.ldg, a.n, c1, r0);
        put_tile<T>(tH, G, a.ldg, a.n, c0, r0 + 1);
        put_tile<T>(tI, G, a.ldg, a.n, c1, r0 + 1);
        clear_tile<T>(G, a.ldg, a.n, r0, c0);
        clear_tile<T>(G, a.ldg, a.n, r0, c1);
        clear_tile<T>(G, a.ldg, a.n, r0 + 1, c0);
        clear_tile<T>(G, a.ldg, a.n, r0 + 1, c1);
        mm32_2<T>(tA, tF, tF, tG, tG, true, false, -1.0, true);  // G_r0,r0 -= R_.,r0^H R_.,r0
        mm32_2<T>(tB, tF, tH, tG, tI, true, false, -1.0, true);  // G_r0,r0+1 -= R_.,r0^H R_.,r0+1
        mm32_2<T>(tC, tH, tH, tI, tI, true, false, -1.0, true);  // G_r0+1,r0+1 -= R_.,r0+1^H R_.,r0+1
        int* rp[4] = {&gs[c0 * nt + r0], &gs[c1 * nt + r0], &gs[c0 * nt + r0 + 1], &gs[c1 * nt + r0 + 1]};
        const int rv[4] = {K + 1, K + 1, K + 1, K + 1};
        publish_n(4, rp, rv);
      }
      stamp(K, 6);
    } else {
      // Complex data, seven tiles: publish the block, form and publish the
      // look-ahead's R tiles (tA, tB, tC, tF), then load the next block into
      // the X tiles and update it; the roles rotate.
      put_tile<T>(tA, G, a.ldg, a.n, c0, c0);
      put_tile<T>(tB, G, a.ldg, a.n, c0, c1);
      put_tile<T>(tC, G, a.ldg, a.n, c1, c1);
      clear_tile<T>(G, a.ldg, a.n, c1, c0);
      put_tile<T>(tX, X, a.ldx, a.n, c0, c0);
      put_tile<T>(tE, X, a.ldx, a.n, c0, c1);
      put_tile<T>(tD, X, a.ldx, a.n, c1, c1);
      clear_tile<T>(X, a.ldx, a.n, c1, c0);
      if (a.ny > 0) {
        T* Xc = static_cast<T*>(a.Xc);
        put_tile<T>(tX, Xc, a.ldxc, a.n, c0, c0);
        put_tile<T>(tE, Xc, a.ldxc, a.n, c0, c1);
        put_tile<T>(tD, Xc, a.ldxc, a.n, c1, c1);
        clear_tile<T>(Xc, a.ldxc, a.n, c1, c0);
      }
      int* fp[6] = {&zs[c0 * nt + c0], &zs[c0 * nt + c1], &zs[c1 * nt + c1],
                    &gs[c0 * nt + c0], &gs[c0 * nt + c1], &gs[c1 * nt + c1]};
      const int fv[6] = {1, 1, 1, K + 1, K + 1, K + 1};
      publish_n(6, fp, fv);
      if (a.ny > 0 && threadIdx.x == 0) atomicAdd(a.xdone + K, 1);
      stamp(K, 6);
      if (r0 < nt) {
        {
          int* const wp[4] = {&gs[c0 * nt + r0], &gs[c1 * nt + r0], &gs[c0 * nt + r0 + 1], &gs[c1 * nt + r0 + 1]};
          const int wv[4] = {K, K, K, K};
          if (!wait_all<4>(wp, wv, ctl)) return;
        }
        load_n<4, T>(tA, TileSrc{G, a.ldg, c0, r0}, tB, TileSrc{G, a.ldg, c1, r0}, tC,
                     TileSrc{G, a.ldg, c0, r0 + 1}, tF, TileSrc{G, a.ldg, c1, r0 + 1}, tF, TileSrc{}, a.n);
        __syncthreads();
        mm32_2<T>(tB, tE, tA, tD, tB, true, false, 1.0, false);  // R_c1,r0 = X01^H G_c0,r0 + X11^H G_c1,r0
        mm32<T>(tA, tX, true, tA, false, 1.0, false);            // R_c0,r0 = X00^H G_c0,r0
        mm32_2<T>(tF, tE, tC, tD, tF, true, false, 1.0, false);  // R_c1,r0+1
        mm32<T>(tC, tX, true, tC, false, 1.0, false);            // R_c0,r0+1
        put_tile<T>(tA, G, a.ldg, a.n, c0, r0);
        put_tile<T>(tB, G, a.ldg, a.n, c1, r0);
        put_tile<T>(tC, G, a.ldg, a.n, c0, r0 + 1);
        put_tile<T>(tF, G, a.ldg, a.n, c1, r0 + 1);
        clear_tile<T>(G, a.ldg, a.n, r0, c0);
        clear_tile<T>(G, a.ldg, a.n, r0, c1);
        clear_tile<T>(G, a.ldg, a.n, r0 + 1, c0);
        clear_tile<T>(G, a.ldg, a.n, r0 + 1, c1);
        int* rp[4] = {&gs[c0 * nt + r0], &gs[c1 * nt + r0], &gs[c0 * nt + r0 + 1], &gs[c1 * nt + r0 + 1]};
        const int rv[4] = {K + 1, K + 1, K + 1, K + 1};
        publish_n(4, rp, rv);
        // the next block (K updates) into the X tiles, then pair K's update
        {
          int* const wp[3] = {&gs[r0 * nt + r0], &gs[r0 * nt + r0 + 1], &gs[(r0 + 1) * nt + r0 + 1]};
          const int wv[3] = {K, K, K};
          if (!wait_all<3>(wp, wv, ctl)) return;
        }
        load_n<3, T>(tX, TileSrc{G, a.ldg, r0, r0}, tE, TileSrc{G, a.ldg, r0, r0 + 1}, tD,
                     TileSrc{G, a.ldg, r0 + 1, r0 + 1}, tD, TileSrc{}, tD, TileSrc{}, a.n);
        __syncthreads();
        mm32_2<T>(tX, tA, tA, tB, tB, true, false, -1.0, true);  // G_r0,r0 -= R_.,r0^H R_.,r0
        mm32_2<T>(tE, tA, tC, tB, tF, true, false, -1.0, true);  // G_r0,r0+1 -= R_.,r0^H R_.,r0+1
        mm32_2<T>(tD, tC, tC, tF, tF, true, false, -1.0, true);  // G_r0+1,r0+1 -= R_.,r0+1^H R_.,r0+1
        T(*oA)[33] = tA;
        T(*oB)[33] = tB;
        T(*oC)[33] = tC;
        tA = tX;
        tB = tE;
        tC = tD;
        tX = oA;
        tE = oB;
        tD = oC;
      }
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) chol_tile_kernel(CholTileArgs a) {
  // the panel's 32 x 32 tiles (ten for real data, seven for complex), six
  // for the worker tasks (a right-product task reuses them)
  __shared__ __align__(16) T lds[kPanelTiles<T>][TB][33];
  static_assert(sizeof(lds) >= kYLdsElems<T> * sizeof(T), "right-product slabs fit the tiles");
  T(*tA)[33] = lds[0];
  T(*tB)[33] = lds[1];
  T(*tC)[33] = lds[2];
  T(*tD)[33] = lds[3];
  T(*tE)[33] = lds[4];
  T(*tX)[33] = lds[5];
  __shared__ __align__(16) T rowbuf[2 * TB];
  __shared__ int task_s;
  if (blockIdx.x == 0) {
    chol_panel<T>(a, lds, rowbuf);
    return;
  }
  T* G = reinterpret_cast<T*>(a.G);
  T* X = reinterpret_cast<T*>(a.X);
  const int nt = a.nt;
  int* ctl = a.ctl;
  int* gs = a.gst;
  int* zs = a.zst;
  int iter = 0;
  for (;;) {
    if (threadIdx.x == 0) task_s = atomicAdd(ctl, 1);
    __syncthreads();
    const int t = task_s;
    if (t >= a.ntasks) break;
    while (a.iter_off[iter + 1] <= t) ++iter;  // tasks are taken in increasing order
    const CholTask tk = chol_task_in_iter(nt, iter, a.ny, t - a.iter_off[iter], a.ysplit);
    const int K = tk.k, c0 = 2 * K, c1 = 2 * K + 1;
    switch (tk.type) {
      case CT_P2: {  // [R_c0j; R_c1j] = X_KK^H [G_c0j; G_c1j]; the mirrored lower tiles of G are zeroed
        const int j = tk.j;
        {
          // the panel's block (R and X published together), G_c0j and G_c1j with K updates
          int* const wp[3] = {&gs[c1 * nt + c1], &gs[c0 * nt + j], &gs[c1 * nt + j]};
          const int wv[3] = {K + 1, K, K};
          if (!wait_all<3>(wp, wv, ctl)) break;
        }
        load_n<5, T>(tA, TileSrc{X, a.ldx, c0, c0}, tB, TileSrc{X, a.ldx, c0, c1}, tC, TileSrc{X, a.ldx, c1, c1},
                     tD, TileSrc{G, a.ldg, c0, j}, tE, TileSrc{G, a.ldg, c1, j}, a.n);
        __syncthreads();
        mm32_2<T>(tE, tB, tD, tC, tE, true, false, 1.0, false);  // R_c1j = X01^H G0 + X11^H G1
        mm32<T>(tD, tA, true, tD, false, 1.0, false);            // R_c0j = X00^H G0
        put_tile<T>(tD, G, a.ldg, a.n, c0, j);
        put_tile<T>(tE, G, a.ldg, a.n, c1, j);
        clear_tile<T>(G, a.ldg, a.n, j, c0);
        clear_tile<T>(G, a.ldg, a.n, j, c1);
        int* fp[2] = {&gs[c0 * nt + j], &gs[c1 * nt + j]};
        const int fv[2] = {K + 1, K + 1};
        publish_n(2, fp, fv);
        break;
      }
      case CT_U2: {  // G_ij -= R_c0i^H R_c0j + R_c1i^H R_c1j
        const int i = tk.i, j = tk.j;
        {
          int* const wp[5] = {&gs[c0 * nt + i], &gs[c1 * nt + i], &gs[c0 * nt + j], &gs[c1 * nt + j],
                              &gs[i * nt + j]};
          const int wv[5] = {K + 1, K + 1, K + 1, K + 1, K};
          if (!wait_all<5>(wp, wv, ctl)) break;
        }
        if (i == j) {
          load_n<3, T>(tA, TileSrc{G, a.ldg, c0, i}, tB, TileSrc{G, a.ldg, c1, i}, tE, TileSrc{G, a.ldg, i, j},
                       tE, TileSrc{}, tE, TileSrc{}, a.n);
          __syncthreads();
          mm32_2<T>(tE, tA, tA, tB, tB, true, false, -1.0, true);
        } else {
          load_n<5, T>(tA, TileSrc{G, a.ldg, c0, i}, tB, TileSrc{G, a.ldg, c1, i}, tC, TileSrc{G, a.ldg, c0, j},
                       tD, TileSrc{G, a.ldg, c1, j}, tE, TileSrc{G, a.ldg, i, j}, a.n);
          __syncthreads();
          mm32_2<T>(tE, tA, tC, tB, tD, true, false, -1.0, true);
        }
        put_tile<T>(tE, G, a.ldg, a.n, i, j);
        int* fp[1] = {&gs[i * nt + j]};
        const int fv[1] = {K + 1};
        publish_n(1, fp, fv);
        break;
      }
      case CT_UQ2: {  // Z_jl -= X_jc0 R_c0l + X_jc1 R_c1l (X_c1c0 = 0); the first update overwrites
        const int jr = tk.i, l = tk.j;
        const int base = jr / 2;  // the pair of row jr: its first update comes from step K = base
        const int fin = K - base + 1;  // X_jr,c0 / X_jr,c1 final value
        const bool first = K == base;
        const bool has0 = jr <= c0;
        {
          // (counters start at 0: a target of 0 is always met)
          int* const wp[5] = {&zs[jr * nt + c0], &zs[jr * nt + c1], &gs[c0 * nt + l], &gs[c1 * nt + l],
                              &zs[jr * nt + l]};
          const int wv[5] = {has0 ? fin : 0, fin, K + 1, K + 1, first ? 0 : K - base};
          if (!wait_all<5>(wp, wv, ctl)) break;
        }
        {
          // tB = X_j,c1, tD = R_c1l, [tE = Z_jl], [tA = X_j,c0, tC = R_c0l]
          const TileSrc xj1{X, a.ldx, jr, c1}, r1l{G, a.ldg, c1, l}, zjl{X, a.ldx, jr, l},
              xj0{X, a.ldx, jr, c0}, r0l{G, a.ldg, c0, l};
          if (first && has0) load_n<4, T>(tB, xj1, tD, r1l, tA, xj0, tC, r0l, tC, TileSrc{}, a.n);
          else if (first) load_n<2, T>(tB, xj1, tD, r1l, tD, TileSrc{}, tD, TileSrc{}, tD, TileSrc{}, a.n);
          else if (has0) load_n<5, T>(tB, xj1, tD, r1l, tE, zjl, tA, xj0, tC, r0l, a.n);
          else load_n<3, T>(tB, xj1, tD, r1l, tE, zjl, tE, TileSrc{}, tE, TileSrc{}, a.n);
        }
        __syncthreads();
        mm32_2<T>(tE, has0 ? tA : nullptr, tC, tB, tD, false, false, -1.0, !first);
        put_tile<T>(tE, X, a.ldx, a.n, jr, l);
        int* fp[1] = {&zs[jr * nt + l]};
        const int fv[1] = {K - base + 1};
        publish_n(1, fp, fv);
        break;
      }
      case CT_PQ2: {  // [X_jc0 X_jc1] = [Z_jc0 Z_jc1] X_KK (j < c0); the mirrored lower tiles of X are zeroed
        const int jr = tk.j;
        const int cnt = K - jr / 2;  // UQ2 updates of Z_jr,c0 and Z_jr,c1
        {
          int* const wp[3] = {&gs[c1 * nt + c1], &zs[jr * nt + c0], &zs[jr * nt + c1]};
          const int wv[3] = {K + 1, cnt, cnt};
          if (!wait_all<3>(wp, wv, ctl)) break;
        }
        load_n<5, T>(tA, TileSrc{X, a.ldx, jr, c0}, tB, TileSrc{X, a.ldx, jr, c1}, tC, TileSrc{X, a.ldx, c0, c0},
                     tD, TileSrc{X, a.ldx, c0, c1}, tE, TileSrc{X, a.ldx, c1, c1}, a.n);
        __syncthreads();
        mm32_2<T>(tB, tA, tD, tB, tE, false, false, 1.0, false);  // X_jc1 = Z_jc0 X01 + Z_jc1 X11
        mm32<T>(tA, tA, false, tC, false, 1.0, false);             // X_jc0 = Z_jc0 X00
        put_tile<T>(tA, X, a.ldx, a.n, jr, c0);
        put_tile<T>(tB, X, a.ldx, a.n, jr, c1);
        clear_tile<T>(X, a.ldx, a.n, c0, jr);
        clear_tile<T>(X, a.ldx, a.n, c1, jr);
        int* fp[2] = {&zs[jr * nt + c0], &zs[jr * nt + c1]};
        const int fv[2] = {cnt + 1, cnt + 1};
        publish_n(2, fp, fv);
        if (a.ny > 0) {
          put_tile<T>(tA, static_cast<T*>(a.Xc), a.ldxc, a.n, jr, c0);
          put_tile<T>(tB, static_cast<T*>(a.Xc), a.ldxc, a.n, jr, c1);
          publish_n(0, nullptr, nullptr);
          if (threadIdx.x == 0) atomicAdd(a.xdone + K, 1);
        }
        break;
      }
      case CT_Y: {  // rows of Y's column block K, once X's column block K is final
        {
          int* const wp[1] = {&a.xdone[K]};
          const int wv[1] = {2 * K + 1};
          if (!wait_all<1>(wp, wv, ctl)) break;
        }
        const long long t_start = (long long)__builtin_amdgcn_s_memrealtime();
        y_task<T>(a, K, tk.j % a.ny, K >= a.ysplit ? 1 + tk.j / a.ny : 0, &lds[0][0][0]);
        if (a.trace != nullptr && threadIdx.x == 0) {
          const long long t_end = (long long)__builtin_amdgcn_s_memrealtime();
          // the pair's last right-product task to finish; every task's start
          // and end after the panel stamps (tests: 8 * 32 + 2 * ntasks entries)
          atomicMax(reinterpret_cast<unsigned long long*>(a.trace + K * 8 + 7), (unsigned long long)t_end);
          a.trace[8 * 32 + 2 * t] = t_start;
          a.trace[8 * 32 + 2 * t + 1] = t_end;
        }
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

// Y[:, c0 : n] = 0 (m rows)
__global__ void zero_cols_kernel(double* Y, int64_t ldy, int m, int c0, int n) {
  const int64_t tot = (int64_t)m * (n - c0);
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256)
    Y[e % m + (c0 + e / m) * ldy] = 0.0;
}

// tiles of the pair-step kernel: ceil(n / 32) rounded up to even
inline int chol_tiles(int n) {
  const int t = std::max(1, cdiv(n, TB));
  return t + (t & 1);
}

}  // namespace

size_t chol_tile_flag_ints(int n) {
  const int nt = chol_tiles(std::min(n, kCholTileMaxN));
  return (size_t)2 * nt * nt + 64;
}

// One persistent launch for n <= kCholTileMaxN.  flags: chol_tile_flag_ints(n)
// ints of device memory (cleared here by a kernel, not a memset, so the call
// can be captured).  info (device, optional): first non-positive pivot + offset.
template <typename T>
int chol_tile_factor(T* G, int64_t ldg, int n, T* X, int64_t ldx, int* flags, int* info, int info_offset,
                     hipStream_t s, int grid, long long* trace, const CholRightProduct* rp) {
  if (n <= 0) return 0;
  if (n > kCholTileMaxN) return -3;
  if (rp != nullptr && rp->m < 0) return -1;
  CholTileArgs a{};
  a.G = G; a.ldg = ldg; a.X = X; a.ldx = ldx;
  a.n = n;
  a.nt = chol_tiles(n);
  if (rp != nullptr && rp->m > 0) {
    if (rp->Xc == nullptr || rp->ldxc < n || rp->ldxc % 16 != 0 || (reinterpret_cast<uintptr_t>(rp->Xc) & 127) != 0)
      return -1;
    a.B = rp->B; a.ldb = rp->ldb; a.Y = rp->Y; a.ldy = rp->ldy;
    a.Xc = rp->Xc; a.ldxc = rp->ldxc;
    a.m = rp->m;
    a.ny = cdiv(rp->m, kYRowsOf<T>);
  }
  // the pairs whose right-product tasks would outlast the elimination: the
  // last quarter (at least three) split their tasks in two
  a.ysplit = 1 << 30;
  if (a.ny > 0 && a.nt / 2 >= 4) a.ysplit = a.nt / 2 - std::max(3, a.nt / 8);
  int off = 0;
  const int nq = a.nt / 2;
  for (int K = 0; K < nq; ++K) {
    a.iter_off[K] = off;
    off += chol_iter_total(a.nt, K, a.ny, a.ysplit);
  }
  a.iter_off[nq] = off;
  a.ntasks = off;
  a.gst = flags;
  a.zst = flags + a.nt * a.nt;
  a.ctl = flags + 2 * a.nt * a.nt;
  a.xdone = a.ctl + 32;  // one counter per pair (at most 32 pairs)
  a.info = info;
  a.info_offset = info_offset;
  a.trace = trace;
  const int nflags = 2 * a.nt * a.nt + 64;
  hipLaunchKernelGGL((chol_zero_ints_kernel<T>), dim3(cdiv(nflags, 256)), dim3(256), 0, s, flags, nflags);
  if (a.ny > 0 && a.ysplit < a.nt / 2) {
    const int c0 = 64 * a.ysplit;  // < n (every pair has a column)
    // complex columns as real ones of twice the length
    constexpr int w = sizeof(T) / sizeof(double);
    hipLaunchKernelGGL(zero_cols_kernel, dim3(1024), dim3(256), 0, s, static_cast<double*>(a.Y), w * a.ldy, w * a.m,
                       c0, n);
  }
  // the panel workgroup plus at least one for the other tasks; up to n ~ 1024
  // the chain dominates and fewer polling workgroups contend less for the
  // coherent loads, beyond it the task throughput needs every CU
  // (with a right product the workgroups that wait for the chain have its
  // tasks to run: every CU)
  if (grid <= 0) grid = (n <= 1280 && a.ny == 0) ? 128 : kCholGridY;
  grid = std::max(2, std::min(grid, a.ntasks + 1));
  hipLaunchKernelGGL((chol_tile_kernel<T>), dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

template int chol_tile_factor<double>(double*, int64_t, int, double*, int64_t, int*, int*, int, hipStream_t, int,
                                      long long*, const CholRightProduct*);
template int chol_tile_factor<zd>(zd*, int64_t, int, zd*, int64_t, int*, int*, int, hipStream_t, int, long long*,
                                  const CholRightProduct*);

// Host-side check of the task order (test hook) for ntiles tiles (rounded up
// to T, even): every
// task once, every dependency (the waits of the kernel, with the panel's
// iteration K run just before the tasks of iteration K) satisfied by tasks
// earlier in the order, and the final counters as documented.  Returns the
// task count, or -1.
int chol_tile_order_check(int ntiles, int ny, int ysplit) {
  if (ntiles <= 0 || ntiles > kCholTileMaxNt || ny < 0) return -1;
  const int T = ntiles + (ntiles & 1);
  std::vector<int> gst(T * T, 0), zst(T * T, 0), xd(T / 2, 0), yd(T / 2, 0);
  auto G = [&](int r, int c) -> int& { return gst[r * T + c]; };
  auto Z = [&](int r, int c) -> int& { return zst[r * T + c]; };
  int total = 0;
  for (int K = 0; K < T / 2; ++K) {
    const int c0 = 2 * K, c1 = 2 * K + 1;
    // the panel's iteration K (its block carries K updates: K = 0, or the
    // look-ahead of iteration K - 1 checked below)
    if (G(c0, c0) != K || G(c0, c1) != K || G(c1, c1) != K) return -1;
    G(c0, c0) = G(c0, c1) = G(c1, c1) = K + 1;
    Z(c0, c0) = Z(c0, c1) = Z(c1, c1) = 1;
    ++xd[K];
    const int r0 = c0 + 2;
    if (r0 < T) {
      // the look-ahead waits only for queue tasks of earlier iterations
      for (int r = c0; r <= c1; ++r)
        for (int c = r0; c <= r0 + 1; ++c) {
          if (G(r, c) != K) return -1;
          G(r, c) = K + 1;
        }
      if (G(r0, r0) != K || G(r0, r0 + 1) != K || G(r0 + 1, r0 + 1) != K) return -1;
      G(r0, r0) = G(r0, r0 + 1) = G(r0 + 1, r0 + 1) = K + 1;
    }
    const int cnt = chol_iter_total(T, K, ny, ysplit);
    for (int u = 0; u < cnt; ++u) {
      const CholTask t = chol_task_in_iter(T, K, ny, u, ysplit);
      if (t.k != K) return -1;
      switch (t.type) {
        case CT_P2: {
          const int j = t.j;
          if (j < c1 + 1 || j >= T || G(c1, c1) < K + 1 || G(c0, j) < K || G(c1, j) < K) return -1;
          if (G(c0, j) != K || G(c1, j) != K) return -1;  // written once
          G(c0, j) = G(c1, j) = K + 1;
          break;
        }
        case CT_U2: {
          const int i = t.i, j = t.j;
          if (!(c1 < i && i <= j && j < T)) return -1;
          if (G(c0, i) < K + 1 || G(c1, i) < K + 1 || G(c0, j) < K + 1 || G(c1, j) < K + 1) return -1;
          if (G(i, j) != K) return -1;
          G(i, j) = K + 1;
          break;
        }
        case CT_UQ2: {
          const int jr = t.i, l = t.j, base = jr / 2, fin = K - base + 1;
          if (!(jr <= c1 && c1 < l && l < T)) return -1;
          if (jr <= c0 && Z(jr, c0) < fin) return -1;
          if (Z(jr, c1) < fin || G(c0, l) < K + 1 || G(c1, l) < K + 1) return -1;
          if (Z(jr, l) != K - base) return -1;
          Z(jr, l) = K - base + 1;
          break;
        }
        case CT_PQ2: {
          const int jr = t.j, c = K - jr / 2;
          if (!(jr < c0) || G(c1, c1) < K + 1 || Z(jr, c0) != c || Z(jr, c1) != c) return -1;
          Z(jr, c0) = Z(jr, c1) = c + 1;
          ++xd[K];
          break;
        }
        case CT_Y: {
          // the counter's target: X's column block K final (every tile of
          // rows 0 .. c1 in columns c0, c1 at its final version)
          const int halves = K >= ysplit ? 2 : 1;
          if (t.j < 0 || t.j >= halves * ny || xd[K] != 2 * K + 1) return -1;
          for (int r = 0; r <= c1; ++r)
            if ((r <= c0 && Z(r, c0) != c0 / 2 - r / 2 + 1) || Z(r, c1) != c1 / 2 - r / 2 + 1) return -1;
          if (yd[K] != t.j) return -1;  // every row block (and half) once, in order
          ++yd[K];
          break;
        }
        default:
          return -1;
      }
      ++total;
    }
  }
  for (int r = 0; r < T; ++r)
    for (int c = r; c < T; ++c)
      if (G(r, c) != r / 2 + 1 || Z(r, c) != c / 2 - r / 2 + 1) return -1;
  for (int K = 0; K < T / 2; ++K)
    if (yd[K] != (K >= ysplit ? 2 : 1) * ny) return -1;
  return total;
}

}  // namespace mak
