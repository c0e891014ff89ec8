// Reduced-precision GEMM on FP64 data for the mixed-precision polar bulk:
//   C = alpha op(A) op(B) + beta C with A, B, C in FP64 memory, the product
//   computed on FP32 MFMA (v_mfma_f32_16x16x4_f32) or BF16 MFMA
//   (v_mfma_f32_16x16x32_bf16), both accumulating in FP32; alpha, beta and the
//   accumulation with C happen in FP64.
//
// Two steps:
//   1. pack_lo_kernel rounds op(A) (M x K) and op(B)^T (N x K) once into
//      row-major, k-contiguous copies in the compute type (rows padded to a
//      multiple of 128, k to a multiple of 64, padding zero), so the GEMM reads
//      each operand row as 128-byte runs with 16-byte loads and needs no bounds
//      checks in its main loop;
//   2. gemm_lo_kernel: 128 x 128 tile, 4 waves of 64 x 64, one 128-byte k-slab
//      per step (32 FP32 or 64 BF16 values), the next slab loaded into
//      registers while the MFMAs of the current one run.
//
// A lane of either MFMA holds 8 consecutive k values of one row per step:
// the BF16 MFMA takes k = 8 (l / 16) .. + 7 natively; for FP32 the sum over k
// is order-free, so MFMA e of a 32-wide slab uses k = 8 (l / 16) + e for both
// operands and every fragment is two 16-byte LDS reads.  The operands are
// swapped in the MFMA so lane l holds ROW l % 16 of C (128-byte column stores).
#include <algorithm>

#include "mak_common.hpp"
#include "mak_kernels.hpp"

namespace mak {

namespace {

template <int PREC> struct LoType;
template <> struct LoType<32> { using T = float; };
template <> struct LoType<16> { using T = __bf16; };

constexpr int kLoRows = 128;  // tile rows / row padding of the packed copies
constexpr int kLoK = 64;      // k padding of the packed copies

inline int64_t lo_rows(int r) { return cdiv64(std::max(r, 0), kLoRows) * kLoRows; }
inline int64_t lo_kpad(int k) { return cdiv64(std::max(k, 0), kLoK) * kLoK; }

// D (Rp x Kp, row-major, ld Kp) = round(src(r, k)) for r < R, k < K, zero
// elsewhere; src(r, k) = S[k + r ld] (kcontig) or S[r + k ld].  One 64 x 64
// block per workgroup, read along the source's contiguous index and written
// along k through an LDS transpose.
template <typename TL>
__device__ inline void pack_lo_body(const double* S, int64_t ld, int kcontig, int R, int K, TL* D, int64_t Kp,
                                    float (*tile)[65]) {
  const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  if (kcontig) {
    for (int q = ty; q < 64; q += 4) {
      const int r = r0 + q, k = k0 + tx;
      tile[q][tx] = (r < R && k < K) ? (float)S[k + (int64_t)r * ld] : 0.f;
    }
  } else {
    for (int q = ty; q < 64; q += 4) {
      const int r = r0 + tx, k = k0 + q;
      tile[tx][q] = (r < R && k < K) ? (float)S[r + (int64_t)k * ld] : 0.f;
    }
  }
  __syncthreads();
  for (int q = ty; q < 64; q += 4) D[(int64_t)(r0 + q) * Kp + k0 + tx] = (TL)tile[q][tx];
}
template <typename TL>
__global__ void __launch_bounds__(256) pack_lo_kernel(const double* S, int64_t ld, int kcontig, int R, int K, TL* D,
                                                      int64_t Kp, const int* gate, int want) {
  __shared__ float tile[64][65];
  if (gate_closed(gate, want)) return;
  pack_lo_body<TL>(S, ld, kcontig, R, K, D, Kp, tile);
}
// Both operands of a product in one launch (blockIdx.z: 0 = A, R = M rows;
// 1 = B, R = N rows), into BF16 when *g16 == 1, else FP32 when *g32 == 1,
// else nothing; A at (TL*) base, B at (TL*) base + boff.
__global__ void __launch_bounds__(256) pack_lo_dual_kernel(const double* A, int64_t lda, int acontig, int M,
                                                           const double* B, int64_t ldb, int bcontig, int N, int K,
                                                           void* base, int64_t boff, int64_t Kp, int ablocks,
                                                           int bblocks, const int* g16, const int* g32) {
  __shared__ float tile[64][65];
  const bool h = *g16 == 1;
  if (!h && *g32 != 1) return;
  const bool isB = blockIdx.z == 1;
  if ((int)blockIdx.y >= (isB ? bblocks : ablocks)) return;
  const double* S = isB ? B : A;
  const int64_t ld = isB ? ldb : lda;
  const int kc = isB ? bcontig : acontig, R = isB ? N : M;
  const int64_t off = isB ? boff : 0;
  if (h) pack_lo_body<__bf16>(S, ld, kc, R, K, reinterpret_cast<__bf16*>(base) + off, Kp, tile);
  else pack_lo_body<float>(S, ld, kc, R, K, reinterpret_cast<float*>(base) + off, Kp, tile);
}

template <typename TL>
void pack_lo(const double* S, int64_t ld, int kcontig, int R, int K, TL* D, int64_t Rp, int64_t Kp,
             hipStream_t s, const int* gate, int want) {
  if (Rp <= 0 || Kp <= 0) return;
  dim3 grid((unsigned)(Kp / 64), (unsigned)(Rp / 64));
  hipLaunchKernelGGL((pack_lo_kernel<TL>), grid, dim3(256), 0, s, S, ld, kcontig, R, K, D, Kp, gate, want);
}

// Complex data through its real embedding, with the real and imaginary parts
// interleaved along k (k' = 2 k + part): for C = A B (A M x K, B K x N
// complex), C_re and C_im side by side (M x 2N) = [A packed: A(i, k).re and
// .im at k' = 2k, 2k + 1] x [B embedded: 2K x 2N, (2k + p, j) = the entry
// of [[B.re, B.im], [-B.im, B.re]] in block row p, column j].  A k-clip for
// upper triangular B stays a prefix of k' (k <= j).
//
// A (M x K complex, ld in complex entries) -> D (Rp x Kp, k' contiguous):
// one 64-row x 32-column block per workgroup, read down the columns (16-byte
// steps) and written along k' through LDS.
template <typename TL>
__device__ inline void pack_cplx_a_body(const double* S, int64_t ld, int R, int K, TL* D, int64_t Kp,
                                        float (*tre)[65], float (*tim)[65]) {
  const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int q = ty; q < 32; q += 4) {
    const int r = r0 + tx, k = k0 + q;
    float re = 0.f, im = 0.f;
    if (r < R && k < K) {
      const double* e = S + 2 * (r + (int64_t)k * ld);
      re = (float)e[0];
      im = (float)e[1];
    }
    tre[q][tx] = re;
    tim[q][tx] = im;
  }
  __syncthreads();
  for (int q = ty; q < 64; q += 4) {
    const float v = (tx & 1) ? tim[tx >> 1][q] : tre[tx >> 1][q];
    D[(int64_t)(r0 + q) * Kp + 2 * k0 + tx] = (TL)v;
  }
}
// B (K x N complex) -> D (Np x Kp): row j < 2N, column k' < 2K holds the
// embedded entry above (zero beyond); reads along k are contiguous.
template <typename TL>
__device__ inline void pack_cplx_b_body(const double* S, int64_t ld, int K, int N, TL* D, int64_t Kp) {
  const int j0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int kp = k0 + tx, k = kp >> 1, pk = kp & 1;
  for (int q = ty; q < 64; q += 4) {
    const int j = j0 + q;
    float v = 0.f;
    if (j < 2 * N && k < K) {
      const int pj = j >= N ? 1 : 0, c = j - pj * N;
      const double x = S[2 * (k + (int64_t)c * ld) + (pk == pj ? 0 : 1)];
      v = (float)((pk == 1 && pj == 0) ? -x : x);
    }
    D[(int64_t)j * Kp + kp] = (TL)v;
  }
}
// Both complex packs in one launch (blockIdx.z: 0 = A, 1 = B), into BF16 when
// *g16 == 1, else FP32 when *g32 == 1, else nothing; A at (TL*) base, B at
// (TL*) base + boff.
__global__ void __launch_bounds__(256) pack_cplx_dual_kernel(const double* A, int64_t lda, int M, const double* B,
                                                             int64_t ldb, int K, int N, void* base, int64_t boff,
                                                             int64_t Kp, int ablocks, int bblocks, const int* g16,
                                                             const int* g32) {
  __shared__ float tre[32][65], tim[32][65];
  const bool h = *g16 == 1;
  if (!h && *g32 != 1) return;
  if (blockIdx.z == 0) {
    if ((int)blockIdx.y >= ablocks) return;
    if (h) pack_cplx_a_body<__bf16>(A, lda, M, K, reinterpret_cast<__bf16*>(base), Kp, tre, tim);
    else pack_cplx_a_body<float>(A, lda, M, K, reinterpret_cast<float*>(base), Kp, tre, tim);
  } else {
    if ((int)blockIdx.y >= bblocks) return;
    if (h) pack_cplx_b_body<__bf16>(B, ldb, K, N, reinterpret_cast<__bf16*>(base) + boff, Kp);
    else pack_cplx_b_body<float>(B, ldb, K, N, reinterpret_cast<float*>(base) + boff, Kp);
  }
}

// native 16-byte vector (HIP's uint4 is a struct, which keeps arrays of it in
// scratch memory)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct LoGemmArgs {
  int M, N;
  int64_t Kp;            // packed k extent (multiple of 64)
  const void* Ap;        // op(A), Mp x Kp, k contiguous
  const void* Bp;        // op(B)^T, Np x Kp, k contiguous (null: at Ap + bOff elements)
  int64_t bOff;
  double* C;
  int64_t ldc;
  const double* Cin;     // the beta term's matrix (C itself by default)
  int64_t ldcin;
  double alpha, beta;
  int tri;
  int upper;             // only the tiles on or above the diagonal (Hermitian results)
  const int* gate;       // run only if *gate == gate_want (when non-null)
  int gate_want;
  int cplx_n;            // > 0: complex C (N = 2 cplx_n real columns, C_re then C_im) from the
                         // real embedding; C and Cin are then complex arrays (ld in complex entries)
  int cplx_upperB;       // complex B upper triangular: k' clipped to 2 (max column of the tile + 1)
};

// Tile order: XCD-contiguous ranges of tiles, except under triangular hints,
// where the work per tile varies and the round-robin deal over the XCDs keeps
// them balanced; the index whose K range grows is reversed (longest first).
// For op(B) upper triangular with an even number of tile columns and at most
// two workgroups per CU (256 CUs), the first half of the launch takes the
// column blocks longest first and the second half shortest first, so
// workgroups b and b + total / 2, which share a CU, carry equal work together.
constexpr int kLoCus = 256;
__device__ inline void lo_tile_coords(int& tm, int& tn, int tri) {
  const int gx = gridDim.x, gy = gridDim.y;
  const int total = gx * gy;
  int L = blockIdx.x + blockIdx.y * gx;
  if ((total & 7) == 0 && tri == 0) L = (L & 7) * (total >> 3) + (L >> 3);
  tm = L % gx;
  tn = L / gx;
  if (tri & TRI_A_LOWER) tm = gx - 1 - tm;
  if (tri & TRI_B_UPPER) {
    if ((gy & 1) == 0 && !(tri & TRI_A_LOWER) && total <= 2 * kLoCus) tn = tn < gy / 2 ? gy - 1 - tn : tn - gy / 2;
    else tn = gy - 1 - tn;
  }
}

// As, Bs: kLoRows x 9 16-byte vectors each (8 per row + 16 bytes pad)
template <int PREC>
__device__ inline void gemm_lo_body(const LoGemmArgs& g, u32x4 (*As)[9], u32x4 (*Bs)[9]) {
  using TL = typename LoType<PREC>::T;
  constexpr int BK = 128 / (int)sizeof(TL);  // k values per 128-byte slab
  constexpr int TM = 4, TN = 4;              // 16 x 16 blocks per wave (64 x 64)
  const void* gBp = g.Bp != nullptr ? g.Bp : reinterpret_cast<const TL*>(g.Ap) + g.bOff;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  int tmi, tni;
  lo_tile_coords(tmi, tni, g.tri);
  if (g.upper && tmi > tni) return;  // strictly below the diagonal: not wanted
  const int m0 = tmi * kLoRows, n0 = tni * kLoRows;
  // K range: the hints bound the nonzero terms of a tile; the slab-aligned
  // range covers them (the packed copies are zero beyond K)
  int64_t kbeg = 0, kend = g.Kp;
  if (g.tri & TRI_A_LOWER) kend = std::min<int64_t>(kend, m0 + kLoRows);
  if (g.tri & TRI_B_UPPER) kend = std::min<int64_t>(kend, n0 + kLoRows);
  if (g.tri & TRI_A_UPPER) kbeg = std::max<int64_t>(kbeg, m0);
  if (g.tri & TRI_B_LOWER) kbeg = std::max<int64_t>(kbeg, n0);
  if (g.cplx_upperB) {
    // the tile's complex columns: c = j (C_re half) or j - cplx_n (C_im half)
    const int n = g.cplx_n, jlast = n0 + kLoRows - 1;
    const int cmax = (jlast < n) ? jlast : (n0 >= n ? std::min(jlast - n, n - 1) : n - 1);
    kend = std::min<int64_t>(kend, 2 * (int64_t)(cmax + 1));
  }
  kbeg = kbeg / BK * BK;
  kend = std::min<int64_t>(g.Kp, (kend + BK - 1) / BK * BK);

  // loads: thread t moves 16-byte vector (t % 8) of rows t / 8 + 32 u
  const int lr = tid >> 3, lv = tid & 7;
  const char* Ab = reinterpret_cast<const char*>(g.Ap) + ((int64_t)(m0 + lr) * g.Kp) * (int64_t)sizeof(TL) + lv * 16;
  const char* Bb = reinterpret_cast<const char*>(gBp) + ((int64_t)(n0 + lr) * g.Kp) * (int64_t)sizeof(TL) + lv * 16;
  const int64_t rstride = 32 * g.Kp * (int64_t)sizeof(TL);
  u32x4 pa[4], pb[4];
  auto fetch = [&](int64_t k0) {
    const int64_t off = k0 * (int64_t)sizeof(TL);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      pa[u] = *reinterpret_cast<const u32x4*>(Ab + u * rstride + off);
      pb[u] = *reinterpret_cast<const u32x4*>(Bb + u * rstride + off);
    }
  };

  v4f acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  const int fr = lane & 15, fq = lane >> 4;
  if (kbeg < kend) fetch(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      As[lr + 32 * u][lv] = pa[u];
      Bs[lr + 32 * u][lv] = pb[u];
    }
    __syncthreads();
    if (k0 + BK < kend) fetch(k0 + BK);
    if constexpr (PREC == 16) {
#pragma unroll
      for (int st = 0; st < 2; ++st) {
        v8bf af[TM], bf[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const u32x4 v = As[wm * 64 + i * 16 + fr][4 * st + fq];
          af[i] = __builtin_bit_cast(v8bf, v);
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const u32x4 v = Bs[wn * 64 + j * 16 + fr][4 * st + fq];
          bf[j] = __builtin_bit_cast(v8bf, v);
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = mfma_bf16(bf[j], af[i], acc[i][j]);
      }
    } else {
      float af[TM][8], bf[TN][8];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const u32x4 v0 = As[wm * 64 + i * 16 + fr][2 * fq], v1 = As[wm * 64 + i * 16 + fr][2 * fq + 1];
        const f32x4 f0 = __builtin_bit_cast(f32x4, v0), f1 = __builtin_bit_cast(f32x4, v1);
        for (int e = 0; e < 4; ++e) { af[i][e] = f0[e]; af[i][4 + e] = f1[e]; }
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const u32x4 v0 = Bs[wn * 64 + j * 16 + fr][2 * fq], v1 = Bs[wn * 64 + j * 16 + fr][2 * fq + 1];
        const f32x4 f0 = __builtin_bit_cast(f32x4, v0), f1 = __builtin_bit_cast(f32x4, v1);
        for (int e = 0; e < 4; ++e) { bf[j][e] = f0[e]; bf[j][4 + e] = f1[e]; }
      }
#pragma unroll
      for (int e = 0; e < 8; ++e)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = mfma_f32(bf[j][e], af[i][e], acc[i][j]);
    }
    __syncthreads();
  }
  const int rowb = m0 + wm * 64 + fr, colb = n0 + wn * 64;
  const double alpha = g.alpha, beta = g.beta;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rowb + i * 16, col = colb + j * 16 + mfma_f32_row(lane, r);
        if (row < g.M && col < g.N && (!g.upper || row <= col)) {
          int64_t co = row + (int64_t)col * g.ldc, ci = row + (int64_t)col * g.ldcin;
          if (g.cplx_n > 0) {
            const int part = col >= g.cplx_n ? 1 : 0, c = col - part * g.cplx_n;
            co = 2 * (row + (int64_t)c * g.ldc) + part;
            ci = 2 * (row + (int64_t)c * g.ldcin) + part;
          }
          const double v = alpha * (double)acc[i][j][r];
          g.C[co] = (beta != 0.0) ? fma(beta, g.Cin[ci], v) : v;
        }
      }
}

template <int PREC>
__global__ void __launch_bounds__(256, 2) gemm_lo_kernel(LoGemmArgs g) {
  __shared__ u32x4 As[kLoRows][9];
  __shared__ u32x4 Bs[kLoRows][9];
  if (gate_closed(g.gate, g.gate_want)) return;
  gemm_lo_body<PREC>(g, As, Bs);
}
// BF16 when *g16 == 1, else FP32 when *g32 == 1 (operands packed by the dual
// packs in that precision), else nothing.
__global__ void __launch_bounds__(256, 2) gemm_lo_dual_kernel(LoGemmArgs g, const int* g16, const int* g32) {
  __shared__ u32x4 As[kLoRows][9];
  __shared__ u32x4 Bs[kLoRows][9];
  if (*g16 == 1) gemm_lo_body<16>(g, As, Bs);
  else if (*g32 == 1) gemm_lo_body<32>(g, As, Bs);
}

template <int PREC>
void gemm_lo(int opA, int opB, const GemmArgs& g, void* ws, hipStream_t s) {
  using TL = typename LoType<PREC>::T;
  const int64_t Mp = lo_rows(g.M), Np = lo_rows(g.N), Kp = lo_kpad(g.K);
  TL* Ap = reinterpret_cast<TL*>(ws);
  TL* Bp = Ap + Mp * Kp;
  // op(A)(i, k): A[i + k lda] (N) or A[k + i lda] (T / C, real data)
  pack_lo<TL>(reinterpret_cast<const double*>(g.A), g.lda, opA != OP_N, g.M, g.K, Ap, Mp, Kp, s, g.gate,
              g.gate_want);
  // op(B)^T(j, k) = op(B)(k, j): B[k + j ldb] (N) or B[j + k ldb] (T / C)
  pack_lo<TL>(reinterpret_cast<const double*>(g.B), g.ldb, opB == OP_N, g.N, g.K, Bp, Np, Kp, s, g.gate,
              g.gate_want);
  LoGemmArgs a{};
  a.M = g.M; a.N = g.N; a.Kp = Kp;
  a.Ap = Ap; a.Bp = Bp;
  a.C = reinterpret_cast<double*>(g.C); a.ldc = g.ldc;
  a.Cin = g.Cin != nullptr ? reinterpret_cast<const double*>(g.Cin) : a.C;
  a.ldcin = g.Cin != nullptr ? g.ldcin : g.ldc;
  a.alpha = g.alpha_re; a.beta = g.beta_re;
  a.tri = g.tri;
  a.upper = g.upperC;
  a.gate = g.gate;
  a.gate_want = g.gate_want;
  dim3 grid((unsigned)(Mp / kLoRows), (unsigned)(Np / kLoRows), 1);
  hipLaunchKernelGGL((gemm_lo_kernel<PREC>), grid, dim3(256), 0, s, a);
}

// The BF16 and FP32 forms of gemm_lo (and of the complex product through the
// real embedding) in one set of launches:
// BF16 when *g16 == 1, else FP32 when *g32 == 1, else nothing (for a choice
// made on the device, so the form not taken costs no launches).  ws holds the
// FP32 size of the packed operands.
void gemm_lo_dual(int opA, int opB, const GemmArgs& g, const int* g16, const int* g32, void* ws, hipStream_t s) {
  const int64_t Mp = lo_rows(g.M), Np = lo_rows(g.N), Kp = lo_kpad(g.K);
  if (Mp <= 0 || Np <= 0 || Kp <= 0) return;
  const int ab = (int)(Mp / 64), bb = (int)(Np / 64);
  // op(A)(i, k): A[i + k lda] (N) or A[k + i lda] (T / C, real data);
  // op(B)^T(j, k) = op(B)(k, j): B[k + j ldb] (N) or B[j + k ldb] (T / C)
  hipLaunchKernelGGL(pack_lo_dual_kernel, dim3((unsigned)(Kp / 64), (unsigned)std::max(ab, bb), 2), dim3(256), 0, s,
                     reinterpret_cast<const double*>(g.A), g.lda, opA != OP_N ? 1 : 0, g.M,
                     reinterpret_cast<const double*>(g.B), g.ldb, opB == OP_N ? 1 : 0, g.N, g.K, ws, Mp * Kp, Kp,
                     ab, bb, g16, g32);
  LoGemmArgs a{};
  a.M = g.M; a.N = g.N; a.Kp = Kp;
  a.Ap = ws; a.Bp = nullptr; a.bOff = Mp * Kp;
  a.C = reinterpret_cast<double*>(g.C); a.ldc = g.ldc;
  a.Cin = g.Cin != nullptr ? reinterpret_cast<const double*>(g.Cin) : a.C;
  a.ldcin = g.Cin != nullptr ? g.ldcin : g.ldc;
  a.alpha = g.alpha_re; a.beta = g.beta_re;
  a.tri = g.tri;
  a.upper = g.upperC;
  hipLaunchKernelGGL(gemm_lo_dual_kernel, dim3((unsigned)(Mp / kLoRows), (unsigned)(Np / kLoRows), 1), dim3(256), 0,
                     s, a, g16, g32);
}
void gemm_lo_cplx_dual(const GemmArgs& g, const int* g16, const int* g32, void* ws, hipStream_t s) {
  const int64_t Mp = lo_rows(g.M), Np = lo_rows(2 * g.N), Kp = lo_kpad(2 * g.K);
  const int ab = (int)(Mp / 64), bb = (int)(Np / 64);
  hipLaunchKernelGGL(pack_cplx_dual_kernel, dim3((unsigned)(Kp / 64), (unsigned)std::max(ab, bb), 2), dim3(256), 0,
                     s, reinterpret_cast<const double*>(g.A), g.lda, g.M, reinterpret_cast<const double*>(g.B), g.ldb,
                     g.K, g.N, ws, Mp * Kp, Kp, ab, bb, g16, g32);
  LoGemmArgs a{};
  a.M = g.M; a.N = 2 * g.N; a.Kp = Kp;
  a.Ap = ws; a.Bp = nullptr; a.bOff = Mp * Kp;
  a.C = reinterpret_cast<double*>(g.C); a.ldc = g.ldc;
  a.Cin = g.Cin != nullptr ? reinterpret_cast<const double*>(g.Cin) : a.C;
  a.ldcin = g.Cin != nullptr ? g.ldcin : g.ldc;
  a.alpha = g.alpha_re; a.beta = g.beta_re;
  a.cplx_n = g.N;
  a.cplx_upperB = (g.tri & TRI_B_UPPER) ? 1 : 0;
  hipLaunchKernelGGL(gemm_lo_dual_kernel, dim3((unsigned)(Mp / kLoRows), (unsigned)(Np / kLoRows), 1), dim3(256), 0,
                     s, a, g16, g32);
}

// Y = X (M x N), when the gate is open.
__global__ void copy_gated_kernel(const double* X, int64_t ldx, double* Y, int64_t ldy, int M, int N,
                                  const int* gate, int want) {
  if (gate_closed(gate, want)) return;
  const int64_t tot = (int64_t)M * N;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) {
    const int64_t r = e % M, c = e / M;
    Y[r + c * ldy] = X[r + c * ldx];
  }
}

}  // namespace

size_t gemm_prec_workspace_bytes(int M, int N, int K) {
  return (size_t)(4 * (lo_rows(M) + lo_rows(N)) * lo_kpad(K));
}

// Precision-dispatched GEMM: 64 -> FP64 MFMA, 32 / 16 -> FP32 / BF16 MFMA
// (real data only; complex data always runs in FP64).  ws holds
// gemm_prec_workspace_bytes(M, N, K) bytes for the packed operand copies.
template <typename T> void gemm_prec(int prec, int opA, int opB, GemmArgs g, void* ws, hipStream_t s) {
  if constexpr (ScalarTraits<T>::is_complex) {
    gemm<T>(opA, opB, g, s);
  } else {
    if (g.M <= 0 || g.N <= 0) return;
    if (prec == 64) {
      if (g.Cin != nullptr && g.beta_re != 0.0 && g.Cin != g.C) {
        // the FP64 kernels take the beta term from C: copy Cin there first
        const int64_t tot = (int64_t)g.M * g.N;
        hipLaunchKernelGGL(copy_gated_kernel, dim3((unsigned)std::min<int64_t>(cdiv64(tot, 256), 4096)), dim3(256), 0,
                           s, reinterpret_cast<const double*>(g.Cin), g.ldcin, reinterpret_cast<double*>(g.C), g.ldc,
                           g.M, g.N, g.gate, g.gate_want);
      }
      gemm<T>(opA, opB, g, s);
      return;
    }
    if (prec == 16) gemm_lo<16>(opA, opB, g, ws, s);
    else gemm_lo<32>(opA, opB, g, ws, s);
  }
}
template void gemm_prec<double>(int, int, int, GemmArgs, void*, hipStream_t);
template void gemm_prec<zd>(int, int, int, GemmArgs, void*, hipStream_t);

size_t gemm_lo_complex_workspace_bytes(int M, int N, int K) { return gemm_prec_workspace_bytes(M, 2 * N, 2 * K); }

int gemm_lo_reduced(const GemmArgs& g, int opA, int opB, const int* g16, const int* g32, void* ws, hipStream_t s,
                    bool cplx) {
  if (g.M <= 0 || g.N <= 0) return 0;
  if (cplx) {
    if (opA != OP_N || opB != OP_N || g.alpha_im != 0.0 || g.beta_im != 0.0 || g.batch > 1 || g.ksplit > 1 ||
        (g.tri & ~TRI_B_UPPER) != 0 || g.upperC || g.unitLowerA)
      return -1;
    gemm_lo_cplx_dual(g, g16, g32, ws, s);
  } else {
    if (g.batch > 1 || g.ksplit > 1 || g.unitLowerA) return -1;
    gemm_lo_dual(opA, opB, g, g16, g32, ws, s);
  }
  return 0;
}

}  // namespace mak
