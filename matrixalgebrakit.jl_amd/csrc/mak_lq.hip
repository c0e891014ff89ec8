// LQ and null-space factorisations on the QR path:
//   lq_compact! / lq_full!  through the QR of A^H (LQViaTransposedQR,
//                           src/implementations/lq.jl:283-310): A^H = Qt Rt,
//                           so A = Rt^H Qt^H, L = Rt^H, Q = Qt^H;
//   qr_null!                N = trailing m - min(m, n) columns of the full Q
//                           of the Householder QR (H_1 ... H_k [0; I],
//                           src/implementations/qr.jl:232-263);
//   lq_null!                N^H from the QR null space of A^H (lq.jl:312-318).
// The conjugate transposes are LDS-tiled copies; the factorisations are the
// QR driver's (mak_qr.hip), run on workspace copies.
#include <algorithm>

#include "mak_common.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

namespace {

inline size_t al256(size_t b) { return ((b + 255) / 256) * 256; }

// B (n x m) = A^H, A m x n; one 32 x 32 block per workgroup, staged in LDS so
// that both the reads and the writes run along columns.
template <typename T>
__global__ void __launch_bounds__(256) adjoint_kernel(int m, int n, const T* A, int64_t lda, T* B, int64_t ldb) {
  __shared__ T tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;  // block of A: rows r0.., cols c0..
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int q = ty; q < 32; q += 8) {
    const int r = r0 + tx, c = c0 + q;
    if (r < m && c < n) tile[q][tx] = A[r + (int64_t)c * lda];
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int br = c0 + tx, bc = r0 + q;  // B(br, bc) = conj(A(bc, br))
    if (br < n && bc < m) B[br + (int64_t)bc * ldb] = s_conj(tile[tx][q]);
  }
}

template <typename T> void adjoint_into(int m, int n, const T* A, int64_t lda, T* B, int64_t ldb, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL((adjoint_kernel<T>), dim3(cdiv(m, 32), cdiv(n, 32)), dim3(256), 0, s, m, n, A, lda, B, ldb);
}

// Workspace of the LQ of an m x n matrix: A^H (n x m), Qt (n x qc), Rt (rr x
// m) and the QR workspace of the n x m problem.
struct LqLayout {
  size_t offAt, offQt, offRt, offQr, bytes;
};
template <typename T> LqLayout lq_layout(int m, int n, int full) {
  const int k = std::min(m, n), qc = full ? n : k, rr = full ? n : k;
  LqLayout L{};
  size_t off = 0;
  auto take = [&](size_t b) { size_t o = off; off += al256(b); return o; };
  L.offAt = take((size_t)n * m * sizeof(T));
  L.offQt = take((size_t)n * qc * sizeof(T));
  L.offRt = take((size_t)rr * m * sizeof(T));
  L.offQr = take(qr_workspace_bytes<T>(n, m, full));
  L.bytes = off;
  return L;
}

// Workspace of the null space of an m x n matrix (QR form): the full Q (m x m)
// and the QR workspace; the LQ form adds A^H (its QR is of the n x m A^H).
struct NullLayout {
  size_t offAt, offQf, offQr, bytes;
};
template <typename T> NullLayout null_layout(int rows, int cols, int transposed) {
  NullLayout L{};
  size_t off = 0;
  auto take = [&](size_t b) { size_t o = off; off += al256(b); return o; };
  L.offAt = transposed ? take((size_t)rows * cols * sizeof(T)) : 0;
  L.offQf = take((size_t)rows * rows * sizeof(T));
  L.offQr = take(qr_workspace_bytes<T>(rows, cols, 1));
  L.bytes = off;
  return L;
}

}  // namespace

template <typename T> size_t lq_workspace_bytes(int m, int n, int full) {
  return lq_layout<T>(std::max(m, 0), std::max(n, 0), full).bytes;
}

// lq_{compact,full}!(A, (L, Q)): L m x k (m x n full) lower triangular, or
// null (not computed); Q k x n (n x n full) with orthonormal rows.  A is read
// (the QR runs on its conjugate transpose in the workspace).  positive: L has
// a real non-negative diagonal (the QR's positive gauge, transposed).
template <typename T>
int lq_driver(hipStream_t s, int full, int m, int n, const T* A, int64_t lda, T* L, int64_t ldl, T* Q, int64_t ldq,
              int positive, void* work, size_t work_bytes) {
  if (m < 0) return -2;
  if (n < 0) return -3;
  if ((int64_t)m * n > 0 && A == nullptr) return -4;
  if (lda < std::max(1, m)) return -5;
  const int k = std::min(m, n);
  const int lcols = full ? n : k, qrows = full ? n : k;
  if (L != nullptr && ldl < std::max(1, m)) return -7;
  if ((int64_t)qrows * n > 0 && Q == nullptr) return -8;
  if (ldq < std::max(1, qrows)) return -9;
  if (ranges_overlap<T>(A, lda, m, n, Q, ldq, qrows, n)) return -8;
  if (L != nullptr && ranges_overlap<T>(A, lda, m, n, L, ldl, m, lcols)) return -6;
  const size_t need = lq_workspace_bytes<T>(m, n, full);
  if (work == nullptr || work_bytes < need) return -11;
  char* ws = reinterpret_cast<char*>(work);
  const LqLayout Ly = lq_layout<T>(m, n, full);
  T* At = reinterpret_cast<T*>(ws + Ly.offAt);
  T* Qt = reinterpret_cast<T*>(ws + Ly.offQt);
  T* Rt = reinterpret_cast<T*>(ws + Ly.offRt);
  const int rr = full ? n : k;
  adjoint_into<T>(m, n, A, lda, At, std::max(1, n), s);
  const int st = qr_driver<T>(s, full, n, m, At, std::max(1, n), Qt, std::max(1, n), L ? Rt : nullptr,
                              std::max(1, rr), positive, ws + Ly.offQr, need - Ly.offQr, nullptr);
  if (st != 0) return st < 0 ? 1 : st;  // the inner call's arguments are this driver's own
  adjoint_into<T>(n, qrows, Qt, std::max(1, n), Q, ldq, s);         // Q = Qt^H (qrows x n)
  if (L) adjoint_into<T>(rr, m, Rt, std::max(1, rr), L, ldl, s);     // L = Rt^H (m x rr)
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

template <typename T> size_t null_workspace_bytes(int m, int n, int lq) {
  m = std::max(m, 0);
  n = std::max(n, 0);
  return lq ? null_layout<T>(n, m, 1).bytes : null_layout<T>(m, n, 0).bytes;
}

// qr_null!(A, N): N m x (m - k); lq_null!(A, Nh): Nh (n - k) x n.  A is
// destroyed (qr_null!, as the reference's geqrt!) or read (lq_null!).
template <typename T>
int null_driver(hipStream_t s, int lq, int m, int n, T* A, int64_t lda, T* N, int64_t ldn, void* work,
                size_t work_bytes) {
  if (m < 0) return -2;
  if (n < 0) return -3;
  if ((int64_t)m * n > 0 && A == nullptr) return -4;
  if (lda < std::max(1, m)) return -5;
  const int k = std::min(m, n);
  const int rows = lq ? n : m, cols = lq ? m : n;  // the QR problem (A^H for lq)
  const int nn = rows - k;                          // null-space dimension
  const int nrows = lq ? nn : m, ncols = lq ? n : nn;
  if ((int64_t)nrows * ncols > 0 && N == nullptr) return -6;
  if (ldn < std::max(1, nrows)) return -7;
  if (ranges_overlap<T>(A, lda, m, n, N, ldn, nrows, ncols)) return -6;
  const size_t need = null_workspace_bytes<T>(m, n, lq);
  if (work == nullptr || work_bytes < need) return -9;
  if (nn == 0) return 0;
  char* ws = reinterpret_cast<char*>(work);
  const NullLayout Ly = null_layout<T>(rows, cols, lq);
  T* Qf = reinterpret_cast<T*>(ws + Ly.offQf);
  T* Aq = A;
  int64_t ldaq = lda;
  if (lq) {
    Aq = reinterpret_cast<T*>(ws + Ly.offAt);
    ldaq = std::max(1, rows);
    adjoint_into<T>(m, n, A, lda, Aq, ldaq, s);
  }
  // the full Q without gauge: its trailing columns are H_1 ... H_k [0; I]
  const int st = qr_driver<T>(s, 1, rows, cols, Aq, ldaq, Qf, std::max(1, rows), nullptr, 1, 0, ws + Ly.offQr,
                              need - Ly.offQr, nullptr);
  if (st != 0) return st < 0 ? 1 : st;
  const T* Qn = Qf + (int64_t)k * rows;  // columns k .. rows-1
  if (lq) adjoint_into<T>(rows, nn, Qn, rows, N, ldn, s);
  else copy_mat<T>(rows, nn, Qn, rows, N, ldn, s);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

template size_t lq_workspace_bytes<double>(int, int, int);
template size_t lq_workspace_bytes<zd>(int, int, int);
template size_t null_workspace_bytes<double>(int, int, int);
template size_t null_workspace_bytes<zd>(int, int, int);
template int lq_driver<double>(hipStream_t, int, int, int, const double*, int64_t, double*, int64_t, double*, int64_t,
                               int, void*, size_t);
template int lq_driver<zd>(hipStream_t, int, int, int, const zd*, int64_t, zd*, int64_t, zd*, int64_t, int, void*,
                           size_t);
template int null_driver<double>(hipStream_t, int, int, int, double*, int64_t, double*, int64_t, void*, size_t);
template int null_driver<zd>(hipStream_t, int, int, int, zd*, int64_t, zd*, int64_t, void*, size_t);

}  // namespace mak

extern "C" {

int mak_lq_workspace_size(int dtype, int full, int m, int n, size_t* bytes) {
  if (bytes == nullptr) return -5;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (dtype == MAK_F64) *bytes = mak::lq_workspace_bytes<double>(m, n, full);
  else if (dtype == MAK_C64) *bytes = mak::lq_workspace_bytes<mak::zd>(m, n, full);
  else return -1;
  return 0;
}

int mak_lq_compact_d(void* s, int m, int n, const double* A, int lda, double* L, int ldl, double* Q, int ldq,
                     int positive, void* work, size_t work_bytes) {
  return mak::lq_driver<double>((hipStream_t)s, 0, m, n, A, lda, L, ldl, Q, ldq, positive, work, work_bytes);
}
int mak_lq_compact_z(void* s, int m, int n, const void* A, int lda, void* L, int ldl, void* Q, int ldq, int positive,
                     void* work, size_t work_bytes) {
  return mak::lq_driver<mak::zd>((hipStream_t)s, 0, m, n, (const mak::zd*)A, lda, (mak::zd*)L, ldl, (mak::zd*)Q, ldq,
                                 positive, work, work_bytes);
}
int mak_lq_full_d(void* s, int m, int n, const double* A, int lda, double* L, int ldl, double* Q, int ldq,
                  int positive, void* work, size_t work_bytes) {
  return mak::lq_driver<double>((hipStream_t)s, 1, m, n, A, lda, L, ldl, Q, ldq, positive, work, work_bytes);
}
int mak_lq_full_z(void* s, int m, int n, const void* A, int lda, void* L, int ldl, void* Q, int ldq, int positive,
                  void* work, size_t work_bytes) {
  return mak::lq_driver<mak::zd>((hipStream_t)s, 1, m, n, (const mak::zd*)A, lda, (mak::zd*)L, ldl, (mak::zd*)Q, ldq,
                                 positive, work, work_bytes);
}

int mak_null_workspace_size(int dtype, int lq, int m, int n, size_t* bytes) {
  if (bytes == nullptr) return -5;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (dtype == MAK_F64) *bytes = mak::null_workspace_bytes<double>(m, n, lq);
  else if (dtype == MAK_C64) *bytes = mak::null_workspace_bytes<mak::zd>(m, n, lq);
  else return -1;
  return 0;
}

int mak_qr_null_d(void* s, int m, int n, double* A, int lda, double* N, int ldn, void* work, size_t work_bytes) {
  return mak::null_driver<double>((hipStream_t)s, 0, m, n, A, lda, N, ldn, work, work_bytes);
}
int mak_qr_null_z(void* s, int m, int n, void* A, int lda, void* N, int ldn, void* work, size_t work_bytes) {
  return mak::null_driver<mak::zd>((hipStream_t)s, 0, m, n, (mak::zd*)A, lda, (mak::zd*)N, ldn, work, work_bytes);
}
int mak_lq_null_d(void* s, int m, int n, const double* A, int lda, double* Nh, int ldnh, void* work,
                  size_t work_bytes) {
  return mak::null_driver<double>((hipStream_t)s, 1, m, n, const_cast<double*>(A), lda, Nh, ldnh, work, work_bytes);
}
int mak_lq_null_z(void* s, int m, int n, const void* A, int lda, void* Nh, int ldnh, void* work, size_t work_bytes) {
  return mak::null_driver<mak::zd>((hipStream_t)s, 1, m, n, (mak::zd*)const_cast<void*>(A), lda, (mak::zd*)Nh, ldnh,
                                   work, work_bytes);
}

}  // extern "C"
