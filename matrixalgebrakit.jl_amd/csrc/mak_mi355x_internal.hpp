// Internal declarations shared by the library's host drivers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "mak_mi355x.h"
#include "mak_common.hpp"

namespace mak {

// Does [p, p + span) intersect [q, q + span2)?  (element spans of two
// column-major matrices)
template <typename T>
inline bool ranges_overlap(const T* p, int64_t ld, int rows, int cols, const T* q, int64_t ldq, int qrows,
                           int qcols) {
  if (p == nullptr || q == nullptr || rows <= 0 || cols <= 0 || qrows <= 0 || qcols <= 0) return false;
  const T* pe = p + (cols - 1) * ld + rows;
  const T* qe = q + (qcols - 1) * ldq + qrows;
  return p < qe && q < pe;
}

template <typename T> size_t qr_workspace_bytes(int m, int n, int full);
template <typename T> void geqrf_blocked(T* A, int64_t lda, int m, int n, char* ws, size_t qcols, hipStream_t s);
template <typename T>
void orgqr_blocked(const T* A, int64_t lda, int m, int n, T* Q, int64_t ldq, int qcols, char* ws, hipStream_t s);
template <typename T>
int qr_driver(hipStream_t s, int full, int m, int n, T* A, int64_t lda, T* Q, int64_t ldq, T* R, int64_t ldr,
              int positive, void* work, size_t work_bytes, T* Tout);
template <typename T> void qr_record(const char* ws, int m, int n, int full, int out[9]);

// CholeskyQR2 panel with Householder reconstruction (mak_cholqr.hip): factors
// the H x nbc panel Ap (columns j0.. of the matrix) in place into compact WY
// form (T to Tpan) unless the device-side check sets the flag status[0], in
// which case Ap is untouched (and T not written) and status[2] = j0 for the
// first rejection.  status[1] is the kernels' ticket counter (zero between
// launches).  With sticky = 0 the flag is cleared for an accepted panel, so
// the caller's flag-gated Householder launches factor exactly the rejected
// panels; with sticky = 1 a set flag stays set (and skips every later panel),
// for a caller that checks it on the host.  cw: cholqr_workspace_elems(H).
// nbc_next > 0 (sticky mode only): also form the Gram of the next panel's
// nbc_next columns (over the same rows, before this panel's trailing update);
// carried = 1: take this panel's first Gram from the previous panel's call
// (which had nbc_next > 0) instead of a pass over the panel, once the previous
// trailing update has run; carried = 2: as 1, with R1 and R1^{-1} already
// formed from it by that trailing update (GemmArgs::cGw, at the addresses
// cholqr_carry_buffers gives).  For real data the next panel's Gram is left as
// cholqr_gram_blocks(H) partials, which that trailing update's split-K product
// sums into Gw (GemmArgs::rPart).
size_t cholqr_workspace_elems(int m);
template <typename T>
void cholqr_panel(T* Ap, int64_t lda, int H, int nbc, T* Tpan, T* cw, int* status, int j0, int sticky,
                  int carried, int nbc_next, hipStream_t s);
template <typename T> void cholqr_carry_buffers(T* cw, T** Gw, T** partN, T** R1, T** R1inv);
// Workgroups (and Gram partials) of the panel kernels for an H-row panel.
int cholqr_gram_blocks(int H);
void zero_ints(int* p, int n, hipStream_t s);

// Whole-matrix CholeskyQR2 for qr_compact!(positive = true) (mak_cholqr2.hip):
// eligible for m >= n, n <= kCholTileMaxN.  cholqr2_enqueue reads A only and
// writes Q / R (R may be null) when the device-side check accepts; the status
// word (cholqr2_status: [0] = 1 accepted, [2] Cholesky info, [4..5] ||E||_F
// as a double) is valid once the stream reaches it.  cholqr2_scratch: its
// m x n scratch (free after the call; the in-place Q mode reuses it).
bool cholqr2_eligible(int m, int n);
template <typename T> size_t cholqr2_workspace_bytes(int m, int n);
template <typename T>
void cholqr2_enqueue(hipStream_t s, int m, int n, const T* A, int64_t lda, T* Q, int64_t ldq, T* R, int64_t ldr,
                     char* ws);
template <typename T> int* cholqr2_status(char* ws, int m, int n);
template <typename T> T* cholqr2_scratch(char* ws, int m, int n);

template <typename T> size_t polar_workspace_bytes(int m, int n);
template <typename T>
int polar_driver(hipStream_t s, int right, int m, int n, const T* A, int64_t lda, T* W, int64_t ldw, T* P,
                 int64_t ldp, int precision, int maxiter, double l0, void* work, size_t work_bytes,
                 mak_polar_info* info);

}  // namespace mak
