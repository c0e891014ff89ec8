// Blocked Householder QR driver (qr_compact! / qr_full!) and its C-ABI.
//
// Mirrors MatrixAlgebraKit's qr_householder! (src/implementations/qr.jl:132-188):
// factor A in place into compact WY form (V strictly below the diagonal, R on
// and above, one 32x32 T per panel, like LAPACK geqrt), form Q by applying the
// block reflectors to the identity backwards (orgqr order), copy triu(A) into
// R and apply the positive gauge (src/common/gauge.jl:16-25).
//
// Per panel of 32 columns:
//   * rows <= node capacity: one geqr2 node (registers + LDS, one workgroup);
//   * taller: TSQR tree of geqr2 nodes, root finish (LU of Q - D), expansion;
//   * trailing update C <- C - V (T^H (V^H C)) as split-K MFMA GEMM, a small
//     reduction with T^H, and an MFMA rank-32 update.
#include <algorithm>
#include <vector>

#include "mak_common.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

namespace {

constexpr int NB = 32;

struct Level {
  int H;       // stacked rows
  int nodes;
  int chunk;   // 0 = balanced (leaves)
  size_t offX, offQ, offT;  // workspace offsets (bytes); X/Q unused at level 0
};

struct TreePlan {
  std::vector<Level> levels;  // levels[0] = leaves
  size_t offV1 = 0, offM = 0, offU = 0;
  size_t bytes = 0;           // bytes used by the tree (relative to its base)
};

template <typename T> TreePlan plan_tree(int H, int nbc, size_t base) {
  TreePlan p;
  const int cap = geqr2_node_capacity<T>();
  const int fan = cap / NB;
  size_t off = base;
  auto take = [&](size_t elems) {
    size_t o = off;
    off += ((elems * sizeof(T) + 255) / 256) * 256;
    return o;
  };
  Level l0{H, cdiv(H, cap), 0, 0, 0, 0};
  l0.offT = take((size_t)l0.nodes * NB * NB);
  p.levels.push_back(l0);
  p.offV1 = take((size_t)l0.nodes * NB * NB);
  while (p.levels.back().nodes > 1) {
    const Level& prev = p.levels.back();
    Level l;
    l.H = prev.nodes * nbc;
    l.chunk = fan * nbc;
    l.nodes = cdiv(prev.nodes, fan);
    l.offX = take((size_t)l.H * nbc);
    l.offQ = take((size_t)l.H * nbc);
    l.offT = take((size_t)l.nodes * NB * NB);
    p.levels.push_back(l);
  }
  p.offM = take(NB * NB);
  p.offU = take(NB * NB);
  p.bytes = off - base;
  return p;
}

// Panels at least this tall go through the CholeskyQR2 path first.
constexpr int kCholQrMinRows = 64;

template <typename T> struct QrLayout {
  size_t offTpanels = 0;  // NB*NB per panel
  size_t offCq = 0, offFlags = 0;  // CholeskyQR2 scratch; flag + counter
  size_t offTree = 0;
  size_t offWp = 0, offW2 = 0;
  size_t wpElems = 0, w2Elems = 0;
  // 128-column block reflectors of Q's formation (kGroup panels):
  size_t offVc = 0, offS = 0, offTg = 0, offWg = 0;
  size_t offTs = 0;   // T of every pair of groups (kSGB x kSGB each)
  size_t offQs = 0;   // m x min(m, n) scratch: Q of the in-place mode (Q === A)
  size_t offCq2 = 0;  // whole-matrix CholeskyQR2 (qr_compact!, m >= n), 0 when absent
  size_t bytes = 0;
};

inline size_t align256(size_t b) { return ((b + 255) / 256) * 256; }

// Q is formed by block reflectors of kGroup panels (kGB columns); on large
// problems pairs of groups are merged into one reflector of kSGB columns
// (half the passes over Q, and K = 256 for its update).
constexpr int kGroup = 4;
constexpr int kGB = kGroup * NB;
constexpr int kSGB = 2 * kGB;
constexpr int64_t kSuperGroupMinQ = int64_t(1) << 23;  // m * qcols from which pairs are merged

// Split-K slices (rows x N x ksplit elements) of the K = m products with
// `rows` rows, maximised over the N <= ncmax that occur (ksplit * N peaks at a
// column count below ncmax, so every N is tried).
size_t wp_block_elems(int m, int ncmax, int rows) {
  size_t most = 0;
  for (int N = 1; N <= std::max(ncmax, kGB); ++N) {
    int ks, kc;
    splitk_plan(N, m, 128, ks, kc);
    most = std::max(most, (size_t)ks * N);
  }
  return (size_t)rows * most;
}

template <typename T> QrLayout<T> qr_layout(int m, int n, int qcols) {
  QrLayout<T> L;
  const int minmn = std::min(m, n);
  const int npan = cdiv(std::max(minmn, 1), NB);
  size_t off = 0;
  L.offTpanels = off;
  off += align256((size_t)npan * NB * NB * sizeof(T));
  if (m >= kCholQrMinRows) {
    L.offCq = off;
    off += align256(cholqr_workspace_elems(m) * sizeof(T));
    L.offFlags = off;
    off += 256;
  }
  L.offTree = off;
  if (m > geqr2_node_capacity<T>()) {
    TreePlan tp = plan_tree<T>(m, NB, 0);
    off += align256(tp.bytes);
  }
  const int ncmax = std::max(n, qcols);
  int ks, kc;
  splitk_plan(ncmax, m, 64, ks, kc);
  L.wpElems = (size_t)NB * ncmax * std::max(ks, 1);
  // ksplit grows as N shrinks: bound ksplit * N over every N <= ncmax
  L.wpElems = std::max(L.wpElems, (size_t)NB * (size_t)(kSplitTargetWgs * 64 + ncmax));
  if (minmn > NB) L.wpElems = std::max(L.wpElems, wp_block_elems(m, ncmax, minmn > kGB ? kSGB : kGB));
  L.offWp = off;
  off += align256(L.wpElems * sizeof(T));
  L.w2Elems = (size_t)NB * ncmax;
  if (minmn > NB) L.w2Elems = std::max(L.w2Elems, (size_t)(minmn > kGB ? kSGB : kGB) * ncmax);
  L.offW2 = off;
  off += align256(L.w2Elems * sizeof(T));
  if (minmn > NB) {
    // V of every panel (unit lower), then per group of kGroup panels S = V^H V
    // and the merged T, so all groups' T are merged in one launch
    const size_t ngr = (size_t)cdiv(minmn, kGB);
    L.offVc = off;
    off += align256((size_t)m * ngr * kGB * sizeof(T));
    L.offS = off;
    off += align256(ngr * kGB * kGB * sizeof(T));
    L.offTg = off;
    off += align256(ngr * kGB * kGB * sizeof(T));
    L.offWg = off;
    off += align256((size_t)(minmn > kGB ? kSGB : kGB) * ncmax * sizeof(T));
    if (minmn > kGB) {
      L.offTs = off;
      off += align256((size_t)cdiv((int)ngr, 2) * kSGB * kSGB * sizeof(T));
    }
  }
  L.offQs = off;
  off += align256((size_t)m * minmn * sizeof(T));
  if (qcols == minmn && cholqr2_eligible(m, n)) {
    L.offCq2 = off;
    off += align256(cholqr2_workspace_bytes<T>(m, n));
  }
  L.bytes = off;
  return L;
}

// How panels are factored:
//   kHouseholderOnly  Householder node / TSQR launches only;
//   kCholGated        CholeskyQR2 first, then the Householder launches gated on
//                     its flag (they exit at once unless the panel was
//                     rejected): no host synchronisation, so a stream under
//                     graph capture can record it;
//   kCholHostChecked  CholeskyQR2 alone with a sticky flag that also skips every
//                     later panel; the caller reads the flag once on the host
//                     and refactors from the first rejected panel.
enum PanelMode { kHouseholderOnly = 0, kCholGated = 1, kCholHostChecked = 2 };

// carried / nbc_next: the carried first Gram (cholqr_panel), host-checked
// CholeskyQR2 panels only.
template <typename T>
void factor_panel(T* A, int64_t lda, int j0, int m, int nbc, char* ws, const QrLayout<T>& L, int mode,
                  hipStream_t s, int carried = 0, int nbc_next = 0) {
  T* Ap = A + j0 + (int64_t)j0 * lda;
  const int H = m - j0;
  T* Tpan = reinterpret_cast<T*>(ws + L.offTpanels) + (size_t)(j0 / NB) * NB * NB;
  const int cap = geqr2_node_capacity<T>();
  int* status = L.offFlags ? reinterpret_cast<int*>(ws + L.offFlags) : nullptr;
  const int* gate = nullptr;
  int want = 1;
  if (mode != kHouseholderOnly && status != nullptr && H >= kCholQrMinRows) {
    const bool hc = mode == kCholHostChecked;
    cholqr_panel<T>(Ap, lda, H, nbc, Tpan, reinterpret_cast<T*>(ws + L.offCq), status, j0, hc ? 1 : 0,
                    hc ? carried : 0, hc ? nbc_next : 0, s);
    if (mode == kCholHostChecked) return;
    gate = status;  // the Householder launches below run only for a rejected panel
  } else if (mode == kCholHostChecked && status != nullptr) {
    gate = status;  // short panel: skipped once an earlier panel was rejected
    want = 0;
  }
  if (H <= cap) {
    Geqr2NodeArgs g{};
    g.X = Ap; g.ldx = lda; g.H = H; g.nodes = 1; g.chunk = 0; g.nbc = nbc;
    g.Tout = Tpan; g.tstride = NB * NB;
    g.gate = gate; g.gate_want = want;
    launch_geqr2_node<T>(g, cap, s);
    return;
  }
  char* tb = ws + L.offTree;
  TreePlan tp = plan_tree<T>(H, nbc, 0);
  auto ptr = [&](size_t off) { return reinterpret_cast<T*>(tb + off); };
  const int top = (int)tp.levels.size() - 1;
  // factor every level bottom-up
  for (int lev = 0; lev <= top; ++lev) {
    const Level& lv = tp.levels[lev];
    Geqr2NodeArgs g{};
    if (lev == 0) { g.X = Ap; g.ldx = lda; }
    else { g.X = ptr(lv.offX); g.ldx = lv.H; }
    g.H = lv.H; g.nodes = lv.nodes; g.chunk = lv.chunk; g.nbc = nbc;
    g.Tout = ptr(lv.offT); g.tstride = NB * NB;
    if (lev < top) { g.Rnext = ptr(tp.levels[lev + 1].offX); g.ldr = tp.levels[lev + 1].H; }
    if (lev == 0) g.V1out = ptr(tp.offV1);
    g.gate = gate; g.gate_want = want;
    launch_geqr2_node<T>(g, cap, s);
  }
  // root finish
  {
    const Level& lt = tp.levels[top];
    TsqrRootArgs r{};
    r.Xtop = ptr(lt.offX); r.ldxtop = lt.H; r.htop = lt.H; r.nbc = nbc;
    r.Ttop = ptr(lt.offT);
    r.nchain = top;
    for (int lev = 0; lev < top; ++lev) {
      if (lev == 0) { r.chainV1[0] = ptr(tp.offV1); r.chainLd[0] = NB; }
      else { r.chainV1[lev] = ptr(tp.levels[lev].offX); r.chainLd[lev] = tp.levels[lev].H; }
      r.chainT[lev] = ptr(tp.levels[lev].offT);
    }
    r.Mtop = ptr(tp.offM); r.Tout = Tpan; r.Uout = ptr(tp.offU); r.Apanel = Ap; r.lda = lda;
    r.gate = gate; r.gate_want = want;
    launch_tsqr_root_finish<T>(r, s);
  }
  // expansion top-down
  for (int lev = top - 1; lev >= 0; --lev) {
    const Level& lv = tp.levels[lev];
    TsqrExpandArgs e{};
    e.H = lv.H; e.nodes = lv.nodes; e.chunk = lv.chunk; e.nbc = nbc;
    if (lev == 0) { e.X = Ap; e.ldx = lda; e.V1base = ptr(tp.offV1); e.v1stride = NB * NB; e.v1ld = NB; }
    else { e.X = ptr(lv.offX); e.ldx = lv.H; e.V1base = ptr(lv.offX); e.v1stride = lv.chunk; e.v1ld = lv.H; }
    e.Tnodes = ptr(lv.offT);
    if (lev == top - 1) {
      const Level& lt = tp.levels[top];
      e.Mtop = ptr(tp.offM); e.Xtop = ptr(lt.offX); e.ldxtop = lt.H;
    } else {
      e.Qparent = ptr(tp.levels[lev + 1].offQ); e.ldqp = tp.levels[lev + 1].H;
    }
    if (lev == 0) e.Umat = ptr(tp.offU);
    else { e.Qout = ptr(lv.offQ); e.ldqo = lv.H; }
    const int maxrows = (lev == 0) ? cdiv(lv.H, lv.nodes) + 1 : lv.chunk;
    e.gate = gate; e.gate_want = want;
    launch_tsqr_expand<T>(e, maxrows, s);
  }
}

// C (rows x ncols) <- C - V op(T) (V^H C); V is the unit-lower panel at Vp.
template <typename T>
// carry_nbc > 0 (real data, host-checked CholeskyQR2): the update also forms
// the next panel's R1 from its carried Gram (GemmArgs::cGw).
void apply_wy(const T* Vp, int64_t ldv, int rows, int nbc, const T* Tm, int transT, T* C, int64_t ldc,
              int ncols, char* ws, const QrLayout<T>& L, hipStream_t s, int carry_nbc = 0, int carry_j0 = 0) {
  if (ncols <= 0 || rows <= 0) return;
  T* Wp = reinterpret_cast<T*>(ws + L.offWp);
  T* W2 = reinterpret_cast<T*>(ws + L.offW2);
  int ks, kc;
  splitk_plan(ncols, rows, 64, ks, kc);
  const int64_t slice = (int64_t)nbc * ncols;
  {
    GemmArgs g{};
    g.M = nbc; g.N = ncols; g.K = rows;
    g.A = Vp; g.lda = ldv; g.B = C; g.ldb = ldc; g.C = Wp; g.ldc = nbc;
    g.alpha_re = 1.0; g.beta_re = 0.0;
    g.ksplit = ks; g.kchunk = kc; g.splitStride = slice; g.batch = 1; g.unitLowerA = 1;
    if (carry_nbc > 0 && !ScalarTraits<T>::is_complex) {  // sum the next panel's Gram partials on the side
      T *Gw, *partN, *R1, *R1i;
      cholqr_carry_buffers<T>(reinterpret_cast<T*>(ws + L.offCq), &Gw, &partN, &R1, &R1i);
      g.rPart = partN; g.rOut = Gw; g.rN = cholqr_gram_blocks(rows);
    }
    gemm<T>(OP_C, OP_N, g, s);
  }
  wy_reduce<T>(Wp, ks, slice, Tm, transT, nbc, ncols, W2, s);
  {
    GemmArgs g{};
    g.M = rows; g.N = ncols; g.K = nbc;
    g.A = Vp; g.lda = ldv; g.B = W2; g.ldb = nbc; g.C = C; g.ldc = ldc;
    g.alpha_re = -1.0; g.beta_re = 1.0;
    g.ksplit = 1; g.kchunk = nbc; g.batch = 1; g.unitLowerA = 1;
    if (carry_nbc > 0 && !ScalarTraits<T>::is_complex) {
      T *Gw, *partN, *R1, *R1i;
      cholqr_carry_buffers<T>(reinterpret_cast<T*>(ws + L.offCq), &Gw, &partN, &R1, &R1i);
      g.cGw = Gw; g.cR1 = R1; g.cR1i = R1i;
      g.cStatus = reinterpret_cast<int*>(ws + L.offFlags);
      g.cJ0 = carry_j0; g.cNbc = carry_nbc;
    }
    gemm<T>(OP_N, OP_N, g, s);
  }
}

}  // namespace

template <typename T> size_t qr_workspace_bytes(int m, int n, int full) {
  const int minmn = std::min(m, n);
  const int qcols = full ? m : minmn;
  return qr_layout<T>(m, n, qcols).bytes;
}
template size_t qr_workspace_bytes<double>(int, int, int);
template size_t qr_workspace_bytes<zd>(int, int, int);

// Panels from column `from` on, each followed by its trailing update.
template <typename T>
void geqrf_panels(T* A, int64_t lda, int m, int n, char* ws, const QrLayout<T>& L, int from, int mode,
                  hipStream_t s) {
  const int minmn = std::min(m, n);
  const bool chol = mode == kCholHostChecked && L.offFlags != 0;
  int carried = 0;
  for (int j0 = from; j0 < minmn; j0 += NB) {
    const int nbc = std::min(NB, minmn - j0);
    // consecutive CholeskyQR2 panels pass the next panel's first Gram along
    const int nbc_next = (chol && m - j0 >= kCholQrMinRows && j0 + nbc < minmn && m - j0 - nbc >= kCholQrMinRows)
                             ? std::min(NB, minmn - j0 - nbc) : 0;
    factor_panel<T>(A, lda, j0, m, nbc, ws, L, mode, s, carried, nbc_next);
    // real data: the trailing update also factors the next panel's carried
    // Gram (2); complex: the next finish kernel does (1)
    carried = nbc_next > 0 ? (ScalarTraits<T>::is_complex ? 1 : 2) : 0;
    const int nc = n - j0 - nbc;
    if (nc > 0) {
      const T* Tpan = reinterpret_cast<const T*>(ws + L.offTpanels) + (size_t)(j0 / NB) * NB * NB;
      apply_wy<T>(A + j0 + (int64_t)j0 * lda, lda, m - j0, nbc, Tpan, 1, A + j0 + (int64_t)(j0 + nbc) * lda, lda,
                  nc, ws, L, s, carried == 2 ? nbc_next : 0, j0 + nbc);
    }
  }
}

// Enqueues the factorisation.  Outside graph capture the panels run in
// kCholHostChecked mode and the function returns true: the caller may enqueue
// more work that reads the factors, and must then call geqrf_finish, which
// synchronises the stream once to read the CholeskyQR2 flag.  Under capture
// the panels run in kCholGated mode (nothing to check; returns false).
//
// In kCholHostChecked mode the trailing updates are not gated: the T blocks
// start at zero and a skipped panel leaves its T at zero, which makes every
// update with it an exact no-op (C - V 0 V^H C = C).
template <typename T>
bool geqrf_enqueue(T* A, int64_t lda, int m, int n, char* ws, size_t qcols, hipStream_t s) {
  const int minmn = std::min(m, n);
  QrLayout<T> L = qr_layout<T>(m, n, (int)qcols);
  int* status = L.offFlags ? reinterpret_cast<int*>(ws + L.offFlags) : nullptr;
  int mode = kHouseholderOnly;
  if (status != nullptr) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) cs = hipStreamCaptureStatusActive;
    mode = cs == hipStreamCaptureStatusNone ? kCholHostChecked : kCholGated;
    zero_ints(status, 3, s);  // flag, Gram ticket counter, first rejected column
    if (mode == kCholHostChecked) {
      const int npan = cdiv(minmn, NB);
      zero_ints(reinterpret_cast<int*>(ws + L.offTpanels), (int)((size_t)npan * NB * NB * sizeof(T) / sizeof(int)), s);
    }
  }
  geqrf_panels<T>(A, lda, m, n, ws, L, 0, mode, s);
  return mode == kCholHostChecked;
}

// Reads the flag (synchronising s).  Panels before the first rejected one are
// final and the trailing matrix is updated up to it; every later panel was
// skipped, so the factorisation resumes there with Householder TSQR.  Returns
// 1 when it refactored (work enqueued after geqrf_enqueue must be redone), 0
// when every panel was accepted, -1 when the flag could not be read or is
// inconsistent (a HIP error: the factorisation is not known to be valid).
template <typename T>
int geqrf_finish(T* A, int64_t lda, int m, int n, char* ws, size_t qcols, hipStream_t s) {
  const int minmn = std::min(m, n);
  QrLayout<T> L = qr_layout<T>(m, n, (int)qcols);
  if (!L.offFlags) return 0;
  int h[3] = {0, 0, 0};
  if (hipMemcpyAsync(h, ws + L.offFlags, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -1;
  if (h[0] == 0) return 0;
  if (h[2] < 0 || h[2] >= minmn || h[2] % NB != 0) return -1;
  geqrf_panels<T>(A, lda, m, n, ws, L, h[2], kHouseholderOnly, s);
  return 1;
}

template <typename T>
void geqrf_blocked(T* A, int64_t lda, int m, int n, char* ws, size_t qcols, hipStream_t s) {
  if (geqrf_enqueue<T>(A, lda, m, n, ws, qcols, s)) geqrf_finish<T>(A, lda, m, n, ws, qcols, s);
}

// Columns from which the whole-matrix CholeskyQR2 path is tried first.
constexpr int kCholQr2MinN = 64;

// out (kg x N, ld kg) = V^H X for V = Vc (rows x kg, ld ldv): split-K
// product into the Wp slices, then their sum.
template <typename T>
void vh_times(const T* Vc, int64_t ldv, int rows, int kg, const T* X, int64_t ldx, int N, T* out, char* ws,
              const QrLayout<T>& L, hipStream_t s) {
  T* Wp = reinterpret_cast<T*>(ws + L.offWp);
  int ks, kc;
  splitk_plan(N, rows, 128, ks, kc);
  GemmArgs g{};
  g.M = kg; g.N = N; g.K = rows;
  g.A = Vc; g.lda = ldv; g.B = X; g.ldb = ldx; g.C = Wp; g.ldc = kg;
  g.alpha_re = 1.0; g.beta_re = 0.0;
  g.ksplit = ks; g.kchunk = kc; g.splitStride = (int64_t)kg * N; g.batch = 1;
  gemm<T>(OP_C, OP_N, g, s);
  sum_slices<T>(Wp, ks, (int64_t)kg * N, kg, N, out, kg, s);
}

// C (rows x ncols) <- C - V op(Tg) (V^H C) for a block reflector with V = Vc
// (unit lower, rows x kg, ld ldv) and its kg x kg T (ld kg): three GEMMs with
// K = rows, kg, kg.
template <typename T>
void apply_block_wy(const T* Vc, int64_t ldv, int rows, int kg, const T* Tg, int transT, T* C, int64_t ldc,
                    int ncols, char* ws, const QrLayout<T>& L, hipStream_t s) {
  if (ncols <= 0 || rows <= 0) return;
  T* Wg = reinterpret_cast<T*>(ws + L.offWg);
  T* W2 = reinterpret_cast<T*>(ws + L.offW2);
  vh_times<T>(Vc, ldv, rows, kg, C, ldc, ncols, Wg, ws, L, s);
  {
    GemmArgs g{};
    g.M = kg; g.N = ncols; g.K = kg;
    g.A = Tg; g.lda = kg; g.B = Wg; g.ldb = kg; g.C = W2; g.ldc = kg;
    g.alpha_re = 1.0; g.beta_re = 0.0;
    g.ksplit = 1; g.kchunk = kg; g.batch = 1;
    gemm<T>(transT ? OP_C : OP_N, OP_N, g, s);
  }
  {
    GemmArgs g{};
    g.M = rows; g.N = ncols; g.K = kg;
    g.A = Vc; g.lda = ldv; g.B = W2; g.ldb = kg; g.C = C; g.ldc = ldc;
    g.alpha_re = -1.0; g.beta_re = 1.0;
    g.ksplit = 1; g.kchunk = kg; g.batch = 1;
    gemm<T>(OP_N, OP_N, g, s);
  }
}

// Q = H_0 ... H_{npan-1} [I; 0], applied backwards in groups of kGroup panels
// (block reflectors of kGB columns).  Before the backward sweep: V of every
// panel is copied out unit lower in one launch, S = V^H V per group, and the
// T of every group is merged in one launch.
template <typename T>
void orgqr_blocked(const T* A, int64_t lda, int m, int n, T* Q, int64_t ldq, int qcols, char* ws, hipStream_t s) {
  const int minmn = std::min(m, n);
  QrLayout<T> L = qr_layout<T>(m, n, qcols);
  set_identity<T>(Q, ldq, m, qcols, s);
  const int npan = cdiv(minmn, NB);
  const int ngr = cdiv(npan, kGroup);
  const T* Tpanels = reinterpret_cast<const T*>(ws + L.offTpanels);
  T* Vc = reinterpret_cast<T*>(ws + L.offVc);
  T* Sg = reinterpret_cast<T*>(ws + L.offS);
  T* Tg = reinterpret_cast<T*>(ws + L.offTg);
  if (npan > 1) {
    copy_unit_lower<T>(A, lda, m, minmn, Vc, m, s);
    // the full groups' S in one batched split-K product over all m rows (the
    // rows above a group's first column are zero in Vc)
    const int nfull = minmn / kGB;
    if (nfull > 0) {
      int ks, kc;
      splitk_plan(kGB * nfull, m, 128, ks, kc);
      T* Wp = reinterpret_cast<T*>(ws + L.offWp);
      GemmArgs g{};
      g.M = kGB; g.N = kGB; g.K = m;
      g.A = Vc; g.lda = m; g.strideA = (int64_t)kGB * m;
      g.B = Vc; g.ldb = m; g.strideB = (int64_t)kGB * m;
      g.C = Wp; g.ldc = kGB; g.strideC = (int64_t)kGB * kGB;
      g.alpha_re = 1.0; g.beta_re = 0.0;
      g.ksplit = ks; g.kchunk = kc; g.splitStride = (int64_t)kGB * kGB * nfull; g.batch = nfull;
      gemm<T>(OP_C, OP_N, g, s);
      sum_slices<T>(Wp, ks, (int64_t)kGB * kGB * nfull, kGB, kGB * nfull, Sg, kGB, s);
    }
    for (int g = nfull; g < ngr; ++g) {
      const int p0 = g * kGroup, np = std::min(kGroup, npan - p0);
      if (np < 2) continue;
      const int j0 = p0 * NB;
      const int kg = std::min(minmn, (p0 + np) * NB) - j0;
      const T* Vg = Vc + j0 + (int64_t)j0 * m;
      vh_times<T>(Vg, m, m - j0, kg, Vg, m, kg, Sg + (size_t)g * kGB * kGB, ws, L, s);
    }
    wy_merge_groups<T>(Sg, Tpanels, npan, minmn, Tg, s);
  }
  // Pairs of groups (2G, 2G + 1) as one reflector of kg1 + kg2 columns:
  //   (I - V1 T1 V1^H)(I - V2 T2 V2^H) = I - [V1 V2] [T1  -T1 (V1^H V2) T2; 0  T2] [V1 V2]^H
  // (V1^H V2 over the rows where V2 is nonzero, split-K; the two small
  // products through the W2 scratch).
  const int nsg = cdiv(ngr, 2);
  const bool pairs = L.offTs != 0 && ngr >= 2 && (int64_t)m * qcols >= kSuperGroupMinQ;
  T* Ts = L.offTs != 0 ? reinterpret_cast<T*>(ws + L.offTs) : nullptr;
  auto group_cols = [&](int g) { return std::min(minmn, (g + 1) * kGB) - g * kGB; };
  if (pairs) {
    T* W2 = reinterpret_cast<T*>(ws + L.offW2);
    for (int G = 0; 2 * G + 1 < ngr; ++G) {
      const int g1 = 2 * G, g2 = g1 + 1;
      const int kg1 = kGB, kg2 = group_cols(g2), kgs = kg1 + kg2;
      const int j1 = g1 * kGB, j2 = g2 * kGB;
      T* Tsg = Ts + (size_t)G * kSGB * kSGB;  // ld kgs
      T* S12 = W2;                            // kg1 x kg2
      T* X12 = W2 + (size_t)kGB * kGB;        // kg1 x kg2
      const T* T1 = Tg + (size_t)g1 * kGB * kGB;  // ld kg1
      const T* T2 = Tg + (size_t)g2 * kGB * kGB;  // ld kg2
      vh_times<T>(Vc + j2 + (int64_t)j1 * m, m, m - j2, kg1, Vc + j2 + (int64_t)j2 * m, m, kg2, S12, ws, L, s);
      {
        GemmArgs g{};
        g.M = kg1; g.N = kg2; g.K = kg1;
        g.A = T1; g.lda = kg1; g.B = S12; g.ldb = kg1; g.C = X12; g.ldc = kg1;
        g.alpha_re = 1.0; g.beta_re = 0.0;
        g.ksplit = 1; g.kchunk = kg1; g.batch = 1;
        gemm<T>(OP_N, OP_N, g, s);
      }
      {
        GemmArgs g{};
        g.M = kg1; g.N = kg2; g.K = kg2;
        g.A = X12; g.lda = kg1; g.B = T2; g.ldb = kg2; g.C = Tsg + (int64_t)kg1 * kgs; g.ldc = kgs;
        g.alpha_re = -1.0; g.beta_re = 0.0;
        g.ksplit = 1; g.kchunk = kg2; g.batch = 1;
        gemm<T>(OP_N, OP_N, g, s);
      }
      copy_mat<T>(kg1, kg1, T1, kg1, Tsg, kgs, s);
      copy_mat<T>(kg2, kg2, T2, kg2, Tsg + kg1 + (int64_t)kg1 * kgs, kgs, s);
      fill_zero<T>(Tsg + kg1, kgs, kg2, kg1, s);
    }
  }
  auto apply_group = [&](int g) {
    const int p0 = g * kGroup, np = std::min(kGroup, npan - p0);
    const int j0 = p0 * NB;
    const int kg = std::min(minmn, (p0 + np) * NB) - j0;
    T* C = Q + j0 + (int64_t)j0 * ldq;
    if (np == 1)
      apply_wy<T>(A + j0 + (int64_t)j0 * lda, lda, m - j0, kg, Tpanels + (size_t)p0 * NB * NB, 0, C, ldq,
                  qcols - j0, ws, L, s);
    else
      apply_block_wy<T>(Vc + j0 + (int64_t)j0 * m, m, m - j0, kg, Tg + (size_t)g * kGB * kGB, 0, C, ldq,
                        qcols - j0, ws, L, s);
  };
  for (int G = nsg - 1; G >= 0; --G) {
    const int g1 = 2 * G, g2 = g1 + 1;
    if (pairs && g2 < ngr) {
      const int j0 = g1 * kGB, kgs = kGB + group_cols(g2);
      apply_block_wy<T>(Vc + j0 + (int64_t)j0 * m, m, m - j0, kgs, Ts + (size_t)G * kSGB * kSGB, 0,
                        Q + j0 + (int64_t)j0 * ldq, ldq, qcols - j0, ws, L, s);
    } else {
      if (g2 < ngr) apply_group(g2);
      apply_group(g1);
    }
  }
}

// Status words of the last call that used this workspace (read by the test
// suite through maktest_qr_record): the CholeskyQR2 panel flags (flag, ticket,
// first rejected column) and the whole-matrix CholeskyQR2 status.
template <typename T> void qr_record(const char* ws, int m, int n, int full, int out[9]) {
  const int minmn = std::min(m, n);
  QrLayout<T> L = qr_layout<T>(m, n, full ? m : minmn);
  int h[3] = {0, 0, 0};
  int st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (L.offFlags) hipMemcpy(h, ws + L.offFlags, sizeof(h), hipMemcpyDeviceToHost);
  if (L.offCq2) hipMemcpy(st, cholqr2_status<T>(const_cast<char*>(ws) + L.offCq2, m, n), sizeof(st),
                          hipMemcpyDeviceToHost);
  out[0] = st[0];                  // 1: the whole-matrix CholeskyQR2 result was taken
  out[1] = h[0] ? h[2] : -1;       // first panel handed to Householder TSQR (-1: none)
  out[2] = st[2];                  // Cholesky info of the whole-matrix path
  out[3] = L.offCq2 != 0;          // the whole-matrix path exists for this shape
  out[4] = st[4];                  // ||E||_F of the whole-matrix path (a double in out[4..5])
  out[5] = st[5];
  out[6] = st[3];                  // 1: its final product took the FP32 correction
  out[7] = st[6];                  // 1: its final product ran on FP64
  out[8] = st[7];                  // 1: its final product took the BF16 correction
}

template <typename T>
int qr_driver(hipStream_t s, int full, int m, int n, T* A, int64_t lda, T* Q, int64_t ldq, T* R, int64_t ldr,
              int positive, void* work, size_t work_bytes, T* Tout) {
  if (m < 0) return -2;
  if (n < 0) return -3;
  if ((int64_t)m * n > 0 && A == nullptr) return -4;
  if (lda < std::max(1, m)) return -5;
  const int minmn = std::min(m, n);
  const int qcols = full ? m : minmn;
  const int rrows = full ? m : minmn;
  if ((int64_t)m * qcols > 0 && Q == nullptr) return -6;
  if (ldq < std::max(1, m)) return -7;
  if (R != nullptr && ldr < std::max(1, rrows)) return -9;
  // Q === A: only for a tall input without R and without the positive gauge,
  // as in the reference (src/implementations/qr.jl:150-153); any other
  // overlap of Q with A is invalid
  const bool inplace = Q != nullptr && Q == A && ldq == lda && qcols == n && m >= n;
  if (inplace && (R != nullptr || positive)) return -6;
  if (!inplace && ranges_overlap<T>(A, lda, m, n, Q, ldq, m, qcols)) return -6;
  const size_t need = qr_workspace_bytes<T>(m, n, full);
  if (work == nullptr || work_bytes < need) return -11;
  char* ws = reinterpret_cast<char*>(work);
  QrLayout<T> L = qr_layout<T>(m, n, qcols);
  if (L.offCq2) zero_ints(cholqr2_status<T>(ws + L.offCq2, m, n), 8, s);
  if (L.offFlags) zero_ints(reinterpret_cast<int*>(ws + L.offFlags), 3, s);
  if (minmn == 0) {
    set_identity<T>(Q, ldq, m, qcols, s);
    if (R) qr_extract_r<T>(A, lda, R, ldr, rrows, n, 0, 0, s);
    return hipGetLastError() == hipSuccess ? 0 : 1;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) cs = hipStreamCaptureStatusActive;
  // qr_compact!(positive = true): whole-matrix CholeskyQR2 first (A is only
  // read; a rejected input goes on to the Householder path below)
  if (L.offCq2 && positive && !inplace && Tout == nullptr && n >= kCholQr2MinN && cs == hipStreamCaptureStatusNone) {
    char* cw = ws + L.offCq2;
    cholqr2_enqueue<T>(s, m, n, A, lda, Q, ldq, R, ldr, cw);
    int st = 0;
    if (hipMemcpyAsync(&st, cholqr2_status<T>(cw, m, n), sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return 1;
    if (st == 1) return hipGetLastError() == hipSuccess ? 0 : 1;
  }
  T* Qd = inplace ? reinterpret_cast<T*>(ws + L.offQs) : Q;
  const int64_t ldqd = inplace ? m : ldq;
  // Q, R and the tape are formed before the CholeskyQR2 flag is read, so the
  // one synchronisation does not leave the GPU waiting for the host; they are
  // formed again in the (rare) case that a panel was refactored.  In the
  // in-place mode Q replaces A's factors, so it is formed only after the check.
  auto outputs = [&]() {
    if (Tout) {  // keep the compact WY factors' T blocks (32 x 32 per panel) for the adjoint
      const int npan = cdiv(minmn, NB);
      hipMemcpyAsync(Tout, ws, sizeof(T) * NB * NB * npan, hipMemcpyDeviceToDevice, s);
    }
    if (R) qr_extract_r<T>(A, lda, R, ldr, rrows, n, minmn, positive, s);
    orgqr_blocked<T>(A, lda, m, n, Qd, ldqd, qcols, ws, s);
    if (positive) qr_gauge_q<T>(A, lda, Qd, ldqd, m, minmn, s);
  };
  const bool check = geqrf_enqueue<T>(A, lda, m, n, ws, qcols, s);
  if (!inplace) outputs();
  if (check) {
    const int f = geqrf_finish<T>(A, lda, m, n, ws, qcols, s);
    if (f < 0) return 1;
    if (f == 1 && !inplace) outputs();
  }
  if (inplace) {
    outputs();
    copy_mat<T>(m, qcols, Qd, ldqd, Q, ldq, s);
  }
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

template void geqrf_blocked<double>(double*, int64_t, int, int, char*, size_t, hipStream_t);
template void geqrf_blocked<zd>(zd*, int64_t, int, int, char*, size_t, hipStream_t);
template void orgqr_blocked<double>(const double*, int64_t, int, int, double*, int64_t, int, char*, hipStream_t);
template void orgqr_blocked<zd>(const zd*, int64_t, int, int, zd*, int64_t, int, char*, hipStream_t);
template int qr_driver<double>(hipStream_t, int, int, int, double*, int64_t, double*, int64_t, double*, int64_t,
                               int, void*, size_t, double*);
template int qr_driver<zd>(hipStream_t, int, int, int, zd*, int64_t, zd*, int64_t, zd*, int64_t, int, void*,
                           size_t, zd*);
template void qr_record<double>(const char*, int, int, int, int*);
template void qr_record<zd>(const char*, int, int, int, int*);

}  // namespace mak

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int mak_qr_workspace_size(int dtype, int full, int m, int n, size_t* bytes) {
  if (bytes == nullptr) return -5;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (dtype == MAK_F64) *bytes = mak::qr_workspace_bytes<double>(m, n, full);
  else if (dtype == MAK_C64) *bytes = mak::qr_workspace_bytes<mak::zd>(m, n, full);
  else return -1;
  return 0;
}

int mak_qr_compact_d(void* s, int m, int n, double* A, int lda, double* Q, int ldq, double* R, int ldr,
                     int positive, void* work, size_t work_bytes) {
  return mak::qr_driver<double>((hipStream_t)s, 0, m, n, A, lda, Q, ldq, R, ldr, positive, work, work_bytes, nullptr);
}
int mak_qr_compact_z(void* s, int m, int n, void* A, int lda, void* Q, int ldq, void* R, int ldr,
                     int positive, void* work, size_t work_bytes) {
  return mak::qr_driver<mak::zd>((hipStream_t)s, 0, m, n, (mak::zd*)A, lda, (mak::zd*)Q, ldq, (mak::zd*)R, ldr, positive,
                                 work, work_bytes, nullptr);
}
int mak_qr_full_d(void* s, int m, int n, double* A, int lda, double* Q, int ldq, double* R, int ldr,
                  int positive, void* work, size_t work_bytes) {
  return mak::qr_driver<double>((hipStream_t)s, 1, m, n, A, lda, Q, ldq, R, ldr, positive, work, work_bytes, nullptr);
}
int mak_qr_full_z(void* s, int m, int n, void* A, int lda, void* Q, int ldq, void* R, int ldr, int positive,
                  void* work, size_t work_bytes) {
  return mak::qr_driver<mak::zd>((hipStream_t)s, 1, m, n, (mak::zd*)A, lda, (mak::zd*)Q, ldq, (mak::zd*)R, ldr, positive,
                                 work, work_bytes, nullptr);
}

int mak_qr_compact_wy_d(void* s, int m, int n, double* A, int lda, double* Q, int ldq, double* R, int ldr,
                        double* Tout, int positive, void* work, size_t work_bytes) {
  if (Tout == nullptr) return -10;
  return mak::qr_driver<double>((hipStream_t)s, 0, m, n, A, lda, Q, ldq, R, ldr, positive, work, work_bytes, Tout);
}
int mak_qr_compact_wy_z(void* s, int m, int n, void* A, int lda, void* Q, int ldq, void* R, int ldr, void* Tout,
                        int positive, void* work, size_t work_bytes) {
  if (Tout == nullptr) return -10;
  return mak::qr_driver<mak::zd>((hipStream_t)s, 0, m, n, (mak::zd*)A, lda, (mak::zd*)Q, ldq, (mak::zd*)R, ldr,
                                 positive, work, work_bytes, (mak::zd*)Tout);
}

}  // extern "C"
