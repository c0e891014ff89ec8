/*
 * mak_mi355x.h -- C ABI of the MI355X (gfx950) QR / polar backend for
 * MatrixAlgebraKit.jl.
 *
 * Conventions (shared by every entry point):
 *   - All matrix pointers are DEVICE pointers (HIP), column-major with an
 *     explicit leading dimension, like LAPACK and Julia's Matrix
 *     (MatrixAlgebraKit src/yalapack.jl:179,316-317).
 *   - Complex matrices are ComplexF64, interleaved (re, im): the layout of
 *     Julia's ComplexF64 and numpy's complex128.  They are passed as void*.
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *     Work is enqueued on that stream (the polar's blocked Cholesky, n > 2048,
 *     also forks work onto a per-device look-ahead stream and makes the
 *     caller's stream wait for it with an event before it returns); no call
 *     synchronises the device or other streams.  Some calls wait for their
 *     own stream to read a small status word back to the host:
 *       mak_qr_*           once per call (the CholeskyQR2 acceptance check),
 *                          twice when the whole-matrix CholeskyQR2 result is
 *                          rejected; under stream capture never (the gated
 *                          form is enqueued instead);
 *       mak_*_polar_*      once for ||A||_F, then once per QDWH iteration
 *                          (convergence and the Cholesky status) and once
 *                          more per QR-based iteration (the closed-form test);
 *       mak_*_pullback_*   once (the rank / convergence counters);
 *       mak_qr_batched_execute is fully asynchronous (one graph launch).
 *   - Calls are reentrant: all state lives in the caller's workspace (the
 *     batched plans own their graph and stream; the polar's blocked Cholesky,
 *     n > 2048, forks onto one look-ahead stream per device, created with two
 *     events at its first use and joined back before the call returns);
 *     concurrent calls need distinct workspaces.
 *   - `work` is a caller-owned device workspace of at least the size returned
 *     by the matching *_workspace_size query (the library allocates nothing;
 *     this mirrors LAPACK's lwork query, src/yalapack.jl:180-197).
 *   - Return value: 0 on success; -i when argument i (1-based) is invalid,
 *     mirroring LAPACK's info / chkargsok
 *     (ext/MatrixAlgebraKitAMDGPUExt/yarocsolver.jl:88-89); > 0 for a numerical
 *     failure or a HIP launch error.
 *   - Optional outputs (R, P) are skipped when their pointer is NULL: the
 *     reference's "empty R / empty P" convention
 *     (src/implementations/qr.jl:149, src/implementations/polar.jl:64,76).
 */
#ifndef MAK_MI355X_H
#define MAK_MI355X_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types */
#define MAK_F64 0 /* Float64 */
#define MAK_C64 1 /* ComplexF64 */

/* ------------------------------------------------------------------ QR --- */

/* Device workspace bytes for mak_qr_{compact,full}_{d,z} on an m x n input.
 * full = 0 for qr_compact!, 1 for qr_full!.
 * Replaces the lwork = -1 query of YALAPACK.geqrt!/gemqrt!
 * (src/yalapack.jl:303-333, :882-928). */
int mak_qr_workspace_size(int dtype, int full, int m, int n, size_t* bytes);

/* qr_compact!(A, (Q, R), Householder(; positive)) for Float64 / ComplexF64.
 * A (m x n) is destroyed (it holds the compact WY factors afterwards),
 * Q is m x min(m,n), R is min(m,n) x n or NULL (R not computed).
 * positive != 0 gives R a real non-negative diagonal (gauge fix).
 * The factors are FP64-accurate; on the whole-matrix CholeskyQR2 path the
 * final correction term Q1 (W2^{-1} - I), of norm ~ ||Q1^H Q1 - I||, runs on
 * BF16 (that norm <= 5e-14) or FP32 (<= 1e-10) MFMA with an error bound below
 * Q's own FP64 rounding, else on FP64.
 * Replaces qr_compact!(A, QR, ::Householder) -> qr_householder!
 * (src/implementations/qr.jl:110-113, :132-188). */
int mak_qr_compact_d(void* stream, int m, int n, double* A, int lda, double* Q, int ldq, double* R, int ldr,
                     int positive, void* work, size_t work_bytes);
int mak_qr_compact_z(void* stream, int m, int n, void* A, int lda, void* Q, int ldq, void* R, int ldr,
                     int positive, void* work, size_t work_bytes);

/* qr_full!(A, (Q, R), Householder(; positive)): Q is m x m, R is m x n or NULL.
 * Replaces qr_full!(A, QR, ::Householder) (src/implementations/qr.jl:106-109). */
int mak_qr_full_d(void* stream, int m, int n, double* A, int lda, double* Q, int ldq, double* R, int ldr,
                  int positive, void* work, size_t work_bytes);
int mak_qr_full_z(void* stream, int m, int n, void* A, int lda, void* Q, int ldq, void* R, int ldr,
                  int positive, void* work, size_t work_bytes);

/* qr_compact! as the primal of the reverse rule: like mak_qr_compact_{d,z}, and
 * additionally keeps the compact WY factors the adjoint re-forms Q from: on
 * return A holds V (strictly below the diagonal) and R' (Householder signs),
 * Tout (32 x 32 x ceil(min(m,n)/32), column-major per panel) the T blocks.
 * Replaces the primal call inside rrule(::typeof(qr_compact!), A, QR, alg)
 * (ext/MatrixAlgebraKitChainRulesCoreExt.jl:26-43). */
int mak_qr_compact_wy_d(void* stream, int m, int n, double* A, int lda, double* Q, int ldq, double* R, int ldr,
                        double* Tout, int positive, void* work, size_t work_bytes);
int mak_qr_compact_wy_z(void* stream, int m, int n, void* A, int lda, void* Q, int ldq, void* R, int ldr, void* Tout,
                        int positive, void* work, size_t work_bytes);

/* qr_compact! without a workspace argument.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as mak_qr_compact_d; the workspace is a static
 * 64 MiB device array per GPU, held until the process ends (problems whose
 * mak_qr_workspace_size exceeds 64 MiB return 7; Float64 1024 x 512 needs
 * about 47 MiB).  The call allocates nothing.  Calls on one GPU share the
 * array: calls that overlap in time on one GPU (other streams or threads)
 * race, so the caller serialises them. */
int mak_qr_compact_noalloc_d(void* stream, int m, int n, double* A, int lda, double* Q, int ldq, double* R, int ldr,
                             int positive);

/* ------------------------------------------------------ LQ, null spaces --- */

/* Device workspace bytes of the LQ entry points (full = 0 / 1). */
int mak_lq_workspace_size(int dtype, int full, int m, int n, size_t* bytes);

/* lq_compact!(A, (L, Q), alg) / lq_full!(A, (L, Q), alg) through the QR of
 * A^H, as LQViaTransposedQR: A^H = Qt Rt gives L = Rt^H (m x min(m,n), m x n
 * for full; lower triangular; NULL = not computed) and Q = Qt^H (min(m,n) x n,
 * n x n for full) with orthonormal rows.  A is only read.  positive = 1: L
 * has a real non-negative diagonal.  Q and L must not overlap A.
 * Replaces lq_via_qr! (src/implementations/lq.jl:283-310) behind
 * lq_compact! / lq_full! (:285-296). */
int mak_lq_compact_d(void* stream, int m, int n, const double* A, int lda, double* L, int ldl, double* Q, int ldq,
                     int positive, void* work, size_t work_bytes);
int mak_lq_compact_z(void* stream, int m, int n, const void* A, int lda, void* L, int ldl, void* Q, int ldq,
                     int positive, void* work, size_t work_bytes);
int mak_lq_full_d(void* stream, int m, int n, const double* A, int lda, double* L, int ldl, double* Q, int ldq,
                  int positive, void* work, size_t work_bytes);
int mak_lq_full_z(void* stream, int m, int n, const void* A, int lda, void* L, int ldl, void* Q, int ldq,
                  int positive, void* work, size_t work_bytes);

/* Device workspace bytes of qr_null! (lq = 0) / lq_null! (lq = 1). */
int mak_null_workspace_size(int dtype, int lq, int m, int n, size_t* bytes);

/* qr_null!(A, N, alg): N (m x (m - min(m,n))) = H_1 ... H_k [0; I], the
 * trailing columns of the Householder QR's full Q (no gauge); an orthonormal
 * basis of the complement of range(A).  A is destroyed.
 * Replaces qr_null_householder! (src/implementations/qr.jl:114-117,236-263). */
int mak_qr_null_d(void* stream, int m, int n, double* A, int lda, double* N, int ldn, void* work,
                  size_t work_bytes);
int mak_qr_null_z(void* stream, int m, int n, void* A, int lda, void* N, int ldn, void* work, size_t work_bytes);

/* lq_null!(A, Nh, alg): Nh ((n - min(m,n)) x n) = the conjugate transpose of
 * the QR null space of A^H; orthonormal rows spanning the complement of the
 * row space of A.  A is only read.
 * Replaces lq_null_via_qr! (src/implementations/lq.jl:297-301,312-318). */
int mak_lq_null_d(void* stream, int m, int n, const double* A, int lda, double* Nh, int ldnh, void* work,
                  size_t work_bytes);
int mak_lq_null_z(void* stream, int m, int n, const void* A, int lda, void* Nh, int ldnh, void* work,
                  size_t work_bytes);

/* --------------------------------------------------------------- polar --- */

/* Diagnostics of one polar call (optional output, may be NULL). */
typedef struct mak_polar_info {
  int iterations;   /* QDWH iterations run */
  int qr_steps;     /* iterations in QR form (QR of [sqrt(c) X; I]): c > 100,
                       or c > 1e5 at precision 32 / 16 */
  int chol_steps;   /* iterations in Cholesky form */
  int ns_steps;     /* FP64 Newton-Schulz refinement steps (precision 32 / 16) */
  double alpha;     /* ||A||_F, the scaling of X_0 */
  double l0;        /* lower bound used for sigma_min(X_0) */
  double l_final;   /* lower bound after the last iteration */
  double last_diff; /* ||X_k - X_{k-1}||_F / ||X_k||_F of the last iteration;
                       precision 32 / 16: ||I - X^H X||_F before the last
                       Newton-Schulz step */
  int completed;    /* 1: A was rank deficient and the partial isometry of the
                       iteration was completed to an isometry (a second run) */
  int poly_steps;   /* precision 32 / 16: steps of the polynomial phase that
                       replaces the Cholesky-form iterations (c <= 100) */
} mak_polar_info;

/* Device workspace bytes for the polar entry points on an m x n input. */
int mak_polar_workspace_size(int dtype, int m, int n, size_t* bytes);

/* left_polar!(A, (W, P), alg) for m >= n: A = W P, W m x n with orthonormal
 * columns, P = herm(W^H A) n x n Hermitian positive semidefinite, or NULL
 * (P not computed).  A is left untouched.
 *   precision: 64 (FP64 throughout), 32 or 16 (reduced-precision bulk): the
 *   ill-conditioned QR-form iterations keep their Gram, Cholesky and first
 *   CholeskyQR pass in FP64 (second pass on FP32 MFMA); Cholesky-form
 *   iterations run in FP64 while c > 100; from c <= 100 on, cubic polynomial
 *   steps X (alpha I - beta X^H X) on FP32 MFMA (BF16 for the small late
 *   corrections at 16) take over, followed by FP64 Newton-Schulz refinement
 *   (FP64 Gram).  P = herm(W^H A) is then an FP32 product.  Complex data
 *   only supports 64.
 *   maxiter <= 0 selects 20; l0 outside (0, 1) selects 1e-10.
 *   W is an isometry also for rank-deficient A (as the SVD route's U V^H):
 *   the iteration's partial isometry is completed (info->completed), and
 *   A = 0 gives W = I (m x n), P = 0.
 * Replaces left_polar!(A, WP, ::PolarViaSVD / ::PolarNewton)
 * (src/implementations/polar.jl:59-70, :99-111). */
int mak_left_polar_d(void* stream, int m, int n, const double* A, int lda, double* W, int ldw, double* P, int ldp,
                     int precision, int maxiter, double l0, void* work, size_t work_bytes, mak_polar_info* info);
int mak_left_polar_z(void* stream, int m, int n, const void* A, int lda, void* W, int ldw, void* P, int ldp,
                     int precision, int maxiter, double l0, void* work, size_t work_bytes, mak_polar_info* info);

/* right_polar!(A, (P, Wh), alg) for n >= m: A = P Wh, P m x m (or NULL),
 * Wh m x n with orthonormal rows.
 * Replaces right_polar!(A, PWh, ::PolarViaSVD / ::PolarNewton)
 * (src/implementations/polar.jl:71-82, :113-125). */
int mak_right_polar_d(void* stream, int m, int n, const double* A, int lda, double* P, int ldp, double* Wh, int ldw,
                      int precision, int maxiter, double l0, void* work, size_t work_bytes, mak_polar_info* info);
int mak_right_polar_z(void* stream, int m, int n, const void* A, int lda, void* P, int ldp, void* Wh, int ldw,
                      int precision, int maxiter, double l0, void* work, size_t work_bytes, mak_polar_info* info);

/* ----------------------------------------------------- SVD via polar --- */

/* Device workspace bytes of mak_svd_polar_{d,z}. */
int mak_svd_polar_workspace_size(int dtype, int m, int n, size_t* bytes);

/* svd_compact!(A, (U, S, Vh), SVDViaPolar()) (U, Vh non-NULL) or
 * svd_vals!(A, S, SVDViaPolar()) (U = Vh = NULL): A = W P by FP64 QDWH
 * (m >= n; the right polar A = P Wh for m < n), then the Hermitian
 * eigendecomposition of P (rocSOLVER syevd / heevd, opened at the first
 * call), S descending (real, on the device, min(m,n)), U m x k, Vh k x n, and
 * the gauge of gaugefix!(svd_compact!, U, Vh): the first entry of largest
 * magnitude of every column of U is real and positive.  A is only read.
 * Returns 3 when rocSOLVER cannot be opened; info as for the polar calls.
 * Replaces gesvdp! behind svd_compact! / svd_vals! with SVDViaPolar
 * (src/interface/decompositions.jl:173-183; src/implementations/svd.jl:
 * 161,184-206; ext/MatrixAlgebraKitCUDAExt/yacusolver.jl:101-180). */
int mak_svd_polar_d(void* stream, int m, int n, const double* A, int lda, double* S, double* U, int ldu, double* Vh,
                    int ldvh, int maxiter, void* work, size_t work_bytes, mak_polar_info* info);
int mak_svd_polar_z(void* stream, int m, int n, const void* A, int lda, double* S, void* U, int ldu, void* Vh,
                    int ldvh, int maxiter, void* work, size_t work_bytes, mak_polar_info* info);

/* ------------------------------------------------------- pushforwards --- */

/* Device workspace bytes of the polar pushforwards (right = 0 / 1). */
int mak_polar_pushforward_workspace_size(int dtype, int right, int m, int n, size_t* bytes);

/* left_polar_pushforward!(dA, A, (W, P), (dW, dP)) for m >= n: dP = W^H dA,
 * K from P K + K P = dP - dP^H, dW = dA P^{-1} - W W^H dA P^{-1} + W K,
 * dP -= K P; the Sylvester solve and P^{-1} through the eigendecomposition of
 * P (exact for any positive definite P).  dW (m x n) and dP (n x n) are
 * overwritten.  Returns 2 when P is not positive definite.
 * Replaces src/pushforwards/polar.jl:1-12. */
int mak_left_polar_pushforward_d(void* stream, int m, int n, const double* W, int ldw, const double* P, int ldp,
                                 const double* dA, int ldda, double* dW, int lddw, double* dP, int lddp, void* work,
                                 size_t work_bytes);
int mak_left_polar_pushforward_z(void* stream, int m, int n, const void* W, int ldw, const void* P, int ldp,
                                 const void* dA, int ldda, void* dW, int lddw, void* dP, int lddp, void* work,
                                 size_t work_bytes);
/* right_polar_pushforward!(dA, A, (P, Wh), (dP, dWh)) for m <= n: Wh, dA, dWh
 * m x n, P, dP m x m.  Replaces src/pushforwards/polar.jl:14-25. */
int mak_right_polar_pushforward_d(void* stream, int m, int n, const double* Wh, int ldwh, const double* P, int ldp,
                                  const double* dA, int ldda, double* dWh, int lddwh, double* dP, int lddp,
                                  void* work, size_t work_bytes);
int mak_right_polar_pushforward_z(void* stream, int m, int n, const void* Wh, int ldwh, const void* P, int ldp,
                                  const void* dA, int ldda, void* dWh, int lddwh, void* dP, int lddp, void* work,
                                  size_t work_bytes);

/* ----------------------------------------------------------- adjoints --- */

/* Device workspace of the reverse rules for `batch` matrices of size m x n:
 * kind 0 = mak_qr_pullback_*, kind 1 = mak_left_polar_pullback_*,
 * kind 2 = mak_qr_full_cotangent_*. */
int mak_pullback_workspace_size(int kind, int dtype, int batch, int m, int n, size_t* bytes);

/* qr_pullback!(dA, A, (Q, R), (dQ, dR)) for `batch` equally sized matrices
 * (matrix b at pointer + b * stride), accumulating into dA (dA += ...).  Q is
 * re-formed from the compact WY factors (A_wy, T) kept by
 * mak_qr_compact_wy_*, with the gauge signs when `positive` matches the
 * primal's positive = true; R is the primal's R (min(m,n) x n, only its upper
 * triangle is read); p is the rank (qr_rank, src/pullbacks/qr.jl:1-2).
 * dQ (m x min(m,n)) / dR may be NULL (absent cotangent).  gauge_out
 * (device, `batch` doubles, may be NULL) receives the largest magnitude of the
 * gauge-dependent cotangent entries, dQ[:, p:min(m,n)] and the upper triangle
 * of dR[p:, p:] (the caller warns when it exceeds its gauge tolerance, as
 * src/pullbacks/qr.jl:50-51); it is written asynchronously.
 * Replaces qr_pullback! (src/pullbacks/qr.jl:73-120). */
int mak_qr_pullback_d(void* stream, int batch, int m, int n, int p, int positive, const double* A, int lda,
                      long long strideA, const double* T, long long strideT, const double* R, int ldr,
                      long long strideR, const double* dQ, int lddq, long long strideDQ, const double* dR, int lddr,
                      long long strideDR, double* dA, int ldda, long long strideDA, void* work, size_t work_bytes,
                      double* gauge_out);
int mak_qr_pullback_z(void* stream, int batch, int m, int n, int p, int positive, const void* A, int lda,
                      long long strideA, const void* T, long long strideT, const void* R, int ldr, long long strideR,
                      const void* dQ, int lddq, long long strideDQ, const void* dR, int lddr, long long strideDR,
                      void* dA, int ldda, long long strideDA, void* work, size_t work_bytes, double* gauge_out);

/* The cotangent preparation of a qr_full! primal (ΔQ m x m): dQ1 (m x
 * min(m,n)) = the compact cotangent mak_qr_pullback_* takes and gauge_out
 * (device, `batch` doubles, may be NULL) the largest magnitude of its
 * gauge-dependent part.  Rank p = min(m,n): dQ3 = dQ[:, k:m] is split into
 * its part along Q1, which moves into dQ1 -= Q3 (Q1^H dQ3)^H, and the rest
 * (gauge); p < min(m,n): dQ1 = dQ[:, :k], gauge over dQ[:, p:m].  Q is the
 * full Q (m x m).  Replaces check_and_prepare_qr_cotangents
 * (src/pullbacks/qr.jl:4-53) for full Q. */
int mak_qr_full_cotangent_d(void* stream, int batch, int m, int n, int p, const double* Q, int ldq, long long strideQ,
                            const double* dQ, int lddq, long long strideDQ, double* dQ1, int ldd1, long long strideD1,
                            double* gauge_out, void* work, size_t work_bytes);
int mak_qr_full_cotangent_z(void* stream, int batch, int m, int n, int p, const void* Q, int ldq, long long strideQ,
                            const void* dQ, int lddq, long long strideDQ, void* dQ1, int ldd1, long long strideD1,
                            double* gauge_out, void* work, size_t work_bytes);

/* left_polar_pullback!(dA, A, (W, P), (dW, dP)) for `batch` equally sized
 * matrices (m >= n), accumulating into dA.  The Sylvester equation
 * P C + C P = M - M^H is solved by preconditioned conjugate gradients
 * (maxiter <= 0: 200, tol <= 0: 1e-14 relative residual); when CG has not
 * converged after 48 steps (ill-conditioned P), the batch is solved exactly
 * through the eigendecomposition of P (rocSOLVER, opened at first use),
 * which also gives P^{-1}.  iters receives the CG step count (may be NULL).
 * Returns 6 when CG did not reach tol and rocSOLVER is not available, 2 when
 * P is not positive definite.
 * Replaces left_polar_pullback! (src/pullbacks/polar.jl:7-29). */
int mak_left_polar_pullback_d(void* stream, int batch, int m, int n, const double* W, int ldw, long long strideW,
                              const double* P, int ldp, long long strideP, const double* dW, int lddw,
                              long long strideDW, const double* dP, int lddp, long long strideDP, double* dA,
                              int ldda, long long strideDA, int maxiter, double tol, void* work, size_t work_bytes,
                              int* iters);
int mak_left_polar_pullback_z(void* stream, int batch, int m, int n, const void* W, int ldw, long long strideW,
                              const void* P, int ldp, long long strideP, const void* dW, int lddw,
                              long long strideDW, const void* dP, int lddp, long long strideDP, void* dA, int ldda,
                              long long strideDA, int maxiter, double tol, void* work, size_t work_bytes,
                              int* iters);

/* ------------------------------------------------------- batched QR --- */

/* A plan for qr_compact!(A_i, (Q_i, R_i)) over `count` independent problems of
 * arbitrary sizes m_i x n_i (Float64 or ComplexF64, all of one type).  All
 * size, leading-dimension and pointer arrays are HOST arrays; the matrices are
 * device memory that must stay valid for the plan's lifetime.  Plan creation
 * builds the task lists in the caller's workspace, optionally autotunes the
 * apply kernel (tune != 0: the workspace then also holds a scratch copy of the
 * shapes; results are cached per shape set for the process), and captures the
 * whole launch sequence in a hipGraph.  A_i is overwritten (compact WY form),
 * R may be NULL (no R) or hold NULL entries.
 * No counterpart in the reference, which has no batched API: it replaces a
 * caller's loop of qr_compact! over blocks (src/implementations/qr.jl:110-113). */
typedef struct mak_qr_batched_plan mak_qr_batched_plan;

int mak_qr_batched_workspace_size(int dtype, int count, const int* m, const int* n, int tune, size_t* bytes);
int mak_qr_batched_plan_create(mak_qr_batched_plan** plan, int dtype, int count, const int* m, const int* n,
                               void* const* A, const int* lda, void* const* Q, const int* ldq, void* const* R,
                               const int* ldr, int positive, int tune, void* work, size_t work_bytes);
/* Replays the captured graph on `stream` (asynchronous). */
int mak_qr_batched_execute(mak_qr_batched_plan* plan, void* stream);
/* Chosen chunk width, persistent workgroups per CU, launches in the graph and
 * problems routed to the single-matrix driver (any pointer may be NULL). */
int mak_qr_batched_plan_info(const mak_qr_batched_plan* plan, int* chunk_cols, int* grid_per_cu, int* launches,
                             int* fallback_problems);
int mak_qr_batched_plan_destroy(mak_qr_batched_plan* plan);
/* One-shot form: create (no tuning unless cached), execute, synchronise the
 * stream, destroy. */
int mak_qr_compact_batched(void* stream, int dtype, int count, const int* m, const int* n, void* const* A,
                           const int* lda, void* const* Q, const int* ldq, void* const* R, const int* ldr,
                           int positive, void* work, size_t work_bytes);

#ifdef __cplusplus
}
#endif

#endif /* MAK_MI355X_H */
