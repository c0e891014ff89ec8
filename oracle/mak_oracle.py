"""CPU oracle for the QR / polar path of MatrixAlgebraKit.jl.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The product path (the HIP library and
its Python host layer) never calls into it.

Each function restates the reference's algorithm on the CPU with numpy/scipy.
The reference is pure Julia; its arithmetic lives in LAPACK, called through
``src/yalapack.jl``.  The same LAPACK routines are reached here through
``scipy.linalg.lapack`` (scipy's bundled OpenBLAS LAPACK):

* ``qr_compact`` / ``qr_full`` / ``qr_null``  follow
  ``src/implementations/qr.jl:132-188`` and ``:236-263`` (blocked path,
  ``dgeqrt`` with nb = min(m, n, 36), ``src/yalapack.jl:88-89,303-333``, then
  ``dgemqrt('L','N')`` applied to the identity, ``src/yalapack.jl:882-928``,
  then ``triu`` and the sign gauge of ``src/common/gauge.jl:16-25`` with
  ``sign_safe`` from ``src/common/safemethods.jl:10-11``).
* ``householder_native`` follows ``_householder!``
  (``src/common/householder.jl:35-67``).
* ``left_polar_svd`` / ``right_polar_svd`` follow
  ``src/implementations/polar.jl:59-95`` with ``gesdvd!``
  (``src/yalapack.jl:2173-2202``: ``gesdd`` job 'S', ``gesvd`` fallback).
* ``left_polar_newton`` / ``right_polar_newton`` follow
  ``src/implementations/polar.jl:99-220``.
* ``qr_pullback``, ``remove_qr_gauge_dependence``, ``left_polar_pullback``,
  ``right_polar_pullback`` follow ``src/pullbacks/qr.jl:1-120,162-189`` and
  ``src/pullbacks/polar.jl:7-59``; ``_sylvester`` is Julia's
  ``sylvester(A, B, C)`` (solves ``A X + X B + C = 0``,
  ``src/common/pullbacks.jl:15``).

Pinning (what checks each function against something other than itself) is
listed per function in ``tests/test_oracle.py``; every function here has at
least one independent check there (closed form, second library route,
invariant, or finite differences).
"""
from __future__ import annotations

import warnings

import numpy as np
import scipy.linalg as sla
import scipy.linalg.lapack as lapack

DEFAULT_QR_BLOCKSIZE = 36  # src/yalapack.jl:88


def _lp(dtype, name):
    """Pick the d/z LAPACK routine for a dtype (src/yalapack.jl:92-98 eltype loop)."""
    prefix = {np.dtype(np.float64): "d", np.dtype(np.complex128): "z",
              np.dtype(np.float32): "s", np.dtype(np.complex64): "c"}[np.dtype(dtype)]
    if prefix in "sd" and name in ("ungqr", "unmqr"):
        name = "or" + name[2:]
    return getattr(lapack, prefix + name)


def eps(dtype) -> float:
    return float(np.finfo(np.dtype(dtype)).eps)


def defaulttol(A) -> float:
    """``defaulttol(x) = eps(real(T))^(2/3)`` (src/common/defaults.jl:31)."""
    return eps(np.real(np.zeros(1, dtype=A.dtype)).dtype) ** (2.0 / 3.0)


def sign_safe(x):
    """``sign_safe`` (src/common/safemethods.jl:10-11): real -> +-1 with 0 -> +1;
    complex -> z/|z| with 0 -> 1."""
    x = np.asarray(x)
    if np.iscomplexobj(x):
        a = np.abs(x)
        out = np.ones_like(x)
        nz = a != 0
        out[nz] = x[nz] / a[nz]
        return out
    return np.where(x < 0, -1.0, 1.0).astype(x.dtype)


def default_qr_blocksize(A) -> int:
    """``default_qr_blocksize(A) = min(size(A)..., 36)`` (src/yalapack.jl:88-89)."""
    return min(A.shape[0], A.shape[1], DEFAULT_QR_BLOCKSIZE)


# ---------------------------------------------------------------------------
# QR (src/implementations/qr.jl:132-188)
# ---------------------------------------------------------------------------
def geqrt(A, nb):
    """``YALAPACK.geqrt!`` (src/yalapack.jl:303-333): blocked compact-WY QR.
    Returns (A with V below / R on+above the diagonal, T nb x min(m,n))."""
    f = _lp(A.dtype, "geqrt")
    a, t, info = f(nb, np.asfortranarray(A))
    if info != 0:
        raise ValueError(f"geqrt info={info}")
    return a, t


def gemqrt_left_notrans(V, T, C):
    """``YALAPACK.gemqrt!('L','N',V,T,C)`` (src/yalapack.jl:882-928): C <- Q C."""
    f = _lp(V.dtype, "gemqrt")
    c, info = f(np.asfortranarray(V), np.asfortranarray(T), np.asfortranarray(C),
                side="L", trans="N")
    if info != 0:
        raise ValueError(f"gemqrt info={info}")
    return c


def gaugefix_qr(Q, R, Rd):
    """``gaugefix!(qr_householder!, Q, R, Rd)`` (src/common/gauge.jl:16-25):
    Q[:, j] *= s_j and R[j, :] *= conj(s_j), s = sign_safe(Rd)."""
    s = sign_safe(Rd)
    k = len(s)
    Q[:, :k] *= s[None, :]
    if R is not None:
        R[:k, :] *= np.conj(s)[:, None]
    return Q, R


def qr_householder(A, kind="compact", positive=True, blocksize=0, compute_r=True):
    """Restatement of ``qr_householder!(::LAPACK, A, Q, R)``
    (src/implementations/qr.jl:132-188), blocked branch (blocksize > 1), which is
    what the LAPACK driver uses by default for a plain matrix (:137).

    ``kind`` is "compact" (Q m x min(m,n), R min(m,n) x n) or "full"
    (Q m x m, R m x n), matching ``initialize_output`` (:63-75).
    """
    A = np.array(A, order="F", copy=True)
    if A.dtype not in (np.float64, np.complex128, np.float32, np.complex64):
        A = A.astype(np.float64)
    m, n = A.shape
    minmn = min(m, n)
    qcols = minmn if kind == "compact" else m
    rrows = minmn if kind == "compact" else m
    Q = np.zeros((m, qcols), dtype=A.dtype, order="F")
    R = np.zeros((rrows, n), dtype=A.dtype, order="F") if compute_r else None
    if minmn == 0:
        Q[:, :] = np.eye(m, qcols, dtype=A.dtype)
        return Q, R
    if blocksize <= 0:
        blocksize = default_qr_blocksize(A)
    nb = min(minmn, blocksize)
    a, t = geqrt(A, nb)                                   # :158-161
    Q[:, :] = np.eye(m, qcols, dtype=A.dtype)             # one!(Q)
    Q = gemqrt_left_notrans(a[:, :minmn], t, Q)           # :163
    Rd = np.diagonal(a)[:minmn].copy()                    # diagview(A), :179
    if compute_r:
        R[:minmn, :] = np.triu(a[:minmn, :])              # :180-181 uppertriangular!
    if positive:                                          # :182-184
        gaugefix_qr(Q, R, Rd)
    return Q, R


def qr_compact(A, positive=True, blocksize=0):
    """``qr_compact!(A, (Q, R), Householder())`` (src/implementations/qr.jl:110-113)."""
    return qr_householder(A, "compact", positive, blocksize)


def qr_full(A, positive=True, blocksize=0):
    """``qr_full!(A, (Q, R), Householder())`` (src/implementations/qr.jl:106-109)."""
    return qr_householder(A, "full", positive, blocksize)


def qr_null(A, blocksize=0):
    """``qr_null_householder!(::LAPACK, A, N)`` (src/implementations/qr.jl:236-263):
    N = Q[:, minmn+1:m] from applying the block reflectors to [0; I]."""
    A = np.array(A, order="F", copy=True)
    m, n = A.shape
    minmn = min(m, n)
    N = np.zeros((m, m - minmn), dtype=A.dtype, order="F")
    N[minmn:, :] = np.eye(m - minmn, dtype=A.dtype)
    if minmn == 0 or m == minmn:
        return N
    if blocksize <= 0:
        blocksize = default_qr_blocksize(A)
    nb = min(minmn, blocksize)
    a, t = geqrt(A, nb)
    return gemqrt_left_notrans(a[:, :minmn], t, N)


def lq_via_qr(A, full=False, positive=True):
    """``lq_via_qr!`` (src/implementations/lq.jl:303-310) behind
    ``lq_compact!`` / ``lq_full!(A, LQ, ::LQViaTransposedQR)`` (:283-296):
    the QR of A^H, conjugate-transposed: L = R^H, Q = Q_t^H."""
    At = np.conj(np.asarray(A)).T.copy()
    Qt, Rt = qr_full(At, positive) if full else qr_compact(At, positive)
    return np.conj(Rt).T.copy(), np.conj(Qt).T.copy()


def lq_null(A):
    """``lq_null_via_qr!`` (src/implementations/lq.jl:312-318): N^H from the QR
    null space of A^H."""
    return np.conj(qr_null(np.conj(np.asarray(A)).T.copy())).T.copy()


def householder_native(x, i=0):
    """``_householder!(v, i)`` (src/common/householder.jl:35-67), 0-based i.

    Returns (beta, v, nu) such that (I - beta v v^H) x = nu e_i with nu = ||x||
    real and non-negative, v[i] = 1.
    """
    v = np.array(x, copy=True)
    T = v.dtype
    sigma = float(np.sum(np.abs(v[:i]) ** 2) + np.sum(np.abs(v[i + 1:]) ** 2))
    vi = v[i]
    nu = np.sqrt(abs(vi) ** 2 + sigma)
    if sigma == 0 and vi == nu:
        return T.type(0), v, nu
    if np.real(vi) < 0:
        vi = vi - nu
    else:
        vi = ((vi - np.conj(vi)) * nu - sigma) / (np.conj(vi) + nu)
    v = v / vi
    v[i] = 1
    beta = -np.conj(vi) / nu
    return beta, v, nu


def qr_native(A):
    """``qr_householder!(::Native, A, Q, R)`` (src/implementations/qr.jl:189-230):
    column-by-column ``_householder!`` with positive diagonal, Q built backwards."""
    A = np.array(A, dtype=np.result_type(A, np.float64), copy=True)
    m, n = A.shape
    minmn = min(m, n)
    R = np.zeros((minmn, n), dtype=A.dtype)
    betas = []
    for j in range(minmn):
        R[:j, j] = A[:j, j]
        beta, v, nu = householder_native(A[j:, j], 0)
        A[j:, j] = v
        R[j, j] = nu
        if beta != 0:
            w = v.conj() @ A[j:, j + 1:]
            A[j:, j + 1:] -= beta * np.outer(v, w)
        betas.append(beta)
    for j in range(minmn, n):
        R[:, j] = A[:minmn, j]
    Q = np.eye(m, minmn, dtype=A.dtype)
    for j in range(minmn - 1, -1, -1):
        v = A[j:, j].copy()
        v[0] = 1
        b = np.conj(betas[j])
        if b != 0:
            Q[j:, :] -= b * np.outer(v, v.conj() @ Q[j:, :])
    return Q, R


# ---------------------------------------------------------------------------
# Polar (src/implementations/polar.jl)
# ---------------------------------------------------------------------------
def gesdvd(A):
    """``YALAPACK.gesdvd!`` (src/yalapack.jl:2173-2202): gesdd job 'S' with gesvd
    fallback when gesdd reports info > 0."""
    A = np.array(A, order="F", copy=True)
    try:
        return sla.svd(A, full_matrices=False, lapack_driver="gesdd")
    except np.linalg.LinAlgError:
        return sla.svd(A, full_matrices=False, lapack_driver="gesvd")


def gaugefix_svd_compact(U, Vh):
    """``gaugefix!(svd_compact!, U, Vh)`` (src/common/gauge.jl:72-80): the
    first entry of largest magnitude of every column of U made real positive
    (``_argmaxabs`` keeps the earlier entry on ties, :11-13)."""
    U = np.array(U, copy=True)
    Vh = np.array(Vh, copy=True)
    for j in range(U.shape[1]):
        i = int(np.argmax(np.abs(U[:, j])))   # first maximiser
        x = U[i, j]
        sgn = x / abs(x) if x != 0 else 0.0
        U[:, j] *= np.conj(sgn)
        Vh[j, :] *= sgn
    return U, Vh


def svd_compact(A, fixgauge=True):
    """``svd_compact!(A, (U, S, Vh), alg)`` through ``gesdvd`` with the gauge
    fix (src/implementations/svd.jl:184-189)."""
    U, S, Vh = gesdvd(A)
    if fixgauge:
        U, Vh = gaugefix_svd_compact(U, Vh)
    return U, S, Vh


def left_polar_svd(A, compute_p=True):
    """``left_polar!(A, (W, P), PolarViaSVD)`` (src/implementations/polar.jl:59-70)."""
    m, n = A.shape
    if m < n:
        raise ValueError(f"input matrix needs at least as many rows ({m}) as columns ({n})")
    U, S, Vh = gesdvd(A)
    W = U @ Vh
    if not compute_p:
        return W, None
    X = np.sqrt(S)[:, None] * Vh          # lmul!(sqrt.(S), Vh)
    P = X.conj().T @ X                     # _mul_herm!, exactly Hermitian (syrk)
    P = (P + P.conj().T) / 2
    return W, P


def right_polar_svd(A, compute_p=True):
    """``right_polar!(A, (P, Wh), PolarViaSVD)`` (src/implementations/polar.jl:71-82)."""
    m, n = A.shape
    if n < m:
        raise ValueError(f"input matrix needs at least as many columns ({n}) as rows ({m})")
    U, S, Vh = gesdvd(A)
    Wh = U @ Vh
    if not compute_p:
        return None, Wh
    X = U * np.sqrt(S)[None, :]
    P = X @ X.conj().T
    P = (P + P.conj().T) / 2
    return P, Wh


def project_hermitian(A):
    """``project_hermitian!`` (src/implementations/projections.jl:61-64): (A + A^H)/2."""
    return (A + A.conj().T) / 2


def _avgdiff(A, B):
    """``_avgdiff!`` (src/implementations/polar.jl:209-220)."""
    return (A + B) / 2, B - A


def _norm_inf(X):
    """Julia's ``norm(X, Inf)`` of an array: the largest entry in absolute value
    (LinearAlgebra's ``norm`` treats a matrix as a vector; ``opnorm`` would be
    the row-sum norm)."""
    return float(np.max(np.abs(X))) if X.size else 0.0


def left_polar_newton(A, tol=None, maxiter=10, compute_p=True, info=None):
    """``left_polar!(A, WP, PolarNewton)`` + ``left_polar_newton!``
    (src/implementations/polar.jl:99-111,128-166).  The convergence measure is
    ``norm(Rᴴinv, Inf)`` (:148,156), the max-abs entry of the last difference.
    ``info`` (a dict, optional) receives the iteration count and that measure."""
    A = np.array(A, copy=True)
    m, n = A.shape
    if tol is None:
        tol = defaulttol(A)
    A0 = A.copy()
    if m > n:
        Q, R = qr_compact(A)
        Rc = R.copy()
        Rhinv = sla.solve_triangular(Rc, np.eye(n, dtype=A.dtype), trans="C", lower=False)
    else:
        Q = None
        R = A.copy()
        Rc = R.copy()
        Rhinv = np.linalg.inv(Rc).conj().T
    R = Rc

    def step(R, Rhinv):
        gamma = np.sqrt(np.linalg.norm(Rhinv) / np.linalg.norm(R))
        return _avgdiff(R * gamma, Rhinv / gamma)

    R, Rhinv = step(R, Rhinv)
    i = 1
    conv = _norm_inf(Rhinv)
    while i < maxiter and conv > tol:
        Rhinv = np.linalg.inv(R).conj().T
        R, Rhinv = step(R, Rhinv)
        conv = _norm_inf(Rhinv)
        i += 1
    if info is not None:
        info["iterations"] = i
        info["conv"] = conv
    if conv > tol:
        warnings.warn("`left_polar!` via Newton iteration did not converge within "
                      f"{maxiter} iterations (final residual: {conv})")
    W = Q @ R if Q is not None else R
    if not compute_p:
        return W, None
    return W, project_hermitian(W.conj().T @ A0)


def right_polar_newton(A, tol=None, maxiter=10, compute_p=True):
    """``right_polar!(A, PWh, PolarNewton)`` (src/implementations/polar.jl:113-125,
    168-206), restated through the conjugate transpose (LQ of A = QR of A^H)."""
    W, P = left_polar_newton(np.conj(A).T, tol=tol, maxiter=maxiter, compute_p=compute_p)
    return (None if P is None else P.conj().T), W.conj().T


# ---------------------------------------------------------------------------
# Matrix properties (src/common/matrixproperties.jl:49-54)
# ---------------------------------------------------------------------------
def left_polar_pushforward(dA, A, WP):
    """``left_polar_pushforward!`` (src/pushforwards/polar.jl:1-12): (dW, dP)."""
    W, P = WP
    dP = W.conj().T @ dA
    K = _sylvester(P, P, dP.conj().T - dP)
    dW = dA @ np.linalg.inv(P)
    WhdAiP = W.conj().T @ dW
    dW = dW - W @ WhdAiP + W @ K
    dP = dP - K @ P
    return dW, dP


def right_polar_pushforward(dA, A, PWh):
    """``right_polar_pushforward!`` (src/pushforwards/polar.jl:14-25): (dWh, dP)."""
    P, Wh = PWh
    dP = dA @ Wh.conj().T
    K = _sylvester(P, P, dP.conj().T - dP)
    dWh = np.linalg.inv(P) @ dA
    iPdAW = dWh @ Wh.conj().T
    dWh = dWh - iPdAW @ Wh + K @ Wh
    dP = dP - P @ K
    return dWh, dP


def isisometric(A, atol=0.0, rtol=None):
    """``is_left_isometric``: ||A^H A - I||_F <= max(atol, rtol ||A^H A||_F)."""
    if rtol is None:
        rtol = defaulttol(A)
    P = A.conj().T @ A
    nP = np.linalg.norm(P)
    P = P - np.eye(P.shape[0])
    return np.linalg.norm(P) <= max(atol, rtol * nP)


# ---------------------------------------------------------------------------
# Pullbacks (src/pullbacks/qr.jl, src/pullbacks/polar.jl)
# ---------------------------------------------------------------------------
def default_pullback_rank_atol(R):
    """``eps(norm(R, Inf))^(3/4)`` (src/common/defaults.jl:57)."""
    nrm = np.max(np.abs(R)) if R.size else 0.0
    return float(np.spacing(nrm)) ** 0.75 if nrm > 0 else 0.0


def default_pullback_gauge_atol(X):
    """``default_pullback_gauge_atol`` (src/common/defaults.jl:39)."""
    if X is None:
        return 0.0
    nrm = np.max(np.abs(X)) if X.size else 0.0
    return float(np.spacing(nrm)) ** 0.75 if nrm > 0 else 0.0


def qr_rank(R, rank_atol=None):
    """``qr_rank`` (src/pullbacks/qr.jl:1-2): last index with |R_ii| >= atol."""
    if rank_atol is None:
        rank_atol = default_pullback_rank_atol(R)
    d = np.abs(np.diagonal(R))
    idx = np.nonzero(d >= rank_atol)[0]
    return int(idx[-1] + 1) if len(idx) else 0


def remove_qr_gauge_dependence(dQ, dR, A, Q, R, rank_atol=None):
    """``remove_qr_gauge_dependence!`` (src/pullbacks/qr.jl:162-189)."""
    dQ = dQ.copy()
    dR = dR.copy()
    r = qr_rank(R, rank_atol)
    m, n = A.shape
    minmn = min(m, n)
    Q1 = Q[:, :r]
    dQ[:, r:minmn] = 0
    if r == minmn:
        dQ3 = dQ[:, minmn:]
        dQ[:, minmn:] = Q1 @ (Q1.conj().T @ dQ3)
    else:
        dQ[:, minmn:] = 0
    sub = dR[r:minmn, r:]
    k = min(sub.shape)
    sub[np.arange(k), np.arange(k)] = 0
    if r < minmn:
        iu = np.triu_indices(sub.shape[0], 0, sub.shape[1])
        sub[iu] = 0
    dR[r:minmn, r:] = sub
    return dQ, dR


def check_and_prepare_qr_cotangents(Q, R, dQ, dR, p, gauge_atol=None):
    """``check_and_prepare_qr_cotangents`` (src/pullbacks/qr.jl:4-53)."""
    m, n = Q.shape[0], R.shape[1]
    minmn = min(m, n)
    if gauge_atol is None:
        gauge_atol = default_pullback_gauge_atol(dQ)
    dgauge = 0.0
    Q1 = Q[:, :p]
    dQ1 = np.zeros_like(Q1)
    if dQ is not None:
        if dQ.shape != Q.shape:
            raise ValueError("ΔQ must have the same size as Q")
        dQ1[:, :] = dQ[:, :p]
        if p == minmn:
            dQ3 = dQ[:, minmn:].copy()
            Q3 = Q[:, minmn:]
            Q1hdQ3 = Q1.conj().T @ dQ3
            dQ3 -= Q1 @ Q1hdQ3
            dgauge_q = np.max(np.abs(dQ3)) if dQ3.size else 0.0
            dQ1 -= Q3 @ Q1hdQ3.conj().T
        else:
            d23 = dQ[:, p:]
            dgauge_q = np.max(np.abs(d23)) if d23.size else 0.0
        dgauge = max(dgauge, dgauge_q)
    if dR is not None:
        if dR.shape != R.shape:
            raise ValueError("ΔR must have the same size as R")
        dR11 = np.triu(dR[:p, :p])
        dR12 = dR[:p, p:n]
        if p < minmn:
            d22 = dR[p:minmn, p:n]
            up = np.triu(d22)
            dgauge = max(dgauge, np.max(np.abs(up)) if up.size else 0.0)
    else:
        dR11 = dR12 = None
    if dgauge > gauge_atol:
        warnings.warn(f"`qr` cotangents sensitive to gauge choice: (|Δgauge| = {dgauge})")
    return dQ1, dR11, dR12


def qr_pullback(dA, A, QR, dQR, rank_atol=None, gauge_atol=None):
    """``qr_pullback!`` (src/pullbacks/qr.jl:73-120): returns dA + pullback."""
    Q, R = QR
    dQ, dR = dQR
    dA = np.array(dA, copy=True)
    m, n = Q.shape[0], R.shape[1]
    p = qr_rank(R, rank_atol)
    Q1 = Q[:, :p]
    R11 = np.triu(R[:p, :p])
    R12 = R[:p, p:n]
    dQ1, dR11, dR12 = check_and_prepare_qr_cotangents(Q, R, dQ, dR, p, gauge_atol)
    if dR is not None and n > p:
        dQ1 = dQ1 - Q1 @ (dR12 @ R12.conj().T)
        dA[:, p:n] += Q1 @ dR12
    M = np.zeros((p, p), dtype=R.dtype)
    if dR is not None:
        M += dR11 @ R11.conj().T
    M -= Q1.conj().T @ dQ1
    iu = np.triu_indices(p, 1)
    M.T[iu] = np.conj(M[iu])              # copyltu: lower <- conj(upper)
    if np.iscomplexobj(M):
        M[np.arange(p), np.arange(p)] = np.real(np.diagonal(M))
    dQ1 = dQ1 + Q1 @ M
    # dA1 += dQ1 / R11^H
    # X R11^H = dQ1  <=>  R11 X^H = dQ1^H (LinearAlgebra._rdiv!, :118)
    X = sla.solve_triangular(R11, dQ1.conj().T, lower=False).conj().T
    dA[:, :p] += X
    return dA


def _sylvester(A, B, C):
    """Julia ``sylvester(A, B, C)``: X solving A X + X B + C = 0
    (src/common/pullbacks.jl:15)."""
    return sla.solve_sylvester(A, B, -C)


def left_polar_pullback(dA, A, WP, dWP):
    """``left_polar_pullback!`` (src/pullbacks/polar.jl:7-29)."""
    W, P = WP
    dW, dP = dWP
    dA = np.array(dA, copy=True)
    if dP is not None:
        dP = project_hermitian(dP)
    M = np.zeros_like(P)
    if dW is not None:
        M += W.conj().T @ dW
    if dP is not None:
        M -= dP @ P
    C = _sylvester(P, P, M.conj().T - M)
    if dP is not None:
        C = C + dP
    dA += W @ C
    if dW is not None:
        dWP = np.linalg.solve(P.T, dW.T).T          # dW / P
        dWP = dWP - W @ (W.conj().T @ dWP)
        dA += dWP
    return dA


def right_polar_pullback(dA, A, PWh, dPWh):
    """``right_polar_pullback!`` (src/pullbacks/polar.jl:37-59)."""
    P, Wh = PWh
    dP, dWh = dPWh
    dA = np.array(dA, copy=True)
    if dP is not None:
        dP = project_hermitian(dP)
    M = np.zeros_like(P)
    if dWh is not None:
        M += dWh @ Wh.conj().T
    if dP is not None:
        M -= P @ dP
    C = _sylvester(P, P, M.conj().T - M)
    if dP is not None:
        C = C + dP
    dA += C @ Wh
    if dWh is not None:
        X = np.linalg.solve(P, dWh)                 # P \ dWh
        X = X - (X @ Wh.conj().T) @ Wh
        dA += X
    return dA


# ---------------------------------------------------------------------------
# Seeded inputs for the measured configurations (builder-owned generator)
# ---------------------------------------------------------------------------
def gaussian(m, n, dtype=np.float64, seed=0):
    """iid N(0,1) entries (complex: real and imaginary parts iid N(0,1)), from
    numpy's PCG64 with a fixed seed, column-major."""
    rng = np.random.default_rng(seed)
    if np.dtype(dtype).kind == "c":
        re = rng.standard_normal((m, n))
        im = rng.standard_normal((m, n))
        return np.asfortranarray((re + 1j * im).astype(dtype))
    return np.asfortranarray(rng.standard_normal((m, n)).astype(dtype))


def haar_orthogonal(n, seed):
    """Q of a Gaussian with positive-diagonal R (the construction of
    ``instantiate_unitary``, test/testsuite/TestSuite.jl:85)."""
    Q, _ = qr_compact(gaussian(n, n, np.float64, seed))
    return Q


def conditioned_matrix(n, cond=1e6, seed=0):
    """A = U diag(sigma) V^T, U and V Haar-orthogonal, sigma log-spaced 1 .. 1/cond."""
    U = haar_orthogonal(n, seed)
    V = haar_orthogonal(n, seed + 1)
    s = np.logspace(0, -np.log10(cond), n)
    return np.asfortranarray((U * s[None, :]) @ V.T), s


def batched_shapes(count=4096, seed=0, mmin=16, mmax=1024, nmin=8):
    """m = floor(exp(U(ln 16, ln 1025))) clipped to [16, 1024], n ~ U{8..m}."""
    rng = np.random.default_rng(seed)
    m = np.floor(np.exp(rng.uniform(np.log(mmin), np.log(mmax + 1), count))).astype(np.int64)
    m = np.clip(m, mmin, mmax)
    n = np.array([rng.integers(nmin, mi + 1) for mi in m], dtype=np.int64)
    return m, n


def geqrf_flops(m, n):
    """LAPACK geqrf count: 2 n^2 (m - n/3) for m >= n, 2 m^2 (n - m/3) otherwise."""
    m = np.asarray(m, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    return np.where(m >= n, 2 * n * n * (m - n / 3), 2 * m * m * (n - m / 3))


def orgqr_flops(m, ncols, k):
    """LAPACK orgqr count for an m x ncols Q from k reflectors:
    4 m ncols k - 2 (m + ncols) k^2 + 4 k^3 / 3."""
    m = np.asarray(m, dtype=np.float64)
    ncols = np.asarray(ncols, dtype=np.float64)
    k = np.asarray(k, dtype=np.float64)
    return 4 * m * ncols * k - 2 * (m + ncols) * k * k + 4 * k ** 3 / 3


def qr_flops(m, n, complex_=False, kind="compact"):
    """Algorithmic flops of qr_compact (geqrf + orgqr; 4 m n^2 - 4 n^3/3 for
    m >= n) or qr_full (4 m^2 n - 2 m n^2 + 2 n^3 / 3); complex counts x4."""
    m = np.asarray(m, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    k = np.minimum(m, n)
    ncols = k if kind == "compact" else m
    f = geqrf_flops(m, n) + orgqr_flops(m, ncols, k)
    return f * (4 if complex_ else 1)
