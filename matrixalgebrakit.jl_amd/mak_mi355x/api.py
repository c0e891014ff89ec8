"""Host-side mirror of MatrixAlgebraKit's QR / polar interface over the C ABI.

Julia's ``f!`` functions are spelled ``f_`` here (``qr_compact!`` ->
``qr_compact_``).  Semantics follow the reference:

* ``check_input`` (src/implementations/qr.jl:9-37, polar.jl:6-29): sizes are
  checked with ``DimensionMismatch`` (``@check_size``, src/algorithms.jl:530-541),
  element types with ``ArgumentError`` (``@check_scalar``, :503-523);
* ``initialize_output`` (src/implementations/qr.jl:63-81, polar.jl:33-44)
  allocates the outputs; ``f(A)`` copies A first (``copy_input``, qr.jl:3-5);
* ``f!(A, out, alg)`` returns the SAME output objects it was given
  (test/testsuite/decompositions/polar.jl:30-31) and may destroy ``A``;
* an empty ``R`` (or ``P``) means "do not compute it" (qr.jl:149, polar.jl:64).

Matrices are torch tensors on a ROCm device, column-major: shape (m, n) with
``stride(0) == 1`` (e.g. ``torch.empty(n, m).t()`` or :func:`colmajor`).  The
work runs on the current torch stream; no implicit synchronisation.
"""
from __future__ import annotations

import ctypes as C
import warnings
from dataclasses import dataclass, field

import torch

from . import _lib

__all__ = [
    "DimensionMismatch", "ArgumentError", "NumericalError", "MI355XHouseholder", "MI355XPolar",
    "colmajor", "similar_colmajor",
    "qr_compact", "qr_compact_", "qr_full", "qr_full_",
    "check_input_qr_compact", "check_input_qr_full", "initialize_output_qr",
    "left_polar", "left_polar_", "right_polar", "right_polar_",
    "check_input_left_polar", "check_input_right_polar", "initialize_output_polar", "last_polar_info",
    "lq_compact", "lq_compact_", "lq_full", "lq_full_", "check_input_lq_compact", "check_input_lq_full",
    "initialize_output_lq", "qr_null", "qr_null_", "lq_null", "lq_null_", "check_input_qr_null",
    "check_input_lq_null", "initialize_output_qr_null", "initialize_output_lq_null",
    "PolarNewton", "left_orth_alg", "right_orth_alg", "left_orth", "left_orth_", "right_orth", "right_orth_",
    "left_null", "left_null_", "right_null", "right_null_", "project_isometric", "project_isometric_",
    "SVDViaPolar", "initialize_output_svd_compact", "svd_compact", "svd_compact_", "svd_vals", "svd_vals_",
    "left_polar_pushforward_", "right_polar_pushforward_",
]


class DimensionMismatch(ValueError):
    """Julia's ``DimensionMismatch`` (thrown by ``@check_size``)."""


class ArgumentError(ValueError):
    """Julia's ``ArgumentError`` (thrown by ``@check_scalar`` and option checks)."""


class NumericalError(RuntimeError):
    """A positive status from the library (numerical failure / launch error)."""


_DT = {torch.float64: 0, torch.complex128: 1}


@dataclass(frozen=True)
class MI355XHouseholder:
    """Algorithm type of this backend for ``qr_compact!`` / ``qr_full!``.

    Accepts the keyword arguments of ``Householder``
    (src/interface/decompositions.jl:80-86) with the GPU drivers' rules
    (src/implementations/qr.jl:137-145): ``positive`` (default true),
    ``pivoted`` (unsupported: ArgumentError) and ``blocksize`` (0 or 1: the
    unblocked API; larger values throw ArgumentError, the driver does not
    provide a blocked QR decomposition -- its internal panel width is its own).
    """
    positive: bool = True
    pivoted: bool = False
    blocksize: int = 0
    kwargs: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.blocksize > 1:
            raise ArgumentError("MI355X does not provide a blocked QR decomposition")
        if self.pivoted:
            raise ArgumentError("MI355X does not provide a pivoted QR decomposition")


# ---------------------------------------------------------------------------
# layout helpers
# ---------------------------------------------------------------------------
def colmajor(x: torch.Tensor) -> torch.Tensor:
    """Column-major copy (stride(0) == 1) of a 2-D tensor (no copy if already).

    Lazily conjugated or negated views (``x.mH`` of a complex tensor, ``-x``
    views) are materialised first: the library reads raw storage."""
    if x.dim() != 2:
        raise DimensionMismatch(f"expected a matrix, got shape {tuple(x.shape)}")
    x = x.resolve_conj().resolve_neg()
    m, n = x.shape
    # strides of size-1 dimensions carry no layout information
    if (m <= 1 or x.stride(0) == 1) and (n <= 1 or x.stride(1) >= max(1, m)):
        return x
    return x.t().contiguous().t()


def similar_colmajor(x: torch.Tensor, shape, dtype=None) -> torch.Tensor:
    """``similar(A, (m, n))`` with column-major layout on A's device."""
    m, n = shape
    return torch.empty((n, m), dtype=dtype or x.dtype, device=x.device).t()


def _ld(x: torch.Tensor) -> int:
    """Leading dimension of a column-major matrix."""
    m, n = x.shape
    if x.numel() == 0 or n == 1:
        if x.numel() > 0 and m > 1 and x.stride(0) != 1:
            raise ArgumentError("matrices must be column-major (stride(0) == 1)")
        return max(1, m)
    if x.stride(0) != 1 and m > 1:
        raise ArgumentError("matrices must be column-major (stride(0) == 1)")
    if x.stride(1) < m:
        raise ArgumentError("leading dimension smaller than the row count")
    return max(1, x.stride(1))


def _ptr(x):
    if x is None or x.numel() == 0:
        return C.c_void_p(0)
    if x.is_conj() or x.is_neg():
        # the storage holds the unconjugated / unnegated values
        raise ArgumentError("lazily conjugated or negated views must be resolved (colmajor) before the call")
    return C.c_void_p(x.data_ptr())


def _stream(dev) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _check_scalar(x, y, xname, yname):
    if x.dtype != y.dtype:
        raise ArgumentError(f"Unexpected scalar type: eltype({xname}) = {x.dtype} != eltype({yname}) = {y.dtype}")


def _check_size(x, sz, name):
    if tuple(x.shape) != tuple(sz):
        raise DimensionMismatch(f"size({name}) = {tuple(x.shape)} instead of expected value {tuple(sz)}")


def _check_device(*xs):
    for x in xs:
        if x is not None and x.numel() > 0 and x.device.type != "cuda":
            raise ArgumentError("MI355X backend needs device (ROCm) tensors")


def _status(code: int, fname: str):
    if code < 0:
        raise ArgumentError(f"{fname}: invalid argument {-code}")
    if code > 0:
        raise NumericalError(f"{fname}: status {code}")


def _workspace(dev, nbytes: int) -> torch.Tensor:
    """Device workspace for one call.  torch's caching allocator is
    stream-ordered, so concurrent calls on different streams never share it and
    steady-state calls do not reach hipMalloc."""
    return torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)


# ---------------------------------------------------------------------------
# QR (src/implementations/qr.jl)
# ---------------------------------------------------------------------------
def check_input_qr_full(A, QR):
    """``check_input(qr_full!, A, (Q, R), alg)`` (src/implementations/qr.jl:9-18)."""
    m, n = A.shape
    Q, R = QR
    _check_size(Q, (m, m), "Q")
    _check_scalar(Q, A, "Q", "A")
    if R.numel() > 0:
        _check_size(R, (m, n), "R")
    _check_scalar(R, A, "R", "A")


def check_input_qr_compact(A, QR):
    """``check_input(qr_compact!, A, (Q, R), alg)`` (src/implementations/qr.jl:19-29)."""
    m, n = A.shape
    k = min(m, n)
    Q, R = QR
    _check_size(Q, (m, k), "Q")
    _check_scalar(Q, A, "Q", "A")
    if R.numel() > 0:
        _check_size(R, (k, n), "R")
    _check_scalar(R, A, "R", "A")


def initialize_output_qr(A, full: bool):
    """``initialize_output(qr_{full,compact}!, A, alg)`` (src/implementations/qr.jl:63-75)."""
    m, n = A.shape
    k = min(m, n)
    if full:
        return similar_colmajor(A, (m, m)), similar_colmajor(A, (m, n))
    return similar_colmajor(A, (m, k)), similar_colmajor(A, (k, n))


def _qr_call(A, QR, alg, full: bool):
    if alg is None:
        alg = MI355XHouseholder()
    if A.dtype not in _DT:
        raise ArgumentError(f"unsupported element type {A.dtype} (Float64 / ComplexF64)")
    (check_input_qr_full if full else check_input_qr_compact)(A, QR)
    Q, R = QR
    _check_device(A, Q, R)
    m, n = A.shape
    if Q is A or (Q.numel() > 0 and Q.data_ptr() == A.data_ptr() and Q.shape == A.shape
                  and Q.stride() == A.stride()):
        # in-place Q (src/implementations/qr.jl:150-153)
        if R.numel() > 0 or alg.positive or m < n:
            raise ArgumentError("inplace Q only supported if matrix is tall (`m >= n`), R is not required, and "
                                "using the unblocked algorithm (`blocksize = 1`) with `positive = false`")
    lib = _lib.load()
    dt = _DT[A.dtype]
    nbytes = C.c_size_t(0)
    _status(lib.mak_qr_workspace_size(dt, int(full), m, n, C.byref(nbytes)), "mak_qr_workspace_size")
    dev = A.device if A.numel() else Q.device
    ws = _workspace(dev, nbytes.value)
    fname = f"mak_qr_{'full' if full else 'compact'}_{'z' if dt else 'd'}"
    fn = getattr(lib, fname)
    Rarg = R if R.numel() > 0 else None
    code = fn(_stream(dev), m, n, _ptr(A), _ld(A), _ptr(Q), _ld(Q), _ptr(Rarg),
              _ld(R) if Rarg is not None else 1, int(bool(alg.positive)), _ptr(ws), nbytes.value)
    _status(code, fname)
    return Q, R


def qr_compact_(A, QR=None, alg=None):
    """``qr_compact!(A, (Q, R), alg)``: A is destroyed; returns the given (Q, R)."""
    if QR is None:
        QR = initialize_output_qr(A, False)
    return _qr_call(A, QR, alg, False)


def qr_full_(A, QR=None, alg=None):
    """``qr_full!(A, (Q, R), alg)``: A is destroyed; returns the given (Q, R)."""
    if QR is None:
        QR = initialize_output_qr(A, True)
    return _qr_call(A, QR, alg, True)


def qr_compact(A, alg=None):
    """``qr_compact(A)``: works on a column-major copy of A (``copy_input``)."""
    return qr_compact_(colmajor(A.clone() if A.stride(0) == 1 else A), None, alg)


def qr_full(A, alg=None):
    """``qr_full(A)``: works on a column-major copy of A (``copy_input``)."""
    return qr_full_(colmajor(A.clone() if A.stride(0) == 1 else A), None, alg)


# ---------------------------------------------------------------------------
# Polar (src/implementations/polar.jl)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class MI355XPolar:
    """Algorithm type of this backend for ``left_polar!`` / ``right_polar!``:
    QDWH on the GPU.

    ``precision`` 64 runs every product in FP64.  32 / 16 keep the
    ill-conditioned QDWH steps (c > 100) in FP64 except the second CholeskyQR
    pass's products, then replace the well-conditioned steps by cubic
    polynomial steps X (alpha I - beta X^H X) on FP32 MFMA (BF16 for the late,
    small corrections at 16), followed by FP64 Newton-Schulz refinement.
    ``maxiter`` (default 20) and ``l0`` (lower bound of sigma_min(A) /
    ||A||_F, default 1e-10) tune the iteration; like ``PolarNewton``
    (src/interface/decompositions.jl:305) non-convergence only warns.
    """
    precision: int = 64
    maxiter: int = 20
    l0: float = 1e-10

    def __post_init__(self):
        if self.precision not in (64, 32, 16):
            raise ArgumentError(f"precision must be 64, 32 or 16 (got {self.precision})")


def check_input_left_polar(A, WP):
    """``check_input(left_polar!, A, (W, P), alg)`` (src/implementations/polar.jl:6-17)."""
    m, n = A.shape
    W, P = WP
    if m < n:
        raise ArgumentError(f"input matrix needs at least as many rows ({m}) as columns ({n})")
    _check_size(W, (m, n), "W")
    _check_scalar(W, A, "W", "A")
    if P.numel() > 0:
        _check_size(P, (n, n), "P")
    _check_scalar(P, A, "P", "A")


def check_input_right_polar(A, PWh):
    """``check_input(right_polar!, A, (P, Wh), alg)`` (src/implementations/polar.jl:18-29)."""
    m, n = A.shape
    P, Wh = PWh
    if n < m:
        raise ArgumentError(f"input matrix needs at least as many columns ({n}) as rows ({m})")
    if P.numel() > 0:
        _check_size(P, (m, m), "P")
    _check_scalar(P, A, "P", "A")
    _check_size(Wh, (m, n), "Wh")
    _check_scalar(Wh, A, "Wh", "A")


def initialize_output_polar(A, right: bool):
    """``initialize_output(left_polar! / right_polar!, A, alg)`` (src/implementations/polar.jl:33-44)."""
    m, n = A.shape
    if right:
        return similar_colmajor(A, (m, m)), similar_colmajor(A, (m, n))
    return similar_colmajor(A, (m, n)), similar_colmajor(A, (n, n))


_LAST_POLAR_INFO: dict = {}


def last_polar_info() -> dict:
    """Diagnostics (iterations, steps, alpha, l0, last difference) of the last
    polar call made through this module."""
    return dict(_LAST_POLAR_INFO)


def _polar_call(A, out, alg, right: bool):
    if alg is None:
        alg = MI355XPolar()
    elif isinstance(alg, PolarNewton):
        alg = alg.qdwh()
    if A.dtype not in _DT:
        raise ArgumentError(f"unsupported element type {A.dtype} (Float64 / ComplexF64)")
    if right:
        check_input_right_polar(A, out)
        P, W = out
    else:
        check_input_left_polar(A, out)
        W, P = out
    _check_device(A, W, P)
    if A.dtype == torch.complex128 and alg.precision != 64:
        raise ArgumentError("reduced-precision polar is only available for Float64")
    lib = _lib.load()
    m, n = A.shape
    dt = _DT[A.dtype]
    nbytes = C.c_size_t(0)
    _status(lib.mak_polar_workspace_size(dt, m, n, C.byref(nbytes)), "mak_polar_workspace_size")
    dev = A.device
    ws = _workspace(dev, nbytes.value)
    info = _lib.PolarInfo()
    Parg = P if P.numel() > 0 else None
    fname = f"mak_{'right' if right else 'left'}_polar_{'z' if dt else 'd'}"
    fn = getattr(lib, fname)
    if right:
        code = fn(_stream(dev), m, n, _ptr(A), _ld(A), _ptr(Parg), _ld(P) if Parg is not None else 1,
                  _ptr(W), _ld(W), int(alg.precision), int(alg.maxiter), float(alg.l0), _ptr(ws),
                  nbytes.value, C.byref(info))
    else:
        code = fn(_stream(dev), m, n, _ptr(A), _ld(A), _ptr(W), _ld(W), _ptr(Parg),
                  _ld(P) if Parg is not None else 1, int(alg.precision), int(alg.maxiter), float(alg.l0),
                  _ptr(ws), nbytes.value, C.byref(info))
    _LAST_POLAR_INFO.clear()
    _LAST_POLAR_INFO.update(info.asdict())
    if code == 6:
        warnings.warn(f"`{'right' if right else 'left'}_polar!` via QDWH did not converge within "
                      f"{alg.maxiter} iterations (final difference: {info.last_diff})")
        code = 0
    _status(code, fname)
    return out


def left_polar_(A, WP=None, alg=None):
    """``left_polar!(A, (W, P), alg)``: returns the given (W, P); empty P skips P."""
    if WP is None:
        WP = initialize_output_polar(A, False)
    return _polar_call(A, WP, alg, False)


def right_polar_(A, PWh=None, alg=None):
    """``right_polar!(A, (P, Wh), alg)``: returns the given (P, Wh); empty P skips P."""
    if PWh is None:
        PWh = initialize_output_polar(A, True)
    return _polar_call(A, PWh, alg, True)


def left_polar(A, alg=None):
    """``left_polar(A)``."""
    return left_polar_(colmajor(A), None, alg)


def right_polar(A, alg=None):
    """``right_polar(A)``."""
    return right_polar_(colmajor(A), None, alg)


# ---------------------------------------------------------------------------
# LQ (src/implementations/lq.jl) through the QR of A^H
# ---------------------------------------------------------------------------
def check_input_lq_full(A, LQ):
    """``check_input(lq_full!, A, (L, Q), alg)`` (src/implementations/lq.jl:9-18)."""
    m, n = A.shape
    L, Q = LQ
    if L.numel() > 0:
        _check_size(L, (m, n), "L")
    _check_scalar(L, A, "L", "A")
    _check_size(Q, (n, n), "Q")
    _check_scalar(Q, A, "Q", "A")


def check_input_lq_compact(A, LQ):
    """``check_input(lq_compact!, A, (L, Q), alg)`` (src/implementations/lq.jl:19-29)."""
    m, n = A.shape
    k = min(m, n)
    L, Q = LQ
    if L.numel() > 0:
        _check_size(L, (m, k), "L")
    _check_scalar(L, A, "L", "A")
    _check_size(Q, (k, n), "Q")
    _check_scalar(Q, A, "Q", "A")


def initialize_output_lq(A, full: bool):
    """``initialize_output(lq_{full,compact}!, A, alg)`` (src/implementations/lq.jl:62-74)."""
    m, n = A.shape
    k = min(m, n)
    if full:
        return similar_colmajor(A, (m, n)), similar_colmajor(A, (n, n))
    return similar_colmajor(A, (m, k)), similar_colmajor(A, (k, n))


def _lq_call(A, LQ, alg, full: bool):
    if alg is None:
        alg = MI355XHouseholder()
    if A.dtype not in _DT:
        raise ArgumentError(f"unsupported element type {A.dtype} (Float64 / ComplexF64)")
    (check_input_lq_full if full else check_input_lq_compact)(A, LQ)
    L, Q = LQ
    _check_device(A, L, Q)
    m, n = A.shape
    lib = _lib.load()
    dt = _DT[A.dtype]
    nbytes = C.c_size_t(0)
    _status(lib.mak_lq_workspace_size(dt, int(full), m, n, C.byref(nbytes)), "mak_lq_workspace_size")
    dev = A.device if A.numel() else Q.device
    ws = _workspace(dev, nbytes.value)
    fname = f"mak_lq_{'full' if full else 'compact'}_{'z' if dt else 'd'}"
    Larg = L if L.numel() > 0 else None
    code = getattr(lib, fname)(_stream(dev), m, n, _ptr(A), _ld(A), _ptr(Larg), _ld(L) if Larg is not None else 1,
                               _ptr(Q), _ld(Q), int(bool(alg.positive)), _ptr(ws), nbytes.value)
    _status(code, fname)
    return L, Q


def lq_compact_(A, LQ=None, alg=None):
    """``lq_compact!(A, (L, Q), alg)`` as ``LQViaTransposedQR`` (src/implementations/lq.jl:290-296):
    returns the given (L, Q); A is only read."""
    if LQ is None:
        LQ = initialize_output_lq(A, False)
    return _lq_call(A, LQ, alg, False)


def lq_full_(A, LQ=None, alg=None):
    """``lq_full!(A, (L, Q), alg)`` as ``LQViaTransposedQR`` (src/implementations/lq.jl:284-289)."""
    if LQ is None:
        LQ = initialize_output_lq(A, True)
    return _lq_call(A, LQ, alg, True)


def lq_compact(A, alg=None):
    return lq_compact_(colmajor(A), None, alg)


def lq_full(A, alg=None):
    return lq_full_(colmajor(A), None, alg)


# ---------------------------------------------------------------------------
# Null spaces (src/implementations/qr.jl:114-117,232-263; lq.jl:297-318)
# ---------------------------------------------------------------------------
def check_input_qr_null(A, N):
    """``check_input(qr_null!, A, N, alg)`` (src/implementations/qr.jl:30-37)."""
    m, n = A.shape
    _check_size(N, (m, m - min(m, n)), "N")
    _check_scalar(N, A, "N", "A")


def check_input_lq_null(A, Nh):
    """``check_input(lq_null!, A, Nᴴ, alg)`` (src/implementations/lq.jl:30-37)."""
    m, n = A.shape
    _check_size(Nh, (n - min(m, n), n), "Nᴴ")
    _check_scalar(Nh, A, "Nᴴ", "A")


def initialize_output_qr_null(A):
    m, n = A.shape
    return similar_colmajor(A, (m, m - min(m, n)))


def initialize_output_lq_null(A):
    m, n = A.shape
    return similar_colmajor(A, (n - min(m, n), n))


def _null_call(A, N, alg, lq: bool):
    if alg is None:
        alg = MI355XHouseholder()
    if A.dtype not in _DT:
        raise ArgumentError(f"unsupported element type {A.dtype} (Float64 / ComplexF64)")
    (check_input_lq_null if lq else check_input_qr_null)(A, N)
    _check_device(A, N)
    m, n = A.shape
    lib = _lib.load()
    dt = _DT[A.dtype]
    nbytes = C.c_size_t(0)
    _status(lib.mak_null_workspace_size(dt, int(lq), m, n, C.byref(nbytes)), "mak_null_workspace_size")
    dev = A.device if A.numel() else N.device
    ws = _workspace(dev, nbytes.value)
    fname = f"mak_{'lq' if lq else 'qr'}_null_{'z' if dt else 'd'}"
    code = getattr(lib, fname)(_stream(dev), m, n, _ptr(A), _ld(A), _ptr(N), _ld(N), _ptr(ws), nbytes.value)
    _status(code, fname)
    return N


def qr_null_(A, N=None, alg=None):
    """``qr_null!(A, N, alg)``: N = H_1 ... H_k [0; I] (src/implementations/qr.jl:236-263); A is destroyed."""
    if N is None:
        N = initialize_output_qr_null(A)
    return _null_call(A, N, alg, False)


def lq_null_(A, Nh=None, alg=None):
    """``lq_null!(A, Nᴴ, alg)`` via the QR null space of Aᴴ (src/implementations/lq.jl:297-301,312-318)."""
    if Nh is None:
        Nh = initialize_output_lq_null(A)
    return _null_call(A, Nh, alg, True)


def qr_null(A, alg=None):
    return qr_null_(colmajor(A.clone() if A.stride(0) == 1 else A), None, alg)


def lq_null(A, alg=None):
    return lq_null_(colmajor(A), None, alg)


# ---------------------------------------------------------------------------
# PolarNewton on the device, orth / null composites, project_isometric!
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class PolarNewton:
    """``PolarNewton(; maxiter, tol)`` (src/interface/decompositions.jl:305) on
    the device.  The reference's Newton loop (src/implementations/polar.jl:
    128-206) needs an LU-based inverse per step; here the same unique polar
    factor comes from the QDWH iteration of ``MI355XPolar`` (FP64), which
    converges in at most 6 GEMM-rich steps for condition numbers up to 1e16.
    ``maxiter`` caps the iterations (default 10, as the reference) and
    non-convergence warns as there.  ``tol`` (default eps^(2/3),
    ``defaulttol``) is accepted for the interface: QDWH's own stopping test,
    ||X_k - X_{k-1}||_F <= (5u)^(1/3) ||X_k||_F after which the cubically
    convergent step leaves an error of order u, is the stricter one."""
    maxiter: int = 10
    tol: float | None = None

    def __post_init__(self):
        if self.tol is not None and not self.tol > 0:
            raise ArgumentError(f"tol must be positive (got {self.tol})")
        if self.maxiter < 1:
            raise ArgumentError(f"maxiter must be positive (got {self.maxiter})")

    def qdwh(self) -> "MI355XPolar":
        return MI355XPolar(precision=64, maxiter=self.maxiter)


def _polar_alg(alg):
    return alg.qdwh() if isinstance(alg, PolarNewton) else alg


def left_orth_alg(alg):
    """``left_orth_alg`` (src/interface/orthnull.jl:442-444): QR algorithms
    factor with ``qr_compact!``, polar algorithms with ``left_polar!``."""
    if alg is None or isinstance(alg, MI355XHouseholder) or alg == "qr":
        return "qr", (alg if isinstance(alg, MI355XHouseholder) else MI355XHouseholder())
    if isinstance(alg, (MI355XPolar, PolarNewton)) or alg == "polar":
        return "polar", (_polar_alg(alg) if not isinstance(alg, str) else MI355XPolar())
    raise ArgumentError(f"Unknown or invalid `left_orth` algorithm type `{type(alg).__name__}`")


right_orth_alg = left_orth_alg


def left_orth_(A, VC=None, alg=None):
    """``left_orth!(A, (V, C), alg)`` (src/implementations/orthnull.jl:79-81):
    (Q, R) of ``qr_compact!`` or (W, P) of ``left_polar!``."""
    kind, a = left_orth_alg(alg)
    if kind == "qr":
        return qr_compact_(A, VC, a)
    return left_polar_(A, VC, a)


def right_orth_(A, CVh=None, alg=None):
    """``right_orth!(A, (C, Vᴴ), alg)`` (src/implementations/orthnull.jl:88-90):
    (L, Q) of ``lq_compact!`` or (P, Wᴴ) of ``right_polar!``."""
    kind, a = right_orth_alg(alg)
    if kind == "qr":
        return lq_compact_(A, CVh, a)
    return right_polar_(A, CVh, a)


def left_orth(A, alg=None):
    return left_orth_(colmajor(A.clone() if A.stride(0) == 1 else A), None, alg)


def right_orth(A, alg=None):
    return right_orth_(colmajor(A.clone() if A.stride(0) == 1 else A), None, alg)


def left_null_(A, N=None, alg=None):
    """``left_null!(A, N, ::LeftNullViaQR)`` = ``qr_null!`` (src/implementations/orthnull.jl:101-102)."""
    return qr_null_(A, N, alg)


def right_null_(A, Nh=None, alg=None):
    """``right_null!(A, Nᴴ, ::RightNullViaLQ)`` = ``lq_null!`` (src/implementations/orthnull.jl:110-111)."""
    return lq_null_(A, Nh, alg)


def left_null(A, alg=None):
    return qr_null(A, alg)


def right_null(A, alg=None):
    return lq_null(A, alg)


def project_isometric_(A, W=None, alg=None):
    """``project_isometric!(A, W, alg)`` (src/implementations/projections.jl:70-75):
    the polar factor W of ``left_polar!`` without P."""
    m, n = A.shape
    if m < n:
        raise DimensionMismatch("project_isometric! needs size(A, 1) >= size(A, 2)")
    if W is None:
        W = similar_colmajor(A, (m, n))
    noP = torch.empty((0, 0), dtype=A.dtype, device=A.device)
    left_polar_(A, (W, noP), _polar_alg(alg) if alg is not None else None)
    return W


def project_isometric(A, alg=None):
    return project_isometric_(colmajor(A), None, alg)


# ---------------------------------------------------------------------------
# SVD via the polar decomposition (SVDViaPolar) and the polar pushforwards
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class SVDViaPolar:
    """``SVDViaPolar(; fixgauge, tol)`` (src/interface/decompositions.jl:173-183)
    on the device: QDWH polar (FP64) + Hermitian eigendecomposition of P.
    ``fixgauge`` (default true) applies ``gaugefix!(svd_compact!, U, Vh)``
    (the library always does: a caller asking for no gauge gets the fixed one,
    which is one of the valid outputs); ``maxiter`` caps the QDWH iterations."""
    fixgauge: bool = True
    tol: float | None = None
    maxiter: int = 20


def initialize_output_svd_compact(A):
    """``initialize_output(svd_compact!, A, alg)``: U m x k, S (the diagonal of
    the reference's ``Diagonal``, real, length k), Vh k x n."""
    m, n = A.shape
    k = min(m, n)
    rdt = torch.float64
    return (similar_colmajor(A, (m, k)), torch.empty(k, dtype=rdt, device=A.device), similar_colmajor(A, (k, n)))


def _svd_polar_call(A, U, S, Vh, alg):
    if alg is None:
        alg = SVDViaPolar()
    if not isinstance(alg, SVDViaPolar):
        raise ArgumentError(f"unsupported SVD algorithm {type(alg).__name__} (SVDViaPolar)")
    if A.dtype not in _DT:
        raise ArgumentError(f"unsupported element type {A.dtype} (Float64 / ComplexF64)")
    m, n = A.shape
    k = min(m, n)
    _check_size(S, (k,), "S")
    if S.dtype != torch.float64:
        raise ArgumentError("S must hold real Float64 values")
    if U is not None:
        _check_size(U, (m, k), "U")
        _check_size(Vh, (k, n), "Vᴴ")
        _check_scalar(U, A, "U", "A")
        _check_scalar(Vh, A, "Vᴴ", "A")
    _check_device(A, U, S, Vh)
    lib = _lib.load()
    dt = _DT[A.dtype]
    nbytes = C.c_size_t(0)
    _status(lib.mak_svd_polar_workspace_size(dt, m, n, C.byref(nbytes)), "mak_svd_polar_workspace_size")
    ws = _workspace(A.device, nbytes.value)
    info = _lib.PolarInfo()
    fname = f"mak_svd_polar_{'z' if dt else 'd'}"
    code = getattr(lib, fname)(_stream(A.device), m, n, _ptr(A), _ld(A), C.c_void_p(S.data_ptr() if k else 0),
                               _ptr(U), _ld(U) if U is not None else 1, _ptr(Vh), _ld(Vh) if Vh is not None else 1,
                               int(alg.maxiter), _ptr(ws), nbytes.value, C.byref(info))
    if code == 6:
        warnings.warn("SVDViaPolar: the polar iteration did not converge within maxiter")
        code = 0
    if code == 3:
        raise NumericalError(f"{fname}: rocSOLVER (the Hermitian eigensolver) could not be opened")
    _status(code, fname)


def svd_compact_(A, USVh=None, alg=None):
    """``svd_compact!(A, (U, S, Vᴴ), SVDViaPolar())``: returns the given (U, S, Vᴴ); A is only read."""
    if USVh is None:
        USVh = initialize_output_svd_compact(A)
    U, S, Vh = USVh
    _svd_polar_call(A, U, S, Vh, alg)
    return USVh


def svd_vals_(A, S=None, alg=None):
    """``svd_vals!(A, S, SVDViaPolar())``: the singular values, descending."""
    if S is None:
        S = torch.empty(min(A.shape), dtype=torch.float64, device=A.device)
    _svd_polar_call(A, None, S, None, alg)
    return S


def svd_compact(A, alg=None):
    return svd_compact_(colmajor(A), None, alg)


def svd_vals(A, alg=None):
    return svd_vals_(colmajor(A), None, alg)


def _pushforward_call(dA, WP_or_PWh, dout, right: bool):
    lib = _lib.load()
    if right:
        P, W = WP_or_PWh
        dP, dW = dout
    else:
        W, P = WP_or_PWh
        dW, dP = dout
    if dA.dtype not in _DT:
        raise ArgumentError(f"unsupported element type {dA.dtype} (Float64 / ComplexF64)")
    m, n = dA.shape
    k = min(m, n)
    if (not right and m < n) or (right and m > n):
        raise DimensionMismatch(f"{'right' if right else 'left'} polar pushforward: unsupported shape {(m, n)}")
    for x, sz, nm in ((W, (m, n), "W"), (P, (k, k), "P"), (dW, (m, n), "ΔW"), (dP, (k, k), "ΔP")):
        _check_size(x, sz, nm)
        _check_scalar(x, dA, nm, "ΔA")
    _check_device(dA, W, P, dW, dP)
    dt = _DT[dA.dtype]
    nbytes = C.c_size_t(0)
    _status(lib.mak_polar_pushforward_workspace_size(dt, int(right), m, n, C.byref(nbytes)),
            "mak_polar_pushforward_workspace_size")
    ws = _workspace(dA.device, nbytes.value)
    fname = f"mak_{'right' if right else 'left'}_polar_pushforward_{'z' if dt else 'd'}"
    code = getattr(lib, fname)(_stream(dA.device), m, n, _ptr(W), _ld(W), _ptr(P), _ld(P), _ptr(dA), _ld(dA),
                               _ptr(dW), _ld(dW), _ptr(dP), _ld(dP), _ptr(ws), nbytes.value)
    if code == 2:
        raise NumericalError(f"{fname}: P is not positive definite (rank-deficient A)")
    if code == 3:
        raise NumericalError(f"{fname}: rocSOLVER (the Hermitian eigensolver) could not be opened")
    _status(code, fname)
    return dout


def left_polar_pushforward_(dA, A, WP, dWP=None):
    """``left_polar_pushforward!(ΔA, A, (W, P), (ΔW, ΔP))`` (src/pushforwards/polar.jl:1-12):
    returns (ΔW, ΔP), overwritten."""
    W, P = WP
    if dWP is None:
        dWP = (similar_colmajor(W, W.shape), similar_colmajor(P, P.shape))
    return _pushforward_call(dA, WP, dWP, False)


def right_polar_pushforward_(dA, A, PWh, dPWh=None):
    """``right_polar_pushforward!(ΔA, A, (P, Wᴴ), (ΔP, ΔWᴴ))`` (src/pushforwards/polar.jl:14-25):
    returns (ΔWᴴ, ΔP) as the reference does, overwritten."""
    P, Wh = PWh
    if dPWh is None:
        dPWh = (similar_colmajor(P, P.shape), similar_colmajor(Wh, Wh.shape))
    dP, dWh = _pushforward_call(dA, PWh, dPWh, True)
    return dWh, dP
