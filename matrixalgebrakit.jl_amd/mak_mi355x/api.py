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

    ``precision`` 64 runs every product in FP64; 32 / 16 run the bulk products
    of the iteration in FP32 / BF16 (FP32 accumulation) with FP64 Gram
    matrices, Cholesky factorisations, weights and first CholeskyQR pass of
    the QR iterations, followed by FP64 Newton-Schulz refinement.  ``maxiter`` (default 20) and ``l0`` (lower bound
    of sigma_min(A) / ||A||_F, default 1e-10) tune the iteration; like
    ``PolarNewton`` (src/interface/decompositions.jl:305) non-convergence only
    warns.
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
