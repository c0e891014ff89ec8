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
from dataclasses import dataclass, field

import torch

from . import _lib

__all__ = [
    "DimensionMismatch", "ArgumentError", "MI355XHouseholder", "colmajor", "similar_colmajor",
    "qr_compact", "qr_compact_", "qr_full", "qr_full_",
    "check_input_qr_compact", "check_input_qr_full", "initialize_output_qr",
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
    (src/interface/decompositions.jl:80-86): ``positive`` (default true),
    ``pivoted`` (unsupported: ArgumentError, as for the GPU drivers,
    src/implementations/qr.jl:142-143) and ``blocksize`` (the panel width is
    fixed at 32 columns by the kernels; any value is accepted).
    """
    positive: bool = True
    pivoted: bool = False
    blocksize: int = 0
    kwargs: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.pivoted:
            raise ArgumentError("MI355X does not provide a pivoted QR decomposition")


# ---------------------------------------------------------------------------
# layout helpers
# ---------------------------------------------------------------------------
def colmajor(x: torch.Tensor) -> torch.Tensor:
    """Column-major copy (stride(0) == 1) of a 2-D tensor (no copy if already)."""
    if x.dim() != 2:
        raise DimensionMismatch(f"expected a matrix, got shape {tuple(x.shape)}")
    if x.stride(0) == 1 and x.stride(1) >= max(1, x.shape[0]):
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
    return C.c_void_p(x.data_ptr()) if x is not None and x.numel() > 0 else C.c_void_p(0)


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
    lib = _lib.load()
    m, n = A.shape
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
