"""Reverse rules of qr_compact! and left_polar! / right_polar! on the GPU.

Mirrors MatrixAlgebraKit's pullback API (declared public in
src/MatrixAlgebraKit.jl:69-75):

* ``qr_pullback_(dA, A, (Q, R), (dQ, dR), tape=...)`` adds the pullback of
  ``qr_compact(A; positive=true)`` to dA (src/pullbacks/qr.jl:73-120).  Q is
  re-formed on the device from the compact WY factors kept by
  :func:`qr_compact_rrule` (the primal of the rule, cf.
  ext/MatrixAlgebraKitChainRulesCoreExt.jl:26-43) instead of being read back.
* ``left_polar_pullback_(dA, A, (W, P), (dW, dP))`` and
  ``right_polar_pullback_(dA, A, (P, Wh), (dP, dWh))``
  (src/pullbacks/polar.jl:7-59).
* Absent cotangents are ``None`` (``iszerotangent``, src/common/pullbacks.jl:9-12).
* Batched forms take 3-D tensors (batch, m, n) whose matrices are column-major
  (see :func:`colmajor_batch`).
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
from dataclasses import dataclass

import torch

from . import _lib
from .api import (_DT, ArgumentError, DimensionMismatch, MI355XHouseholder, _check_device, _ld, _ptr, _status,
                  _stream, _workspace, colmajor, initialize_output_qr)

__all__ = ["QRTape", "qr_compact_rrule", "qr_rank", "qr_pullback_", "qr_pullback_batched_",
           "left_polar_pullback_", "left_polar_pullback_batched_", "right_polar_pullback_",
           "colmajor_batch", "empty_colmajor_batch"]

NB = 32


def colmajor_batch(x: torch.Tensor) -> torch.Tensor:
    """(batch, m, n) tensor whose matrices are column-major and packed (lazily
    conjugated / negated views are materialised)."""
    x = x.resolve_conj().resolve_neg()
    b, m, n = x.shape
    if x.stride(1) == 1 and x.stride(2) == m and x.stride(0) == m * n:
        return x
    return x.transpose(1, 2).contiguous().transpose(1, 2)


def empty_colmajor_batch(b, m, n, dtype, device):
    return torch.empty((b, n, m), dtype=dtype, device=device).transpose(1, 2)


def _bld(x):
    """leading dimension and batch stride of a colmajor_batch tensor"""
    b, m, n = x.shape
    if x.numel() and (x.stride(1) != 1 and m > 1):
        raise ArgumentError("batched matrices must be column-major")
    return max(1, x.stride(2) if n > 1 else m), x.stride(0)


@dataclass
class QRTape:
    """What the reverse rule keeps from the primal: the compact WY factors
    (V strictly below the diagonal of ``Awy``; one 32 x 32 T per panel)."""
    Awy: torch.Tensor
    T: torch.Tensor
    positive: bool = True


def qr_compact_rrule(A, alg=None):
    """Primal of ``rrule(qr_compact!, A, QR, alg)``: returns (Q, R, tape); A is
    copied first (``copy_input``) so the caller's A is untouched."""
    if alg is None:
        alg = MI355XHouseholder()
    if A.dtype not in _DT:
        raise ArgumentError(f"unsupported element type {A.dtype}")
    Awy = colmajor(A.clone() if A.stride(0) == 1 else A)
    Q, R = initialize_output_qr(Awy, False)
    _check_device(Awy, Q, R)
    m, n = Awy.shape
    k = min(m, n)
    T = torch.zeros((max(1, -(-k // NB)) * NB * NB,), dtype=A.dtype, device=A.device)
    lib = _lib.load()
    dt = _DT[A.dtype]
    nbytes = C.c_size_t(0)
    _status(lib.mak_qr_workspace_size(dt, 0, m, n, C.byref(nbytes)), "mak_qr_workspace_size")
    ws = _workspace(Awy.device, nbytes.value)
    fname = f"mak_qr_compact_wy_{'z' if dt else 'd'}"
    code = getattr(lib, fname)(_stream(Awy.device), m, n, _ptr(Awy), _ld(Awy), _ptr(Q), _ld(Q), _ptr(R), _ld(R),
                               _ptr(T), int(bool(alg.positive)), _ptr(ws), nbytes.value)
    _status(code, fname)
    return Q, R, QRTape(Awy, T, bool(alg.positive))


def _spacing(x: float) -> float:
    return math.ulp(x) if x > 0 else 0.0


def qr_rank(R: torch.Tensor, rank_atol=None) -> int:
    """``qr_rank`` (src/pullbacks/qr.jl:1-2): last index with |R_ii| >= rank_atol,
    rank_atol = eps(norm(R, Inf))^(3/4) (src/common/defaults.jl:57)."""
    if rank_atol is None:
        nrm = R.abs().max().item() if R.numel() else 0.0
        rank_atol = _spacing(nrm) ** 0.75
    d = torch.diagonal(R).abs().cpu()
    idx = torch.nonzero(d >= rank_atol)
    return int(idx[-1].item() + 1) if len(idx) else 0


def qr_pullback_batched_(dA, tapeA, T, R, dQ, dR, p=None, rank_atol=None, positive=True):
    """Batched ``qr_pullback!``: every argument is (batch, ., .) column-major;
    ``tapeA`` / ``T`` are the compact WY factors of the primal (computed with
    ``positive``)."""
    batch, m, n = dA.shape
    k = min(m, n)
    if R.shape != (batch, k, n):
        raise DimensionMismatch(f"size(R) = {tuple(R.shape)} instead of {(batch, k, n)}")
    if dQ is not None and dQ.shape != (batch, m, k):
        raise DimensionMismatch("ΔQ must have the same size as Q")
    if dR is not None and dR.shape != R.shape:
        raise DimensionMismatch("ΔR must have the same size as R")
    if p is None:
        ps = {qr_rank(R[b], rank_atol) for b in range(batch)}
        if len(ps) > 1:
            for b in range(batch):      # mixed ranks: one call per matrix
                qr_pullback_batched_(dA[b:b + 1], tapeA[b:b + 1], T[b:b + 1], R[b:b + 1],
                                     None if dQ is None else dQ[b:b + 1],
                                     None if dR is None else dR[b:b + 1], None, rank_atol, positive)
            return dA
        p = ps.pop()
    lib = _lib.load()
    dt = _DT[dA.dtype]
    nbytes = C.c_size_t(0)
    _status(lib.mak_pullback_workspace_size(0, dt, batch, m, n, C.byref(nbytes)), "mak_pullback_workspace_size")
    ws = _workspace(dA.device, nbytes.value)
    lda, sa = _bld(tapeA)
    ldr, sr = _bld(R)
    ldda, sda = _bld(dA)
    if dQ is not None:
        lddq, sdq = _bld(dQ)
    else:
        lddq, sdq = m, 0
    if dR is not None:
        lddr, sdr = _bld(dR)
    else:
        lddr, sdr = k, 0
    st = T.stride(0) if T.dim() == 2 else T.numel() // batch
    fname = f"mak_qr_pullback_{'z' if dt else 'd'}"
    code = getattr(lib, fname)(_stream(dA.device), batch, m, n, p, int(bool(positive)), _ptr(tapeA), lda, sa, _ptr(T), st, _ptr(R), ldr,
                               sr, _ptr(dQ), lddq, sdq, _ptr(dR), lddr, sdr, _ptr(dA), ldda, sda, _ptr(ws),
                               nbytes.value)
    _status(code, fname)
    return dA


def qr_pullback_(dA, A, QR, dQR, tape: QRTape | None = None, rank_atol=None):
    """``qr_pullback!(ΔA, A, (Q, R), (ΔQ, ΔR))``: ΔA += pullback.  Without a
    tape the compact WY factors are recomputed from A on the device."""
    Q, R = QR
    dQ, dR = dQR
    m, n = dA.shape
    if (m, n) != (Q.shape[0], R.shape[1]):
        raise DimensionMismatch(f"size of ΔA ({tuple(dA.shape)}) does not match size of Q*R ({m}, {R.shape[1]})")
    if tape is None:
        _, _, tape = qr_compact_rrule(A)
    T = tape.T.reshape(1, -1)
    dAc = colmajor(dA)
    qr_pullback_batched_(dAc.unsqueeze(0), tape.Awy.unsqueeze(0), T, colmajor(R).unsqueeze(0),
                         None if dQ is None else colmajor(dQ).unsqueeze(0),
                         None if dR is None else colmajor(dR).unsqueeze(0), None, rank_atol, tape.positive)
    if dAc is not dA:
        dA.copy_(dAc)
    return dA


def left_polar_pullback_batched_(dA, W, P, dW, dP, maxiter=200, tol=1e-14):
    batch, m, n = dA.shape
    if W.shape != (batch, m, n) or P.shape != (batch, n, n):
        raise DimensionMismatch("W / P sizes do not match ΔA")
    lib = _lib.load()
    dt = _DT[dA.dtype]
    nbytes = C.c_size_t(0)
    _status(lib.mak_pullback_workspace_size(1, dt, batch, m, n, C.byref(nbytes)), "mak_pullback_workspace_size")
    ws = _workspace(dA.device, nbytes.value)
    ldw, sw = _bld(W)
    ldp, sp = _bld(P)
    ldda, sda = _bld(dA)
    lddw, sdw = _bld(dW) if dW is not None else (m, 0)
    lddp, sdp = _bld(dP) if dP is not None else (n, 0)
    iters = C.c_int(0)
    fname = f"mak_left_polar_pullback_{'z' if dt else 'd'}"
    code = getattr(lib, fname)(_stream(dA.device), batch, m, n, _ptr(W), ldw, sw, _ptr(P), ldp, sp, _ptr(dW), lddw,
                               sdw, _ptr(dP), lddp, sdp, _ptr(dA), ldda, sda, int(maxiter), float(tol), _ptr(ws),
                               nbytes.value, C.byref(iters))
    if code == 6:
        warnings.warn(f"Sylvester solve of the polar pullback did not converge within {maxiter} iterations")
        code = 0
    _status(code, fname)
    return dA, iters.value


def left_polar_pullback_(dA, A, WP, dWP, maxiter=200, tol=1e-14):
    """``left_polar_pullback!(ΔA, A, (W, P), (ΔW, ΔP))``: ΔA += pullback."""
    W, P = WP
    dW, dP = dWP
    m, n = dA.shape
    if W.shape != (m, n) or P.shape != (n, n):
        raise DimensionMismatch(f"size of ΔA ({tuple(dA.shape)}) does not match W ({tuple(W.shape)}) and "
                                f"P ({tuple(P.shape)})")
    if dW is not None and dW.shape != W.shape:
        raise DimensionMismatch("ΔW must have the same size as W")
    if dP is not None and dP.shape != P.shape:
        raise DimensionMismatch("ΔP must have the same size as P")
    dAc = colmajor(dA)
    left_polar_pullback_batched_(dAc.unsqueeze(0), colmajor(W).unsqueeze(0), colmajor(P).unsqueeze(0),
                                 None if dW is None else colmajor(dW).unsqueeze(0),
                                 None if dP is None else colmajor(dP).unsqueeze(0), maxiter, tol)
    if dAc is not dA:
        dA.copy_(dAc)
    return dA


def right_polar_pullback_(dA, A, PWh, dPWh, maxiter=200, tol=1e-14):
    """``right_polar_pullback!(ΔA, A, (P, Wh), (ΔP, ΔWh))``: ΔA += pullback,
    computed as the left rule of A^H = Wh^H P (cotangents ΔWh^H and ΔP) and
    conjugate-transposed back."""
    P, Wh = PWh
    dP, dWh = dPWh
    m, n = dA.shape
    dB = colmajor(torch.zeros((n, m), dtype=dA.dtype, device=dA.device))
    W = colmajor(Wh.mH)
    dW = None if dWh is None else colmajor(dWh.mH)
    dPp = None if dP is None else colmajor(dP.mH)
    left_polar_pullback_(dB, None, (W, colmajor(P)), (dW, dPp), maxiter, tol)
    dA.add_(dB.mH)
    return dA
