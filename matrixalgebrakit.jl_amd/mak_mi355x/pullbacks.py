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
import warnings
from dataclasses import dataclass

import torch

from . import _lib
from .api import (_DT, ArgumentError, DimensionMismatch, MI355XHouseholder, _check_device, _ld, _ptr, _status,
                  _stream, _workspace, colmajor, initialize_output_qr)

__all__ = ["QRTape", "qr_compact_rrule", "qr_rank", "qr_ranks", "qr_pullback_", "qr_pullback_batched_",
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


def qr_rank(R: torch.Tensor, rank_atol=None) -> int:
    """``qr_rank`` (src/pullbacks/qr.jl:1-2): last index with |R_ii| >= rank_atol,
    rank_atol = eps(norm(R, Inf))^(3/4) (src/common/defaults.jl:57)."""
    return int(qr_ranks(R.unsqueeze(0), rank_atol)[0])


def _spacing_t(x: torch.Tensor) -> torch.Tensor:
    """Julia's eps(x) for non-negative float64 tensors (0 for x = 0)."""
    _, e = torch.frexp(x)
    return torch.where(x > 0, torch.ldexp(torch.ones_like(x), e - 53), torch.zeros_like(x))


def _maxabs_batch(X: torch.Tensor) -> torch.Tensor:
    b = X.shape[0]
    if X.numel() == 0:
        return torch.zeros(b, dtype=torch.float64, device=X.device)
    return X.abs().reshape(b, -1).amax(dim=1).to(torch.float64)


def qr_ranks(R3: torch.Tensor, rank_atol=None):
    """``qr_rank`` of every matrix of a (batch, k, n) stack, on the device:
    one host transfer for the whole batch."""
    b, k, n = R3.shape
    kk = min(k, n)
    if kk == 0:
        return [0] * b
    d = torch.diagonal(R3, dim1=1, dim2=2).abs().to(torch.float64)
    if rank_atol is None:
        atol = _spacing_t(_maxabs_batch(R3)) ** 0.75
    else:
        atol = torch.full((b,), float(rank_atol), dtype=torch.float64, device=R3.device)
    pos = torch.arange(1, kk + 1, device=R3.device, dtype=torch.int64)
    p = torch.where(d >= atol[:, None], pos[None, :], torch.zeros_like(pos)[None, :]).amax(dim=1)
    return [int(v) for v in p.cpu().tolist()]


def qr_pullback_batched_(dA, tapeA, T, R, dQ, dR, p=None, rank_atol=None, positive=True, gauge_atol=None, Q=None):
    """Batched ``qr_pullback!``: every argument is (batch, ., .) column-major;
    ``tapeA`` / ``T`` are the compact WY factors of the primal (computed with
    ``positive``).  ``dQ`` is (batch, m, min(m,n)), or (batch, m, m) for a
    ``qr_full!`` primal, whose full Q must then be passed as ``Q``.  The
    gauge-dependent part of the cotangents is measured on the device and, as
    in src/pullbacks/qr.jl:50-51, a warning is issued when it exceeds
    ``gauge_atol`` (default eps(norm(ΔQ, Inf))^(3/4) per matrix)."""
    batch, m, n = dA.shape
    k = min(m, n)
    if R.shape != (batch, k, n):
        raise DimensionMismatch(f"size(R) = {tuple(R.shape)} instead of {(batch, k, n)}")
    full = dQ is not None and tuple(dQ.shape) == (batch, m, m) and m > k
    if dQ is not None and not full and dQ.shape != (batch, m, k):
        raise DimensionMismatch("ΔQ must have the same size as Q")
    if full and (Q is None or tuple(Q.shape) != (batch, m, m)):
        raise DimensionMismatch("a qr_full! cotangent ΔQ (m x m) needs the full Q")
    if dR is not None and dR.shape != R.shape:
        raise DimensionMismatch("ΔR must have the same size as R")
    ps = [p] * batch if p is not None else qr_ranks(R, rank_atol)
    if len(set(ps)) > 1:
        for b in range(batch):      # mixed ranks: one call per matrix
            qr_pullback_batched_(dA[b:b + 1], tapeA[b:b + 1], T[b:b + 1], R[b:b + 1],
                                 None if dQ is None else dQ[b:b + 1], None if dR is None else dR[b:b + 1],
                                 ps[b], rank_atol, positive, gauge_atol, None if Q is None else Q[b:b + 1])
        return dA
    p = ps[0] if batch else 0
    lib = _lib.load()
    dt = _DT[dA.dtype]
    dev = dA.device
    gauge = torch.zeros(max(batch, 1), dtype=torch.float64, device=dev) if (full or p < k) else None
    if full:
        # check_and_prepare_qr_cotangents for the full Q (src/pullbacks/qr.jl:19-26)
        nb = C.c_size_t(0)
        _status(lib.mak_pullback_workspace_size(2, dt, batch, m, n, C.byref(nb)), "mak_pullback_workspace_size")
        fws = _workspace(dev, nb.value)
        gfull = torch.zeros(max(batch, 1), dtype=torch.float64, device=dev)
        dQ1 = empty_colmajor_batch(batch, m, k, dA.dtype, dev)
        ldq, sq = _bld(Q)
        lddq, sdq = _bld(dQ)
        ld1, sd1 = _bld(dQ1)
        fname = f"mak_qr_full_cotangent_{'z' if dt else 'd'}"
        code = getattr(lib, fname)(_stream(dev), batch, m, n, p, _ptr(Q), ldq, sq, _ptr(dQ), lddq, sdq, _ptr(dQ1),
                                   ld1, sd1, C.c_void_p(gfull.data_ptr()), _ptr(fws), nb.value)
        _status(code, fname)
        dQ_full, dQ = dQ, dQ1
    nbytes = C.c_size_t(0)
    _status(lib.mak_pullback_workspace_size(0, dt, batch, m, n, C.byref(nbytes)), "mak_pullback_workspace_size")
    ws = _workspace(dev, nbytes.value)
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
    code = getattr(lib, fname)(_stream(dev), batch, m, n, p, int(bool(positive)), _ptr(tapeA), lda, sa, _ptr(T), st,
                               _ptr(R), ldr, sr, _ptr(dQ), lddq, sdq, _ptr(dR), lddr, sdr, _ptr(dA), ldda, sda,
                               _ptr(ws), nbytes.value, C.c_void_p(gauge.data_ptr() if gauge is not None else 0))
    _status(code, fname)
    # a full-rank compact cotangent has no gauge-dependent entries: nothing to check
    if batch and (full or p < k):
        if full:
            # p = k: the gauge part of a full cotangent is the one the preparation measured
            gauge = torch.maximum(gauge, gfull)
            dQ = dQ_full
        if gauge_atol is None:
            atol = (_spacing_t(_maxabs_batch(dQ)) ** 0.75 if dQ is not None
                    else torch.zeros(batch, dtype=torch.float64, device=dev))
        else:
            atol = torch.full((batch,), float(gauge_atol), dtype=torch.float64, device=dev)
        g, a = torch.stack([gauge[:batch], atol]).cpu().tolist()
        worst = max(g)
        if any(gi > ai for gi, ai in zip(g, a)):
            warnings.warn(f"`qr` cotangents sensitive to gauge choice: (|Δgauge| = {worst})")
    return dA


def qr_pullback_(dA, A, QR, dQR, tape: QRTape | None = None, rank_atol=None, gauge_atol=None):
    """``qr_pullback!(ΔA, A, (Q, R), (ΔQ, ΔR))``: ΔA += pullback.  Q / ΔQ may be
    the compact (m x min(m,n)) or the full (m x m) factor.  Without a tape the
    compact WY factors are recomputed from A on the device."""
    Q, R = QR
    dQ, dR = dQR
    m, n = dA.shape
    if (m, n) != (Q.shape[0], R.shape[1]):
        raise DimensionMismatch(f"size of ΔA ({tuple(dA.shape)}) does not match size of Q*R ({m}, {R.shape[1]})")
    if dQ is not None and tuple(dQ.shape) != tuple(Q.shape):
        raise DimensionMismatch("ΔQ must have the same size as Q")
    k = min(m, n)
    if R.shape[0] != k:
        # qr_full!'s R (m x n): its rows beyond min(m, n) are zero, and so is
        # the part of a valid cotangent there
        R = R[:k]
        dR = None if dR is None else dR[:k]
    if tape is None:
        _, _, tape = qr_compact_rrule(A)
    T = tape.T.reshape(1, -1)
    full = Q.shape[1] == m and m > k
    dAc = colmajor(dA)
    qr_pullback_batched_(dAc.unsqueeze(0), tape.Awy.unsqueeze(0), T, colmajor(R).unsqueeze(0),
                         None if dQ is None else colmajor(dQ).unsqueeze(0),
                         None if dR is None else colmajor(dR).unsqueeze(0), None, rank_atol, tape.positive,
                         gauge_atol, colmajor(Q).unsqueeze(0) if full else None)
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
