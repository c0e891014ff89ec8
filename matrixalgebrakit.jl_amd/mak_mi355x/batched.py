"""Batched qr_compact! over many independent, unevenly sized problems.

The reference has no batched API; a tensor-network caller loops
``qr_compact!`` over its blocks (src/implementations/qr.jl:110-113 per
block).  :class:`BatchedQRPlan` runs all blocks through one captured hipGraph
(``mak_qr_batched_plan_create`` / ``mak_qr_batched_execute``); every block
gets exactly what ``qr_compact!(A_i, (Q_i, R_i), Householder(; positive))``
gives it, and ``A_i`` is overwritten like the single-matrix call.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .api import (_DT, ArgumentError, MI355XHouseholder, _ld, _status, _stream, _workspace, colmajor,
                  initialize_output_qr)

__all__ = ["BatchedQRPlan", "qr_compact_batched", "block_shapes"]


def block_shapes(count: int, seed: int = 0, mmin: int = 16, mmax: int = 1024, nmin: int = 8):
    """Block shapes of a block-sparse tensor-network workload: m log-uniform in
    [mmin, mmax], n uniform in [nmin, m] (numpy PCG64, fixed seed)."""
    rng = np.random.default_rng(seed)
    m = np.floor(np.exp(rng.uniform(math.log(mmin), math.log(mmax + 1), size=count))).astype(np.int64)
    m = np.clip(m, mmin, mmax)
    n = np.array([rng.integers(min(nmin, mi), mi + 1) for mi in m], dtype=np.int64)
    return [(int(a), int(b)) for a, b in zip(m, n)]


def _arr(vals, ctype):
    return (ctype * max(len(vals), 1))(*vals)


class BatchedQRPlan:
    """Plan for ``qr_compact!`` on every matrix of ``As`` (2-D device tensors of
    one dtype, column-major).  Outputs ``Qs`` / ``Rs`` are allocated when not
    given; ``Rs=False`` skips R.  ``tune=True`` lets the library time its apply
    kernel candidates on a scratch copy of the shapes first (cached per shape
    set for the process)."""

    def __init__(self, As, Qs=None, Rs=None, alg: MI355XHouseholder | None = None, tune: bool = False):
        if alg is None:
            alg = MI355XHouseholder()
        if not As:
            raise ArgumentError("empty batch")
        dtype = As[0].dtype
        if dtype not in _DT:
            raise ArgumentError(f"unsupported element type {dtype}")
        dev = As[0].device
        self.As = []
        for A in As:
            if A.dtype != dtype or A.device != dev or A.dim() != 2:
                raise ArgumentError("all blocks must be matrices of one dtype on one device")
            _ld(A)                     # raises unless column-major
            self.As.append(A)
        if Qs is None:
            Qs = [initialize_output_qr(A, False)[0] for A in self.As]
        if Rs is None:
            Rs = [initialize_output_qr(A, False)[1] for A in self.As]
        self.Qs, self.Rs = Qs, (Rs if Rs is not False else None)
        cnt = len(self.As)
        m = [A.shape[0] for A in self.As]
        n = [A.shape[1] for A in self.As]
        ld = _ld
        self._m, self._n = _arr(m, C.c_int), _arr(n, C.c_int)
        self._A = _arr([A.data_ptr() for A in self.As], C.c_void_p)
        self._lda = _arr([ld(A) for A in self.As], C.c_int)
        self._Q = _arr([Q.data_ptr() for Q in self.Qs], C.c_void_p)
        self._ldq = _arr([ld(Q) for Q in self.Qs], C.c_int)
        if self.Rs is not None:
            self._R = _arr([R.data_ptr() if R.numel() else None for R in self.Rs], C.c_void_p)
            self._ldr = _arr([ld(R) for R in self.Rs], C.c_int)
        lib = _lib.load()
        self._lib = lib
        dt = _DT[dtype]
        nbytes = C.c_size_t(0)
        _status(lib.mak_qr_batched_workspace_size(dt, cnt, self._m, self._n, int(tune), C.byref(nbytes)),
                "mak_qr_batched_workspace_size")
        self._ws = _workspace(dev, nbytes.value)
        torch.cuda.synchronize(dev)
        h = C.c_void_p(None)
        code = lib.mak_qr_batched_plan_create(C.byref(h), dt, cnt, self._m, self._n, self._A, self._lda, self._Q,
                                              self._ldq, self._R if self.Rs is not None else None,
                                              self._ldr if self.Rs is not None else None, int(bool(alg.positive)),
                                              int(tune), self._ws.data_ptr(), nbytes.value)
        _status(code, "mak_qr_batched_plan_create")
        self._h = h
        self.device = dev

    def execute(self):
        """Factor every block (asynchronous on torch's current stream)."""
        if self._h is None:
            raise ArgumentError("plan is closed")
        _status(self._lib.mak_qr_batched_execute(self._h, _stream(self.device)), "mak_qr_batched_execute")
        return self.Qs, self.Rs

    def info(self) -> dict:
        vals = [C.c_int(0) for _ in range(4)]
        _status(self._lib.mak_qr_batched_plan_info(self._h, *[C.byref(v) for v in vals]), "mak_qr_batched_plan_info")
        return dict(zip(("chunk_cols", "grid_per_cu", "launches", "fallback_problems"), (v.value for v in vals)))

    def close(self):
        if self._h is not None:
            torch.cuda.synchronize(self.device)
            self._lib.mak_qr_batched_plan_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


def qr_compact_batched(As, alg: MI355XHouseholder | None = None, tune: bool = False):
    """``[qr_compact(A) for A in As]`` in one graph launch (inputs are copied
    first, as ``qr_compact`` does)."""
    Ac = [colmajor(A.clone() if A.stride(0) == 1 else A) for A in As]
    plan = BatchedQRPlan(Ac, alg=alg, tune=tune)
    Qs, Rs = plan.execute()
    plan.close()
    return Qs, Rs
