"""CPU-side checks of the C ABI and the host layer (no GPU calls)."""
import ctypes as C

import pytest
import torch

import mak_mi355x as mak
from mak_mi355x import _lib


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _lib.declared_symbols()
    assert declared, "no declarations parsed from include/mak_mi355x.h"
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"
    # the ctypes signature table covers exactly the declared API
    assert set(declared) == set(_lib.SIGNATURES)


def test_workspace_query_is_host_only_and_validates():
    lib = _lib.load()
    nb = C.c_size_t(0)
    assert lib.mak_qr_workspace_size(0, 0, 8192, 1024, C.byref(nb)) == 0
    assert nb.value > 0
    full = C.c_size_t(0)
    assert lib.mak_qr_workspace_size(0, 1, 8192, 1024, C.byref(full)) == 0
    assert full.value >= nb.value
    z = C.c_size_t(0)
    assert lib.mak_qr_workspace_size(1, 0, 8192, 1024, C.byref(z)) == 0
    assert z.value > nb.value
    assert lib.mak_qr_workspace_size(7, 0, 10, 10, C.byref(nb)) == -1
    assert lib.mak_qr_workspace_size(0, 0, -1, 10, C.byref(nb)) == -3
    assert lib.mak_qr_workspace_size(0, 0, 10, -1, C.byref(nb)) == -4
    assert lib.mak_qr_workspace_size(0, 0, 10, 10, None) == -5


def test_entry_points_reject_bad_arguments_before_touching_the_device():
    lib = _lib.load()
    null = C.c_void_p(0)
    # m < 0, n < 0, lda < m: rejected by argument checks (no launch happens)
    assert lib.mak_qr_compact_d(null, -1, 3, null, 1, null, 1, null, 1, 1, null, 0) == -2
    assert lib.mak_qr_compact_d(null, 3, -1, null, 3, null, 3, null, 1, 1, null, 0) == -3
    assert lib.mak_qr_compact_d(null, 5, 3, C.c_void_p(8), 4, C.c_void_p(8), 5, null, 1, 1, null, 0) == -5
    assert lib.mak_qr_compact_d(null, 5, 3, C.c_void_p(8), 5, C.c_void_p(8), 5, null, 1, 1, null, 0) == -11


def test_host_check_input_errors_mirror_reference():
    A = torch.zeros(6, 4, dtype=torch.float64)
    Q = torch.zeros(6, 3, dtype=torch.float64)
    R = torch.zeros(4, 4, dtype=torch.float64)
    with pytest.raises(mak.DimensionMismatch):
        mak.check_input_qr_compact(A, (Q, R))
    Q = torch.zeros(6, 4, dtype=torch.float64)
    mak.check_input_qr_compact(A, (Q, R))
    with pytest.raises(mak.ArgumentError):
        mak.check_input_qr_compact(A, (Q.to(torch.complex128), R))
    with pytest.raises(mak.DimensionMismatch):
        mak.check_input_qr_full(A, (Q, R))
    # empty R is accepted (R not computed)
    mak.check_input_qr_compact(A, (Q, torch.zeros(0, 0, dtype=torch.float64)))
    with pytest.raises(mak.ArgumentError):
        mak.MI355XHouseholder(pivoted=True)
    with pytest.raises(mak.ArgumentError):
        mak.qr_compact_(A, (Q, R))   # host tensors are refused before any launch


def test_initialize_output_layout():
    A = torch.zeros(0, 0, dtype=torch.float64)
    A = mak.similar_colmajor(torch.zeros(1, dtype=torch.float64), (7, 5))
    assert A.stride(0) == 1 and A.stride(1) == 7
    Q, R = mak.initialize_output_qr(A, False)
    assert Q.shape == (7, 5) and R.shape == (5, 5) and Q.stride(0) == 1
    Q, R = mak.initialize_output_qr(A, True)
    assert Q.shape == (7, 7) and R.shape == (7, 5)


def test_workspace_queries_accept_empty_shapes():
    """Host-side size queries (no GPU work) for empty factorisations."""
    lib = _lib.load()
    b = C.c_size_t(0)
    for dt in (0, 1):
        for full in (0, 1):
            for m, n in ((4, 0), (0, 5), (0, 0), (1, 1)):
                assert lib.mak_qr_workspace_size(dt, full, m, n, C.byref(b)) == 0
        ms = (C.c_int * 5)(0, 4, 1, 3000, 700)
        ns = (C.c_int * 5)(5, 0, 1, 40, 700)
        assert lib.mak_qr_batched_workspace_size(dt, 5, ms, ns, 1, C.byref(b)) == 0 and b.value > 0
