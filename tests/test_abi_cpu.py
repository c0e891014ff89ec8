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
    assert full.value > 0
    z = C.c_size_t(0)
    assert lib.mak_qr_workspace_size(1, 0, 8192, 1024, C.byref(z)) == 0
    assert z.value > nb.value
    assert lib.mak_qr_workspace_size(7, 0, 10, 10, C.byref(nb)) == -1
    assert lib.mak_qr_workspace_size(0, 0, -1, 10, C.byref(nb)) == -3
    assert lib.mak_qr_workspace_size(0, 0, 10, -1, C.byref(nb)) == -4
    assert lib.mak_qr_workspace_size(0, 0, 10, 10, None) == -5


def test_noalloc_documented_case_fits_the_static_scratch():
    """The allocation-free qr_compact! entry runs on a 64 MiB static device
    array: the case the header quotes (Float64 1024 x 512, about 47 MiB) and
    the GPU test's second, smaller call fit it, the 8192 x 1024 headline does
    not (that call returns 7)."""
    lib = _lib.load()
    cap = 64 << 20
    for (m, n), lo, hi in (((1024, 512), 40 << 20, cap), ((300, 70), 0, cap), ((8192, 1024), cap, None)):
        nb = C.c_size_t(0)
        assert lib.mak_qr_workspace_size(0, 0, m, n, C.byref(nb)) == 0
        assert nb.value >= lo and (hi is None or nb.value <= hi), (m, n, nb.value)


def test_entry_points_reject_bad_arguments_before_touching_the_device():
    lib = _lib.load()
    null = C.c_void_p(0)
    # m < 0, n < 0, lda < m: rejected by argument checks (no launch happens)
    assert lib.mak_qr_compact_d(null, -1, 3, null, 1, null, 1, null, 1, 1, null, 0) == -2
    assert lib.mak_qr_compact_d(null, 3, -1, null, 3, null, 3, null, 1, 1, null, 0) == -3
    assert lib.mak_qr_compact_d(null, 5, 3, C.c_void_p(8), 4, C.c_void_p(8), 5, null, 1, 1, null, 0) == -5
    assert lib.mak_qr_compact_d(null, 5, 3, C.c_void_p(8), 5, C.c_void_p(4096), 5, null, 1, 1, null, 0) == -11


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


@pytest.mark.parametrize("ny,ysplit", [(0, 1 << 30), (1, 1 << 30), (128, 1 << 30), (64, 13), (3, 0)])
@pytest.mark.parametrize("nt", [1, 2, 3, 7, 32, 64])
def test_chol_tile_task_order(nt, ny, ysplit):
    """The persistent Cholesky kernel's task order (host code in the library):
    every task appears once and every dependency precedes it (the panel's
    step K, with its look-ahead to the next pair, runs before the tasks of
    step K and waits only for tasks of earlier steps), so a workgroup waiting
    on a task it took can always be satisfied.  Tiles T = nt rounded up to
    even; per pair step K with rem = T - 2K - 2 tiles after the pair and
    b = max(rem - 2, 0): b P2 (the panel forms the next pair's two), the next
    pair's 2b U2 right of its diagonal block (the panel updates that block),
    2K PQ2, (2K + 2) rem UQ2, the bulk b(b+1)/2 U2 and ny right-product tasks
    (2 ny from pair ysplit on, one per half of the contraction), each after
    X's column block K is final."""
    lib = _lib.load()
    lib.maktest_chol_order_check.restype = C.c_int
    lib.maktest_chol_order_check.argtypes = [C.c_int, C.c_int, C.c_int]
    T = nt + (nt & 1)
    expect = 0
    for k in range(T // 2):
        rem = T - 2 * k - 2
        b = max(rem - 2, 0)
        expect += b + 2 * b + 2 * k + (2 * k + 2) * rem + b * (b + 1) // 2
        expect += ny * (2 if k >= ysplit else 1)
    assert lib.maktest_chol_order_check(nt, ny, ysplit) == expect
    assert lib.maktest_chol_order_check(0, ny, ysplit) == -1
    assert lib.maktest_chol_order_check(65, ny, ysplit) == -1
    assert lib.maktest_chol_order_check(nt, -1, ysplit) == -1


def test_qr_q_aliasing_rules():
    """Q === A is accepted only for a tall input with no R and positive = false
    (src/implementations/qr.jl:150-153); other Q / A overlaps are invalid
    argument 6.  The checks run before any device work, so fake device
    pointers are enough."""
    lib = _lib.load()
    A = C.c_void_p(0x100000)
    R = C.c_void_p(0x90000000)
    m, n = 100, 50
    # in place with R requested, or with the positive gauge: -6
    assert lib.mak_qr_compact_d(None, m, n, A, m, A, m, R, n, 0, None, 0) == -6
    assert lib.mak_qr_compact_d(None, m, n, A, m, A, m, None, 1, 1, None, 0) == -6
    # in place, no R, positive = false: accepted up to the workspace check (-11)
    assert lib.mak_qr_compact_d(None, m, n, A, m, A, m, None, 1, 0, None, 0) == -11
    # Q overlapping A without being A: -6
    Qp = C.c_void_p(0x100000 + 8 * 30)
    assert lib.mak_qr_compact_d(None, m, n, A, m, Qp, m, None, 1, 0, None, 0) == -6
    # disjoint: only the missing workspace is reported
    Qd = C.c_void_p(0x100000 + 8 * m * n)
    assert lib.mak_qr_compact_d(None, m, n, A, m, Qd, m, R, n, 1, None, 0) == -11
    # wide A cannot be its own Q
    assert lib.mak_qr_compact_d(None, 40, 50, A, 40, A, 40, None, 1, 0, None, 0) == -6


def test_householder_options():
    """blocksize > 1 and pivoted throw ArgumentError, as for the reference's
    GPU drivers (src/implementations/qr.jl:139-143)."""
    mak.MI355XHouseholder(blocksize=1)
    with pytest.raises(mak.ArgumentError, match="blocked"):
        mak.MI355XHouseholder(blocksize=32)
    with pytest.raises(mak.ArgumentError, match="pivoted"):
        mak.MI355XHouseholder(pivoted=True)


def test_lq_and_null_entry_points_validate_before_the_device():
    lib = _lib.load()
    null = C.c_void_p(0)
    b = C.c_size_t(0)
    for dt in (0, 1):
        for full in (0, 1):
            for m, n in ((4, 0), (0, 5), (3, 9), (9, 3)):
                assert lib.mak_lq_workspace_size(dt, full, m, n, C.byref(b)) == 0 and b.value > 0
        for lq in (0, 1):
            assert lib.mak_null_workspace_size(dt, lq, 9, 3, C.byref(b)) == 0 and b.value > 0
    assert lib.mak_lq_workspace_size(5, 0, 3, 3, C.byref(b)) == -1
    assert lib.mak_lq_workspace_size(0, 0, -1, 3, C.byref(b)) == -3
    assert lib.mak_null_workspace_size(0, 0, 3, -1, C.byref(b)) == -4
    assert lib.mak_null_workspace_size(0, 0, 3, 3, None) == -5
    # lq_compact!: m < 0, n < 0, lda < m, ldq < k, Q overlapping A, no workspace
    assert lib.mak_lq_compact_d(null, -1, 3, null, 1, null, 1, null, 1, 1, null, 0) == -2
    assert lib.mak_lq_compact_d(null, 3, -1, null, 3, null, 3, null, 1, 1, null, 0) == -3
    assert lib.mak_lq_compact_d(null, 3, 5, C.c_void_p(8), 2, null, 3, C.c_void_p(4096), 3, 1, null, 0) == -5
    assert lib.mak_lq_compact_d(null, 3, 5, C.c_void_p(8), 3, null, 3, C.c_void_p(4096), 2, 1, null, 0) == -9
    assert lib.mak_lq_compact_d(null, 3, 5, C.c_void_p(8), 3, null, 3, C.c_void_p(16), 3, 1, null, 0) == -8
    assert lib.mak_lq_full_z(null, 3, 5, C.c_void_p(16), 3, null, 3, C.c_void_p(1 << 20), 5, 1, null, 0) == -11
    # qr_null! / lq_null!: ldn too small, N overlapping A, no workspace
    assert lib.mak_qr_null_d(null, 9, 3, C.c_void_p(8), 9, C.c_void_p(1 << 20), 8, null, 0) == -7
    assert lib.mak_qr_null_d(null, 9, 3, C.c_void_p(8), 9, C.c_void_p(16), 9, null, 0) == -6
    assert lib.mak_qr_null_z(null, 9, 3, C.c_void_p(16), 9, C.c_void_p(1 << 20), 9, null, 0) == -9
    assert lib.mak_lq_null_d(null, 3, 9, C.c_void_p(8), 3, C.c_void_p(1 << 20), 5, null, 0) == -7


def test_lq_null_orth_host_checks_mirror_reference():
    A = torch.zeros(3, 7, dtype=torch.float64)
    L, Q = mak.initialize_output_lq(A, False)
    assert L.shape == (3, 3) and Q.shape == (3, 7) and Q.stride(0) == 1
    L, Q = mak.initialize_output_lq(A, True)
    assert L.shape == (3, 7) and Q.shape == (7, 7)
    with pytest.raises(mak.DimensionMismatch):
        mak.check_input_lq_compact(A, (L, Q))
    mak.check_input_lq_full(A, (torch.zeros(0, 0, dtype=torch.float64), Q))
    assert mak.initialize_output_qr_null(A).shape == (3, 0)
    assert mak.initialize_output_lq_null(A).shape == (4, 7)
    with pytest.raises(mak.DimensionMismatch):
        mak.check_input_lq_null(A, torch.zeros(3, 7, dtype=torch.float64))
    with pytest.raises(mak.ArgumentError):
        mak.check_input_qr_null(A.t(), torch.zeros(7, 4, dtype=torch.complex128))
    assert mak.left_orth_alg(None)[0] == "qr"
    assert mak.left_orth_alg(mak.PolarNewton())[0] == "polar"
    assert mak.left_orth_alg(mak.MI355XPolar(precision=32))[1].precision == 32
    with pytest.raises(mak.ArgumentError):
        mak.left_orth_alg(3.0)
    with pytest.raises(mak.ArgumentError):
        mak.PolarNewton(tol=-1.0)
