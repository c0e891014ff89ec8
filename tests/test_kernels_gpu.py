"""Kernel-level parity: one TSQR level of geqr2 nodes against the oracle's
LAPACK geqrt (scipy) on every node's rows.

Each node's block is factored independently, so the factored block (V below,
R on and above the diagonal) and the 32 x 32 T of node i must match
geqrt(block_i, nb=32) to 1e-12 relative (Householder QR is backward stable
and the reflectors are unique under the xLARFG convention).  Both node
kernel layouts are checked, and against each other.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import mak_mi355x as mak
from mak_mi355x import _lib
from oracle import mak_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def hooks():
    lib = _lib.load()
    lib.maktest_set_node_variant.restype = C.c_int
    lib.maktest_set_node_variant.argtypes = [C.c_int]
    lib.maktest_geqr2_level.restype = C.c_int
    lib.maktest_geqr2_level.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                        C.c_void_p]
    return lib


def run_level(lib, A, nodes, nbc, variant, dev):
    X = mak.colmajor(torch.from_numpy(np.asfortranarray(A)).to(dev))
    T = torch.zeros((nodes, 32, 32), dtype=X.dtype, device=dev)
    prev = lib.maktest_set_node_variant(variant)
    try:
        code = lib.maktest_geqr2_level(mak.api._stream(dev), 1 if X.is_complex() else 0, A.shape[0], nodes, nbc,
                                       X.data_ptr(), X.stride(1), T.data_ptr())
        torch.cuda.synchronize()
    finally:
        lib.maktest_set_node_variant(prev)
    assert code == 0
    return X.cpu().numpy(), T.cpu().numpy().transpose(0, 2, 1)   # T[i] column-major


CASES = [(1, 1, 1), (5, 1, 3), (31, 1, 32), (32, 1, 32), (33, 1, 32), (100, 1, 7), (256, 1, 32), (512, 1, 32),
         (4000, 8, 32), (3001, 7, 20)]


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("variant", [1, 2, 3])
@pytest.mark.parametrize("H,nodes,nbc", CASES, ids=[f"{h}r-{nd}n-{c}c" for h, nd, c in CASES])
def test_node_level_matches_geqrt(H, nodes, nbc, variant, dt, dev):
    cap = 256 if dt == np.complex128 else 512
    if -(-H // nodes) > cap:
        pytest.skip("node taller than the kernel's capacity")
    lib = hooks()
    A = O.gaussian(H, nbc, dt, seed=H + 7 * nodes + nbc)
    X, T = run_level(lib, A, nodes, nbc, variant, dev)
    for i in range(nodes):
        r0, r1 = (i * H) // nodes, ((i + 1) * H) // nodes
        a_ref, t_ref = O.geqrt(A[r0:r1], min(32, r1 - r0, nbc))
        k = min(r1 - r0, nbc)
        assert rel(X[r0:r1], a_ref) < 1e-12, (i, rel(X[r0:r1], a_ref))
        assert rel(np.triu(T[i][:k, :k]), np.triu(t_ref[:k, :k])) < 1e-12, i


def test_node_variants_agree_at_full_capacity(dev):
    lib = hooks()
    A = O.gaussian(16 * 512, 32, np.float64, seed=3)
    X1, T1 = run_level(lib, A, 16, 32, 1, dev)
    X2, T2 = run_level(lib, A, 16, 32, 2, dev)
    assert rel(X2, X1) < 1e-13 and rel(T2, T1) < 1e-13


# ------------------------------------------------------------- Cholesky ----

def potrf_hook():
    lib = _lib.load()
    lib.maktest_potrf.restype = C.c_int
    lib.maktest_potrf.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_void_p]
    return lib


def run_potrf(Zh, dev):
    lib = potrf_hook()
    n = Zh.shape[0]
    Z = mak.colmajor(torch.from_numpy(np.asfortranarray(Zh)).to(dev))
    Y = mak.colmajor(torch.full((n, n), 7.0, dtype=Z.dtype, device=dev))
    tmp = torch.empty(n * n // 4 + 64 * n + 1, dtype=Z.dtype, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    code = lib.maktest_potrf(mak.api._stream(dev), 1 if Z.is_complex() else 0, n, Z.data_ptr(), Z.stride(1),
                             Y.data_ptr(), Y.stride(1), tmp.data_ptr(), info.data_ptr())
    torch.cuda.synchronize()
    assert code == 0
    return Z.cpu().numpy(), Y.cpu().numpy(), int(info.item())


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 130, 200, 1000])
def test_potrf_with_inverse(n, dt, dev):
    """W upper with W^H W = Z (numpy's Cholesky factor, conjugate-transposed)
    and Y = W^{-1}; 1e-12 relative for these well-conditioned Z."""
    G = O.gaussian(2 * n + 3, n, dt, seed=n)
    Zh = G.conj().T @ G
    Wref = np.linalg.cholesky(Zh).conj().T
    W, Y, info = run_potrf(Zh, dev)
    assert info == 0
    assert rel(W, Wref) < 1e-12, rel(W, Wref)
    assert np.allclose(np.tril(W, -1), 0) and np.allclose(np.tril(Y, -1), 0)
    assert np.linalg.norm(Y @ W - np.eye(n)) / np.sqrt(n) < 1e-12


def test_potrf_reports_first_nonpositive_minor(dev):
    n = 150
    G = O.gaussian(n, n, np.float64, seed=2)
    Zh = G.T @ G
    Zh[100:, :] = 0.0
    Zh[:, 100:] = 0.0          # leading 100 x 100 is PD, minor 101 is singular
    _, _, info = run_potrf(Zh, dev)
    assert info == 101
