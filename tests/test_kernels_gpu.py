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
    lib.maktest_geqr2_level.restype = C.c_int
    lib.maktest_geqr2_level.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                        C.c_void_p, C.c_int]
    return lib


def run_level(lib, A, nodes, nbc, variant, dev):
    X = mak.colmajor(torch.from_numpy(np.asfortranarray(A)).to(dev))
    T = torch.zeros((nodes, 32, 32), dtype=X.dtype, device=dev)
    code = lib.maktest_geqr2_level(mak.api._stream(dev), 1 if X.is_complex() else 0, A.shape[0], nodes, nbc,
                                   X.data_ptr(), X.stride(1), T.data_ptr(), variant)
    torch.cuda.synchronize()
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
@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 63, 64, 65, 130, 200, 1000, 1024, 2048, 2100, 3000])
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


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("n,bad", [(150, 100), (150, 64), (70, 33), (2300, 2090)])
def test_potrf_reports_first_nonpositive_minor(n, bad, dt, dev):
    """info = j + 1 for the first minor that is singular (tile kernel alone and
    inside the blocked loop; in the first or second tile of a pair step, where
    the panel stops before or after its look-ahead)."""
    G = O.gaussian(n, n, dt, seed=2)
    Zh = G.conj().T @ G
    Zh[bad:, :] = 0.0
    Zh[:, bad:] = 0.0          # leading bad x bad is PD, minor bad + 1 is singular
    _, _, info = run_potrf(Zh, dev)
    assert info == bad + 1


def chol_tile_hook():
    lib = _lib.load()
    lib.maktest_chol_tile_flag_ints.restype = C.c_size_t
    lib.maktest_chol_tile_flag_ints.argtypes = [C.c_int]
    lib.maktest_chol_tile.restype = C.c_int
    lib.maktest_chol_tile.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                      C.c_void_p, C.c_void_p, C.c_int]
    return lib


@pytest.mark.parametrize("grid", [2, 3, 64, 0])
@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_chol_tile_any_grid(grid, dt, dev):
    """The persistent tile kernel gives the same factors with any number of
    workgroups (one workgroup runs every task in order; waits never block it),
    and its strict lower parts are zero."""
    n = 300
    lib = chol_tile_hook()
    G = O.gaussian(2 * n, n, dt, seed=7)
    Zh = G.conj().T @ G
    Wref = np.linalg.cholesky(Zh).conj().T
    Z = mak.colmajor(torch.from_numpy(np.asfortranarray(Zh)).to(dev))
    Y = mak.colmajor(torch.full((n, n), 5.0, dtype=Z.dtype, device=dev))
    flags = torch.full((lib.maktest_chol_tile_flag_ints(n),), -9, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    code = lib.maktest_chol_tile(mak.api._stream(dev), 1 if Z.is_complex() else 0, n, Z.data_ptr(), Z.stride(1),
                                 Y.data_ptr(), Y.stride(1), flags.data_ptr(), info.data_ptr(), grid)
    torch.cuda.synchronize()
    assert code == 0 and int(info.item()) == 0
    W, X = Z.cpu().numpy(), Y.cpu().numpy()
    assert rel(W, Wref) < 1e-12
    assert np.all(np.tril(W, -1) == 0) and np.all(np.tril(X, -1) == 0)
    assert np.linalg.norm(X @ W - np.eye(n)) / np.sqrt(n) < 1e-12
    # abort flag clear, every task taken
    fl = flags.cpu().numpy()
    nt = -(-n // 32)
    assert fl[2 * nt * nt + 1] == 0


@pytest.mark.parametrize("grid", [2, 3, 0])
@pytest.mark.parametrize("n,m", [(300, 300), (300, 1000), (1000, 4099), (1024, 8192), (37, 70)])
def test_chol_tile_right_product(n, m, grid, dev):
    """Y = B X computed inside the Cholesky launch (task by task, as X's
    column blocks become final) equals B @ X of the launch's own X, for any
    workgroup count and ragged m, n; the factors are those of the plain
    launch."""
    lib = chol_tile_hook()
    lib.maktest_chol_tile_product.restype = C.c_int
    lib.maktest_chol_tile_product.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                              C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                              C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    Bh = O.gaussian(m, n, np.float64, seed=11 + n + m)
    Zh = Bh.T @ Bh
    Z = mak.colmajor(torch.from_numpy(np.asfortranarray(Zh)).to(dev))
    Z0 = Z.clone()
    X = mak.colmajor(torch.full((n, n), 5.0, dtype=torch.float64, device=dev))
    X0 = X.clone()
    B = mak.colmajor(torch.from_numpy(np.asfortranarray(Bh)).to(dev))
    Y = mak.colmajor(torch.full((m, n), 7.0, dtype=torch.float64, device=dev))
    ldxc = -(-n // 16) * 16
    Xc = torch.full((ldxc * n,), 3.0, dtype=torch.float64, device=dev)
    nfl = lib.maktest_chol_tile_flag_ints(n)
    flags = torch.full((nfl,), -9, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    st = mak.api._stream(dev)
    code = lib.maktest_chol_tile_product(st, n, Z.data_ptr(), Z.stride(1), X.data_ptr(), X.stride(1),
                                         flags.data_ptr(), info.data_ptr(), grid, m, B.data_ptr(), B.stride(1),
                                         Y.data_ptr(), Y.stride(1), Xc.data_ptr(), ldxc, None)
    torch.cuda.synchronize()
    assert code == 0 and int(info.item()) == 0
    # the copy holds X (upper part and the diagonal pairs' lower tiles as written)
    Xcp = Xc.cpu().numpy().reshape(n, ldxc).T[:n, :]
    assert np.array_equal(np.triu(Xcp), X.cpu().numpy())
    # misaligned scratch or leading dimension is rejected
    assert lib.maktest_chol_tile_product(st, n, Z.data_ptr(), Z.stride(1), X.data_ptr(), X.stride(1),
                                         flags.data_ptr(), info.data_ptr(), grid, m, B.data_ptr(), B.stride(1),
                                         Y.data_ptr(), Y.stride(1), Xc.data_ptr(), ldxc + 1, None) == -1
    Xh, Yh = X.cpu().numpy(), Y.cpu().numpy()
    assert np.linalg.norm(Yh - Bh @ Xh) / np.linalg.norm(Bh @ Xh) < 1e-13
    # Y = B R^{-1} has orthonormal columns (Z = B^T B, well conditioned)
    assert np.linalg.norm(Yh.T @ Yh - np.eye(n)) < 1e-11
    # the factors are the plain launch's
    flags.fill_(-9)
    code = lib.maktest_chol_tile(st, 0, n, Z0.data_ptr(), Z0.stride(1), X0.data_ptr(), X0.stride(1),
                                 flags.data_ptr(), info.data_ptr(), grid)
    torch.cuda.synchronize()
    assert code == 0
    assert np.array_equal(X0.cpu().numpy(), Xh) and np.array_equal(Z0.cpu().numpy(), Z.cpu().numpy())


# ------------------------------------------- FP32 / BF16 MFMA products ----

def gemm_prec_hook():
    lib = _lib.load()
    lib.maktest_gemm_prec_ws.restype = C.c_size_t
    lib.maktest_gemm_prec_ws.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.maktest_gemm_prec.restype = C.c_int
    lib.maktest_gemm_prec.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                      C.c_double, C.c_double, C.c_int, C.c_void_p]
    return lib


def round_to(x, prec):
    """x rounded as the kernels round it: to FP32, then (BF16) to the nearest
    BF16 with ties to even."""
    f = np.asarray(x, dtype=np.float64).astype(np.float32)
    if prec == 16:
        b = f.view(np.uint32).astype(np.uint64)
        b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
        f = b.astype(np.uint32).view(np.float32)
    return f.astype(np.float64)


def run_gemm_prec(prec, opA, opB, A, B, Cin, alpha, beta, tri, dev):
    lib = gemm_prec_hook()
    M, N = Cin.shape
    K = A.shape[1] if opA == 0 else A.shape[0]
    At = mak.colmajor(torch.from_numpy(np.asfortranarray(A)).to(dev))
    Bt = mak.colmajor(torch.from_numpy(np.asfortranarray(B)).to(dev))
    Ct = mak.colmajor(torch.from_numpy(np.asfortranarray(Cin)).to(dev))
    ws = torch.full((lib.maktest_gemm_prec_ws(M, N, K) // 4 + 1,), float("nan"), dtype=torch.float32, device=dev)
    code = lib.maktest_gemm_prec(mak.api._stream(dev), prec, opA, opB, M, N, K, At.data_ptr(), max(At.stride(1), 1),
                                 Bt.data_ptr(), max(Bt.stride(1), 1), Ct.data_ptr(), max(Ct.stride(1), 1),
                                 alpha, beta, tri, ws.data_ptr())
    torch.cuda.synchronize()
    assert code == 0
    return Ct.cpu().numpy()


PREC_SHAPES = [(1, 1, 1), (37, 129, 65), (130, 70, 200), (256, 384, 96), (300, 17, 3), (64, 64, 0)]


@pytest.mark.parametrize("prec", [32, 16])
@pytest.mark.parametrize("opA,opB", [(0, 0), (0, 1), (1, 0), (1, 1)], ids=["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("M,N,K", PREC_SHAPES, ids=[f"{m}x{n}x{k}" for m, n, k in PREC_SHAPES])
def test_gemm_prec_exact_on_small_integers(M, N, K, opA, opB, prec, dev):
    """Small-integer operands are exact in BF16 and FP32 and every partial sum
    stays below 2^24, so the FP32-accumulated product must equal the exact one
    bit for bit: a wrong k pairing, row/column index or transposition fails."""
    rng = np.random.default_rng(M * 1000 + N * 10 + K + opA * 3 + opB)
    A = rng.integers(-4, 5, size=(M, K) if opA == 0 else (K, M)).astype(np.float64)
    B = rng.integers(-4, 5, size=(K, N) if opB == 0 else (N, K)).astype(np.float64)
    Cin = rng.integers(-8, 9, size=(M, N)).astype(np.float64)
    opa = A if opA == 0 else A.T
    opb = B if opB == 0 else B.T
    want = -2.0 * (opa @ opb) + 0.5 * Cin
    got = run_gemm_prec(prec, opA, opB, A, B, Cin, -2.0, 0.5, 0, dev)
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize("prec,tol", [(32, 1e-5), (16, 1e-5)])
def test_gemm_prec_gaussian_vs_rounded_inputs(prec, tol, dev):
    """Gaussian operands: the result matches the FP64 product of the rounded
    inputs to FP32 accumulation accuracy (relative Frobenius error <= tol);
    beta = 0 overwrites NaN in C."""
    M, N, K = 520, 390, 700
    A = O.gaussian(M, K, np.float64, seed=11)
    B = O.gaussian(N, K, np.float64, seed=12)
    Cin = np.full((M, N), np.nan)
    want = 1.25 * (round_to(A, prec) @ round_to(B, prec).T)
    got = run_gemm_prec(prec, 0, 1, A, B, Cin, 1.25, 0.0, 0, dev)
    assert rel(got, want) < tol, rel(got, want)


@pytest.mark.parametrize("prec", [32, 16])
@pytest.mark.parametrize("tri", [2 | 8, 1 | 4], ids=["AupperBlower", "AlowerBupper"])
def test_gemm_prec_triangular_hints(tri, prec, dev):
    """With exact zeros on the hinted side of op(A) and op(B) the clipped K
    ranges give the full product (small integers: exact)."""
    n = 300
    rng = np.random.default_rng(tri + prec)
    A = rng.integers(-3, 4, size=(n, n)).astype(np.float64)
    B = rng.integers(-3, 4, size=(n, n)).astype(np.float64)
    if tri & 2:
        A, B = np.triu(A), np.tril(B)
    else:
        A, B = np.tril(A), np.triu(B)
    Cin = np.zeros((n, n))
    got = run_gemm_prec(prec, 0, 0, A, B, Cin, 1.0, 0.0, tri, dev)
    assert np.array_equal(got, A @ B)
