"""Parity of the HIP polar path (QDWH) with the CPU oracle (the reference's
PolarViaSVD path: gesdd -> W = U V^H, P = (sqrt(S) V^H)^H (sqrt(S) V^H)).

The polar decomposition of a full-rank matrix is unique, so W and P are
compared elementwise.  Tolerances (relative Frobenius):
  * FP64, Gaussian inputs (condition <= ~1e4):  W, P within 1e-11 of the oracle;
  * FP64, condition 1e6: W within 1e-9 (the perturbation bound of the polar
    factor is ~ cond * u, so both the oracle and this path sit ~1e-11 from the
    exact W), backward error ||A - W P|| / ||A|| <= 1e-14;
  * FP32 / BF16 bulk (precision 32 / 16) + FP64 Newton-Schulz: orthogonality
    ||W^H W - I||_F <= 1e-12, backward error <= 1e-5 (FP32) / 1e-2 (BF16 in the
    Cholesky iterations); the W error is bounded below by cond * u_bulk and is
    reported, not asserted.
"""
from pathlib import Path

import numpy as np
import pytest
import torch

import mak_mi355x as mak
from oracle import mak_oracle as O

pytestmark = pytest.mark.gpu
GOLD = np.load(Path(__file__).parent / "golden" / "golden_small.npz")


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def to_dev(A, dev):
    return mak.colmajor(torch.from_numpy(np.asfortranarray(A)).to(dev))


def run_left(A, dev, alg=None, with_p=True):
    At = to_dev(A, dev)
    W, P = mak.initialize_output_polar(At, False)
    if not with_p:
        P = torch.empty((0, 0), dtype=At.dtype, device=dev)
    W2, P2 = mak.left_polar_(At, (W, P), alg)
    assert W2 is W and P2 is P
    torch.cuda.synchronize()
    return W.cpu().numpy(), (P.cpu().numpy() if with_p else None)


SIZES = [(54, 37), (54, 54), (20, 1), (300, 200), (512, 512), (1000, 130)]


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", SIZES, ids=[f"{m}x{n}" for m, n in SIZES])
def test_left_polar_matches_oracle(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=500 + m + n)
    W, P = run_left(A, dev)
    Wo, Po = O.left_polar_svd(A)
    assert rel(W, Wo) < 1e-11, rel(W, Wo)
    assert rel(P, Po) < 1e-11, rel(P, Po)
    assert O.isisometric(W)
    assert np.allclose(P, P.conj().T)


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(37, 54), (54, 54), (130, 1000)])
def test_right_polar_matches_oracle(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=600 + m + n)
    At = to_dev(A, dev)
    P, Wh = mak.initialize_output_polar(At, True)
    P2, Wh2 = mak.right_polar_(At, (P, Wh))
    assert P2 is P and Wh2 is Wh
    torch.cuda.synchronize()
    Po, Who = O.right_polar_svd(A)
    assert rel(Wh.cpu().numpy(), Who) < 1e-11
    assert rel(P.cpu().numpy(), Po) < 1e-11


@pytest.mark.parametrize("key", [k[:-2] for k in GOLD.files if k.startswith("lpolar_") and k.endswith("_A")])
def test_polar_golden_fixtures(key, dev):
    W, P = run_left(GOLD[key + "_A"], dev)
    assert rel(W, GOLD[key + "_W"]) < 1e-11 and rel(P, GOLD[key + "_P"]) < 1e-11


def test_left_polar_without_P_and_input_untouched(dev):
    A = O.gaussian(90, 60, np.float64, seed=3)
    At = to_dev(A, dev)
    A0 = At.clone()
    W = mak.similar_colmajor(At, (90, 60))
    noP = torch.empty((0, 0), dtype=At.dtype, device=dev)
    W2, P2 = mak.left_polar_(At, (W, noP))
    torch.cuda.synchronize()
    assert W2 is W and P2 is noP
    assert torch.equal(At, A0)
    Wo, _ = O.left_polar_svd(A)
    assert rel(W.cpu().numpy(), Wo) < 1e-11


def test_polar_argument_errors(dev):
    A = to_dev(O.gaussian(10, 20, np.float64, seed=1), dev)
    with pytest.raises(mak.ArgumentError):
        mak.left_polar_(A)          # m < n
    B = to_dev(O.gaussian(20, 10, np.float64, seed=1), dev)
    with pytest.raises(mak.ArgumentError):
        mak.right_polar_(B)         # n < m
    W, P = mak.initialize_output_polar(B, False)
    with pytest.raises(mak.DimensionMismatch):
        mak.left_polar_(B, (W, P[:5, :5]))
    with pytest.raises(mak.ArgumentError):
        mak.MI355XPolar(precision=8)
    C = to_dev(O.gaussian(20, 10, np.complex128, seed=1), dev)
    with pytest.raises(mak.ArgumentError):
        mak.left_polar_(C, None, mak.MI355XPolar(precision=32))


def test_polar_conditioned_1e6_fp64(dev):
    A, _ = O.conditioned_matrix(512, 1e6, seed=0)
    W, P = run_left(A, dev)
    info = mak.last_polar_info()
    Wo, Po = O.left_polar_svd(A)
    bwd = np.linalg.norm(A - W @ P) / np.linalg.norm(A)
    orth = np.linalg.norm(W.T @ W - np.eye(512))
    print(f"cond 1e6 512^2 FP64: relW {rel(W, Wo):.2e} relP {rel(P, Po):.2e} bwd {bwd:.2e} orth {orth:.2e} "
          f"info {info}", flush=True)
    assert rel(W, Wo) < 1e-9 and rel(P, Po) < 1e-12
    assert bwd < 1e-14 and orth < 1e-13
    assert info["qr_steps"] >= 1 and info["iterations"] <= 8


@pytest.mark.parametrize("prec,bwd_tol", [(32, 1e-5), (16, 1e-2)])
def test_polar_mixed_precision(prec, bwd_tol, dev):
    A, _ = O.conditioned_matrix(512, 1e6, seed=0)
    W, P = run_left(A, dev, mak.MI355XPolar(precision=prec))
    info = mak.last_polar_info()
    Wo, _ = O.left_polar_svd(A)
    bwd = np.linalg.norm(A - W @ P) / np.linalg.norm(A)
    orth = np.linalg.norm(W.T @ W - np.eye(512))
    print(f"cond 1e6 512^2 precision {prec}: relW {rel(W, Wo):.2e} bwd {bwd:.2e} orth {orth:.2e} info {info}",
          flush=True)
    assert orth < 1e-12 and bwd < bwd_tol
    assert info["ns_steps"] >= 2


@pytest.mark.parametrize("prec,tol", [(32, 1e-4), (16, 5e-2)])
@pytest.mark.parametrize("m,n", [(54, 37), (300, 170), (1000, 130)], ids=["54x37", "300x170", "1000x130"])
def test_polar_mixed_precision_ragged(m, n, prec, tol, dev):
    """Sizes off every tile multiple (packed FP32 / BF16 operand copies, the
    Hermitian shift of the Newton-Schulz steps): W within tol of the SVD
    oracle's (Gaussian inputs, modest condition numbers), orthogonal to FP64
    accuracy after the FP64 Newton-Schulz steps."""
    A = O.gaussian(m, n, np.float64, seed=m + n)
    W, P = run_left(A, dev, mak.MI355XPolar(precision=prec))
    Wo, _ = O.left_polar_svd(A)
    assert rel(W, Wo) < tol, rel(W, Wo)
    assert np.linalg.norm(W.T @ W - np.eye(n)) < 1e-12
    assert np.linalg.norm(A - W @ P) / np.linalg.norm(A) < tol


def test_polar_full_size_gaussian_properties(dev):
    n = 4096
    A = O.gaussian(n, n, np.float64, seed=7)
    W, P = run_left(A, dev)
    info = mak.last_polar_info()
    bwd = np.linalg.norm(A - W @ P) / np.linalg.norm(A)
    orth = np.linalg.norm(W.T @ W - np.eye(n)) / np.sqrt(n)
    lmin = np.linalg.eigvalsh(P).min()
    print(f"4096^2 Gaussian FP64: bwd {bwd:.2e} orth/sqrt(n) {orth:.2e} min eig P {lmin:.2e} info {info}", flush=True)
    assert bwd < 1e-13 and orth < 1e-13 and lmin > 0


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_left_polar_padded_w_leaves_parent_storage(dt, dev):
    """W given as a view W = big[:m, :n] of a larger column-major buffer (ld >
    m): the result matches the packed call and the parent's other rows are
    untouched (every scratch product lives in the workspace)."""
    m, n = 300, 200
    A = O.gaussian(m, n, dt, seed=77)
    big = mak.colmajor(torch.full((m + 37, n), 3.25, dtype=torch.float64 if dt == np.float64 else torch.complex128,
                                  device=dev))
    W = big[:m, :]
    assert W.stride(0) == 1 and W.stride(1) == m + 37
    At = to_dev(A, dev)
    P = mak.initialize_output_polar(At, False)[1]
    mak.left_polar_(At, (W, P))
    torch.cuda.synchronize()
    Wo, Po = O.left_polar_svd(A)
    assert rel(W.cpu().numpy(), Wo) < 1e-11 and rel(P.cpu().numpy(), Po) < 1e-11
    assert torch.all(big[m:, :] == 3.25)


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n,r", [(300, 100, 80), (120, 120, 100), (200, 64, 1)])
def test_left_polar_rank_deficient_is_isometric(m, n, r, dt, dev):
    """Rank-deficient A: the SVD route (src/implementations/polar.jl:59-70)
    returns an isometry W = U V^H; QDWH's partial isometry is completed, so
    W^H W = I, A = W P, and P matches the oracle's (unique) P."""
    A = O.gaussian(m, r, dt, seed=m + r) @ O.gaussian(r, n, dt, seed=n + r)
    W, P = run_left(A, dev)
    assert O.isisometric(W)
    assert np.linalg.norm(A - W @ P) / np.linalg.norm(A) < 1e-12
    _, Po = O.left_polar_svd(A)
    assert rel(P, Po) < 1e-10
    assert mak.last_polar_info()["completed"] == 1


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_polar_zero_matrix(dt, dev):
    """A = 0: W = I (m x n), P = 0 -- what the SVD route returns (U = I, V = I)."""
    W, P = run_left(np.zeros((50, 30), dtype=dt), dev)
    assert np.array_equal(W, np.eye(50, 30)) and np.all(P == 0)
    At = to_dev(np.zeros((30, 50), dtype=dt), dev)
    P, Wh = mak.right_polar_(At)
    torch.cuda.synchronize()
    assert np.array_equal(Wh.cpu().numpy(), np.eye(30, 50)) and np.all(P.cpu().numpy() == 0)


@pytest.fixture(scope="module")
def cond1e6_4096():
    """A = U diag(s) V^T, 4096 x 4096, U and V Haar-orthogonal (QR of a
    Gaussian with the sign fix), s log-spaced from 1 to 1e-6: the mixed-
    precision polar configuration.  Returns A and the exact polar factor U V^T."""
    n = 4096
    rng = np.random.default_rng(20261203)

    def haar(k):
        Q, R = np.linalg.qr(rng.standard_normal((k, k)))
        return Q * np.sign(np.diagonal(R))[None, :]

    U, V = haar(n), haar(n)
    s = np.logspace(0, -6, n)
    return (U * s[None, :]) @ V.T, U @ V.T


@pytest.mark.parametrize("prec,bwd", [(64, 1e-13), (32, 1e-5), (16, 1e-2)])
def test_left_polar_4096_cond1e6_mixed(prec, bwd, cond1e6_4096, dev):
    """left_polar! on the 4096^2, cond 1e6 configuration with FP64 / FP32 /
    BF16 bulk products + FP64 Newton-Schulz: orthogonality ||W^T W - I||_F
    <= 1e-12 in every precision, backward error ||A - W P||_F / ||A||_F <=
    1e-13 / 1e-5 / 1e-2, P symmetric.  The distance of W to the exact U V^T
    is bounded below by cond * u_bulk (the polar factor's condition number
    is ~1e6 here); it is reported, and for FP64 asserted <= 1e-9."""
    A, Wex = cond1e6_4096
    At = to_dev(A, dev)
    W, P = mak.initialize_output_polar(At, False)
    mak.left_polar_(At, (W, P), mak.MI355XPolar(precision=prec))
    torch.cuda.synchronize()
    info = mak.last_polar_info()
    eye = torch.eye(At.shape[1], dtype=At.dtype, device=dev)
    orth = torch.linalg.norm(W.T @ W - eye).item()
    bwd_err = (torch.linalg.norm(At - W @ P) / torch.linalg.norm(At)).item()
    relW = (torch.linalg.norm(W - torch.from_numpy(Wex).to(dev)) / np.sqrt(At.shape[1])).item()
    sym = (torch.linalg.norm(P - P.T) / torch.linalg.norm(P)).item()
    print(f"left_polar 4096 cond1e6 prec {prec}: |W^TW-I|_F {orth:.2e}, |A-WP|/|A| {bwd_err:.2e}, "
          f"|W-UV^T|/sqrt(n) {relW:.2e}, iterations {info['iterations']} (QR {info['qr_steps']}, "
          f"Chol {info['chol_steps']}, NS {info['ns_steps']})", flush=True)
    assert orth <= 1e-12
    assert bwd_err <= bwd
    assert sym <= 1e-14
    if prec == 64:
        assert relW <= 1e-9
