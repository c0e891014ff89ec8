"""Parity of the HIP LQ, null-space and orth / projection composites with the
CPU oracle.

  * lq_compact! / lq_full! (LQViaTransposedQR, src/implementations/lq.jl:
    283-310): with positive = true the LQ of a full-rank matrix is unique, so L
    and the first min(m, n) rows of Q compare elementwise (1e-12); the trailing
    rows of a full Q are fixed only up to a rotation and are checked through
    unitarity and A = L Q;
  * qr_null! / lq_null! (src/implementations/qr.jl:236-263, lq.jl:312-318):
    the bases are fixed only up to a rotation; the orthogonal projectors
    N N^H compare with the oracle's (1e-12), with orthonormality and
    A^H N = 0 (A Nh^H = 0) at 1e-12 relative;
  * left_orth! / right_orth! / left_null! / right_null! / project_isometric!
    (src/implementations/orthnull.jl:79-111, projections.jl:70-75) and
    PolarNewton on the device, against the oracle's QR / LQ / polar
    (Newton and SVD routes) at 1e-11.
"""
import numpy as np
import pytest
import torch

import mak_mi355x as mak
from oracle import mak_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(54, 37), (37, 54), (54, 54), (300, 70), (70, 300), (1, 9), (9, 1)]
DTS = [np.float64, np.complex128]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def to_dev(A, dev):
    return mak.colmajor(torch.from_numpy(np.asarray(A)).to(dev))


@pytest.mark.parametrize("dt", DTS, ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", SHAPES)
def test_lq_compact_matches_oracle(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=500 + m + 3 * n)
    At = to_dev(A, dev)
    L, Q = mak.lq_compact_(At)
    Lo, Qo = O.lq_via_qr(A)
    L, Q = L.cpu().numpy(), Q.cpu().numpy()
    assert rel(L, Lo) < 1e-12 and rel(Q, Qo) < 1e-12
    assert np.allclose(np.triu(L, 1), 0)
    # A is only read
    assert np.array_equal(At.cpu().numpy(), A)


@pytest.mark.parametrize("dt", DTS, ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", SHAPES)
def test_lq_full_matches_oracle(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=900 + m + 3 * n)
    L, Q = mak.lq_full_(to_dev(A, dev))
    Lo, Qo = O.lq_via_qr(A, full=True)
    L, Q = L.cpu().numpy(), Q.cpu().numpy()
    k = min(m, n)
    assert L.shape == (m, n) and Q.shape == (n, n)
    assert rel(L, Lo) < 1e-12 and rel(Q[:k], Qo[:k]) < 1e-12
    assert rel(Q @ Q.conj().T, np.eye(n)) < 1e-12
    assert rel(L @ Q, A) < 1e-13


def test_lq_without_L_and_argument_errors(dev):
    A = O.gaussian(20, 50, np.float64, seed=4)
    At = to_dev(A, dev)
    Q = mak.similar_colmajor(At, (20, 50))
    L0 = torch.empty((0, 0), dtype=torch.float64, device=dev)
    L, Q2 = mak.lq_compact_(At, (L0, Q))
    assert Q2 is Q and L is L0
    assert rel(Q.cpu().numpy(), O.lq_via_qr(A)[1]) < 1e-12
    with pytest.raises(mak.DimensionMismatch):
        mak.lq_compact_(At, (L0, mak.similar_colmajor(At, (50, 50))))
    # Q overlapping A is an invalid argument of the C ABI
    with pytest.raises(mak.ArgumentError):
        mak.lq_compact_(At, (L0, At[:, :50]))


@pytest.mark.parametrize("dt", DTS, ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(54, 37), (300, 70), (54, 54), (37, 54), (2100, 64)])
def test_qr_null_spans_the_oracle_complement(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=1300 + m + n)
    N = mak.qr_null_(to_dev(A, dev)).cpu().numpy()
    No = O.qr_null(A)
    k = min(m, n)
    assert N.shape == (m, m - k)
    if m > k:
        assert rel(N.conj().T @ N, np.eye(m - k)) < 1e-12
        assert np.linalg.norm(A.conj().T @ N) <= 1e-12 * np.linalg.norm(A) * np.sqrt(m)
        assert rel(N @ N.conj().T, No @ No.conj().T) < 1e-12


@pytest.mark.parametrize("dt", DTS, ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(37, 54), (70, 300), (54, 54), (54, 37)])
def test_lq_null_spans_the_oracle_complement(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=1700 + m + n)
    At = to_dev(A, dev)
    Nh = mak.lq_null_(At).cpu().numpy()
    No = O.lq_null(A)
    k = min(m, n)
    assert Nh.shape == (n - k, n)
    if n > k:
        assert rel(Nh @ Nh.conj().T, np.eye(n - k)) < 1e-12
        assert np.linalg.norm(A @ Nh.conj().T) <= 1e-12 * np.linalg.norm(A) * np.sqrt(n)
        assert rel(Nh.conj().T @ Nh, No.conj().T @ No) < 1e-12
    assert np.array_equal(At.cpu().numpy(), A)


@pytest.mark.parametrize("dt", DTS, ids=["f64", "c64"])
def test_orth_null_composites(dt, dev):
    m, n = 120, 45
    A = O.gaussian(m, n, dt, seed=21)
    # left_orth via QR and via polar
    V, Cm = mak.left_orth(to_dev(A, dev))
    Qo, Ro = O.qr_compact(A)
    assert rel(V.cpu().numpy(), Qo) < 1e-12 and rel(Cm.cpu().numpy(), Ro) < 1e-12
    W, P = mak.left_orth(to_dev(A, dev), mak.MI355XPolar())
    Wo, Po = O.left_polar_svd(A)
    assert rel(W.cpu().numpy(), Wo) < 1e-11 and rel(P.cpu().numpy(), Po) < 1e-11
    # right_orth via LQ and via polar
    B = np.conj(A).T.copy()
    Cr, Vh = mak.right_orth(to_dev(B, dev))
    Lo, Qlo = O.lq_via_qr(B)
    assert rel(Cr.cpu().numpy(), Lo) < 1e-12 and rel(Vh.cpu().numpy(), Qlo) < 1e-12
    Pr, Wh = mak.right_orth(to_dev(B, dev), mak.PolarNewton())
    Pro, Who = O.right_polar_newton(B)
    assert rel(Pr.cpu().numpy(), Pro) < 1e-11 and rel(Wh.cpu().numpy(), Who) < 1e-11
    # left / right null spaces
    N = mak.left_null(to_dev(A, dev)).cpu().numpy()
    assert N.shape == (m, m - n)
    assert rel(N @ N.conj().T, np.eye(m) - Qo @ Qo.conj().T) < 1e-12
    Nh = mak.right_null(to_dev(B, dev)).cpu().numpy()
    assert rel(Nh.conj().T @ Nh, np.eye(m) - Qlo.conj().T @ Qlo) < 1e-12


@pytest.mark.parametrize("dt", DTS, ids=["f64", "c64"])
def test_polar_newton_and_project_isometric(dt, dev):
    A = O.gaussian(80, 50, dt, seed=8)
    Wn, Pn = O.left_polar_newton(A)
    W, P = mak.left_polar(to_dev(A, dev), mak.PolarNewton())
    assert rel(W.cpu().numpy(), Wn) < 1e-11 and rel(P.cpu().numpy(), Pn) < 1e-11
    Wp = mak.project_isometric(to_dev(A, dev)).cpu().numpy()
    assert rel(Wp, Wn) < 1e-11
    with pytest.raises(mak.DimensionMismatch):
        mak.project_isometric(to_dev(A.T.copy(), dev))
