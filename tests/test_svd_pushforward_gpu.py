"""Parity of the SVD via the polar decomposition (SVDViaPolar) and of the
polar pushforwards with the CPU oracle.

  * svd_compact! / svd_vals!: the oracle is gesdd (the reference's default
    SVD route) with gaugefix!(svd_compact!, U, Vh) (src/common/gauge.jl:72-80);
    Gaussian inputs have well-separated singular values, so U, S, Vh compare
    elementwise: S within 1e-12, U and Vh within 1e-10 (relative Frobenius);
  * left/right_polar_pushforward! (src/pushforwards/polar.jl:1-25) against
    the oracle's restatement (itself pinned by central differences) within
    1e-10, including an ill-conditioned P (condition 1e6), where the
    eigendecomposition-based Sylvester solve stays exact.
"""
import numpy as np
import pytest
import torch

import mak_mi355x as mak
from oracle import mak_oracle as O

pytestmark = pytest.mark.gpu

DTS = [np.float64, np.complex128]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def to_dev(A, dev):
    return mak.colmajor(torch.from_numpy(np.asarray(A)).to(dev))


@pytest.mark.parametrize("dt", DTS, ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(54, 37), (37, 54), (54, 54), (300, 70), (70, 300), (1, 6), (6, 1)])
def test_svd_via_polar_matches_oracle(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=2100 + m + 7 * n)
    At = to_dev(A, dev)
    U, S, Vh = mak.svd_compact_(At, None, mak.SVDViaPolar())
    Uo, So, Vho = O.svd_compact(A)
    U, S, Vh = U.cpu().numpy(), S.cpu().numpy(), Vh.cpu().numpy()
    assert rel(S, So) < 1e-12
    assert rel(U, Uo) < 1e-10 and rel(Vh, Vho) < 1e-10
    assert rel((U * S[None, :]) @ Vh, A) < 1e-13
    assert np.array_equal(At.cpu().numpy(), A)   # A is only read
    Sv = mak.svd_vals(At).cpu().numpy()
    assert rel(Sv, So) < 1e-12


def test_svd_via_polar_conditioned_and_errors(dev):
    n = 300
    A, _ = O.conditioned_matrix(n, cond=1e6, seed=5)
    U, S, Vh = mak.svd_compact(to_dev(A, dev))
    So = np.linalg.svd(A, compute_uv=False)
    S = S.cpu().numpy()
    # singular values carry absolute errors of order u ||A||
    assert np.max(np.abs(S - So)) <= 1e-13 * So[0]
    U, Vh = U.cpu().numpy(), Vh.cpu().numpy()
    assert rel(U.T @ U, np.eye(n)) < 1e-12 and rel(Vh @ Vh.T, np.eye(n)) < 1e-12
    assert rel((U * S[None, :]) @ Vh, A) < 1e-13
    At = to_dev(A, dev)
    with pytest.raises(mak.DimensionMismatch):
        mak.svd_compact_(At, (mak.similar_colmajor(At, (n, n - 1)), torch.empty(n, dtype=torch.float64, device=dev),
                              mak.similar_colmajor(At, (n, n))))


@pytest.mark.parametrize("dt", DTS, ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(60, 40), (40, 40), (500, 64)])
def test_left_polar_pushforward_matches_oracle(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=3100 + m + n)
    dA = O.gaussian(m, n, dt, seed=3200 + m + n)
    W, P = O.left_polar_svd(A)
    dWo, dPo = O.left_polar_pushforward(dA, A, (W, P))
    dW, dP = mak.left_polar_pushforward_(to_dev(dA, dev), None, (to_dev(W, dev), to_dev(P, dev)))
    assert rel(dW.cpu().numpy(), dWo) < 1e-10 and rel(dP.cpu().numpy(), dPo) < 1e-10


@pytest.mark.parametrize("dt", DTS, ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(40, 60), (40, 40), (64, 500)])
def test_right_polar_pushforward_matches_oracle(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=4100 + m + n)
    dA = O.gaussian(m, n, dt, seed=4200 + m + n)
    P, Wh = O.right_polar_svd(A)
    dWho, dPo = O.right_polar_pushforward(dA, A, (P, Wh))
    dWh, dP = mak.right_polar_pushforward_(to_dev(dA, dev), None, (to_dev(P, dev), to_dev(Wh, dev)))
    assert rel(dWh.cpu().numpy(), dWho) < 1e-10 and rel(dP.cpu().numpy(), dPo) < 1e-10


def test_polar_pushforward_ill_conditioned_P(dev):
    """P with condition number 1e6: the Sylvester solve through the
    eigendecomposition of P stays accurate (the oracle solves it with
    Bartels-Stewart)."""
    n = 120
    A, _ = O.conditioned_matrix(n, cond=1e6, seed=9)
    dA = O.gaussian(n, n, np.float64, seed=10)
    W, P = O.left_polar_svd(A)
    dWo, dPo = O.left_polar_pushforward(dA, A, (W, P))
    dW, dP = mak.left_polar_pushforward_(to_dev(dA, dev), None, (to_dev(W, dev), to_dev(P, dev)))
    # both sides carry the conditioning of P^{-1} (~1e6 u relative)
    assert rel(dW.cpu().numpy(), dWo) < 1e-8 and rel(dP.cpu().numpy(), dPo) < 1e-8
