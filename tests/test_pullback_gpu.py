"""Parity of the HIP reverse rules (qr_pullback!, left/right_polar_pullback!)
with the CPU oracle's restatement of src/pullbacks/qr.jl and
src/pullbacks/polar.jl, with the golden fixtures, and with central finite
differences of the GPU forward path.

Tolerances (relative Frobenius norm of dA):
  * vs oracle / fixtures: 1e-10 (the rules divide by R11 or P, whose condition
    numbers are <= ~1e3 for these Gaussian inputs, and the Sylvester solve of
    the polar rule stops at a 1e-14 relative residual);
  * vs central differences, h = 1e-6: 1e-6 (truncation O(h^2) plus rounding
    O(u / h) of the forward map).
"""
import numpy as np
import pytest
import torch
from pathlib import Path

import mak_mi355x as mak
from oracle import mak_oracle as O

pytestmark = pytest.mark.gpu
GOLD = np.load(Path(__file__).parent / "golden" / "golden_small.npz")


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def dev_t(A, dev):
    return mak.colmajor(torch.from_numpy(np.asfortranarray(A)).to(dev))


def host(x):
    torch.cuda.synchronize()
    return x.cpu().numpy()


def gpu_qr_pullback(A, dQ, dR, dev, dA0=None):
    At = dev_t(A, dev)
    Q, R, tape = mak.qr_compact_rrule(At)
    dA = dev_t(np.zeros_like(A) if dA0 is None else dA0, dev)
    mak.qr_pullback_(dA, At, (Q, R), (None if dQ is None else dev_t(dQ, dev), None if dR is None else dev_t(dR, dev)),
                     tape=tape)
    return host(dA), host(Q), host(R)


QR_SIZES = [(54, 37), (37, 54), (54, 54), (300, 70), (1025, 64), (700, 300)]


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", QR_SIZES, ids=[f"{m}x{n}" for m, n in QR_SIZES])
@pytest.mark.parametrize("which", ["both", "dQ", "dR"])
def test_qr_pullback_matches_oracle(m, n, dt, which, dev):
    A = O.gaussian(m, n, dt, seed=900 + m + n)
    k = min(m, n)
    dQ = O.gaussian(m, k, dt, seed=901 + m) if which != "dR" else None
    dR = np.triu(O.gaussian(k, n, dt, seed=902 + n)) if which != "dQ" else None
    dA0 = O.gaussian(m, n, dt, seed=903)
    dA, Q, R = gpu_qr_pullback(A, dQ, dR, dev, dA0)
    ref = O.qr_pullback(dA0, A, (Q, R), (dQ, dR))
    assert rel(dA, ref) < 1e-10, rel(dA, ref)


@pytest.mark.parametrize("key", [k[:-2] for k in GOLD.files if k.startswith("qrpb_") and k.endswith("_A")])
def test_qr_pullback_golden(key, dev):
    dA, _, _ = gpu_qr_pullback(GOLD[key + "_A"], GOLD[key + "_dQ"], GOLD[key + "_dR"], dev)
    assert rel(dA, GOLD[key + "_dA"]) < 1e-10


def qr_objective(A, dQ, dR, dev):
    Q, R = mak.qr_compact(dev_t(A, dev))
    return float(np.real(np.vdot(dQ, host(Q)) + np.vdot(dR, host(R))))


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(40, 25), (25, 40), (600, 64)])
def test_qr_pullback_central_differences(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=31)
    k = min(m, n)
    dQ = O.gaussian(m, k, dt, seed=32)
    dR = np.triu(O.gaussian(k, n, dt, seed=33))
    dA, _, _ = gpu_qr_pullback(A, dQ, dR, dev)
    E = O.gaussian(m, n, dt, seed=34)
    h = 1e-6
    fd = (qr_objective(A + h * E, dQ, dR, dev) - qr_objective(A - h * E, dQ, dR, dev)) / (2 * h)
    ad = float(np.real(np.vdot(dA, E)))
    assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1.0), (fd, ad)


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_qr_pullback_rank_deficient(dt, dev):
    m, n, r = 80, 40, 25
    A = O.gaussian(m, r, dt, seed=41) @ O.gaussian(r, n, dt, seed=42)
    At = dev_t(A, dev)
    Q, R, tape = mak.qr_compact_rrule(At)
    Rh = host(R)
    p = mak.qr_rank(R)
    assert p == O.qr_rank(Rh) == r
    dQ = O.gaussian(m, n, dt, seed=43)
    dR = np.triu(O.gaussian(n, n, dt, seed=44))
    dQ, dR = O.remove_qr_gauge_dependence(dQ, dR, A, host(Q), Rh)
    dA = dev_t(np.zeros_like(A), dev)
    mak.qr_pullback_(dA, At, (Q, R), (dev_t(dQ, dev), dev_t(dR, dev)), tape=tape)
    ref = O.qr_pullback(np.zeros_like(A), A, (host(Q), Rh), (dQ, dR))
    assert rel(host(dA), ref) < 1e-9, rel(host(dA), ref)


def test_qr_pullback_nonpositive_gauge(dev):
    """positive = false: the re-formed Q must be the primal's LAPACK-gauge Q."""
    A = O.gaussian(120, 50, np.float64, seed=51)
    At = dev_t(A, dev)
    Q, R, tape = mak.qr_compact_rrule(At, mak.MI355XHouseholder(positive=False))
    Qh, Rh = host(Q), host(R)
    assert (np.diagonal(Rh) < 0).any()
    dQ = O.gaussian(120, 50, np.float64, seed=52)
    dR = np.triu(O.gaussian(50, 50, np.float64, seed=53))
    dA = dev_t(np.zeros_like(A), dev)
    mak.qr_pullback_(dA, At, (Q, R), (dev_t(dQ, dev), dev_t(dR, dev)), tape=tape)
    ref = O.qr_pullback(np.zeros_like(A), A, (Qh, Rh), (dQ, dR))
    assert rel(host(dA), ref) < 1e-10


def test_qr_pullback_batched(dev):
    """A batch of 2048 x 256 problems in one call; members checked against the oracle."""
    batch, m, n = 12, 2048, 256
    As = [O.gaussian(m, n, np.float64, seed=60 + b) for b in range(batch)]
    tapes, Qs, Rs = [], [], []
    for A in As:
        Q, R, tape = mak.qr_compact_rrule(dev_t(A, dev))
        tapes.append(tape); Qs.append(Q); Rs.append(R)
    Awy = mak.colmajor_batch(torch.stack([t.Awy for t in tapes]))
    T = torch.stack([t.T for t in tapes])
    R3 = mak.colmajor_batch(torch.stack(Rs))
    dQs = [O.gaussian(m, n, np.float64, seed=80 + b) for b in range(batch)]
    dRs = [np.triu(O.gaussian(n, n, np.float64, seed=90 + b)) for b in range(batch)]
    dQ3 = mak.colmajor_batch(torch.stack([dev_t(x, dev) for x in dQs]))
    dR3 = mak.colmajor_batch(torch.stack([dev_t(x, dev) for x in dRs]))
    dA3 = mak.empty_colmajor_batch(batch, m, n, torch.float64, dev)
    dA3.zero_()
    mak.qr_pullback_batched_(dA3, Awy, T, R3, dQ3, dR3)
    out = host(dA3)
    for b in (0, 5, batch - 1):
        ref = O.qr_pullback(np.zeros((m, n)), As[b], (host(Qs[b]), host(Rs[b])), (dQs[b], dRs[b]))
        assert rel(out[b], ref) < 1e-10, (b, rel(out[b], ref))


def test_qr_pullback_argument_errors(dev):
    A = dev_t(O.gaussian(30, 20, np.float64, seed=1), dev)
    Q, R, tape = mak.qr_compact_rrule(A)
    dA = dev_t(np.zeros((30, 21)), dev)
    with pytest.raises(mak.DimensionMismatch):
        mak.qr_pullback_(dA, A, (Q, R), (None, None), tape=tape)
    dA = dev_t(np.zeros((30, 20)), dev)
    with pytest.raises(mak.DimensionMismatch):
        mak.qr_pullback_(dA, A, (Q, R), (dev_t(np.zeros((30, 19)), dev), None), tape=tape)
    # absent cotangents: dA unchanged
    mak.qr_pullback_(dA, A, (Q, R), (None, None), tape=tape)
    assert not host(dA).any()


# --------------------------------------------------------------- polar ----

def gpu_left_polar_pullback(A, dW, dP, dev, dA0=None):
    At = dev_t(A, dev)
    W, P = mak.left_polar(At)
    dA = dev_t(np.zeros_like(A) if dA0 is None else dA0, dev)
    mak.left_polar_pullback_(dA, At, (W, P), (None if dW is None else dev_t(dW, dev),
                                              None if dP is None else dev_t(dP, dev)))
    return host(dA), host(W), host(P)


POL_SIZES = [(54, 37), (54, 54), (300, 200), (1000, 130)]


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", POL_SIZES, ids=[f"{m}x{n}" for m, n in POL_SIZES])
@pytest.mark.parametrize("which", ["both", "dW", "dP"])
def test_left_polar_pullback_matches_oracle(m, n, dt, which, dev):
    A = O.gaussian(m, n, dt, seed=700 + m + n)
    dW = O.gaussian(m, n, dt, seed=701) if which != "dP" else None
    dP = O.gaussian(n, n, dt, seed=702) if which != "dW" else None   # not Hermitian: the rule projects it
    dA0 = O.gaussian(m, n, dt, seed=703)
    dA, W, P = gpu_left_polar_pullback(A, dW, dP, dev, dA0)
    ref = O.left_polar_pullback(dA0, A, (W, P), (dW, dP))
    assert rel(dA, ref) < 1e-10, rel(dA, ref)


@pytest.mark.parametrize("key", [k[:-2] for k in GOLD.files if k.startswith("lpolpb_") and k.endswith("_A")])
def test_left_polar_pullback_golden(key, dev):
    dA, _, _ = gpu_left_polar_pullback(GOLD[key + "_A"], GOLD[key + "_dW"], GOLD[key + "_dP"], dev)
    assert rel(dA, GOLD[key + "_dA"]) < 1e-10


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(37, 54), (54, 54), (130, 700)])
def test_right_polar_pullback_matches_oracle(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=750 + m + n)
    At = dev_t(A, dev)
    P, Wh = mak.right_polar(At)
    dP = O.gaussian(m, m, dt, seed=751)
    dWh = O.gaussian(m, n, dt, seed=752)
    dA0 = O.gaussian(m, n, dt, seed=753)
    dA = dev_t(dA0, dev)
    mak.right_polar_pullback_(dA, At, (P, Wh), (dev_t(dP, dev), dev_t(dWh, dev)))
    ref = O.right_polar_pullback(dA0, A, (host(P), host(Wh)), (dP, dWh))
    assert rel(host(dA), ref) < 1e-10, rel(host(dA), ref)


def polar_objective(A, dW, dP, dev):
    W, P = mak.left_polar(dev_t(A, dev))
    return float(np.real(np.vdot(dW, host(W)) + np.vdot(dP, host(P))))


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_left_polar_pullback_central_differences(dt, dev):
    m, n = 60, 40
    A = O.gaussian(m, n, dt, seed=71)
    dW = O.gaussian(m, n, dt, seed=72)
    dP = O.project_hermitian(O.gaussian(n, n, dt, seed=73))
    dA, _, _ = gpu_left_polar_pullback(A, dW, dP, dev)
    E = O.gaussian(m, n, dt, seed=74)
    h = 1e-6
    fd = (polar_objective(A + h * E, dW, dP, dev) - polar_objective(A - h * E, dW, dP, dev)) / (2 * h)
    ad = float(np.real(np.vdot(dA, E)))
    assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1.0), (fd, ad)


def test_left_polar_pullback_batched(dev):
    batch, m, n = 12, 2048, 256
    As = [O.gaussian(m, n, np.float64, seed=120 + b) for b in range(batch)]
    Ws, Ps = [], []
    for A in As:
        W, P = mak.left_polar(dev_t(A, dev))
        Ws.append(W); Ps.append(P)
    W3 = mak.colmajor_batch(torch.stack(Ws))
    P3 = mak.colmajor_batch(torch.stack(Ps))
    dWs = [O.gaussian(m, n, np.float64, seed=140 + b) for b in range(batch)]
    dPs = [O.gaussian(n, n, np.float64, seed=160 + b) for b in range(batch)]
    dW3 = mak.colmajor_batch(torch.stack([dev_t(x, dev) for x in dWs]))
    dP3 = mak.colmajor_batch(torch.stack([dev_t(x, dev) for x in dPs]))
    dA3 = mak.empty_colmajor_batch(batch, m, n, torch.float64, dev)
    dA3.zero_()
    _, iters = mak.left_polar_pullback_batched_(dA3, W3, P3, dW3, dP3)
    out = host(dA3)
    print(f"batched polar pullback {batch}x({m}x{n}): {iters} CG iterations", flush=True)
    for b in (0, 7, batch - 1):
        ref = O.left_polar_pullback(np.zeros((m, n)), As[b], (host(Ws[b]), host(Ps[b])), (dWs[b], dPs[b]))
        assert rel(out[b], ref) < 1e-10, (b, rel(out[b], ref))


@pytest.mark.parametrize("cond,tol", [(1e4, 1e-10), (1e6, 1e-8)])
def test_left_polar_pullback_ill_conditioned(cond, tol, dev):
    """P with condition number 1e4 / 1e6: CG with the diagonal preconditioner
    needs ~sqrt(cond) steps, so past 48 steps the batch is solved through the
    eigendecomposition of P; compared with the oracle's rule (Bartels-Stewart
    Sylvester solve, LU inverse).  Both carry relative errors ~ cond u."""
    n = 200
    A, _ = O.conditioned_matrix(n, cond=cond, seed=31)
    W, P = mak.left_polar(dev_t(A, dev))
    dW = O.gaussian(n, n, np.float64, seed=32)
    dP = O.gaussian(n, n, np.float64, seed=33)
    dA3 = mak.empty_colmajor_batch(1, n, n, torch.float64, dev)
    dA3.zero_()
    _, iters = mak.left_polar_pullback_batched_(dA3, mak.colmajor(W).unsqueeze(0), mak.colmajor(P).unsqueeze(0),
                                                dev_t(dW, dev).unsqueeze(0), dev_t(dP, dev).unsqueeze(0))
    ref = O.left_polar_pullback(np.zeros((n, n)), A, (host(W), host(P)), (dW, dP))
    err = rel(host(dA3)[0], ref)
    print(f"polar pullback cond {cond:.0e}: {iters} CG steps, rel err {err:.2e}", flush=True)
    assert err < tol


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_qr_pullback_full_q_cotangent(dt, dev):
    """qr_full! cotangents (ΔQ m x m): the ΔQ3 part along Q1 enters ΔQ1
    (src/pullbacks/qr.jl:19-26); the result matches the oracle's rule on the
    same full (Q, R).  A ΔQ3 with a component outside Q1's span is gauge
    dependent and warns; one inside it does not."""
    m, n = 90, 40
    A = O.gaussian(m, n, dt, seed=81)
    At = dev_t(A, dev)
    Qf, Rf = mak.qr_full(At)
    Qh, Rh = host(Qf), host(Rf)
    dR = np.triu(O.gaussian(m, n, dt, seed=83))
    dR[n:] = 0
    # gauge-invariant: dQ3 = Q1 X
    dQ = O.gaussian(m, m, dt, seed=82)
    dQ[:, n:] = Qh[:, :n] @ O.gaussian(n, m - n, dt, seed=84)
    dA0 = O.gaussian(m, n, dt, seed=85)
    dA = dev_t(dA0, dev)
    import warnings as _w
    with _w.catch_warnings():
        _w.simplefilter("error")
        mak.qr_pullback_(dA, At, (Qf, Rf), (dev_t(dQ, dev), dev_t(dR, dev)))
    ref = O.qr_pullback(dA0, A, (Qh, Rh), (dQ, dR))
    assert rel(host(dA), ref) < 1e-10, rel(host(dA), ref)
    # gauge-dependent: a random dQ3 warns, and the rule still matches the oracle's
    dQ2 = O.gaussian(m, m, dt, seed=86)
    dA = dev_t(dA0, dev)
    with pytest.warns(UserWarning, match="sensitive to gauge"):
        mak.qr_pullback_(dA, At, (Qf, Rf), (dev_t(dQ2, dev), dev_t(dR, dev)))
    with pytest.warns(UserWarning, match="sensitive to gauge"):
        ref2 = O.qr_pullback(dA0, A, (Qh, Rh), (dQ2, dR))
    assert rel(host(dA), ref2) < 1e-10


def test_qr_pullback_gauge_warning_rank_deficient(dev):
    """Rank-deficient R (p < min(m,n)): cotangent entries on Q[:, p:] or on
    the upper triangle of R[p:, p:] are gauge dependent and warn
    (src/pullbacks/qr.jl:27-29,34-47); zero ones do not."""
    m, n, r = 60, 30, 20
    A = O.gaussian(m, r, np.float64, seed=91) @ O.gaussian(r, n, np.float64, seed=92)
    At = dev_t(A, dev)
    Q, R, tape = mak.qr_compact_rrule(At)
    assert mak.qr_rank(R) == r
    dQ = O.gaussian(m, n, np.float64, seed=93)
    dQ[:, r:] = 0
    dR = np.triu(O.gaussian(n, n, np.float64, seed=94))
    dR[r:] = 0
    import warnings as _w
    dA = dev_t(np.zeros((m, n)), dev)
    with _w.catch_warnings():
        _w.simplefilter("error")
        mak.qr_pullback_(dA, At, (Q, R), (dev_t(dQ, dev), dev_t(dR, dev)), tape=tape)
    ref = O.qr_pullback(np.zeros((m, n)), A, (host(Q), host(R)), (dQ, dR))
    assert rel(host(dA), ref) < 1e-10
    dQ[:, r + 2] = 1.0
    with pytest.warns(UserWarning, match="sensitive to gauge"):
        mak.qr_pullback_(dev_t(np.zeros((m, n)), dev), At, (Q, R), (dev_t(dQ, dev), dev_t(dR, dev)), tape=tape)


def test_qr_ranks_batched_one_transfer(dev):
    """qr_ranks over a stack equals qr_rank per matrix (mixed ranks)."""
    Rs = []
    for b, r in enumerate([5, 8, 8, 3]):
        A = O.gaussian(30, r, np.float64, seed=100 + b) @ O.gaussian(r, 8, np.float64, seed=110 + b)
        Rs.append(mak.qr_compact(dev_t(A, dev))[1])
    R3 = mak.colmajor_batch(torch.stack(Rs))
    assert mak.qr_ranks(R3) == [O.qr_rank(host(R)) for R in Rs] == [5, 8, 8, 3]
