"""Batched qr_compact! (one captured graph for many uneven problems) against
the oracle's per-block qr_compact (dgeqrt + dgemqrt + gauge).

With positive = true the QR of a full-rank block is unique, so Q and R are
compared elementwise: relative Frobenius error <= 1e-12 per block.  At the full
4096-block size every block is checked through ||A - QR|| / ||A|| <= 1e-14,
||Q^H Q - I|| / sqrt(k) <= 1e-14 and a non-negative real diagonal of R, and a
seeded sample of blocks against the oracle.
"""
import numpy as np
import pytest
import torch

import mak_mi355x as mak
from oracle import mak_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make_blocks(shapes, dt, seed, dev):
    host = [O.gaussian(m, n, dt, seed=seed + i) for i, (m, n) in enumerate(shapes)]
    devs = [mak.colmajor(torch.from_numpy(np.asfortranarray(h)).to(dev)) for h in host]
    return host, devs


EDGE = [(1, 1), (5, 1), (1, 5), (32, 32), (33, 33), (64, 32), (31, 64), (100, 7), (256, 256), (257, 100),
        (512, 40), (513, 513), (700, 300), (1024, 64), (1000, 1000), (2, 300), (600, 33), (1100, 40),
        (2100, 37), (0, 5), (4, 0)]


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_batched_edge_shapes_match_oracle(dt, dev):
    host, As = make_blocks(EDGE, dt, 40, dev)
    plan = mak.BatchedQRPlan(As)
    Qs, Rs = plan.execute()
    torch.cuda.synchronize()
    info = plan.info()
    for (m, n), A, Q, R in zip(EDGE, host, Qs, Rs):
        Qo, Ro = O.qr_compact(A)
        assert rel(Q.cpu().numpy(), Qo) < 1e-12, (m, n, rel(Q.cpu().numpy(), Qo))
        assert rel(R.cpu().numpy(), Ro) < 1e-12, (m, n)
    # panels taller than one node use a one-level TSQR inside the batch (two
    # leaves of up to 1024 rows Float64, two to four of up to 256 rows
    # ComplexF64); blocks taller than that (1024 rows ComplexF64, 2048
    # Float64) and empty blocks run through the single-matrix driver
    tall = 1024 if dt == np.complex128 else 2048
    assert info["fallback_problems"] == sum(1 for m, n in EDGE if m > tall or min(m, n) == 0)
    plan.close()


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_batched_random_shapes_match_oracle(dt, dev):
    shapes = mak.block_shapes(150, seed=3, mmax=600)
    host, As = make_blocks(shapes, dt, 7, dev)
    Qs, Rs = mak.qr_compact_batched(As)
    torch.cuda.synchronize()
    for (m, n), A, Ad, Q, R in zip(shapes, host, As, Qs, Rs):
        Qo, Ro = O.qr_compact(A)
        assert rel(Q.cpu().numpy(), Qo) < 1e-12, (m, n)
        assert rel(R.cpu().numpy(), Ro) < 1e-12, (m, n)
        assert np.array_equal(Ad.cpu().numpy(), A)   # qr_compact copies its input


def panel_status(plan):
    """(accepted, rejected) CholeskyQR2 panels of the last execution."""
    import ctypes as C
    from mak_mi355x import _lib
    lib = _lib.load()
    lib.maktest_batched_panel_status.restype = C.c_int
    lib.maktest_batched_panel_status.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    a, r = C.c_int(-1), C.c_int(-1)
    assert lib.maktest_batched_panel_status(plan._h, C.byref(a), C.byref(r)) == 0
    return a.value, r.value


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_batched_cholqr2_panels_taken(dt, dev):
    """Every panel of at least 64 rows is factored by the CholeskyQR2 panel
    task for Gaussian blocks (none falls back), and the factors are the
    oracle's."""
    shapes = mak.block_shapes(200, seed=9, mmax=1000)
    host, As = make_blocks(shapes, dt, 300, dev)
    plan = mak.BatchedQRPlan(As)
    Qs, Rs = plan.execute()
    torch.cuda.synchronize()
    acc, rej = panel_status(plan)
    panels = sum(1 for m, n in shapes for j0 in range(0, min(m, n), 32) if m - j0 >= 64)
    assert (acc, rej) == (panels, 0)
    for i in (0, 50, 120, 199):
        Qo, Ro = O.qr_compact(host[i])
        assert rel(Qs[i].cpu().numpy(), Qo) < 1e-12 and rel(Rs[i].cpu().numpy(), Ro) < 1e-12, shapes[i]
    plan.close()


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_batched_rejected_cholqr2_panels_fall_back(dt, dev):
    """Panels the CholeskyQR2 check rejects -- an exactly zero column, two
    equal columns, a column of magnitude 1e-300 (its squares underflow) --
    are refactored by the gated Householder launches: residual,
    orthogonality and gauge of Householder QR, finite everywhere.  A column
    made dependent by an earlier panel (left as rounding noise by the
    trailing update) is accepted, and the factorisation is still a QR of the
    block.  Gaussian blocks alongside match the oracle."""
    Z = O.gaussian(300, 64, dt, seed=31)
    Z[:, 5] = 0.0
    D = O.gaussian(256, 40, dt, seed=33)
    D[:, 1] = D[:, 0]
    S = O.gaussian(400, 64, dt, seed=32)
    S[:, 7] *= 1e-300
    L = O.gaussian(300, 70, dt, seed=21)
    L[:, 40] = L[:, 3]
    host = [Z, D, S, L, O.gaussian(500, 200, dt, seed=24), O.gaussian(120, 40, dt, seed=25)]
    As = [mak.colmajor(torch.from_numpy(np.asfortranarray(h)).to(dev)) for h in host]
    plan = mak.BatchedQRPlan(As)
    Qs, Rs = plan.execute()
    torch.cuda.synchronize()
    acc, rej = panel_status(plan)
    assert rej >= 3 and acc > 0
    for h, Q, R in zip(host, Qs, Rs):
        Qh, Rh = Q.cpu().numpy(), R.cpu().numpy()
        assert np.all(np.isfinite(Qh)) and np.all(np.isfinite(Rh))
        assert rel(Qh @ Rh, h) < 1e-13
        assert np.linalg.norm(Qh.conj().T @ Qh - np.eye(Qh.shape[1])) < 1e-12
        d = np.diagonal(Rh)
        assert np.all(d.real >= 0) and np.all(d.imag == 0)
    for i in (4, 5):
        Qo, Ro = O.qr_compact(host[i])
        assert rel(Qs[i].cpu().numpy(), Qo) < 1e-12 and rel(Rs[i].cpu().numpy(), Ro) < 1e-12
    plan.close()


def test_batched_plan_replay_and_tuning(dev):
    """The captured graph replays on new data in the same buffers; a tuned plan
    gives the same factors."""
    shapes = mak.block_shapes(300, seed=5, mmax=400)
    host1, As = make_blocks(shapes, np.float64, 100, dev)
    plan = mak.BatchedQRPlan(As, tune=True)
    info = plan.info()
    assert info["chunk_cols"] in (32, 64) and info["grid_per_cu"] in (1, 2, 4) and info["launches"] > 0
    for rep, seed in enumerate((100, 900)):
        host = [O.gaussian(m, n, np.float64, seed=seed + i) for i, (m, n) in enumerate(shapes)]
        for A, h in zip(As, host):
            A.copy_(torch.from_numpy(np.asfortranarray(h)).to(dev))
        Qs, Rs = plan.execute()
        torch.cuda.synchronize()
        for i in (0, 17, 150, 299):
            Qo, Ro = O.qr_compact(host[i])
            assert rel(Qs[i].cpu().numpy(), Qo) < 1e-12 and rel(Rs[i].cpu().numpy(), Ro) < 1e-12, (rep, i)
    plan.close()


def test_batched_without_r_and_nonpositive(dev):
    shapes = [(80, 30), (30, 80), (300, 300)]
    host, As = make_blocks(shapes, np.float64, 11, dev)
    plan = mak.BatchedQRPlan(As, Rs=False, alg=mak.MI355XHouseholder(positive=False))
    Qs, _ = plan.execute()
    torch.cuda.synchronize()
    for A, Q in zip(host, Qs):
        Qo, _ = O.qr_householder(A, "compact", positive=False)
        assert rel(Q.cpu().numpy(), Qo) < 1e-12
    plan.close()


def test_batched_tall_block_with_rejected_cholqr_panel(dev):
    """Blocks taller than two nodes run through the single-matrix driver inside
    the captured graph, where a panel rejected by the CholeskyQR2 check is
    refactored by the flag-gated Householder launches (no host check)."""
    A1 = O.gaussian(2100, 70, np.float64, seed=5)
    A1[:, 40] = A1[:, 3]          # rank-deficient second panel
    A2 = O.gaussian(2300, 40, np.float64, seed=6)
    host = [A1, A2, O.gaussian(300, 20, np.float64, seed=7)]
    As = [mak.colmajor(torch.from_numpy(np.asfortranarray(h)).to(dev)) for h in host]
    plan = mak.BatchedQRPlan(As)
    assert plan.info()["fallback_problems"] == 2
    for rep in range(2):   # the graph replays on the same (restored) inputs
        for A, h in zip(As, host):
            A.copy_(torch.from_numpy(np.asfortranarray(h)).to(dev))
        Qs, Rs = plan.execute()
        torch.cuda.synchronize()
        for h, Q, R in zip(host, Qs, Rs):
            Qh, Rh = Q.cpu().numpy(), R.cpu().numpy()
            assert rel(Qh @ Rh, h) < 1e-13, rep
            assert np.linalg.norm(Qh.T @ Qh - np.eye(Qh.shape[1])) < 1e-12
            assert np.all(np.diagonal(Rh) >= 0)
        Qo, Ro = O.qr_compact(host[1])
        assert rel(Qs[1].cpu().numpy(), Qo) < 1e-12 and rel(Rs[1].cpu().numpy(), Ro) < 1e-12
    plan.close()


@pytest.mark.parametrize("dt", [torch.float64, torch.complex128], ids=["f64", "c64"])
def test_batched_full_size_properties(dt, dev):
    """4096 blocks, m log-uniform in [16, 1024], n uniform in [8, m]."""
    shapes = mak.block_shapes(4096, seed=0)
    g = torch.Generator(device="cpu").manual_seed(1)
    As = []
    for m, n in shapes:
        x = torch.randn((n, m), generator=g, dtype=torch.float64)
        if dt == torch.complex128:
            x = torch.complex(x, torch.randn((n, m), generator=g, dtype=torch.float64))
        As.append(x.to(dev).t())
    A0 = [A.clone() for A in As]
    plan = mak.BatchedQRPlan(As)
    Qs, Rs = plan.execute()
    torch.cuda.synchronize()
    worst_res = worst_orth = 0.0
    for A, Q, R in zip(A0, Qs, Rs):
        k = Q.shape[1]
        worst_res = max(worst_res, (torch.linalg.norm(A - Q @ R) / torch.linalg.norm(A)).item())
        worst_orth = max(worst_orth, (torch.linalg.norm(Q.mH @ Q - torch.eye(k, dtype=dt, device=dev))
                                      / k ** 0.5).item())
        d = torch.diagonal(R)
        if d.is_complex():
            assert (d.imag == 0).all()
            d = d.real
        assert (d >= 0).all()
    print(f"4096 blocks {dt}: worst residual {worst_res:.2e}, worst orthogonality {worst_orth:.2e}, "
          f"{plan.info()}", flush=True)
    assert worst_res < 1e-14 and worst_orth < 1e-14
    rng = np.random.default_rng(2)
    for i in rng.choice(len(shapes), 12, replace=False):
        Qo, Ro = O.qr_compact(A0[i].cpu().numpy())
        assert rel(Qs[i].cpu().numpy(), Qo) < 1e-12 and rel(Rs[i].cpu().numpy(), Ro) < 1e-12, shapes[i]
    plan.close()
