"""Parity of the HIP qr_compact! / qr_full! path with the CPU oracle.

Tolerances: with positive = true the QR of a full-rank matrix is unique, so Q
and R are compared elementwise with the oracle: relative Frobenius error
<= 1e-12 for the Gaussian inputs used here (condition numbers <= ~10^2).
The north-star bound (errors within 10x LAPACK's) is checked at full size on
||A - QR|| / ||A|| and ||Q^H Q - I||.
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


def run_qr_rec(A, dev, full=False, positive=True, wy=False):
    """qr through the C ABI with a workspace kept for the status record
    (maktest_qr_record): returns Q, R and the record
    [whole-matrix CholeskyQR2 taken, first rejected panel column, its
    Cholesky info, whole-matrix path available, ||E||_F (two ints of a
    double), FP32 correction taken, FP64 final product taken, BF16
    correction taken]."""
    import ctypes as C
    from mak_mi355x import _lib
    At = to_dev(A, dev)
    m, n = At.shape
    Q, R = mak.initialize_output_qr(At, full)
    lib = _lib.load()
    lib.maktest_qr_record.restype = C.c_int
    lib.maktest_qr_record.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    dt = 1 if At.is_complex() else 0
    nb = C.c_size_t(0)
    assert lib.mak_qr_workspace_size(dt, int(full), m, n, C.byref(nb)) == 0
    ws = torch.empty(max(nb.value, 1), dtype=torch.uint8, device=dev)
    sfx = "z" if dt else "d"
    args = [mak.api._stream(dev), m, n, mak.api._ptr(At), mak.api._ld(At), mak.api._ptr(Q), mak.api._ld(Q),
            mak.api._ptr(R), mak.api._ld(R)]
    if wy:
        T = torch.zeros(max(1, -(-min(m, n) // 32)) * 32 * 32, dtype=At.dtype, device=dev)
        code = getattr(lib, f"mak_qr_compact_wy_{sfx}")(*args, mak.api._ptr(T), int(positive), mak.api._ptr(ws),
                                                         nb.value)
    else:
        code = getattr(lib, f"mak_qr_{'full' if full else 'compact'}_{sfx}")(*args, int(positive),
                                                                            mak.api._ptr(ws), nb.value)
    torch.cuda.synchronize()
    assert code == 0
    rec = (C.c_int * 9)()
    assert lib.maktest_qr_record(dt, int(full), m, n, C.c_void_p(ws.data_ptr()), rec) == 0
    return Q.cpu().numpy(), R.cpu().numpy(), list(rec)


def rec_norm(rec):
    """||E||_F of the whole-matrix CholeskyQR2 (record entries 4, 5)."""
    return float(np.array(rec[4:6], dtype=np.int32).view(np.float64)[0])


def run_qr(A, dev, full=False, positive=True, with_r=True):
    At = to_dev(A, dev)
    Q, R = mak.initialize_output_qr(At, full)
    if not with_r:
        R = torch.empty((0, 0), dtype=At.dtype, device=dev)
    alg = mak.MI355XHouseholder(positive=positive)
    f = mak.qr_full_ if full else mak.qr_compact_
    Q2, R2 = f(At, (Q, R), alg)
    assert Q2 is Q and R2 is R
    torch.cuda.synchronize()
    return Q.cpu().numpy(), R.cpu().numpy()


SIZES = [(54, 37), (54, 54), (54, 63), (1, 1), (5, 1), (1, 5), (4, 0), (0, 5), (33, 33), (100, 32), (64, 100),
         (1024, 512), (1025, 64), (2100, 64), (3000, 200), (200, 300), (700, 700)]


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", SIZES, ids=[f"{m}x{n}" for m, n in SIZES])
def test_qr_compact_matches_oracle(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=1000 + m + n)
    Q, R = run_qr(A, dev)
    Qo, Ro = O.qr_compact(A)
    assert rel(Q, Qo) < 1e-12, rel(Q, Qo)
    assert rel(R, Ro) < 1e-12, rel(R, Ro)


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(54, 37), (54, 63), (300, 70), (1500, 40)])
def test_qr_full_matches_oracle(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=2000 + m + n)
    Q, R = run_qr(A, dev, full=True)
    Qo, Ro = O.qr_full(A)
    k = min(m, n)
    assert rel(Q[:, :k], Qo[:, :k]) < 1e-12
    assert rel(R, Ro) < 1e-12
    # the trailing columns of a full Q are only fixed up to a unitary rotation:
    # check that Q is unitary and spans the complement
    assert rel(Q.conj().T @ Q, np.eye(m)) < 1e-12


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(54, 37), (2100, 64), (40, 90)])
def test_qr_positive_false_and_no_R(m, n, dt, dev):
    A = O.gaussian(m, n, dt, seed=3000 + m + n)
    Q, R = run_qr(A, dev, positive=False)
    assert rel(Q @ R, A) < 1e-13
    assert O.isisometric(Q)
    assert np.allclose(np.tril(R, -1), 0)
    Qp, Rp = run_qr(A, dev, with_r=False)
    Qo, _ = O.qr_compact(A)
    assert rel(Qp, Qo) < 1e-12


def test_qr_tree_three_levels(dev):
    # real nodes hold 1024 rows and fan in 32: > 32768 rows needs 3 tree levels
    A = O.gaussian(40000, 40, np.float64, seed=5)
    Q, R = run_qr(A, dev)
    Qo, Ro = O.qr_compact(A)
    assert rel(Q, Qo) < 1e-12 and rel(R, Ro) < 1e-12


def test_qr_complex_tree_three_levels(dev):
    # complex nodes hold 256 rows and fan in 8: > 2048 rows needs 3 levels
    A = O.gaussian(2600, 40, np.complex128, seed=6)
    Q, R = run_qr(A, dev)
    Qo, Ro = O.qr_compact(A)
    assert rel(Q, Qo) < 1e-12 and rel(R, Ro) < 1e-12


@pytest.mark.parametrize("key", [k[:-2] for k in GOLD.files if k.startswith("qr_") and k.endswith("_A")])
def test_qr_golden_fixtures(key, dev):
    A = GOLD[key + "_A"]
    Q, R = run_qr(A, dev)
    assert rel(Q, GOLD[key + "_Q"]) < 1e-12 and rel(R, GOLD[key + "_R"]) < 1e-12
    Qf, Rf = run_qr(A, dev, full=True)
    k = min(A.shape)
    assert rel(Qf[:, :k], GOLD[key + "_Qfull"][:, :k]) < 1e-12
    assert rel(Rf, GOLD[key + "_Rfull"]) < 1e-12


def test_qr_rank_deficient_and_special(dev):
    # exact zero columns, repeated columns and a diagonal input
    A = O.gaussian(300, 40, np.float64, seed=9)
    A[:, 5] = 0
    A[:, 7] = A[:, 3]
    Q, R = run_qr(A, dev)
    assert rel(Q @ R, A) < 1e-13 and O.isisometric(Q)
    assert np.all(np.diagonal(R) >= 0)
    d = np.array([3.0, -2.0, 0.5, -0.25])
    Q, R = run_qr(np.diag(d), dev)
    assert np.allclose(np.diagonal(R), np.abs(d)) and np.allclose(np.diagonal(Q), np.sign(d))


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("kind", ["repeated_column", "graded_1e10", "mixed_1e12"])
def test_qr_cholqr_rejected_later_panel(kind, dt, dev):
    """Panels 0-3 are Gaussian; panel 4 (columns 128..159) is made hard for
    CholeskyQR2.  mixed_1e12: its columns span singular values 1 .. 1e-12 in
    mixed directions, so the first Gram is not numerically positive definite
    or the second fails the check, and the driver must refactor from column
    128 with Householder TSQR.  graded_1e10 (column scales 1 .. 1e-10) and an
    exactly repeated column may legitimately pass: diagonal scaling does not
    hurt the Cholesky factor, and a repeated column gives a tiny pivot whose
    Q1 column is normalised rounding noise, orthogonal to the others.  Q is
    not unique for a rank-deficient A, so the checks are backward error,
    orthogonality and the gauge of R."""
    m, n = 2000, 200
    A = O.gaussian(m, n, dt, seed=21)
    if kind == "repeated_column":
        A[:, 150] = A[:, 10]
    elif kind == "graded_1e10":
        A = A * (10.0 ** (-10.0 * np.clip(np.arange(n) - 128, 0, 31) / 31))[None, :]
    else:
        U, _ = np.linalg.qr(O.gaussian(m, 32, dt, seed=22))
        V, _ = np.linalg.qr(O.gaussian(32, 32, dt, seed=23))
        A[:, 128:160] = (U * np.logspace(0, -12, 32)[None, :]) @ V.conj().T
    # the rule's primal (compact WY tape) always takes the panel path
    Q, R, rec = run_qr_rec(A, dev, wy=True)
    rejected = rec[1]
    assert rel(Q @ R, A) < 1e-13
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(n)) < 1e-12
    d = np.diagonal(R)
    assert np.all(d.imag == 0) and np.all(d.real >= 0)
    if kind == "mixed_1e12":
        assert rejected == 128
    else:
        assert rejected in (-1, 128)
    # qr_compact!: the whole-matrix CholeskyQR2 check rejects the mixed case
    # (kappa ~ 1e12) and the panel path takes over; either way the same checks
    Q, R, rec = run_qr_rec(A, dev)
    assert rec[3] == 1
    if kind == "mixed_1e12":
        assert rec[0] == 0 and rec[1] == 128
    assert rel(Q @ R, A) < 1e-13
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(n)) < 1e-12
    d = np.diagonal(R)
    assert np.all(d.imag == 0) and np.all(d.real >= 0)


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(3000, 400), (700, 650), (1200, 1200)])
def test_qr_cholqr_carried_gram_accepts_gaussian_panels(m, n, dt, dev):
    """Every panel after the first takes its first Gram from the previous
    panel's pass (Gw - Rt^H Rt); for Gaussian inputs all of them must pass the
    CholeskyQR2 check (none handed to the Householder path), and the factors
    must match the oracle's unique positive-gauge QR."""
    A = O.gaussian(m, n, dt, seed=m + n)
    Q, R, rec = run_qr_rec(A, dev, wy=True)     # the panel path (compact WY tape)
    assert rec[1] == -1
    Qo, Ro = O.qr_compact(A, positive=True)
    assert rel(Q, Qo) < 1e-12 and rel(R, Ro) < 1e-12


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(64, 64), (100, 64), (1025, 65), (3000, 400), (700, 650), (1200, 1200),
                                 (5000, 1000), (2100, 2048)])
def test_qr_whole_matrix_cholqr2_matches_oracle(m, n, dt, dev):
    """qr_compact!(positive = true) with 64 <= n <= 2048 <= m: the whole-matrix
    CholeskyQR2 path is taken for Gaussian inputs (record[0] == 1) and gives
    the unique positive-gauge QR of the oracle."""
    A = O.gaussian(m, n, dt, seed=7 * m + n)
    Q, R, rec = run_qr_rec(A, dev)
    assert rec[3] == 1 and rec[0] == 1 and rec[2] == 0, rec
    # one final product; Gaussian inputs have ||Q1^H Q1 - I|| small enough
    # for the BF16 (<= 5e-14) or FP32 (<= 1e-10) correction (complex data
    # through its real embedding)
    assert rec[6] + rec[7] + rec[8] == 1
    e = rec_norm(rec)
    assert (rec[8] == 1) == (e <= 5e-14), e
    assert (rec[6] == 1) == (5e-14 < e <= 1e-10), e
    Qo, Ro = O.qr_compact(A, positive=True)
    assert rel(Q, Qo) < 1e-12 and rel(R, Ro) < 1e-12
    assert np.all(np.tril(R, -1) == 0)
    d = np.diagonal(R)
    assert np.all(d.imag == 0) and np.all(d.real > 0)


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("cond", [1e3, 1e5, 1e8, 1e13])
def test_qr_whole_matrix_cholqr2_conditioning(cond, dt, dev):
    """Graded singular values 1 .. 1/cond in random directions: the closed-form
    second pass is accepted while ||Q1^H Q1 - I|| <= 1e-6 (kappa up to about
    1e5) and the input goes on to the Householder path beyond; either way
    the backward error and orthogonality are those of Householder QR.  The
    final product's tier (BF16 / FP32 correction, or FP64) follows ||E||_F
    for real and complex data alike."""
    m, n = 3000, 256
    U, _ = np.linalg.qr(O.gaussian(m, n, dt, seed=31))
    V, _ = np.linalg.qr(O.gaussian(n, n, dt, seed=32))
    A = (U * np.logspace(0, -np.log10(cond), n)[None, :]) @ V.conj().T
    Q, R, rec = run_qr_rec(A, dev)
    if cond <= 1e3:
        assert rec[0] == 1
    if cond >= 1e8:
        assert rec[0] == 0
    # an accepted call takes exactly one final product, the BF16 correction
    # for ||E||_F <= 5e-14, FP32 up to 1e-10 (cond 1e5 here: ||E|| ~ 1e-7, the
    # FP64 product)
    assert rec[6] + rec[7] + rec[8] == rec[0]
    if rec[0]:
        e = rec_norm(rec)
        assert (rec[8] == 1) == (e <= 5e-14), e
        assert (rec[6] == 1) == (5e-14 < e <= 1e-10), e
    if cond == 1e5:
        # accepted for real data (complex data sits at the 1e-6 threshold)
        assert rec[7] == 1 if dt == np.float64 else rec[7] == rec[0]
    assert rel(Q @ R, A) < 1e-14
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(n)) < 1e-13
    d = np.diagonal(R)
    assert np.all(d.imag == 0) and np.all(d.real >= 0)
    if cond <= 1e5:
        Qo, Ro = O.qr_compact(A, positive=True)
        assert rel(R, Ro) < 1e-15 * cond * 100


def test_qr_input_untouched_by_allocating_form(dev):
    A = to_dev(O.gaussian(100, 60, np.float64, seed=3), dev)
    A0 = A.clone()
    Q, R = mak.qr_compact(A)
    torch.cuda.synchronize()
    assert torch.equal(A, A0)
    assert torch.allclose(Q @ R, A0, atol=1e-12)


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_qr_full_size_within_10x_lapack(dt, dev):
    m, n = 8192, 1024
    A = O.gaussian(m, n, dt, seed=42)
    Q, R = run_qr(A, dev)
    Qo, Ro = O.qr_compact(A)
    nA = np.linalg.norm(A)
    res, res_o = np.linalg.norm(A - Q @ R) / nA, np.linalg.norm(A - Qo @ Ro) / nA
    orth = np.linalg.norm(Q.conj().T @ Q - np.eye(n))
    orth_o = np.linalg.norm(Qo.conj().T @ Qo - np.eye(n))
    print(f"{dt.__name__} 8192x1024: |A-QR|/|A| {res:.2e} (LAPACK {res_o:.2e}), "
          f"|Q^HQ-I| {orth:.2e} (LAPACK {orth_o:.2e}), |R-Rref|/|Rref| {rel(R, Ro):.2e}", flush=True)
    assert res <= 10 * res_o and orth <= 10 * orth_o
    assert rel(R, Ro) < 1e-12 and rel(Q, Qo) < 1e-12


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(200, 80), (90, 90), (3000, 300)])
def test_qr_inplace_q(m, n, dt, dev):
    """qr_compact!(A, (A, empty R); positive = false) writes Q over A
    (src/implementations/qr.jl:150-153, 170-171): the same Q as the
    out-of-place call, and the positive gauge or an R in place throw."""
    A = O.gaussian(m, n, dt, seed=m * n)
    alg = mak.MI355XHouseholder(positive=False)
    At = to_dev(A, dev)
    E = torch.empty((0, 0), dtype=At.dtype, device=dev)
    Q2, R2 = mak.qr_compact_(At, (At, E), alg)
    assert Q2 is At and R2 is E
    Ab = to_dev(A, dev)
    Qref, _ = mak.qr_compact_(Ab, (mak.initialize_output_qr(Ab, False)[0], E), alg)
    torch.cuda.synchronize()
    assert torch.equal(At, Qref)
    assert O.isisometric(At.cpu().numpy())
    with pytest.raises(mak.ArgumentError, match="inplace"):
        mak.qr_compact_(At, (At, E), mak.MI355XHouseholder(positive=True))
    with pytest.raises(mak.ArgumentError, match="inplace"):
        R = mak.initialize_output_qr(At, False)[1]
        mak.qr_compact_(At, (At, R), alg)


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
@pytest.mark.parametrize("m,n", [(4000, 200), (3000, 290)])
def test_qr_full_paired_block_reflectors(m, n, dt, dev):
    """qr_full! on shapes whose Q formation merges pairs of 128-column block
    reflectors into one (m * m >= 2^23): 200 columns pair a full group with
    a 72-column one, 290 columns leave the third group (34 columns) unpaired.
    The first n columns of Q and R equal the oracle's compact factors, Q is
    unitary and A = Q R."""
    A = O.gaussian(m, n, dt, seed=5000 + m + n)
    At = to_dev(A, dev)
    Q, R = mak.initialize_output_qr(At, True)
    mak.qr_full_(At.clone(), (Q, R))
    torch.cuda.synchronize()
    Qo, Ro = O.qr_compact(A)
    Rn = R.cpu().numpy()
    assert np.all(Rn[n:, :] == 0)
    assert rel(Rn[:n, :], Ro) < 1e-12 and rel(Q[:, :n].cpu().numpy(), Qo) < 1e-12
    eye = torch.eye(m, dtype=At.dtype, device=dev)
    assert (torch.linalg.norm(Q.mH @ Q - eye) / np.sqrt(m)).item() < 1e-13
    assert (torch.linalg.norm(At - Q @ R) / torch.linalg.norm(At)).item() < 1e-13


@pytest.mark.parametrize("dt", [np.float64, np.complex128], ids=["f64", "c64"])
def test_qr_full_8192x1024(dt, dev):
    """qr_full! on the 8192 x 1024 configuration (Q 8192 x 8192): R and the
    first 1024 columns of Q equal the oracle's unique positive-gauge compact
    factors (<= 1e-12; the remaining columns are fixed by the Householder
    vectors, src/implementations/qr.jl:106-109, and checked through
    unitarity), Q is unitary and the residual / orthogonality are within 10x
    LAPACK's (the oracle's compact factors).  The products of the checks run
    in torch on the device."""
    m, n = 8192, 1024
    A = O.gaussian(m, n, dt, seed=4242)
    At = to_dev(A, dev)
    Q, R = mak.initialize_output_qr(At, True)
    mak.qr_full_(At.clone(), (Q, R))
    torch.cuda.synchronize()
    Qo, Ro = O.qr_compact(A)
    Q1 = Q[:, :n].cpu().numpy()
    Rn = R.cpu().numpy()
    assert np.all(Rn[n:, :] == 0)
    assert rel(Rn[:n, :], Ro) < 1e-12 and rel(Q1, Qo) < 1e-12
    nA = torch.linalg.norm(At)
    res = (torch.linalg.norm(At - Q @ R) / nA).item()
    eye = torch.eye(m, dtype=At.dtype, device=dev)
    orth = torch.linalg.norm(Q.mH @ Q - eye).item()
    del eye
    res_o = np.linalg.norm(A - Qo @ Ro) / np.linalg.norm(A)
    orth_o = np.linalg.norm(Qo.conj().T @ Qo - np.eye(n))
    print(f"qr_full {dt.__name__} 8192x1024: |A-QR|/|A| {res:.2e} (LAPACK compact {res_o:.2e}), "
          f"|Q^HQ-I| {orth:.2e} over 8192 columns (LAPACK compact {orth_o:.2e} over 1024)", flush=True)
    assert res <= 10 * res_o
    # the full Q has 8x the columns of the compact one: the Frobenius bound scales with sqrt(8)
    assert orth <= 10 * orth_o * np.sqrt(m / n)
