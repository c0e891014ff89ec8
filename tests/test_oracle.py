"""Pin the CPU oracle (oracle/mak_oracle.py) against things other than itself.

Per function, what pins it:
  qr_compact / qr_full   - the reference's Native Householder path restated
                           independently (qr_native), R vs the Cholesky factor
                           of A^H A (closed form), Q vs numpy.linalg.qr (a second
                           LAPACK route: geqrf + orgqr) after the sign gauge,
                           diagonal and unitary special cases, golden fixtures
  qr_null                - N^H A = 0 and [Q N] unitary
  householder_native     - H x = ||x|| e_i, H unitary
  left/right_polar_svd   - scipy.linalg.polar (second route), P = sqrtm(A^H A),
                           W^H W = I, P Hermitian positive definite
  left/right_polar_newton- converges to the SVD polar factor
  qr_pullback            - central finite differences (h = 1e-6) of a scalar
                           loss through qr_compact
  left/right_polar_pullback - central finite differences through the polar
  remove_qr_gauge_dependence - projector property (idempotent)
Each pin is also shown to catch a planted mistake (a transposed operand, a
dropped term, a wrong sign) in a copy of the oracle function.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import mak_oracle as O

GOLD = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "golden_small.npz")

SIZES = [(54, 37), (54, 54), (54, 63), (7, 3), (3, 7), (1, 1), (40, 1)]
DTYPES = [np.float64, np.complex128]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", SIZES)
def test_qr_compact_vs_native_householder(m, n, dt):
    A = O.gaussian(m, n, dt, seed=m * 100 + n)
    Q, R = O.qr_compact(A)
    Qn, Rn = O.qr_native(A)
    assert rel(Q, Qn) < 1e-13 and rel(R, Rn) < 1e-13
    assert rel(Q @ R, A) < 1e-14
    assert O.isisometric(Q)
    assert np.all(np.real(np.diagonal(R)) >= 0)
    assert np.allclose(np.imag(np.diagonal(R)), 0)
    assert np.allclose(np.tril(R, -1), 0)


@pytest.mark.parametrize("dt", DTYPES)
def test_qr_R_is_cholesky_factor(dt):
    A = O.gaussian(80, 30, dt, seed=5)
    _, R = O.qr_compact(A)
    L = np.linalg.cholesky(A.conj().T @ A)      # closed form: R = L^H
    assert rel(R, L.conj().T) < 1e-12


@pytest.mark.parametrize("dt", DTYPES)
def test_qr_vs_numpy_geqrf_route(dt):
    A = O.gaussian(70, 45, dt, seed=6)
    Qn, Rn = np.linalg.qr(A)
    s = O.sign_safe(np.diagonal(Rn))
    Qn = Qn * s[None, :]
    Rn = np.conj(s)[:, None] * Rn
    Q, R = O.qr_compact(A)
    assert rel(Q, Qn) < 1e-12 and rel(R, Rn) < 1e-12


def test_qr_diagonal_closed_form():
    d = np.array([3.0, -2.0, 0.0, -0.5])
    A = np.diag(d)
    Q, R = O.qr_compact(A)
    assert np.allclose(np.abs(np.diagonal(R)), np.abs(d))
    assert np.allclose(Q @ R, A)
    # src/implementations/qr.jl:318-333: R = |A|, Q = sign_safe(A) on the diagonal
    assert np.allclose(np.diagonal(Q)[[0, 1, 3]], O.sign_safe(d)[[0, 1, 3]])


def test_qr_unitary_input_is_its_own_Q():
    U = O.haar_orthogonal(30, 3)
    Q, R = O.qr_compact(U)
    assert rel(Q, U) < 1e-13 and rel(R, np.eye(30)) < 1e-13


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", [(54, 37), (54, 63), (10, 10)])
def test_qr_full_and_null(m, n, dt):
    A = O.gaussian(m, n, dt, seed=7)
    Qf, Rf = O.qr_full(A)
    Q, R = O.qr_compact(A)
    k = min(m, n)
    assert Qf.shape == (m, m) and Rf.shape == (m, n)
    assert rel(Qf[:, :k], Q) < 1e-13 and rel(Rf[:k], R) < 1e-13
    assert np.allclose(Rf[k:], 0)
    assert rel(Qf.conj().T @ Qf, np.eye(m)) < 1e-13
    N = O.qr_null(A)
    assert N.shape == (m, m - k)
    if m > k:
        assert np.linalg.norm(N.conj().T @ A) < 1e-12 * np.linalg.norm(A)
        assert rel(np.hstack([Q, N]).conj().T @ np.hstack([Q, N]), np.eye(m)) < 1e-13


@pytest.mark.parametrize("dt", DTYPES)
def test_householder_native(dt):
    x = O.gaussian(9, 1, dt, seed=1)[:, 0]
    beta, v, nu = O.householder_native(x, 0)
    H = np.eye(9) - beta * np.outer(v, v.conj())
    y = H @ x
    assert abs(y[0] - nu) < 1e-13 and np.linalg.norm(y[1:]) < 1e-13
    assert rel(H.conj().T @ H, np.eye(9)) < 1e-13


@pytest.mark.parametrize("key", [k[:-2] for k in GOLD.files if k.startswith("qr_") and k.endswith("_A")])
def test_golden_qr_fixtures(key):
    A = GOLD[key + "_A"]
    Q, R = O.qr_compact(A)
    assert rel(Q, GOLD[key + "_Q"]) < 1e-13 and rel(R, GOLD[key + "_R"]) < 1e-13
    Qf, Rf = O.qr_full(A)
    assert rel(Qf, GOLD[key + "_Qfull"]) < 1e-13 and rel(Rf, GOLD[key + "_Rfull"]) < 1e-13


# ------------------------------------------------------------------- polar
@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", [(54, 37), (54, 54), (20, 1)])
def test_left_polar_svd_vs_scipy_and_sqrtm(m, n, dt):
    A = O.gaussian(m, n, dt, seed=8)
    W, P = O.left_polar_svd(A)
    W2, P2 = sla.polar(A, side="right")
    assert rel(W, W2) < 1e-12 and rel(P, P2) < 1e-12
    assert rel(P, sla.sqrtm(A.conj().T @ A)) < 1e-10
    assert rel(W @ P, A) < 1e-13
    assert O.isisometric(W)
    assert np.allclose(P, P.conj().T) and np.min(np.linalg.eigvalsh(P)) > 0


@pytest.mark.parametrize("dt", DTYPES)
def test_right_polar_svd_vs_scipy(dt):
    A = O.gaussian(37, 54, dt, seed=9)
    P, Wh = O.right_polar_svd(A)
    W2, P2 = sla.polar(A, side="left")
    assert rel(Wh, W2) < 1e-12 and rel(P, P2) < 1e-12
    assert rel(P @ Wh, A) < 1e-13


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", [(54, 37), (54, 54)])
def test_polar_newton_matches_svd(m, n, dt):
    A = O.gaussian(m, n, dt, seed=10)
    W, P = O.left_polar_newton(A)
    Ws, Ps = O.left_polar_svd(A)
    assert rel(W, Ws) < 1e-10 and rel(P, Ps) < 1e-10
    P2, Wh = O.right_polar_newton(A.conj().T.copy())
    assert rel(Wh, Ws.conj().T) < 1e-10


def _newton_orthodiag_conv(Q, d, nsteps):
    """Convergence measures of the scaled Newton iteration on A = Q diag(d)
    (Q orthogonal): every iterate is Q diag(d_k), A^{-H} = Q diag(1/d), the
    Frobenius scaling is gamma = sqrt(||1/d|| / ||d||), so the step is the
    scalar recursion d <- (gamma d + 1/(gamma d)) / 2 with the difference
    Q diag(1/(gamma d) - gamma d).  Returns, per step, the max-abs entry and
    the row-sum norm of that difference."""
    out = []
    for _ in range(nsteps):
        g = np.sqrt(np.linalg.norm(1.0 / d) / np.linalg.norm(d))
        delta = 1.0 / (g * d) - g * d
        D = Q * delta[None, :]
        out.append((np.max(np.abs(D)), np.max(np.sum(np.abs(D), axis=1))))
        d = (g * d + 1.0 / (g * d)) / 2.0
    return out


def test_polar_newton_stops_on_max_abs_entry():
    """The stopping test is Julia's norm(R^{-H} diff, Inf) = max-abs entry
    (src/implementations/polar.jl:148,156), not the row-sum norm: with tol
    between the two measures the iteration must stop on the max-abs one.
    The expected counts come from the scalar recursion on Q diag(d)."""
    n = 40
    Q = O.haar_orthogonal(n, seed=3)
    d = np.logspace(0, -3, n)
    convs = _newton_orthodiag_conv(Q, d, 8)
    # an iteration where the two measures straddle tol
    k = next(i for i, (a, r) in enumerate(convs) if r > 1.5 * a and a < 1e-2)
    tol = np.sqrt(convs[k][0] * convs[k][1])
    expect_maxabs = next(i for i, (a, _) in enumerate(convs) if a <= tol) + 1
    expect_rowsum = next(i for i, (_, r) in enumerate(convs) if r <= tol) + 1
    assert expect_rowsum > expect_maxabs
    info = {}
    W, P = O.left_polar_newton(Q * d[None, :], tol=tol, maxiter=20, info=info)
    assert info["iterations"] == expect_maxabs
    assert abs(info["conv"] - convs[expect_maxabs - 1][0]) <= 1e-6 * convs[expect_maxabs - 1][0] + 1e-15
    # a planted row-sum norm would run longer
    assert info["iterations"] != expect_rowsum


def test_polar_newton_warns_on_maxiter():
    A = O.gaussian(30, 30, np.float64, seed=4)
    with pytest.warns(UserWarning, match="did not converge"):
        O.left_polar_newton(A, maxiter=1)


@pytest.mark.parametrize("key", [k[:-2] for k in GOLD.files if k.startswith("lpolar_") and k.endswith("_A")])
def test_golden_polar_fixtures(key):
    A = GOLD[key + "_A"]
    W, P = O.left_polar_svd(A)
    assert rel(W, GOLD[key + "_W"]) < 1e-12 and rel(P, GOLD[key + "_P"]) < 1e-12


# --------------------------------------------------------------- pullbacks
def _fd_check(f, pb, A, h=1e-6, seed=0):
    """<pullback(cotangent), dA> vs the central difference of <cotangent, f(A)>."""
    rng = np.random.default_rng(seed)
    D = rng.standard_normal(A.shape)
    if np.iscomplexobj(A):
        D = D + 1j * rng.standard_normal(A.shape)
    lhs = np.real(np.vdot(pb, D))
    fp, fm = f(A + h * D), f(A - h * D)
    rhs = (fp - fm) / (2 * h)
    return lhs, rhs


def qr_loss_factory(dQ, dR):
    def loss(A):
        Q, R = O.qr_compact(A)
        return np.real(np.vdot(dQ, Q) + np.vdot(dR, R))
    return loss


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", [(19, 17), (19, 19), (17, 19)])
def test_qr_pullback_finite_differences(m, n, dt):
    A = O.gaussian(m, n, dt, seed=20)
    Q, R = O.qr_compact(A)
    k = min(m, n)
    dQ = O.gaussian(m, k, dt, seed=21)
    dR = np.triu(O.gaussian(k, n, dt, seed=22))
    dQ, dR = O.remove_qr_gauge_dependence(dQ, dR, A, Q, R)
    dA = O.qr_pullback(np.zeros_like(A), A, (Q, R), (dQ, dR))
    lhs, rhs = _fd_check(qr_loss_factory(dQ, dR), dA, A)
    assert abs(lhs - rhs) < 1e-6 * max(1.0, abs(rhs))


def test_qr_pullback_planted_mistake_is_caught():
    """Transposing M's copy (copyltu) in a copy of the oracle breaks the FD check."""
    A = O.gaussian(19, 17, np.float64, seed=20)
    Q, R = O.qr_compact(A)
    dQ = O.gaussian(19, 17, np.float64, seed=21)
    dR = np.triu(O.gaussian(17, 17, np.float64, seed=22))
    p = 17
    M = dR @ R.T - Q.T @ dQ
    M_bad = np.tril(M) + np.tril(M, -1).T            # copies lower -> upper (wrong side)
    dQ1 = dQ + Q @ M_bad
    dA_bad = sla.solve_triangular(R, dQ1.T, lower=False).T
    lhs, rhs = _fd_check(qr_loss_factory(dQ, dR), dA_bad, A)
    assert abs(lhs - rhs) > 1e-3 * max(1.0, abs(rhs))
    del p


def test_remove_qr_gauge_dependence_is_idempotent():
    A = O.gaussian(19, 17, np.float64, seed=3)
    Q, R = O.qr_full(A)
    dQ = O.gaussian(19, 19, np.float64, seed=4)
    dR = O.gaussian(19, 17, np.float64, seed=5)
    a = O.remove_qr_gauge_dependence(dQ, dR, A, Q, R)
    b = O.remove_qr_gauge_dependence(*a, A, Q, R)
    assert np.allclose(a[0], b[0]) and np.allclose(a[1], b[1])


def polar_loss_factory(dW, dP):
    def loss(A):
        W, P = O.left_polar_svd(A)
        return np.real(np.vdot(dW, W) + np.vdot(dP, P))
    return loss


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", [(19, 17), (19, 19)])
def test_left_polar_pullback_finite_differences(m, n, dt):
    A = O.gaussian(m, n, dt, seed=30)
    W, P = O.left_polar_svd(A)
    dW = O.gaussian(m, n, dt, seed=31)
    dP = O.project_hermitian(O.gaussian(n, n, dt, seed=32))
    dA = O.left_polar_pullback(np.zeros_like(A), A, (W, P), (dW, dP))
    lhs, rhs = _fd_check(polar_loss_factory(dW, dP), dA, A)
    assert abs(lhs - rhs) < 1e-6 * max(1.0, abs(rhs))
    # dropping the (I - W W^H) dW P^{-1} term (a plausible mistake) is caught
    Pinv = np.linalg.inv(P)
    bad = dA - (dW - W @ (W.conj().T @ dW)) @ Pinv
    if m > n:
        lhs_b, _ = _fd_check(polar_loss_factory(dW, dP), bad, A)
        assert abs(lhs_b - rhs) > 1e-4 * max(1.0, abs(rhs))


@pytest.mark.parametrize("dt", DTYPES)
def test_right_polar_pullback_finite_differences(dt):
    A = O.gaussian(17, 19, dt, seed=40)
    P, Wh = O.right_polar_svd(A)
    dWh = O.gaussian(17, 19, dt, seed=41)
    dP = O.project_hermitian(O.gaussian(17, 17, dt, seed=42))

    def loss(X):
        P_, Wh_ = O.right_polar_svd(X)
        return np.real(np.vdot(dP, P_) + np.vdot(dWh, Wh_))
    dA = O.right_polar_pullback(np.zeros_like(A), A, (P, Wh), (dP, dWh))
    lhs, rhs = _fd_check(loss, dA, A)
    assert abs(lhs - rhs) < 1e-6 * max(1.0, abs(rhs))


@pytest.mark.parametrize("key", [k[:-2] for k in GOLD.files if k.startswith("qrpb_") and k.endswith("_A")])
def test_golden_qr_pullback_fixtures(key):
    A = GOLD[key + "_A"]
    Q, R = O.qr_compact(A)
    dA = O.qr_pullback(np.zeros_like(A), A, (Q, R), (GOLD[key + "_dQ"], GOLD[key + "_dR"]))
    assert rel(dA, GOLD[key + "_dA"]) < 1e-12


def test_flop_counts():
    assert abs(O.qr_flops(8192, 1024) - (4 * 8192 * 1024 ** 2 - 4 * 1024 ** 3 / 3)) < 1
    assert abs(O.qr_flops(8192, 1024, kind="full")
               - (4 * 8192 ** 2 * 1024 - 2 * 8192 * 1024 ** 2 + 2 * 1024 ** 3 / 3)) < 1


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", [(20, 45), (45, 20), (30, 30)])
def test_lq_via_qr_pinned_by_numpy_and_invariants(m, n, dt):
    """LQ via the transposed QR: A = L Q, L lower triangular with a real
    positive diagonal, Q with orthonormal rows; unique for full-rank A, so it
    equals numpy's LQ (QR of A^H, signs moved to the diagonal)."""
    A = O.gaussian(m, n, dt, seed=31 + m)
    L, Q = O.lq_via_qr(A)
    k = min(m, n)
    assert L.shape == (m, k) and Q.shape == (k, n)
    assert np.allclose(np.triu(L, 1), 0)
    d = np.diag(L)
    assert np.all(np.abs(d.imag) < 1e-14) and np.all(d.real > 0)
    assert rel(L @ Q, A) < 1e-13
    assert rel(Q @ Q.conj().T, np.eye(k)) < 1e-13
    q, r = np.linalg.qr(A.conj().T)
    s = np.diag(r) / np.abs(np.diag(r))
    Ln, Qn = (r * np.conj(s)[:, None]).conj().T, (q * s[None, :]).conj().T
    assert rel(L, Ln) < 1e-12 and rel(Q, Qn) < 1e-12
    Lf, Qf = O.lq_via_qr(A, full=True)
    assert Lf.shape == (m, n) and Qf.shape == (n, n)
    assert rel(Qf @ Qf.conj().T, np.eye(n)) < 1e-13 and rel(Lf @ Qf, A) < 1e-13
    assert rel(Qf[:k], Q) < 1e-12


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", [(40, 12), (12, 40), (20, 20)])
def test_null_spaces_pinned_by_svd(m, n, dt):
    """qr_null / lq_null span the orthogonal complements that the SVD's
    trailing singular vectors span (full-rank A)."""
    A = O.gaussian(m, n, dt, seed=77 + n)
    N = O.qr_null(A)
    Nh = O.lq_null(A)
    k = min(m, n)
    assert N.shape == (m, m - k) and Nh.shape == (n - k, n)
    U, _, Vh = np.linalg.svd(A)
    if m > k:
        assert rel(N.conj().T @ N, np.eye(m - k)) < 1e-13
        assert np.linalg.norm(A.conj().T @ N) < 1e-12 * np.linalg.norm(A)
        Pu = U[:, k:] @ U[:, k:].conj().T
        assert rel(N @ N.conj().T, Pu) < 1e-12
    if n > k:
        assert rel(Nh @ Nh.conj().T, np.eye(n - k)) < 1e-13
        assert np.linalg.norm(A @ Nh.conj().T) < 1e-12 * np.linalg.norm(A)
        Pv = Vh[k:].conj().T @ Vh[k:]
        assert rel(Nh.conj().T @ Nh, Pv) < 1e-12


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", [(40, 25), (25, 40)])
def test_svd_compact_gauge(m, n, dt):
    """svd_compact with gaugefix!: A = U S V^H, orthonormal factors, and the
    first entry of largest magnitude of each U column real positive; equal to
    numpy's SVD after the same gauge (unique for distinct singular values)."""
    A = O.gaussian(m, n, dt, seed=12)
    U, S, Vh = O.svd_compact(A)
    assert rel((U * S[None, :]) @ Vh, A) < 1e-13
    for j in range(U.shape[1]):
        i = int(np.argmax(np.abs(U[:, j])))
        assert abs(U[i, j].imag) < 1e-14 and U[i, j].real > 0
    u, s, vh = np.linalg.svd(A, full_matrices=False)
    sg = np.array([u[int(np.argmax(np.abs(u[:, j]))), j] for j in range(u.shape[1])])
    sg = sg / np.abs(sg)
    assert rel(U, u * np.conj(sg)[None, :]) < 1e-12 and rel(Vh, sg[:, None] * vh) < 1e-12
    assert rel(S, s) < 1e-14


@pytest.mark.parametrize("dt", DTYPES)
def test_polar_pushforwards_match_finite_differences(dt):
    """The polar pushforwards equal central differences of the polar
    decomposition (h = 1e-6) along a random direction."""
    for m, n in [(30, 18), (18, 18)]:
        A = O.gaussian(m, n, dt, seed=40 + m)
        dA = O.gaussian(m, n, dt, seed=41 + m)
        W, P = O.left_polar_svd(A)
        dW, dP = O.left_polar_pushforward(dA, A, (W, P))
        h = 1e-6
        Wp, Pp = O.left_polar_svd(A + h * dA)
        Wm, Pm = O.left_polar_svd(A - h * dA)
        assert rel(dW, (Wp - Wm) / (2 * h)) < 1e-7 and rel(dP, (Pp - Pm) / (2 * h)) < 1e-7
        B, dB = A.conj().T.copy(), dA.conj().T.copy()
        Pr, Wh = O.right_polar_svd(B)
        dWh, dPr = O.right_polar_pushforward(dB, B, (Pr, Wh))
        Prp, Whp = O.right_polar_svd(B + h * dB)
        Prm, Whm = O.right_polar_svd(B - h * dB)
        assert rel(dWh, (Whp - Whm) / (2 * h)) < 1e-7 and rel(dPr, (Prp - Prm) / (2 * h)) < 1e-7


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n,r", [(19, 17, 11), (17, 19, 12)])
def test_qr_pullback_rank_deficient_finite_differences(m, n, r, dt):
    """Rank-deficient branch (p = rank < min(m, n)) of the QR pullback,
    pinned by central differences along the rank-preserving curve
    A(h) = (G + h E) B (I + h F) through A = G B (rank r), whose tangent is
    E B + A F.  The cotangents see only Q[:, :p] and R[:p, :] (gauge
    independent: zero on the remaining columns / rows); dR[:p, p:] is nonzero,
    so the dA[:, p:] += Q1 dR12 and dQ1 -= Q1 dR12 R12^H terms are exercised."""
    G = O.gaussian(m, r, dt, seed=60)
    B = O.gaussian(r, n, dt, seed=61)
    E = O.gaussian(m, r, dt, seed=62)
    F = O.gaussian(n, n, dt, seed=63)
    A = G @ B
    Q, R = O.qr_compact(A)
    k = min(m, n)
    p = O.qr_rank(R)
    assert p == r
    dQ = np.zeros((m, k), dtype=A.dtype)
    dQ[:, :p] = O.gaussian(m, p, dt, seed=64)
    dR = np.zeros((k, n), dtype=A.dtype)
    dR[:p, :] = np.triu(O.gaussian(p, n, dt, seed=65))
    dA = O.qr_pullback(np.zeros_like(A), A, (Q, R), (dQ, dR))

    def loss(h):
        Qh, Rh = O.qr_compact((G + h * E) @ B @ (np.eye(n) + h * F))
        return np.real(np.vdot(dQ[:, :p], Qh[:, :p]) + np.vdot(dR[:p], Rh[:p]))
    h = 1e-6
    fd = (loss(h) - loss(-h)) / (2 * h)
    ad = np.real(np.vdot(dA, E @ B + A @ F))
    assert abs(fd - ad) < 1e-6 * max(1.0, abs(ad)), (fd, ad)
