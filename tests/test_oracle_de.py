"""Pins for the DE-path oracle (no GPU)."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import dlra, dlra_de
from oracle.data import random_factors
from oracle.lra import SVDLike
from oracle.metrics import rel_err


def test_rk4_local_order():
    lam = -3.0
    errs = []
    for h in (0.1, 0.05, 0.025):
        y = dlra_de.rk4(lambda y, t: lam * y, np.array([1.0]), 0.0, h, 1)[0]
        errs.append(abs(y - np.exp(lam * h)))
    # local error ~ h^5: halving h divides the error by ~32
    assert 25 < errs[0] / errs[1] < 40 and 25 < errs[1] / errs[2] < 40


def test_rk4_time_argument():
    # y' = t: exact for polynomial degree <= 4
    y = dlra_de.rk4(lambda y, t: np.array([t ** 3]), np.array([0.0]), 1.0, 0.5, 3)[0]
    assert abs(y - (1.5 ** 4 - 1.0) / 4) < 1e-14


@pytest.mark.parametrize("name", ["primal", "dual", "bug"])
def test_constant_rhs_matches_data_path(name):
    """F(Y, t) = C (constant): every sub-problem is linear in t, RK4 is exact, and the DE step
    equals the data step with increment dt*C."""
    rng = np.random.default_rng(0)
    u = random_factors(rng, 30, 24, 4)
    C = rng.standard_normal((30, 24))
    dt = 0.05
    F = lambda Y, t: C
    de = {"primal": dlra_de.ksl_primal_de_step, "dual": dlra_de.ksl_dual_de_step, "bug": dlra_de.bug_de_step}[name]
    da = {"primal": dlra.ksl_primal_step, "dual": dlra.ksl_dual_step, "bug": dlra.bug_step}[name]
    a = de(F, u.U, u.S, u.V, 0.0, dt, 3)
    b = da(u.U, u.S, u.V, dt * C)
    assert rel_err(*a, *b) < 1e-13


def test_constant_rhs_rank_adaptive_matches_data_path():
    rng = np.random.default_rng(1)
    u = random_factors(rng, 30, 24, 4)
    C = rng.standard_normal((30, 24))
    a = dlra_de.rank_adaptive_bug_de_step(lambda Y, t: C, u.U, u.S, u.V, 0.0, 0.05, 2, 1e-3, 20)
    b = dlra.rank_adaptive_bug_step(u.U, u.S, u.V, 0.05 * C, 1e-3, 20)
    assert a[3] == b[3]
    assert rel_err(*a[:3], *b[:3]) < 1e-13


def test_K_subproblem_against_matrix_exponential():
    """K' = F(K V') V with F(Y) = A Y + Y B' + g h' is linear: compare RK4 to the exact flow."""
    rng = np.random.default_rng(2)
    n, m, r = 20, 16, 3
    A = dlra_de.dirichlet_laplacian(n, 4.0)
    B = dlra_de.dirichlet_laplacian(m, 3.0)
    g, h = rng.standard_normal((n, 1)), rng.standard_normal((m, 1))
    F = dlra_de.LyapunovOperator(A, B, g, h)
    V, _ = np.linalg.qr(rng.standard_normal((m, r)))
    K0 = rng.standard_normal((n, r))
    dt = 0.1
    K = dlra_de.rk4(lambda K, t: F(K @ V.T, t) @ V, K0, 0.0, dt, 200)
    # vec(K)' = (I (x) A + Bt (x) I) vec(K) + vec(g h' V),  Bt = V'BV
    Bt = V.T @ B @ V
    Op = np.kron(np.eye(r), A) + np.kron(Bt, np.eye(n))
    c = (g @ h.T @ V).reshape(-1, order="F")
    Aug = np.zeros((n * r + 1, n * r + 1))
    Aug[: n * r, : n * r] = Op
    Aug[: n * r, -1] = c
    z = sla.expm(dt * Aug) @ np.concatenate([K0.reshape(-1, order="F"), [1.0]])
    exact = z[: n * r].reshape(n, r, order="F")
    assert np.linalg.norm(K - exact) / np.linalg.norm(exact) < 1e-10


def test_laplacian_and_substeps():
    L = dlra_de.dirichlet_laplacian(4)
    assert np.allclose(L, 16 * np.array([[-2, 1, 0, 0], [1, -2, 1, 0], [0, 1, -2, 1], [0, 0, 1, -2]]))
    assert dlra_de.rk4_substeps(4 * 4096 ** 2, 1e-4) == 2685
    ev = np.linalg.eigvalsh(dlra_de.dirichlet_laplacian(64))
    assert -ev.min() < 4 * 64 ** 2


def test_solve_de_driver():
    rng = np.random.default_rng(3)
    n = 24
    u = random_factors(rng, n, n, 3)
    F = dlra_de.LyapunovOperator(dlra_de.dirichlet_laplacian(n, 1.0), dlra_de.dirichlet_laplacian(n, 1.0),
                                 rng.standard_normal((n, 1)), rng.standard_normal((n, 1)))
    prob = dlra_de.MatrixDEProblem(F, u, (0.0, 0.1))
    sol = dlra_de.solve_de(prob, dlra.UnconventionalAlgorithm(), 0.01, 2)
    assert len(sol.Y) == 11
    # rank-r projection of the exact full solution is close for a smooth linear problem
    Y = u.to_matrix()
    for k in range(10):
        Y = dlra_de.rk4(lambda X, t: F(X, t), Y, 0.01 * k, 0.01, 2)
    yf = sol.Y[-1]
    assert np.linalg.norm(yf.to_matrix() - Y) / np.linalg.norm(Y) < 0.5


@pytest.mark.parametrize("name", ["primal", "dual", "strang", "bug"])
def test_scalar_operator_closed_form(name):
    """F(Y) = c Y: K and L flows forward, the KSL S flow backward, so one KSL
    step scales Y by (R(z)^2 R(-z))^N and one BUG step by R(z)^N (z = h c, R the
    RK4 stability polynomial).  A dropped minus sign in the KSL S step fails this."""
    rng = np.random.default_rng(4)
    u = random_factors(rng, 20, 16, 3)
    c, dt, N = -4.5, 0.2, 7
    F = lambda Y, t: c * Y
    R = lambda z: 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    z = dt / N * c
    if name == "primal":
        got, f = dlra_de.ksl_primal_de_step(F, u.U, u.S, u.V, 0.0, dt, N), (R(z) ** 2 * R(-z)) ** N
    elif name == "dual":
        got, f = dlra_de.ksl_dual_de_step(F, u.U, u.S, u.V, 0.0, dt, N), (R(z) ** 2 * R(-z)) ** N
    elif name == "strang":
        a = dlra_de.ksl_primal_de_step(F, u.U, u.S, u.V, 0.0, dt / 2, N)
        got = dlra_de.ksl_dual_de_step(F, *a, dt / 2, dt / 2, N)
        f = (R(z / 2) ** 2 * R(-z / 2)) ** (2 * N)
    else:
        got, f = dlra_de.bug_de_step(F, u.U, u.S, u.V, 0.0, dt, N), R(z) ** N
    assert rel_err(*got, u.U, f * u.S, u.V) < 1e-13


def test_lowrank_evaluation_matches_dense():
    """The factored right-hand sides equal the dense F on the same arguments."""
    import scipy.sparse as sp

    rng = np.random.default_rng(9)
    n, m, r = 26, 21, 4
    A = sp.csr_matrix(dlra_de.dirichlet_laplacian(n, 5.0))
    B = sp.csr_matrix(dlra_de.dirichlet_laplacian(m, 3.0) + np.diag(rng.uniform(-1, 0, m)))
    G, H = rng.standard_normal((n, 2)), rng.standard_normal((m, 2))
    op = dlra_de.LowRankLyapunov(A, B, G, H)
    u = random_factors(rng, n, m, r)
    a = dlra_de.bug_de_step_lowrank(op, u.U, u.S, u.V, 0.0, 0.05, 4)
    b = dlra_de.bug_de_step(op.dense(), u.U, u.S, u.V, 0.0, 0.05, 4)
    assert rel_err(*a, *b) < 1e-13
    for tol, rmax in ((1e-3, 8), (1e-12, 5)):
        a = dlra_de.rank_adaptive_bug_de_step_lowrank(op, u.U, u.S, u.V, 0.0, 0.05, 4, tol, rmax)
        b = dlra_de.rank_adaptive_bug_de_step(op.dense(), u.U, u.S, u.V, 0.0, 0.05, 4, tol, rmax)
        assert a[3] == b[3] and rel_err(*a[:3], *b[:3]) < 1e-13


# --------------------------------------------------------------- quadratic (Burgers) operator


def test_burgers_matrices_are_the_reference_stencils():
    """data_agnostic_approximation.jl:6-29 entry by entry, checked against the closed
    forms they discretise: for rho = sin(2x) (period pi), Delta rho -> -4 nu sin(2x)
    and grad rho -> 2 cos(2x) with O(dx^2) error; rows sum to zero (periodic)."""
    n, nu = 1000, 0.005
    Lap, Grad = dlra_de.burgers_matrices(n, nu)
    dx = np.pi / n
    x = dx / 2 + dx * np.arange(n)
    rho = np.sin(2 * x)
    assert np.abs(Lap @ rho + 4 * nu * rho).max() < 4 * nu * 4 * dx ** 2
    assert np.abs(Grad @ rho - 2 * np.cos(2 * x)).max() < 8 * dx ** 2
    assert np.abs(Lap.sum(axis=1)).max() < 1e-9 and np.abs(Grad.sum(axis=1)).max() < 1e-9
    assert np.allclose(Grad, -Grad.T) and np.allclose(Lap, Lap.T)
    # the reference's Burgers right-hand side
    Y = np.stack([rho, np.cos(x)], 1)
    F = dlra_de.QuadraticOperator(Lap, np.zeros((2, 2)), np.zeros((n, 1)), np.zeros((2, 1)), Grad)
    assert np.allclose(F(Y), Lap @ Y - (Grad @ Y) * Y, rtol=0, atol=1e-12)


def test_burgers_initial_condition():
    """rho0_mat (data_agnostic_approximation.jl:35-44): column (i1, i2) of the m x m
    grid, xi_1 fastest, is ub(x) + 0.5 xi_1 sin(2 pi x) + 0.5 xi_2 sin(3 pi x)."""
    n, m = 50, 4
    R = dlra_de.burgers_initial(n, m)
    assert R.shape == (n, m * m)
    dx = np.pi / n
    x = dx / 2 + dx * np.arange(n)
    xi = np.linspace(-1, 1, m)
    ub = 0.5 * (np.exp(np.cos(x)) - 1.5) * np.sin(x + 2 * np.pi * 0.37)
    col = 1 + m * 2  # xi_1 = xi[1], xi_2 = xi[2]
    assert np.allclose(R[:, col], ub + 0.5 * xi[1] * np.sin(2 * np.pi * x) + 0.5 * xi[2] * np.sin(3 * np.pi * x))


def _tensor(X, Y):
    """T_abc = sum_r X_ra Y_rb Y_rc."""
    return np.einsum("ra,rb,rc->abc", X, Y, Y)


@pytest.mark.parametrize("periodic", [True, False])
def test_quadratic_subproblem_rhs_tensor_forms(periodic):
    """The factored K/L/S right-hand sides of the quadratic operator (the form the
    engine evaluates) against the dense projections F(K V') V, F(U L')' U,
    U' F(U S V') V: N_ic = sum_ab (D K)_ia K_ib T_abc with T = V x V x V,
    Q_jc = sum_ab L_ja L_jb W_abc and P_cd = sum W_efc S_ea S_fb T_abd with
    W = (D U) x U x U."""
    rng = np.random.default_rng(3)
    n, m, r, q = 30, 22, 4, 2
    A = rng.standard_normal((n, n)); A = A + A.T
    B = rng.standard_normal((m, m)); B = B + B.T
    D = np.diag(rng.standard_normal(n)) + np.diag(rng.standard_normal(n - 1), 1) + np.diag(rng.standard_normal(n - 1), -1)
    if periodic:
        D[0, n - 1], D[n - 1, 0] = rng.standard_normal(2)
    G, H = rng.standard_normal((n, q)), rng.standard_normal((m, q))
    F = dlra_de.QuadraticOperator(A, B, G, H, D)
    U, _ = np.linalg.qr(rng.standard_normal((n, r)))
    V, _ = np.linalg.qr(rng.standard_normal((m, r)))
    K, L, S = rng.standard_normal((n, r)), rng.standard_normal((m, r)), rng.standard_normal((r, r))
    T, W = _tensor(V, V), _tensor(D @ U, U)
    Kf = A @ K + K @ (V.T @ B @ V) + G @ (H.T @ V) - np.einsum("ia,ib,abc->ic", D @ K, K, T)
    assert np.allclose(Kf, F(K @ V.T) @ V, rtol=0, atol=1e-11)
    Lf = B @ L + L @ (U.T @ A @ U) + H @ (G.T @ U) - np.einsum("ja,jb,abc->jc", L, L, W)
    assert np.allclose(Lf, F(U @ L.T).T @ U, rtol=0, atol=1e-11)
    Sf = (U.T @ A @ U) @ S + S @ (V.T @ B @ V) + (U.T @ G) @ (H.T @ V) - np.einsum("efc,ea,fb,abd->cd", W, S, S, T)
    assert np.allclose(Sf, U.T @ F(U @ S @ V.T) @ V, rtol=0, atol=1e-11)


def test_api_burgers_samples_match_the_oracle(lri):
    """The package's Burgers initial samples (used by bench.py) equal the oracle's
    restatement of data_agnostic_approximation.jl:35-44."""
    assert np.allclose(lri.burgers_samples(60, 5), dlra_de.burgers_initial(60, 5), rtol=0, atol=1e-15)


# ------------------------------------------------------------------ the Greedy step on a hybrid problem


def test_hybrid_step_zero_dynamics_is_the_polar_factor():
    """With f = 0 the Z integrator leaves Z unchanged and U1 is the orthogonal polar
    factor of X Z (greedy_integrator.jl:79-81), checked against scipy's polar."""
    rng = np.random.default_rng(31)
    n, m, r = 60, 45, 5
    U, _ = np.linalg.qr(rng.standard_normal((n, r)))
    Z = rng.standard_normal((m, r))
    X = rng.standard_normal((n, m))
    U1, Z1 = dlra_de.greedy_hybrid_step(lambda Y, t: np.zeros_like(Y), U, Z, X, 0.0, 0.1, 3)
    assert np.array_equal(Z1, Z)
    W, P = sla.polar(X @ Z)
    assert np.linalg.norm(U1 - W) <= 1e-12
    H = U1.T @ (X @ Z)  # X Z = U1 H with H symmetric positive definite
    assert np.linalg.norm(H - H.T) <= 1e-12 * np.linalg.norm(H) and np.linalg.eigvalsh(0.5 * (H + H.T)).min() > 0


def test_hybrid_step_linear_dynamics_against_the_exponential():
    """F(Y) = A Y + Y B' + g h': the Z ODE f(Z, U, t) = F(U Z')' U is linear,
    Z' = B Z + Z (U'AU)' + h (g'U) (U'U = I), so Z1 = exp of the vectorised
    system; the RK4 result converges to it (and the factored right-hand side of
    the low-rank evaluation gives the same Z1)."""
    rng = np.random.default_rng(32)
    n, m, r = 30, 24, 4
    A = -np.diag(rng.uniform(1, 3, n)) + 0.1 * rng.standard_normal((n, n))
    B = -np.diag(rng.uniform(1, 3, m)) + 0.1 * rng.standard_normal((m, m))
    g, h = rng.standard_normal((n, 1)), rng.standard_normal((m, 1))
    F = dlra_de.LyapunovOperator(A, B, g, h)
    U, _ = np.linalg.qr(rng.standard_normal((n, r)))
    Z = rng.standard_normal((m, r))
    X = rng.standard_normal((n, m))
    dt = 0.3
    _, Z1 = dlra_de.greedy_hybrid_step(F, U, Z, X, 0.0, dt, 200)
    At = U.T @ A @ U
    c = h @ (g.T @ U)
    # vec(B Z + Z At') = (I (x) B + At (x) I) vec(Z)  (column-major vec)
    Lop = np.kron(np.eye(r), B) + np.kron(At, np.eye(m))
    E = sla.expm(dt * Lop)
    z0, cv = Z.flatten(order="F"), c.flatten(order="F")
    z1 = E @ z0 + np.linalg.solve(Lop, (E - np.eye(m * r)) @ cv)
    assert np.linalg.norm(Z1.flatten(order="F") - z1) <= 1e-10 * np.linalg.norm(z1)
    lowrank = dlra_de.LowRankLyapunov(A, B, g, h)
    Z1f = dlra_de.rk4(lowrank.L_rhs(U), Z, 0.0, dt, 200)
    assert np.linalg.norm(Z1f - Z1) <= 1e-12 * np.linalg.norm(Z1)


def test_hybrid_step_scalar_operator_closed_form():
    """F(Y) = c Y: Z' = c Z, so N RK4 steps give R(c h)^N Z with R the RK4 stability
    polynomial; a dropped or transposed factor of the right-hand side fails this."""
    rng = np.random.default_rng(33)
    n, m, r, c, dt, N = 20, 16, 3, -2.0, 0.4, 5
    U, _ = np.linalg.qr(rng.standard_normal((n, r)))
    Z = rng.standard_normal((m, r))
    _, Z1 = dlra_de.greedy_hybrid_step(lambda Y, t: c * Y, U, Z, rng.standard_normal((n, m)), 0.0, dt, N)
    z = c * dt / N
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    assert np.linalg.norm(Z1 - R ** N * Z) <= 1e-14 * np.linalg.norm(Z)


def test_solve_hybrid_driver():
    """Step count and accumulated t of solve (primitives.jl:68-80) for the hybrid problem."""
    rng = np.random.default_rng(34)
    n, m, r = 12, 10, 2
    U, _ = np.linalg.qr(rng.standard_normal((n, r)))
    Z = rng.standard_normal((m, r))
    Xs = rng.standard_normal((n, m))
    prob = dlra_de.MatrixHybridProblem(lambda t: Xs * (1 + t), lambda Y, t: -Y, U, Z, (0.0, 1.0))
    Ys, ts = dlra_de.solve_hybrid(prob, 0.1, 2)
    assert len(Ys) == 11 and abs(ts[-1] - 1.0) < 1e-12
    U1, Z1 = Ys[-1]
    assert np.linalg.norm(U1.T @ U1 - np.eye(r)) <= 1e-12
