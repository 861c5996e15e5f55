"""Pins of the CPU reverse- and forward-mode references of the matrix-ODE BUG step
(oracle/adjoint_de.py): finite differences, forward/reverse duality, the RK4
stability polynomial, and the setup of the differentiation workload."""
import numpy as np
import pytest

from oracle import adjoint_de as ad


def random_op(rng, n, m, q=1, periodic=False, scale=1.0):
    def sym_tri(k, d0, e0):
        T = np.diag(-rng.uniform(*d0, k)) + np.diag(e := rng.uniform(*e0, k - 1), 1) + np.diag(e, -1)
        return T
    A = scale * sym_tri(n, (20, 40), (5, 15))
    B = scale * sym_tri(m, (10, 20), (2, 6))
    if periodic:
        A[0, -1] = A[-1, 0] = 7.0 * scale
    D = np.diag(rng.uniform(-1, 1, n)) + np.diag(rng.uniform(-3, 3, n - 1), 1) + np.diag(rng.uniform(-3, 3, n - 1), -1)
    if periodic:
        D[0, -1], D[-1, 0] = 2.0, -1.5
    return ad.ScaledQuadratic(A, B, rng.standard_normal((n, q)), rng.standard_normal((m, q)), D)


def random_state(rng, n, m, r):
    U = np.linalg.qr(rng.standard_normal((n, r)))[0]
    V = np.linalg.qr(rng.standard_normal((m, r)))[0]
    S = np.diag(np.linspace(1.0, 0.2, r)) + 0.05 * rng.standard_normal((r, r))
    return U, S, V


def pair(x, y):
    return sum(float(np.sum(a * b)) for a, b in zip(x, y))


def test_rk4_reverse_is_the_transposed_stability_polynomial():
    rng = np.random.default_rng(0)
    Lam = rng.standard_normal((5, 5))
    h, N = 0.1, 4
    P = sum(np.linalg.matrix_power(h * Lam, j) / f for j, f in enumerate([1, 1, 2, 6, 24]))
    y0 = rng.standard_normal((5, 2))
    ys = ad.rk4_states(lambda y: Lam @ y, y0, h, N)
    assert np.allclose(ys[-1], np.linalg.matrix_power(P, N) @ y0, rtol=1e-13, atol=1e-13)
    yb = rng.standard_normal((5, 2))
    got = ad.rk4_vjp(lambda y: Lam @ y, lambda x, kb: Lam.T @ kb, ys, h, yb)
    assert np.allclose(got, np.linalg.matrix_power(P.T, N) @ yb, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("periodic", [False, True])
def test_step_vjp_matches_finite_differences(periodic):
    rng = np.random.default_rng(1 + periodic)
    n, m, r, dt, N = 14, 11, 3, 0.01, 3
    F = random_op(rng, n, m, periodic=periodic)
    U, S, V = random_state(rng, n, m, r)
    C = [rng.standard_normal(x.shape) for x in (U, S, V)]  # loss <C, (U1, S1, V1)>
    Ub, Sb, Vb, abar, bbar = ad.bug_de_step_vjp(F, U, S, V, dt, N, *C)

    def loss(U_, S_, V_, a=1.0, b=1.0):
        return pair(C, ad.bug_de_step(F.scaled(a, b), U_, S_, V_, dt, N))

    dirs = [rng.standard_normal(x.shape) for x in (U, S, V)]
    eps = 1e-6
    fd = (loss(*(x + eps * d for x, d in zip((U, S, V), dirs))) - loss(*(x - eps * d for x, d in zip((U, S, V),
                                                                                                     dirs)))) / (2 * eps)
    an = pair((Ub, Sb, Vb), dirs)
    assert abs(fd - an) <= 1e-7 * max(1.0, abs(an))
    fa = (loss(U, S, V, 1 + eps) - loss(U, S, V, 1 - eps)) / (2 * eps)
    fb = (loss(U, S, V, 1, 1 + eps) - loss(U, S, V, 1, 1 - eps)) / (2 * eps)
    assert abs(fa - abar) <= 1e-7 * max(1.0, abs(abar))
    assert abs(fb - bbar) <= 1e-7 * max(1.0, abs(bbar))


def test_forward_and_reverse_mode_are_dual():
    rng = np.random.default_rng(3)
    n, m, r, dt, N = 16, 13, 4, 0.01, 2
    F = random_op(rng, n, m, q=2, periodic=True)
    U, S, V = random_state(rng, n, m, r)
    xd = [rng.standard_normal(x.shape) for x in (U, S, V)]
    ad_, bd_ = 0.7, -0.3
    yb = [rng.standard_normal(x.shape) for x in (U, S, V)]
    J = ad.bug_de_step_jvp(F, U, S, V, dt, N, *xd, ad_, bd_)
    Ub, Sb, Vb, abar, bbar = ad.bug_de_step_vjp(F, U, S, V, dt, N, *yb)
    lhs = pair(yb, J)
    rhs = pair((Ub, Sb, Vb), xd) + abar * ad_ + bbar * bd_
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs))


def test_checkpointed_solve_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    n, m, r, dt, N, steps = 12, 10, 3, 0.01, 2, 5
    F = random_op(rng, n, m, periodic=True)
    U, S, V = random_state(rng, n, m, r)
    lg = ad.mean_column_loss
    val, Ub, Sb, Vb, abar, bbar = ad.bug_de_solve_gradient(F, U, S, V, dt, N, steps, lg, checkpoint=2)
    _, Ub2, Sb2, Vb2, abar2, _ = ad.bug_de_solve_gradient(F, U, S, V, dt, N, steps, lg, checkpoint=7)
    for a, b in ((Ub, Ub2), (Sb, Sb2), (Vb, Vb2)):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-14)
    assert abs(abar - abar2) <= 1e-12 * abs(abar)

    def loss(x, a=1.0):
        return lg(*ad.bug_de_solve(F.scaled(a, 1.0), *x, dt, N, steps))[0]

    dirs = [rng.standard_normal(x.shape) for x in (U, S, V)]
    eps = 1e-6
    fd = (loss([x + eps * d for x, d in zip((U, S, V), dirs)]) - loss([x - eps * d for x, d in zip((U, S, V), dirs)])) / (2 * eps)
    assert abs(fd - pair((Ub, Sb, Vb), dirs)) <= 1e-7 * max(1.0, abs(fd))
    fa = (loss((U, S, V), 1 + eps) - loss((U, S, V), 1 - eps)) / (2 * eps)
    assert abs(fa - abar) <= 1e-7 * max(1.0, abs(fa))
    assert val == pytest.approx(loss((U, S, V)), rel=1e-14)


def test_mean_column_loss_gradient():
    rng = np.random.default_rng(5)
    U, S, V = random_state(rng, 9, 7, 3)
    U = U + 0.1 * rng.standard_normal(U.shape)  # not orthonormal: the general rule
    val, Ub, Sb, Vb = ad.mean_column_loss(U, S, V)
    Y = U @ S @ V.T
    assert val == pytest.approx(float(np.sum(Y.mean(axis=1) ** 2)), rel=1e-14)
    dirs = [rng.standard_normal(x.shape) for x in (U, S, V)]
    eps = 1e-6
    f = lambda t: ad.mean_column_loss(*(x + t * d for x, d in zip((U, S, V), dirs)))[0]  # noqa: E731
    assert (f(eps) - f(-eps)) / (2 * eps) == pytest.approx(pair((Ub, Sb, Vb), dirs), rel=1e-8)


def test_periodic_burgers_and_sine_state():
    n, m, r = 64, 20, 12
    Lap, Grad = ad.periodic_burgers(n, 0.01)
    x = np.arange(n) / n
    s = np.sin(2 * np.pi * 3 * x)
    # spectral exactness of the stencils on a sine: -nu (2 - 2 cos(k dx)) / dx^2 and sin(k dx) / dx
    kdx = 2 * np.pi * 3 / n
    assert np.allclose(Lap @ s, -0.01 * (2 - 2 * np.cos(kdx)) * n * n * s, atol=1e-10)
    assert np.allclose(Grad @ s, np.sin(kdx) * n * np.cos(2 * np.pi * 3 * x), atol=1e-10)
    rng = np.random.default_rng(6)
    xi = rng.standard_normal((8, m))
    U0, S0, V0, Y0 = ad.sine_state(n, xi, r, rng.standard_normal((m, r - 8)))
    assert np.allclose(U0.T @ U0, np.eye(r), atol=1e-13) and np.allclose(V0.T @ V0, np.eye(r), atol=1e-13)
    assert np.allclose(U0 @ S0 @ V0.T, Y0, atol=1e-12)
    assert np.allclose(S0[8:], 0.0, atol=1e-12) and np.allclose(S0[:, 8:], 0.0, atol=1e-12)


def test_package_setup_of_the_differentiation_workload_matches_the_oracle(lri):
    """The package's SineBurgers / periodic_burgers_operator / mean_column_loss (host
    set-up of the differentiation workload) against the restatements above, and the
    xi pullback against central differences of the set-up."""
    n, m, r, K = 64, 24, 12, 8
    rng = np.random.default_rng(7)
    xi = rng.standard_normal((K, m))
    sb = lri.SineBurgers(n, xi, 0.01, r, seed=3)
    U0, S0, V0, _ = ad.sine_state(n, xi, r, np.random.default_rng(3).standard_normal((m, r - K)))
    assert np.allclose(sb.U0, U0, atol=1e-14) and np.allclose(sb.V0, V0, atol=1e-13)
    assert np.allclose(sb.S0, S0, atol=1e-12)
    Lap, Grad = ad.periodic_burgers(n, 0.01)
    op = sb.op
    assert np.allclose(op.A.dense(), Lap, atol=1e-9) and np.allclose(op.D.dense(), Grad, atol=1e-12)
    assert np.allclose(op.B.dense(), 0.0)
    U, S, V = random_state(rng, n, m, r)
    for a, b in zip(lri.mean_column_loss(U, S, V), ad.mean_column_loss(U, S, V)):
        assert np.allclose(a, b, rtol=1e-14, atol=1e-15)
    # xi pullback: d/dt <C, (S0, V0)(xi + t d)>
    Cs, Cv = rng.standard_normal((r, r)), rng.standard_normal((m, r))
    d = rng.standard_normal(xi.shape)
    eps = 1e-6

    def f(t):
        z = lri.SineBurgers(n, xi + t * d, 0.01, r, seed=3)
        return float(np.sum(Cs * z.S0) + np.sum(Cv * z.V0))

    fd = (f(eps) - f(-eps)) / (2 * eps)
    an = float(np.sum(sb.pullback(np.zeros((n, r)), Cs, Cv) * d))
    assert abs(fd - an) <= 1e-7 * max(1.0, abs(an))


@pytest.mark.parametrize("periodic", [False, True])
def test_ksl_step_vjp_matches_finite_differences_and_forward_mode(periodic):
    rng = np.random.default_rng(11 + periodic)
    n, m, r, dt, N = 14, 11, 3, 0.01, 3
    F = random_op(rng, n, m, q=1, periodic=periodic)
    U, S, V = random_state(rng, n, m, r)
    C = [rng.standard_normal(x.shape) for x in (U, S, V)]
    Ub, Sb, Vb, abar, bbar = ad.ksl_primal_de_step_vjp(F, U, S, V, dt, N, *C)

    def loss(U_, S_, V_, a=1.0, b=1.0):
        return pair(C, ad.ksl_primal_de_step(F.scaled(a, b), U_, S_, V_, dt, N))

    dirs = [rng.standard_normal(x.shape) for x in (U, S, V)]
    eps = 1e-6
    fd = (loss(*(x + eps * d for x, d in zip((U, S, V), dirs))) - loss(*(x - eps * d for x, d in zip((U, S, V),
                                                                                                     dirs)))) / (2 * eps)
    an = pair((Ub, Sb, Vb), dirs)
    assert abs(fd - an) <= 1e-7 * max(1.0, abs(an))
    fa = (loss(U, S, V, 1 + eps) - loss(U, S, V, 1 - eps)) / (2 * eps)
    fb = (loss(U, S, V, 1, 1 + eps) - loss(U, S, V, 1, 1 - eps)) / (2 * eps)
    assert abs(fa - abar) <= 1e-7 * max(1.0, abs(abar)) and abs(fb - bbar) <= 1e-7 * max(1.0, abs(bbar))
    J = ad.ksl_primal_de_step_jvp(F, U, S, V, dt, N, *dirs, 0.4, -0.2)
    assert abs(pair(C, J) - (an + 0.4 * abar - 0.2 * bbar)) <= 1e-11 * max(1.0, abs(an))


@pytest.mark.parametrize("periodic", [False, True])
def test_ra_step_vjp_matches_finite_differences(periodic):
    """Rank-adaptive step (rank_adaptive_unconventional.jl:194-233 with the DE
    sub-problems): the augmented form [U0 | Q((I - U0 U0') K1)] gives the reference
    ordering's U1 S1 V1' (dlra_de, Q([K1 U0])); its reverse against central
    differences of the basis-free loss <C, U1 S1 V1'> + <c, sigma> in generic
    directions and along the scales, and against the reference ordering along
    tangents of the orthonormal bases; the rank is held by a clear gap."""
    rng = np.random.default_rng(5 + periodic)
    n, m, r, dt, N = 14, 11, 3, 0.03, 3
    F = random_op(rng, n, m, periodic=periodic)
    U, S, V = random_state(rng, n, m, r)
    tol, rmax = 1e-10, 4
    U1, S1, V1, r1 = ad.ra_bug_de_step(F, U, S, V, dt, N, tol, rmax)
    from oracle import dlra_de
    Ur, Sr, Vr, rr = dlra_de.rank_adaptive_bug_de_step(lambda Y, t: F(Y), U, S, V, 0.0, dt, N, tol, rmax)
    assert r1 == rr == 4
    assert np.abs(U1 @ S1 @ V1.T - Ur @ Sr @ Vr.T).max() <= 1e-14
    C = rng.standard_normal((n, m))
    cs = rng.standard_normal(r1)
    Ub1, Vb1 = C @ V1 @ S1.T, C.T @ U1 @ S1
    Sb1 = U1.T @ C @ V1 + np.diag(cs)
    Ub0, Sb0, Vb0, abar, bbar, r1b = ad.ra_bug_de_step_vjp(F, U, S, V, dt, N, tol, rmax, Ub1, Sb1, Vb1)
    assert r1b == r1

    def loss(step, U_, S_, V_, a=1.0, b=1.0):
        a_, b_, c_, q = step(F.scaled(a, b), U_, S_, V_)
        assert q == r1
        return float(np.sum(C * (a_ @ b_ @ c_.T))) + float(cs @ np.diag(b_))

    aug = lambda F_, U_, S_, V_: ad.ra_bug_de_step(F_, U_, S_, V_, dt, N, tol, rmax)  # noqa: E731
    ref = lambda F_, U_, S_, V_: dlra_de.rank_adaptive_bug_de_step(  # noqa: E731
        lambda Y, t: F_(Y), U_, S_, V_, 0.0, dt, N, tol, rmax)
    eps = 1e-6
    x = (U, S, V)
    D = [rng.standard_normal(a.shape) for a in x]
    fd = (loss(aug, *(a + eps * d for a, d in zip(x, D))) - loss(aug, *(a - eps * d for a, d in zip(x, D)))) / (2 * eps)
    an = pair((Ub0, Sb0, Vb0), D)
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
    fa = (loss(aug, *x, 1 + eps) - loss(aug, *x, 1 - eps)) / (2 * eps)
    fb = (loss(aug, *x, 1, 1 + eps) - loss(aug, *x, 1, 1 - eps)) / (2 * eps)
    assert abs(fa - abar) <= 1e-6 * max(1.0, abs(abar))
    assert abs(fb - bbar) <= 1e-6 * max(1.0, abs(bbar))

    def tangent(X):  # along the orthonormal bases: (I - XX') G + X (W - W')
        G, W = rng.standard_normal(X.shape), rng.standard_normal((X.shape[1], X.shape[1]))
        return G - X @ (X.T @ G) + X @ (W - W.T)

    D = [tangent(U), rng.standard_normal(S.shape), tangent(V)]
    fd = (loss(ref, *(a + eps * d for a, d in zip(x, D))) - loss(ref, *(a - eps * d for a, d in zip(x, D)))) / (2 * eps)
    assert abs(fd - pair((Ub0, Sb0, Vb0), D)) <= 1e-5 * max(1.0, abs(fd))


def test_ra_checkpointed_solve_gradient_matches_finite_differences():
    """Several rank-adaptive steps (the rank grows from 3): the checkpointed reverse
    against central differences of the forward solve, the ranks held by clear gaps."""
    rng = np.random.default_rng(9)
    n, m, r, dt, N, steps = 14, 11, 3, 0.02, 2, 4
    F = random_op(rng, n, m, periodic=True)
    U, S, V = random_state(rng, n, m, r)
    tol, rmax = 1e-3, 5
    lg = ad.mean_column_loss

    def run(F_, x):
        U_, S_, V_ = x
        ranks = []
        for _ in range(steps):
            U_, S_, V_, r1 = ad.ra_bug_de_step(F_, U_, S_, V_, dt, N, tol, rmax)
            ranks.append(r1)
        return (U_, S_, V_), ranks

    (UT, ST, VT), ranks = run(F, (U, S, V))
    assert ranks[0] > r and len(set(ranks)) >= 1
    val, Ub, Sb, Vb, abar, bbar = ad.bug_de_solve_gradient(F, U, S, V, dt, N, steps, lg, checkpoint=3,
                                                           ra=(tol, rmax))
    _, Ub2, Sb2, Vb2, abar2, _ = ad.bug_de_solve_gradient(F, U, S, V, dt, N, steps, lg, checkpoint=1,
                                                          ra=(tol, rmax))
    for a, b in ((Ub, Ub2), (Sb, Sb2), (Vb, Vb2)):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-14)
    assert val == pytest.approx(lg(UT, ST, VT)[0], rel=1e-14)

    def loss(x, a=1.0):
        y, rk = run(F.scaled(a, 1.0), x)
        assert rk == ranks
        return lg(*y)[0]

    def tangent(X):  # along the orthonormal bases: (I - XX') G + X (W - W')
        G, W = rng.standard_normal(X.shape), rng.standard_normal((X.shape[1], X.shape[1]))
        return G - X @ (X.T @ G) + X @ (W - W.T)

    # directions along the orthonormal bases (the composition of several steps is
    # badly conditioned off them, where a solve never goes)
    eps = 1e-6
    for dirs in ([tangent(U), 0 * S, 0 * V], [0 * U, rng.standard_normal(S.shape), 0 * V], [0 * U, 0 * S, tangent(V)]):
        fd = (loss([x + eps * d for x, d in zip((U, S, V), dirs)])
              - loss([x - eps * d for x, d in zip((U, S, V), dirs)])) / (2 * eps)
        assert abs(fd - pair((Ub, Sb, Vb), dirs)) <= 1e-4 * abs(fd) + 1e-9
    fa = (loss((U, S, V), 1 + eps) - loss((U, S, V), 1 - eps)) / (2 * eps)
    assert abs(fa - abar) <= 1e-4 * abs(fa) + 1e-9
