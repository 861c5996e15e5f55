"""The CPU reference gradient of the BUG data step (oracle/adjoint.py) against
central finite differences of the forward step."""
import numpy as np

from oracle import adjoint, dlra
from oracle.data import random_factors


def _dirs(rng, *shapes):
    return [rng.standard_normal(s) for s in shapes]


def test_qr_vjp_against_finite_differences():
    rng = np.random.default_rng(0)
    K = rng.standard_normal((30, 5))
    Qb = rng.standard_normal((30, 5))
    Q, R = np.linalg.qr(K)
    Kb = adjoint.qr_vjp(Q, R, Qb)
    D = rng.standard_normal(K.shape)
    e = 1e-6
    fd = (np.sum(Qb * np.linalg.qr(K + e * D)[0]) - np.sum(Qb * np.linalg.qr(K - e * D)[0])) / (2 * e)
    assert abs(fd - np.sum(Kb * D)) < 1e-7 * max(1.0, abs(fd))


def test_bug_step_vjp_against_finite_differences():
    rng = np.random.default_rng(1)
    n, m, r = 40, 30, 4
    u = random_factors(rng, n, m, r, decay=0.5)
    dA = 1e-1 * rng.standard_normal((n, m))
    Ub1, Sb1, Vb1 = _dirs(rng, (n, r), (r, r), (m, r))
    g = adjoint.bug_step_vjp(u.U, u.S, u.V, dA, Ub1, Sb1, Vb1)

    def f(U, S, V, A):
        U1, S1, V1 = dlra.bug_step(U, S, V, A)
        return np.sum(Ub1 * U1) + np.sum(Sb1 * S1) + np.sum(Vb1 * V1)

    D = _dirs(rng, (n, r), (r, r), (m, r), (n, m))
    e = 1e-6
    fd = (f(u.U + e * D[0], u.S + e * D[1], u.V + e * D[2], dA + e * D[3])
          - f(u.U - e * D[0], u.S - e * D[1], u.V - e * D[2], dA - e * D[3])) / (2 * e)
    an = sum(np.sum(a * d) for a, d in zip(g, D))
    assert abs(fd - an) < 1e-6 * max(1.0, abs(fd))


def linear_loss(C):
    """L = <C, U S V'>: basis-free, so its gradient does not depend on the QR signs."""
    def lg(U, S, V):
        return np.sum(C * (U @ S @ V.T)), C @ V @ S.T, U.T @ C @ V, C.T @ U @ S
    return lg


def test_checkpointed_gradient_against_finite_differences():
    rng = np.random.default_rng(2)
    n, m, r, T = 36, 28, 3, 9
    u = random_factors(rng, n, m, r, decay=0.5)
    snaps = [u.to_matrix()]
    for _ in range(T):
        snaps.append(snaps[-1] + 5e-2 * rng.standard_normal((n, m)))
    C = rng.standard_normal((n, m))
    value, Ub, Sb, Vb = adjoint.bug_solve_gradient(u.U, u.S, u.V, snaps, linear_loss(C), checkpoint=4)

    def run(U, S, V):
        for k in range(T):
            U, S, V = dlra.bug_step(U, S, V, snaps[k + 1] - snaps[k])
        return np.sum(C * (U @ S @ V.T))

    assert abs(run(u.U, u.S, u.V) - value) < 1e-12 * abs(value)
    D = _dirs(rng, (n, r), (r, r), (m, r))
    e = 1e-6
    fd = (run(u.U + e * D[0], u.S + e * D[1], u.V + e * D[2]) - run(u.U - e * D[0], u.S - e * D[1], u.V - e * D[2])) / (2 * e)
    an = np.sum(Ub * D[0]) + np.sum(Sb * D[1]) + np.sum(Vb * D[2])
    assert abs(fd - an) < 1e-6 * max(1.0, abs(fd))
    # checkpoint spacing does not change the gradient
    g1 = adjoint.bug_solve_gradient(u.U, u.S, u.V, snaps, linear_loss(C), checkpoint=1)
    for a, b in zip(g1[1:], (Ub, Sb, Vb)):
        assert np.allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())


def test_qr_jvp_against_finite_differences():
    rng = np.random.default_rng(3)
    K = rng.standard_normal((25, 4))
    Kd = rng.standard_normal(K.shape)
    Q, R = np.linalg.qr(K)
    Qd, Rd = adjoint.qr_jvp(Q, R, Kd)
    e = 1e-6
    Qp, Rp = np.linalg.qr(K + e * Kd)
    Qm, Rm = np.linalg.qr(K - e * Kd)
    assert np.abs((Qp - Qm) / (2 * e) - Qd).max() < 1e-7
    assert np.abs((Rp - Rm) / (2 * e) - Rd).max() < 1e-7


def test_forward_mode_agrees_with_reverse_mode():
    """<ybar, J xdot> (forward mode) = <J' ybar, xdot> (reverse mode), for random
    tangents and cotangents: the CPU forward-mode reference for the reverse step."""
    rng = np.random.default_rng(4)
    n, m, r = 45, 33, 5
    u = random_factors(rng, n, m, r, decay=0.4)
    dA = 1e-1 * rng.standard_normal((n, m))
    xd = _dirs(rng, (n, r), (r, r), (m, r), (n, m))
    yb = _dirs(rng, (n, r), (r, r), (m, r))
    jv = adjoint.bug_step_jvp(u.U, u.S, u.V, dA, *xd)
    vj = adjoint.bug_step_vjp(u.U, u.S, u.V, dA, *yb)
    lhs = sum(np.sum(a * b) for a, b in zip(yb, jv))
    rhs = sum(np.sum(a * b) for a, b in zip(vj, xd))
    assert abs(lhs - rhs) < 1e-11 * max(1.0, abs(lhs))


def test_ksl_primal_vjp_against_finite_differences_and_forward_mode():
    rng = np.random.default_rng(5)
    n, m, r = 38, 29, 4
    u = random_factors(rng, n, m, r, decay=0.5)
    dA = 1e-1 * rng.standard_normal((n, m))
    yb = _dirs(rng, (n, r), (r, r), (m, r))
    g = adjoint.ksl_primal_step_vjp(u.U, u.S, u.V, dA, *yb)

    def f(U, S, V, A):
        U1, S1, V1 = dlra.ksl_primal_step(U, S, V, A)
        return np.sum(yb[0] * U1) + np.sum(yb[1] * S1) + np.sum(yb[2] * V1)

    D = _dirs(rng, (n, r), (r, r), (m, r), (n, m))
    e = 1e-6
    fd = (f(u.U + e * D[0], u.S + e * D[1], u.V + e * D[2], dA + e * D[3])
          - f(u.U - e * D[0], u.S - e * D[1], u.V - e * D[2], dA - e * D[3])) / (2 * e)
    an = sum(np.sum(a * d) for a, d in zip(g, D))
    assert abs(fd - an) < 1e-6 * max(1.0, abs(fd))
    jv = adjoint.ksl_primal_step_jvp(u.U, u.S, u.V, dA, *D)
    lhs = sum(np.sum(a * b) for a, b in zip(yb, jv))
    assert abs(lhs - an) < 1e-11 * max(1.0, abs(lhs))


def _fd_check(rng, step, x, g, yb, e=1e-6):
    """Central difference of <yb, step(x)> along a random direction, and the
    same derivative from the cotangents g of x."""
    D = _dirs(rng, *[a.shape for a in x])

    def f(*y):
        return sum(np.sum(b * o) for b, o in zip(yb, step(*y)))

    fd = (f(*[a + e * d for a, d in zip(x, D)]) - f(*[a - e * d for a, d in zip(x, D)])) / (2 * e)
    an = sum(np.sum(a * d) for a, d in zip(g, D))
    return fd, an


def test_ksl_dual_vjp_against_finite_differences_and_transposition():
    rng = np.random.default_rng(23)
    n, m, r = 35, 27, 4
    u = random_factors(rng, n, m, r, decay=0.5)
    dA = 1e-1 * rng.standard_normal((n, m))
    yb = _dirs(rng, (n, r), (r, r), (m, r))
    g = adjoint.ksl_dual_step_vjp(u.U, u.S, u.V, dA, *yb)
    fd, an = _fd_check(rng, dlra.ksl_dual_step, (u.U, u.S, u.V, dA), g, yb)
    assert abs(fd - an) < 1e-6 * max(1.0, abs(fd))
    # the dual step is the primal step of the transposed problem:
    # dual(U, S, V; dA) = (V1', S1'', U1') for primal(V, S', U; dA') = (U1', S1', V1')
    d1 = dlra.ksl_dual_step(u.U, u.S, u.V, dA)
    p1 = dlra.ksl_primal_step(u.V, u.S.T, u.U, dA.T)
    assert np.allclose(d1[0], p1[2], atol=1e-13) and np.allclose(d1[1], p1[1].T, atol=1e-13)
    assert np.allclose(d1[2], p1[0], atol=1e-13)
    gp = adjoint.ksl_primal_step_vjp(u.V, u.S.T, u.U, dA.T, yb[2], yb[1].T, yb[0])
    for a, b in zip(g, (gp[2], gp[1].T, gp[0], gp[3].T)):
        assert np.linalg.norm(a - b) <= 1e-12 * max(1.0, np.linalg.norm(b))


def test_ksl_strang_vjp_against_finite_differences():
    rng = np.random.default_rng(24)
    n, m, r = 33, 29, 3
    u = random_factors(rng, n, m, r, decay=0.5)
    dA1, dA2 = 1e-1 * rng.standard_normal((n, m)), 1e-1 * rng.standard_normal((n, m))
    yb = _dirs(rng, (n, r), (r, r), (m, r))
    g = adjoint.ksl_strang_step_vjp(u.U, u.S, u.V, dA1, dA2, *yb)

    def strang(U, S, V, a1, a2):
        return dlra.ksl_dual_step(*dlra.ksl_primal_step(U, S, V, a1), a2)

    fd, an = _fd_check(rng, strang, (u.U, u.S, u.V, dA1, dA2), g, yb)
    assert abs(fd - an) < 1e-6 * max(1.0, abs(fd))
    # the Strang step with a zero second increment is the primal step, whose
    # dual half then only re-orthogonalises: its gradient in dA2 still flows
    gz = adjoint.ksl_strang_step_vjp(u.U, u.S, u.V, dA1, 0 * dA2, *yb)
    fd, an = _fd_check(rng, strang, (u.U, u.S, u.V, dA1, 0 * dA2), gz, yb)
    assert abs(fd - an) < 1e-6 * max(1.0, abs(fd))


def test_svd_vjp_matches_finite_differences():
    rng = np.random.default_rng(31)
    q, k = 8, 3
    A = rng.standard_normal((q, q))
    Us, sig, Vst = np.linalg.svd(A)
    C1, C2 = rng.standard_normal((q, q)), rng.standard_normal((q, q))
    c3 = rng.standard_normal(k)

    def loss(A_):  # basis-free in the kept factors: <C1, Uk Sk Vk'> + <c3, sig_k> + <C2, Uk Uk'>
        u, s, vt = np.linalg.svd(A_)
        return (float(np.sum(C1 * ((u[:, :k] * s[:k]) @ vt[:k]))) + float(c3 @ s[:k])
                + float(np.sum(C2 * (u[:, :k] @ u[:, :k].T))))

    Uk, Vk = Us[:, :k], Vst[:k].T
    Ub = C1 @ Vk * sig[:k] + (C2 + C2.T) @ Uk
    Vb = C1.T @ Uk * sig[:k]
    sb = np.einsum("ij,ij->j", Uk, C1 @ Vk) + c3
    Ab = adjoint.svd_vjp(Us, sig, Vst.T, Ub, sb, Vb)
    D = rng.standard_normal((q, q))
    eps = 1e-6
    fd = (loss(A + eps * D) - loss(A - eps * D)) / (2 * eps)
    assert abs(fd - float(np.sum(Ab * D))) <= 1e-7 * max(1.0, abs(fd))


def test_ra_bug_step_vjp_matches_finite_differences():
    """Basis-free loss <C, U1 S1 V1'> + <c, sigma> of the rank-adaptive step (the
    reference's ordering Q([K1 U0]) in the forward) against the reverse in the
    [U0 | Q((I - U0 U0') K1)] form, the truncation rank held by a clear gap."""
    rng = np.random.default_rng(32)
    n, m, r = 30, 24, 4
    U = np.linalg.qr(rng.standard_normal((n, r)))[0]
    V = np.linalg.qr(rng.standard_normal((m, r)))[0]
    S = np.diag([3.0, 2.0, 1.2, 0.7]) + 0.05 * rng.standard_normal((r, r))
    dA = 0.5 * rng.standard_normal((n, 4)) @ rng.standard_normal((4, m)) / np.sqrt(m) + 1e-3 * rng.standard_normal((n, m))
    tol, rmax = 1e-8, 6  # sigma = 3.59 3.31 2.82 2.55 1.65 1.57 | 1.02 0.54: r1 = rmax = 6
    U1, S1, V1, r1 = dlra.rank_adaptive_bug_step(U, S, V, dA, tol, rmax)
    assert r1 == 6
    Ua, Sa, Va, ra_ = adjoint.ra_bug_step_aug(U, S, V, dA, tol, rmax)  # the same result, its own factors
    assert ra_ == r1 and np.allclose(Ua @ Sa @ Va.T, U1 @ S1 @ V1.T, atol=1e-12)
    U1, S1, V1 = Ua, Sa, Va
    C = rng.standard_normal((n, m))
    cs = rng.standard_normal(r1)
    Ub1 = C @ V1 @ S1.T
    Vb1 = C.T @ U1 @ S1
    Sb1 = U1.T @ C @ V1 + np.diag(cs)
    Ub0, Sb0, Vb0, dAb, r1b = adjoint.ra_bug_step_vjp(U, S, V, dA, tol, rmax, Ub1, Sb1, Vb1)
    assert r1b == r1

    def loss(U_, S_, V_, dA_):
        a, b, c, rr = dlra.rank_adaptive_bug_step(U_, S_, V_, dA_, tol, rmax)
        assert rr == r1
        return float(np.sum(C * (a @ b @ c.T))) + float(cs @ np.diag(b))

    def tangent(X):  # directions along orthonormal bases: (I - XX') G + X (W - W')
        G, W = rng.standard_normal(X.shape), rng.standard_normal((X.shape[1], X.shape[1]))
        return G - X @ (X.T @ G) + X @ (W - W.T)

    # the reverse assumes orthonormal U0, V0 (as the engine does); the reference's map and
    # this one agree on orthonormal bases, so they are compared along tangent directions
    dirs = [tangent(U), rng.standard_normal(S.shape), tangent(V), rng.standard_normal(dA.shape)]
    eps = 1e-6
    plus = loss(*(x + eps * d for x, d in zip((U, S, V, dA), dirs)))
    minus = loss(*(x - eps * d for x, d in zip((U, S, V, dA), dirs)))
    fd = (plus - minus) / (2 * eps)
    an = sum(float(np.sum(g * d)) for g, d in zip((Ub0, Sb0, Vb0, dAb), dirs))
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd))
