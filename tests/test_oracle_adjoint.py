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
