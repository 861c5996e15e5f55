"""Parity of the matrix-ODE BUG adjoint (lri_step_de_bug_adjoint) with the CPU
reverse- and forward-mode references of oracle/adjoint_de.py (needs an MI355X).

The step is unconventional_step! (unconventional.jl:133-157) with the default
sub-problem right-hand sides (:51-79) for F(Y) = a A Y + b Y B' + G H' - (D Y) o Y,
fixed RK4 sub-steps, QR with a positive diagonal of R (the engine's and the
oracle's bases then agree column by column, so basis-dependent cotangents are
compared directly).

The engine's sub-problems use U0'U0 = V0'V0 = I (K' = A K + K V0'BV0 + ...), the
oracle's dense F(K V0') V0 keeps V0'V0 in A K V0'V0: the two extensions of the
step off orthonormal bases differ, and so do the cotangents of U0, V0 by a term
U0 Sym (Sym symmetric), normal to the orthonormal bases.  Those are compared on
the tangent space, G - U0 sym(U0'G); the QR backward of the step before removes
the normal part anyway, so solves and every derivative along orthonormal bases
agree.

Tolerances: one step, relative max-norm error of every cotangent and of the two
scale derivatives <= 1e-9 (the engine evaluates the sub-problems in the r^3
tensor form, the oracle densely in n x m); a checkpointed solve <= 1e-8; the
differentiation workload (n = 2048, m = 1024, r = 16) against the CPU
forward-mode directional derivative <= 1e-8, and against central differences of
the engine's own forward solve <= 1e-5 (the differences' truncation error).
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import adjoint_de as ad
from oracle.adjoint import qr_jvp

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


def tangential(X, G):
    """G - X sym(X'G): the part of a cotangent of an orthonormal X along orthonormal bases."""
    XG = X.T @ G
    return G - X @ (0.5 * (XG + XG.T))


def random_operator(lri, n, m, q, periodic, seed):
    rng = np.random.default_rng(seed)
    A = lri.SymTridiagonal(-rng.uniform(20, 40, n), rng.uniform(5, 15, n if periodic else n - 1), periodic)
    B = lri.SymTridiagonal(-rng.uniform(10, 20, m), rng.uniform(2, 6, m - 1))
    D = lri.Tridiagonal(rng.uniform(-3, 3, n), rng.uniform(-1, 1, n), rng.uniform(-3, 3, n), periodic)
    G, H = rng.standard_normal((n, q)), rng.standard_normal((m, q))
    dev = lri.QuadraticOperator(A, B, D, G if q else None, H if q else None)
    host = ad.ScaledQuadratic(sp.csr_matrix(A.dense()), sp.csr_matrix(B.dense()), G, H, sp.csr_matrix(D.dense()))
    return dev, host


def context(lri, dev, U, S, V):
    n, m, r = U.shape[0], V.shape[0], S.shape[0]
    ctx = lri.native.Context(n, m, r)
    ctx.set_state(U, S, V)
    q = dev.G.shape[1]
    ctx.set_operator_lyapunov(dev.A.d, dev.A.e, dev.A.periodic, dev.B.d, dev.B.e, dev.B.periodic,
                              dev.G if q else None, dev.H if q else None)
    ctx.set_operator_quadratic(dev.D.dl, dev.D.dd, dev.D.du, dev.D.periodic)
    return ctx


def random_state(rng, n, m, r):
    U = np.linalg.qr(rng.standard_normal((n, r)))[0]
    V = np.linalg.qr(rng.standard_normal((m, r)))[0]
    S = np.diag(np.linspace(1.0, 0.3, r)) + 0.05 * rng.standard_normal((r, r))
    return U, S, V


@pytest.mark.parametrize("periodic,q,shape,dt", [(True, 1, (50, 37, 4), 0.01), (False, 0, (64, 48, 6), 0.01),
                                                 (True, 2, (300, 200, 16), 0.05), (False, 1, (129, 70, 32), 0.05)])
def test_one_step_adjoint_vs_oracle(lri, periodic, q, shape, dt):
    n, m, r = shape
    say(f"case: matrix-ODE BUG adjoint, one step, n={n} m={m} r={r} q={q} periodic={periodic}")
    dev, host = random_operator(lri, n, m, q, periodic, seed=11)
    rng = np.random.default_rng(12)
    U, S, V = random_state(rng, n, m, r)
    cot = [rng.standard_normal(x.shape) for x in (U, S, V)]
    N = lri.rk4_substeps(dev.spectral_bound(2 * np.abs(U @ S @ V.T).max()), dt)
    with context(lri, dev, U, S, V) as ctx:
        Ub, Sb, Vb, (a, b) = ctx.step_de_bug_adjoint(dt, N, *cot)
        U2, S2, V2 = ctx.get_state()
    assert np.array_equal(U2, U) and np.array_equal(S2, S) and np.array_equal(V2, V)  # state unchanged
    ref = ad.bug_de_step_vjp(host, U, S, V, dt, N, *cot)
    errs = [rel(tangential(U, Ub), tangential(U, ref[0])), rel(Sb, ref[1]),
            rel(tangential(V, Vb), tangential(V, ref[2]))] + [abs(a - ref[3]) / abs(ref[3]),
                                                                 abs(b - ref[4]) / max(abs(ref[4]), 1e-300)]
    say(f"  N={N}: rel err Ub {errs[0]:.1e} Sb {errs[1]:.1e} Vb {errs[2]:.1e} abar {errs[3]:.1e} bbar {errs[4]:.1e}")
    assert max(errs) <= 1e-9


def test_checkpointed_burgers_gradient_vs_oracle(lri):
    """The differentiation workload's problem at a small size: gradient of
    L = ||mean_cols(Y(T))||^2 with respect to xi and nu."""
    n, m, r, K, nu = 128, 32, 12, 8, 0.01
    dt, steps = 1e-4, 6
    xi = np.random.default_rng(21).standard_normal((K, m))
    sb = lri.SineBurgers(n, xi, nu, r, seed=4)
    prob = sb.problem((0.0, steps * dt))
    N = lri.de_substeps(prob, dt)
    say(f"case: Burgers sine gradient n={n} m={m} r={r}, {steps} steps of {N} RK4 sub-steps, checkpoint 4")
    val, (Ub, Sb, Vb), (abar, _) = lri.solve_gradient(prob, lri.UnconventionalAlgorithm(), lri.mean_column_loss, dt,
                                                      checkpoint=4)
    Lap, Grad = ad.periodic_burgers(n, nu)
    host = ad.ScaledQuadratic(sp.csr_matrix(Lap), sp.csr_matrix((m, m)), np.zeros((n, 0)), np.zeros((m, 0)),
                              sp.csr_matrix(Grad))
    u0 = sb.initial()
    rv, rUb, rSb, rVb, rabar, _ = ad.bug_de_solve_gradient(host, u0.U, u0.S, u0.V, dt, N, steps, ad.mean_column_loss,
                                                           checkpoint=4)
    e_xi = rel(sb.pullback(Ub, Sb, Vb), sb.pullback(rUb, rSb, rVb))
    e_nu = abs(abar - rabar) / abs(rabar)
    say(f"  loss {val:.6e} (oracle {rv:.6e}); rel err dL/dxi {e_xi:.1e}, dL/dnu {e_nu:.1e}")
    assert abs(val - rv) <= 1e-10 * abs(rv)
    assert e_xi <= 1e-8 and e_nu <= 1e-8


def workload(lri, steps, seed=0):
    n, m, r, K, nu, dt = 2048, 1024, 16, 8, 1e-2, 1e-4
    xi = np.random.default_rng(seed).standard_normal((K, m))
    sb = lri.SineBurgers(n, xi, nu, r, seed=seed)
    return sb, dt, sb.problem((0.0, steps * dt))


def test_workload_gradient_vs_forward_mode_reference(lri):
    """n = 2048, m = 1024, r = 16 (the differentiation workload's shape) over 3 steps:
    <grad, (d_xi, d_nu)> against the CPU forward-mode (tangent) solve."""
    steps = 3
    sb, dt, prob = workload(lri, steps)
    N = lri.de_substeps(prob, dt)
    say(f"case: differentiation workload shape, {steps} steps, {N} RK4 sub-steps: forward-mode check")
    _, (Ub, Sb, Vb), (abar, _) = lri.solve_gradient(prob, lri.UnconventionalAlgorithm(), lri.mean_column_loss, dt)
    rng = np.random.default_rng(5)
    d_xi = rng.standard_normal(sb.xi.shape)
    d_nu = 0.3
    an = float(np.sum(sb.pullback(Ub, Sb, Vb) * d_xi)) + abar / sb.nu * d_nu
    # tangent of the initial factors along d_xi (U0 fixed): V0 = Q([xi' | omega]) by the
    # tangent QR rule, S0 = C xi V0
    K = sb.xi.shape[0]
    Vd, _ = qr_jvp(sb.V0, sb.R, np.hstack([d_xi.T, np.zeros((sb.m, sb.r - K))]))
    Sd = sb.C @ (d_xi @ sb.V0) + sb.C @ (sb.xi @ Vd)
    Lap, Grad = ad.periodic_burgers(sb.n, 1.0)
    host = ad.ScaledQuadratic(sp.csr_matrix(sb.nu * Lap), sp.csr_matrix((sb.m, sb.m)), np.zeros((sb.n, 0)),
                              np.zeros((sb.m, 0)), sp.csr_matrix(Grad))
    U, S, V = sb.U0, sb.S0, sb.V0
    Ud = np.zeros_like(U)
    for _ in range(steps):  # d/dnu of nu A is A: the scale tangent a-dot = d_nu / nu
        Ud, Sd, Vd = ad.bug_de_step_jvp(host, U, S, V, dt, N, Ud, Sd, Vd, d_nu / sb.nu, 0.0)
        U, S, V = ad.bug_de_step(host, U, S, V, dt, N)
    _, gU, gS, gV = ad.mean_column_loss(U, S, V)
    fm = float(np.sum(gU * Ud) + np.sum(gS * Sd) + np.sum(gV * Vd))
    say(f"  directional derivative: engine adjoint {an:.12e}, CPU forward mode {fm:.12e}")
    assert abs(an - fm) <= 1e-8 * abs(fm)


def test_workload_gradient_vs_finite_differences(lri):
    """The full differentiation workload (200 steps, checkpoint every 20) against central
    differences of the engine's forward solve along random directions in xi and nu."""
    steps = 200
    sb, dt, prob = workload(lri, steps, seed=1)
    N = lri.de_substeps(prob, dt)
    say(f"case: differentiation workload, {steps} steps, {N} RK4 sub-steps, checkpoint 20: finite differences")
    val, (Ub, Sb, Vb), (abar, _) = lri.solve_gradient(prob, lri.UnconventionalAlgorithm(), lri.mean_column_loss, dt,
                                                      checkpoint=20)
    g_xi = sb.pullback(Ub, Sb, Vb)
    g_nu = abar / sb.nu
    rng = np.random.default_rng(6)
    d_xi = rng.standard_normal(sb.xi.shape)
    d_xi /= np.linalg.norm(d_xi)

    def L(xi, nu):
        s = lri.SineBurgers(sb.n, xi, nu, sb.r, seed=1)
        y = lri.solve(s.problem((0.0, steps * dt)), lri.UnconventionalAlgorithm(), dt, substeps=N,
                      save_everystep=False).Y[-1]
        return lri.mean_column_loss(y.U, y.S, y.V)[0]

    assert L(sb.xi, sb.nu) == pytest.approx(val, rel=1e-12)
    h = 1e-4
    fd_xi = (L(sb.xi + h * d_xi, sb.nu) - L(sb.xi - h * d_xi, sb.nu)) / (2 * h)
    hn = 1e-4 * sb.nu
    fd_nu = (L(sb.xi, sb.nu + hn) - L(sb.xi, sb.nu - hn)) / (2 * hn)
    an_xi = float(np.sum(g_xi * d_xi))
    say(f"  loss {val:.6e}; d/dxi: adjoint {an_xi:.10e} FD {fd_xi:.10e}; d/dnu: adjoint {g_nu:.10e} FD {fd_nu:.10e}")
    assert abs(an_xi - fd_xi) <= 1e-5 * abs(fd_xi)
    assert abs(g_nu - fd_nu) <= 1e-5 * abs(fd_nu)


def test_adjoint_argument_errors(lri):
    n, m, r = 40, 30, 3
    dev, _ = random_operator(lri, n, m, 0, True, seed=3)
    U, S, V = random_state(np.random.default_rng(4), n, m, r)
    with lri.native.Context(n, m, r) as ctx:
        ctx.set_state(U, S, V)
        ctx.set_operator_lyapunov(dev.A.d, dev.A.e, True, dev.B.d, dev.B.e, False)
        with pytest.raises(ValueError):
            ctx.step_de_bug_adjoint(0.01, 2, U, S, V)  # no quadratic term set
        ctx.set_operator_quadratic(dev.D.dl, dev.D.dd, dev.D.du, True)
        with pytest.raises(ValueError):
            ctx.step_de_bug_adjoint(0.01, 0, U, S, V)  # substeps < 1
        ctx.step_de_bug_adjoint(0.01, 2, U, S, V)


@pytest.mark.parametrize("periodic,q,shape,dt", [(True, 1, (50, 37, 4), 0.01), (False, 2, (300, 200, 16), 0.05)])
def test_ksl_one_step_adjoint_vs_oracle(lri, periodic, q, shape, dt):
    """lri_step_de_ksl_adjoint against oracle.adjoint_de.ksl_primal_de_step_vjp (V0's cotangent on
    the tangent space: the engine's sub-problems take V0'V0 = I)."""
    n, m, r = shape
    say(f"case: matrix-ODE primal KSL adjoint, one step, n={n} m={m} r={r} q={q} periodic={periodic}")
    dev, host = random_operator(lri, n, m, q, periodic, seed=21)
    rng = np.random.default_rng(22)
    U, S, V = random_state(rng, n, m, r)
    cot = [rng.standard_normal(x.shape) for x in (U, S, V)]
    N = lri.rk4_substeps(dev.spectral_bound(2 * np.abs(U @ S @ V.T).max()), dt)
    with context(lri, dev, U, S, V) as ctx:
        Ub, Sb, Vb, (a, b) = ctx.step_de_ksl_adjoint(dt, N, *cot)
    ref = ad.ksl_primal_de_step_vjp(host, U, S, V, dt, N, *cot)
    errs = [rel(Ub, ref[0]), rel(Sb, ref[1]), rel(tangential(V, Vb), tangential(V, ref[2])),
            abs(a - ref[3]) / abs(ref[3]), abs(b - ref[4]) / max(abs(ref[4]), 1e-300)]
    say(f"  N={N}: rel err Ub {errs[0]:.1e} Sb {errs[1]:.1e} Vb {errs[2]:.1e} abar {errs[3]:.1e} bbar {errs[4]:.1e}")
    assert max(errs) <= 1e-9


def test_ksl_checkpointed_burgers_gradient_vs_oracle(lri):
    n, m, r, K, nu = 128, 32, 12, 8, 0.01
    dt, steps = 1e-4, 6
    xi = np.random.default_rng(23).standard_normal((K, m))
    sb = lri.SineBurgers(n, xi, nu, r, seed=4)
    prob = sb.problem((0.0, steps * dt))
    N = lri.de_substeps(prob, dt)
    say(f"case: Burgers sine gradient, primal KSL, n={n} m={m} r={r}, {steps} steps, checkpoint 4")
    alg = lri.ProjectorSplitting(lri.PrimalLieTrotter())
    val, (Ub, Sb, Vb), (abar, _) = lri.solve_gradient(prob, alg, lri.mean_column_loss, dt, checkpoint=4)
    Lap, Grad = ad.periodic_burgers(n, nu)
    host = ad.ScaledQuadratic(sp.csr_matrix(Lap), sp.csr_matrix((m, m)), np.zeros((n, 0)), np.zeros((m, 0)),
                              sp.csr_matrix(Grad))
    u0 = sb.initial()
    rv, rUb, rSb, rVb, rabar, _ = ad.bug_de_solve_gradient(host, u0.U, u0.S, u0.V, dt, N, steps, ad.mean_column_loss,
                                                           checkpoint=4, ksl=True)
    e_xi = rel(sb.pullback(Ub, Sb, Vb), sb.pullback(rUb, rSb, rVb))
    e_nu = abs(abar - rabar) / abs(rabar)
    say(f"  loss {val:.6e} (oracle {rv:.6e}); rel err dL/dxi {e_xi:.1e}, dL/dnu {e_nu:.1e}")
    assert abs(val - rv) <= 1e-10 * abs(rv)
    assert e_xi <= 1e-8 and e_nu <= 1e-8


def ra_context(lri, dev, U, S, V):
    """A context of capacity 2r (the rank-adaptive step's largest new rank)."""
    n, m, r = U.shape[0], V.shape[0], S.shape[0]
    ctx = lri.native.Context(n, m, 2 * r)
    ctx.set_state(U, S, V)
    q = dev.G.shape[1]
    ctx.set_operator_lyapunov(dev.A.d, dev.A.e, dev.A.periodic, dev.B.d, dev.B.e, dev.B.periodic,
                              dev.G if q else None, dev.H if q else None)
    ctx.set_operator_quadratic(dev.D.dl, dev.D.dd, dev.D.du, dev.D.periodic)
    return ctx


@pytest.mark.parametrize("periodic,q,shape,dt", [(True, 1, (50, 37, 4), 0.01), (False, 0, (129, 70, 8), 0.03),
                                                 (True, 2, (300, 200, 16), 0.05)])
def test_ra_one_step_adjoint_vs_oracle(lri, periodic, q, shape, dt):
    """lri_step_de_bug_ra_adjoint against oracle/adjoint_de.ra_bug_de_step_vjp for the
    basis-free loss <C, U1 S1 V1'> + <c, sigma> (each side's cotangents from its own
    factors, whose singular-vector signs may differ); the rank held at rmax = r + 2."""
    n, m, r = shape
    say(f"case: rank-adaptive matrix-ODE adjoint, one step, n={n} m={m} r={r} q={q} periodic={periodic}")
    dev, host = random_operator(lri, n, m, q, periodic, seed=13)
    rng = np.random.default_rng(14)
    U, S, V = random_state(rng, n, m, r)
    N = lri.rk4_substeps(dev.spectral_bound(2 * np.abs(U @ S @ V.T).max()), dt)
    tol, rmax = 1e-12, r + 2
    with ra_context(lri, dev, U, S, V) as ctx:
        r1 = ctx.step_de_bug_ra(dt, N, tol, rmax)
        U1e, S1e, V1e = ctx.get_state()
    U1o, S1o, V1o, r1o = ad.ra_bug_de_step(host, U, S, V, dt, N, tol, rmax)
    assert r1 == r1o == rmax
    e_fwd = rel(U1e @ S1e @ V1e.T, U1o @ S1o @ V1o.T)
    C = rng.standard_normal((n, m))
    cs = rng.standard_normal(r1)

    def cot(U1, S1, V1):
        return C @ V1 @ S1.T, U1.T @ C @ V1 + np.diag(cs), C.T @ U1 @ S1

    with ra_context(lri, dev, U, S, V) as ctx:
        Ub, Sb, Vb, (a, b) = ctx.step_de_bug_ra_adjoint(dt, N, tol, rmax, *cot(U1e, S1e, V1e))
        U2, S2, V2 = ctx.get_state()
    assert np.array_equal(U2, U) and np.array_equal(S2, S) and np.array_equal(V2, V)  # state unchanged
    ref = ad.ra_bug_de_step_vjp(host, U, S, V, dt, N, tol, rmax, *cot(U1o, S1o, V1o))
    errs = [rel(tangential(U, Ub), tangential(U, ref[0])), rel(Sb, ref[1]),
            rel(tangential(V, Vb), tangential(V, ref[2])), abs(a - ref[3]) / abs(ref[3]),
            abs(b - ref[4]) / max(abs(ref[4]), 1e-300)]
    say(f"  N={N} r1={r1}: forward {e_fwd:.1e}; rel err Ub {errs[0]:.1e} Sb {errs[1]:.1e} Vb {errs[2]:.1e} "
        f"abar {errs[3]:.1e} bbar {errs[4]:.1e}")
    assert e_fwd <= 1e-11
    assert max(errs) <= 1e-8


def test_ra_checkpointed_solve_gradient_vs_oracle(lri):
    """solve_gradient with RankAdaptiveUnconventionalAlgorithm on a matrix ODE (the ranks
    of the forward solve held fixed) against the oracle's checkpointed reverse."""
    n, m, r, dt, steps = 60, 45, 3, 0.02, 5
    tol, rmax = 1e-3, 6
    dev, host = random_operator(lri, n, m, 1, True, seed=15)
    U, S, V = random_state(np.random.default_rng(16), n, m, r)
    prob = lri.MatrixDEProblem(dev, lri.SVDLikeRepresentation(U, S, V), (0.0, steps * dt))
    N = lri.de_substeps(prob, dt)
    say(f"case: rank-adaptive matrix-ODE solve gradient n={n} m={m} r0={r} rmax={rmax}, {steps} steps of {N} "
        f"RK4 sub-steps, checkpoint 2")
    val, (Ub, Sb, Vb), (a, b) = lri.solve_gradient(prob, lri.RankAdaptiveUnconventionalAlgorithm(tol, rmax=rmax),
                                                   lri.mean_column_loss, dt, checkpoint=2)
    rv, rUb, rSb, rVb, ra_, rb_ = ad.bug_de_solve_gradient(host, U, S, V, dt, N, steps, ad.mean_column_loss,
                                                           checkpoint=2, ra=(tol, rmax))
    errs = [rel(tangential(U, Ub), tangential(U, rUb)), rel(Sb, rSb), rel(tangential(V, Vb), tangential(V, rVb)),
            abs(a - ra_) / abs(ra_), abs(b - rb_) / max(abs(rb_), 1e-300)]
    say(f"  loss {val:.6e} (oracle {rv:.6e}); rel err Ub {errs[0]:.1e} Sb {errs[1]:.1e} Vb {errs[2]:.1e} "
        f"abar {errs[3]:.1e} bbar {errs[4]:.1e}")
    assert abs(val - rv) <= 1e-10 * abs(rv)
    assert max(errs) <= 1e-8
