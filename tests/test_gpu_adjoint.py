"""The engine's reverse BUG step (lri_step_bug_adjoint) against the CPU reference
gradient (oracle/adjoint.py) and finite differences (needs an MI355X).

Losses are basis-free (L = <C, U S V'>), so the gradients with respect to the
initial factors do not depend on the QR sign conventions, which differ between
the engine (CholeskyQR, positive R diagonal) and the oracle (Householder).
Tolerance: FP64, relative error of the gradient <= 1e-9; finite differences
(central, step 1e-6) to 1e-6.
"""
import sys

import numpy as np
import pytest

from oracle import adjoint, dlra
from oracle.data import random_factors

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def linear_loss(C):
    def lg(U, S, V):
        return float(np.sum(C * (U @ S @ V.T))), C @ V @ S.T, U.T @ C @ V, C.T @ U @ S
    return lg


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("shape", [(64, 48, 4), (200, 136, 8), (37, 53, 3), (512, 384, 32)])
def test_one_step_gradient_vs_oracle(lri, shape):
    n, m, r = shape
    say(f"case: BUG adjoint one step n={n} m={m} r={r}")
    rng = np.random.default_rng(1)
    u = random_factors(rng, n, m, r, decay=0.3)
    A0 = rng.standard_normal((n, m))
    A1 = A0 + 5e-2 * rng.standard_normal((n, m))
    C = rng.standard_normal((n, m))
    lg = linear_loss(C)
    with lri.native.Context(n, m, r) as ctx:
        ctx.reserve(2, lri.native.F64)
        ctx.upload(0, np.asfortranarray(A0))
        ctx.upload(1, np.asfortranarray(A1))
        ctx.set_state(u.U, u.S, u.V)
        ctx.step_bug(0, 1)
        _, Ub1, Sb1, Vb1 = lg(*ctx.get_state())
        ctx.set_state(u.U, u.S, u.V)
        g = ctx.step_bug_adjoint(0, 1, Ub1, Sb1, Vb1)
        U, S, V = ctx.get_state()
        assert np.array_equal(U, u.U) and np.array_equal(S, u.S)  # the state is left as it was
    o1 = dlra.bug_step(u.U, u.S, u.V, A1 - A0)
    _, oUb, oSb, oVb = lg(*o1)
    go = adjoint.bug_step_vjp(u.U, u.S, u.V, A1 - A0, oUb, oSb, oVb)
    errs = [rel(a, b) for a, b in zip(g, go[:3])]
    say(f"  gradient rel err U {errs[0]:.2e} S {errs[1]:.2e} V {errs[2]:.2e}")
    assert max(errs) <= 1e-9


def test_checkpointed_solve_gradient(lri):
    """solve_gradient over 12 BUG steps (checkpoints every 5) against the oracle's
    checkpointed gradient and a central finite difference of the engine's own solve."""
    n, m, r, T = 160, 120, 5, 12
    rng = np.random.default_rng(2)
    u = random_factors(rng, n, m, r, decay=0.4)
    snaps = [u.to_matrix()]
    for _ in range(T):
        snaps.append(snaps[-1] + 2e-2 * rng.standard_normal((n, m)))
    C = rng.standard_normal((n, m))
    lg = linear_loss(C)
    y0 = lri.SVDLikeRepresentation(u.U, u.S, u.V)
    prob = lri.MatrixDataProblem(snaps, y0)
    value, g = lri.solve_gradient(prob, lri.UnconventionalAlgorithm(), lg, checkpoint=5)
    ov, oUb, oSb, oVb = adjoint.bug_solve_gradient(u.U, u.S, u.V, snaps, lg, checkpoint=5)
    errs = [rel(a, b) for a, b in zip(g, (oUb, oSb, oVb))]
    say(f"  loss {value:.6e} (oracle {ov:.6e}); gradient rel err {max(errs):.2e}")
    assert abs(value - ov) <= 1e-10 * abs(ov)
    assert max(errs) <= 1e-9
    D = [rng.standard_normal(x.shape) for x in (u.U, u.S, u.V)]
    e = 1e-6

    def run(sign):
        y = lri.SVDLikeRepresentation(u.U + sign * e * D[0], u.S + sign * e * D[1], u.V + sign * e * D[2])
        sol = lri.solve(lri.MatrixDataProblem(snaps, y), lri.UnconventionalAlgorithm())
        Y = sol.Y[-1]
        return lg(Y.U, Y.S, Y.V)[0]

    fd = (run(1.0) - run(-1.0)) / (2 * e)
    an = sum(np.sum(a * d) for a, d in zip(g, D))
    say(f"  directional derivative {an:.8e}, finite difference {fd:.8e}")
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd))


def test_gradient_against_cpu_forward_mode(lri):
    """The engine's reverse-mode gradient along random directions against the CPU
    forward-mode reference (tangents propagated through the oracle's steps)."""
    n, m, r, T = 120, 90, 4, 8
    rng = np.random.default_rng(3)
    u = random_factors(rng, n, m, r, decay=0.4)
    snaps = [u.to_matrix()]
    for _ in range(T):
        snaps.append(snaps[-1] + 3e-2 * rng.standard_normal((n, m)))
    C = rng.standard_normal((n, m))
    lg = linear_loss(C)
    y0 = lri.SVDLikeRepresentation(u.U, u.S, u.V)
    _, g = lri.solve_gradient(lri.MatrixDataProblem(snaps, y0), lri.UnconventionalAlgorithm(), lg, checkpoint=3)
    for trial in range(3):
        D = [rng.standard_normal(x.shape) for x in (u.U, u.S, u.V)]
        U, S, V = u.U, u.S, u.V
        Ud, Sd, Vd = D
        for k in range(T):
            dA = snaps[k + 1] - snaps[k]
            Ud, Sd, Vd = adjoint.bug_step_jvp(U, S, V, dA, Ud, Sd, Vd, np.zeros_like(dA))
            U, S, V = dlra.bug_step(U, S, V, dA)
        fwd = np.sum(C * (Ud @ S @ V.T + U @ Sd @ V.T + U @ S @ Vd.T))
        rev = sum(np.sum(a * d) for a, d in zip(g, D))
        say(f"  direction {trial}: forward mode {fwd:.10e}, engine reverse mode {rev:.10e}")
        assert abs(fwd - rev) <= 1e-10 * max(1.0, abs(fwd))


@pytest.mark.parametrize("shape", [(64, 48, 4), (300, 200, 16)])
def test_ksl_one_step_gradient_vs_oracle(lri, shape):
    n, m, r = shape
    say(f"case: KSL adjoint one step n={n} m={m} r={r}")
    rng = np.random.default_rng(4)
    u = random_factors(rng, n, m, r, decay=0.3)
    A0 = rng.standard_normal((n, m))
    A1 = A0 + 5e-2 * rng.standard_normal((n, m))
    lg = linear_loss(rng.standard_normal((n, m)))
    with lri.native.Context(n, m, r) as ctx:
        ctx.reserve(2, lri.native.F64)
        ctx.upload(0, np.asfortranarray(A0))
        ctx.upload(1, np.asfortranarray(A1))
        ctx.set_state(u.U, u.S, u.V)
        ctx.step_ksl(lri.native.PRIMAL, 0, 1)
        _, Ub1, Sb1, Vb1 = lg(*ctx.get_state())
        ctx.set_state(u.U, u.S, u.V)
        g = ctx.step_ksl_adjoint(0, 1, Ub1, Sb1, Vb1)
    o1 = dlra.ksl_primal_step(u.U, u.S, u.V, A1 - A0)
    _, oUb, oSb, oVb = lg(*o1)
    go = adjoint.ksl_primal_step_vjp(u.U, u.S, u.V, A1 - A0, oUb, oSb, oVb)
    errs = [rel(a, b) for a, b in zip(g, go[:3])]
    say(f"  gradient rel err U {errs[0]:.2e} S {errs[1]:.2e} V {errs[2]:.2e}")
    assert max(errs) <= 1e-9


def test_ksl_checkpointed_solve_gradient(lri):
    n, m, r, T = 150, 110, 5, 10
    rng = np.random.default_rng(5)
    u = random_factors(rng, n, m, r, decay=0.4)
    snaps = [u.to_matrix()]
    for _ in range(T):
        snaps.append(snaps[-1] + 2e-2 * rng.standard_normal((n, m)))
    lg = linear_loss(rng.standard_normal((n, m)))
    y0 = lri.SVDLikeRepresentation(u.U, u.S, u.V)
    alg = lri.ProjectorSplitting(lri.PrimalLieTrotter())
    value, g = lri.solve_gradient(lri.MatrixDataProblem(snaps, y0), alg, lg, checkpoint=4)
    ov, *og = adjoint.bug_solve_gradient(u.U, u.S, u.V, snaps, lg, checkpoint=4, ksl=True)
    errs = [rel(a, b) for a, b in zip(g, og)]
    say(f"  KSL loss {value:.6e} (oracle {ov:.6e}); gradient rel err {max(errs):.2e}")
    assert abs(value - ov) <= 1e-10 * abs(ov) and max(errs) <= 1e-9
    D = [rng.standard_normal(x.shape) for x in (u.U, u.S, u.V)]
    e = 1e-6

    def run(sign):
        y = lri.SVDLikeRepresentation(u.U + sign * e * D[0], u.S + sign * e * D[1], u.V + sign * e * D[2])
        Y = lri.solve(lri.MatrixDataProblem(snaps, y), alg).Y[-1]
        return lg(Y.U, Y.S, Y.V)[0]

    fd = (run(1.0) - run(-1.0)) / (2 * e)
    an = sum(np.sum(a * d) for a, d in zip(g, D))
    say(f"  directional derivative {an:.8e}, finite difference {fd:.8e}")
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd))
