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


@pytest.mark.parametrize("shape", [(64, 48, 4), (37, 53, 3), (300, 200, 16), (520, 390, 40)])
def test_ksl_dual_one_step_gradient_vs_oracle(lri, shape):
    """The dual KSL reverse step (the primal one on the context's transposed view)."""
    n, m, r = shape
    say(f"case: dual KSL adjoint one step n={n} m={m} r={r}")
    rng = np.random.default_rng(6)
    u = random_factors(rng, n, m, r, decay=0.3)
    A0 = rng.standard_normal((n, m))
    A1 = A0 + 5e-2 * rng.standard_normal((n, m))
    lg = linear_loss(rng.standard_normal((n, m)))
    with lri.native.Context(n, m, r) as ctx:
        ctx.reserve(2, lri.native.F64)
        ctx.upload(0, np.asfortranarray(A0))
        ctx.upload(1, np.asfortranarray(A1))
        ctx.set_state(u.U, u.S, u.V)
        ctx.step_ksl(lri.native.DUAL, 0, 1)
        _, Ub1, Sb1, Vb1 = lg(*ctx.get_state())
        ctx.set_state(u.U, u.S, u.V)
        g = ctx.step_ksl_dual_adjoint(0, 1, Ub1, Sb1, Vb1)
        U, S, V = ctx.get_state()
        assert np.array_equal(U, u.U) and np.array_equal(S, u.S) and np.array_equal(V, u.V)
        # the context is back in its own orientation: a forward step matches a fresh one
        ctx.step_ksl(lri.native.DUAL, 0, 1)
        again = ctx.get_state()
    o1 = dlra.ksl_dual_step(u.U, u.S, u.V, A1 - A0)
    _, oUb, oSb, oVb = lg(*o1)
    go = adjoint.ksl_dual_step_vjp(u.U, u.S, u.V, A1 - A0, oUb, oSb, oVb)
    errs = [rel(a, b) for a, b in zip(g, go[:3])]
    say(f"  gradient rel err U {errs[0]:.2e} S {errs[1]:.2e} V {errs[2]:.2e}")
    assert max(errs) <= 1e-9
    X1 = again[0] @ again[1] @ again[2].T
    assert rel(X1, o1[0] @ o1[1] @ o1[2].T) <= 1e-10


@pytest.mark.parametrize("shape", [(64, 48, 4), (300, 200, 16)])
def test_ksl_strang_one_step_gradient_vs_oracle(lri, shape):
    n, m, r = shape
    say(f"case: Strang KSL adjoint one step n={n} m={m} r={r}")
    rng = np.random.default_rng(7)
    u = random_factors(rng, n, m, r, decay=0.3)
    A0 = rng.standard_normal((n, m))
    Am = A0 + 3e-2 * rng.standard_normal((n, m))
    A1 = Am + 3e-2 * rng.standard_normal((n, m))
    lg = linear_loss(rng.standard_normal((n, m)))
    with lri.native.Context(n, m, r) as ctx:
        ctx.reserve(3, lri.native.F64)
        for k, A in enumerate((A0, Am, A1)):
            ctx.upload(k, np.asfortranarray(A))
        ctx.set_state(u.U, u.S, u.V)
        ctx.step_ksl_strang(0, 1, 2)
        _, Ub1, Sb1, Vb1 = lg(*ctx.get_state())
        ctx.set_state(u.U, u.S, u.V)
        g = ctx.step_ksl_strang_adjoint(0, 1, 2, Ub1, Sb1, Vb1)
        U, S, V = ctx.get_state()
        assert np.array_equal(U, u.U) and np.array_equal(S, u.S) and np.array_equal(V, u.V)
    o1 = dlra.ksl_dual_step(*dlra.ksl_primal_step(u.U, u.S, u.V, Am - A0), A1 - Am)
    _, oUb, oSb, oVb = lg(*o1)
    go = adjoint.ksl_strang_step_vjp(u.U, u.S, u.V, Am - A0, A1 - Am, oUb, oSb, oVb)
    errs = [rel(a, b) for a, b in zip(g, go[:3])]
    say(f"  gradient rel err U {errs[0]:.2e} S {errs[1]:.2e} V {errs[2]:.2e}")
    assert max(errs) <= 1e-9


def test_ksl_dual_checkpointed_solve_gradient(lri):
    n, m, r, T = 140, 170, 5, 9
    rng = np.random.default_rng(8)
    u = random_factors(rng, n, m, r, decay=0.4)
    snaps = [u.to_matrix()]
    for _ in range(T):
        snaps.append(snaps[-1] + 2e-2 * rng.standard_normal((n, m)))
    lg = linear_loss(rng.standard_normal((n, m)))
    y0 = lri.SVDLikeRepresentation(u.U, u.S, u.V)
    alg = lri.ProjectorSplitting(lri.DualLieTrotter())
    value, g = lri.solve_gradient(lri.MatrixDataProblem(snaps, y0), alg, lg, checkpoint=4)
    ov, *og = adjoint.bug_solve_gradient(u.U, u.S, u.V, snaps, lg, checkpoint=4, ksl="dual")
    errs = [rel(a, b) for a, b in zip(g, og)]
    say(f"  dual KSL loss {value:.6e} (oracle {ov:.6e}); gradient rel err {max(errs):.2e}")
    assert abs(value - ov) <= 1e-10 * abs(ov) and max(errs) <= 1e-9


# ---------------------------------------------------------------------------- rank-adaptive step


def sigma_loss(C, c):
    """<C, U S V'> + <c, diag S> (S diagonal after a rank-adaptive step): basis-free."""
    def lg(U, S, V):
        k = S.shape[0]
        w = np.zeros(k)
        w[:min(k, len(c))] = c[:k]
        return (float(np.sum(C * (U @ S @ V.T))) + float(w @ np.diag(S)), C @ V @ S.T, U.T @ C @ V + np.diag(w),
                C.T @ U @ S)
    return lg


def ra_case(seed, n, m, r):
    rng = np.random.default_rng(seed)
    u = random_factors(rng, n, m, r, decay=0.5)
    A0 = u.to_matrix()
    dA = 0.3 * rng.standard_normal((n, 2 * r)) @ rng.standard_normal((2 * r, m)) / np.sqrt(m)
    return rng, u, A0, A0 + dA


@pytest.mark.parametrize("shape", [(64, 48, 4, 6), (200, 136, 8, 12), (37, 53, 3, 2), (300, 260, 16, 24)])
def test_ra_one_step_gradient_vs_oracle(lri, shape):
    """rankadaptive_unconventional_step! reversed (SVD rule, augmented-basis QR rule) against
    oracle.adjoint.ra_bug_step_vjp; rmax caps the new rank (grown, or shrunk below r)."""
    n, m, r, rmax = shape
    tol = 1e-12
    say(f"case: rank-adaptive BUG adjoint one step n={n} m={m} r={r} rmax={rmax}")
    rng, u, A0, A1 = ra_case(3, n, m, r)
    C = rng.standard_normal((n, m))
    lg = sigma_loss(C, rng.standard_normal(2 * r))
    with lri.native.Context(n, m, max(r, min(rmax, n, m, 2 * r))) as ctx:
        ctx.reserve(2, lri.native.F64)
        ctx.upload(0, np.asfortranarray(A0))
        ctx.upload(1, np.asfortranarray(A1))
        ctx.set_state(u.U, u.S, u.V)
        r1 = ctx.step_bug_ra(0, 1, tol, rmax)
        _, Ub1, Sb1, Vb1 = lg(*ctx.get_state())
        ctx.set_state(u.U, u.S, u.V)
        g = ctx.step_bug_ra_adjoint(0, 1, tol, rmax, Ub1, Sb1, Vb1)
        U, S, V = ctx.get_state()
        assert np.array_equal(U, u.U) and np.array_equal(S, u.S)
    Ua, Sa, Va, ra_ = adjoint.ra_bug_step_aug(u.U, u.S, u.V, A1 - A0, tol, rmax)
    assert ra_ == r1
    _, oUb, oSb, oVb = lg(Ua, Sa, Va)
    go = adjoint.ra_bug_step_vjp(u.U, u.S, u.V, A1 - A0, tol, rmax, oUb, oSb, oVb)
    errs = [rel(a, b) for a, b in zip(g, go[:3])]
    say(f"  r1={r1}: gradient rel err U {errs[0]:.2e} S {errs[1]:.2e} V {errs[2]:.2e}")
    assert max(errs) <= 1e-9


def test_ra_checkpointed_solve_gradient(lri):
    """solve_gradient with RankAdaptiveUnconventionalAlgorithm over 8 steps (checkpoints every
    3), the ranks of the forward solve held, against the oracle's chained reverse steps and a
    central difference of the engine's solve along a direction tangent to the bases."""
    n, m, r, T, rmax, tol = 120, 90, 4, 8, 10, 1e-3
    rng = np.random.default_rng(4)
    u = random_factors(rng, n, m, r, decay=0.5)
    snaps = [u.to_matrix()]
    for _ in range(T):
        snaps.append(snaps[-1] + 0.05 * rng.standard_normal((n, 3)) @ rng.standard_normal((3, m)) / np.sqrt(m)
                     + 1e-3 * rng.standard_normal((n, m)))
    C = rng.standard_normal((n, m))
    lg = linear_loss(C)
    alg = lri.RankAdaptiveUnconventionalAlgorithm(tol, rmax=rmax)
    prob = lri.MatrixDataProblem(snaps, lri.SVDLikeRepresentation(u.U, u.S, u.V))
    value, g = lri.solve_gradient(prob, alg, lg, checkpoint=3)
    states = [(u.U, u.S, u.V)]
    for k in range(T):
        states.append(adjoint.ra_bug_step_aug(*states[-1], snaps[k + 1] - snaps[k], tol, rmax)[:3])
    ov, oUb, oSb, oVb = lg(*states[-1])
    for k in range(T - 1, -1, -1):
        oUb, oSb, oVb, _, _ = adjoint.ra_bug_step_vjp(*states[k], snaps[k + 1] - snaps[k], tol, rmax, oUb, oSb, oVb)
    ranks = [s[1].shape[0] for s in states]
    errs = [rel(a, b) for a, b in zip(g, (oUb, oSb, oVb))]
    say(f"case: rank-adaptive checkpointed gradient, ranks {ranks}: rel err U {errs[0]:.2e} S {errs[1]:.2e} "
        f"V {errs[2]:.2e}")
    assert abs(value - ov) <= 1e-10 * abs(ov)
    assert max(errs) <= 1e-8
    G, W = rng.standard_normal((n, r)), rng.standard_normal((r, r))
    dU = G - u.U @ (u.U.T @ G) + u.U @ (W - W.T)
    dS = rng.standard_normal((r, r))

    def L(t):
        y0 = lri.SVDLikeRepresentation(u.U + t * dU, u.S + t * dS, u.V)
        y = lri.solve(lri.MatrixDataProblem(snaps, y0), alg, save_everystep=False).Y[-1]
        return lg(y.U, y.S, y.V)[0]

    eps = 1e-6
    fd = (L(eps) - L(-eps)) / (2 * eps)
    an = float(np.sum(g[0] * dU) + np.sum(g[1] * dS))
    say(f"  directional derivative: adjoint {an:.10e}, central difference {fd:.10e}")
    assert abs(an - fd) <= 1e-6 * max(1.0, abs(fd))


def test_adjoint_with_device_cotangents(lri):
    """The reverse step reads and writes cotangents in device memory as well (hipMemcpyDefault):
    the same result as with host arrays, bit for bit.  The device buffers come from the HIP
    runtime the library itself runs on."""
    import ctypes

    lri.native.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime already loaded by the library
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    H2D, D2H = 1, 2

    n, m, r = 96, 80, 5
    rng = np.random.default_rng(9)
    u = random_factors(rng, n, m, r, decay=0.3)
    A0 = rng.standard_normal((n, m))
    A1 = A0 + 5e-2 * rng.standard_normal((n, m))
    cot = [np.asfortranarray(rng.standard_normal(s)) for s in ((n, r), (r, r), (m, r))]
    ptrs = []
    try:
        for x in cot:
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), x.nbytes) == 0
            ptrs.append(p)
            assert hip.hipMemcpy(p, x.ctypes.data, x.nbytes, H2D) == 0
        with lri.native.Context(n, m, r) as ctx:
            ctx.reserve(2, lri.native.F64)
            ctx.upload(0, np.asfortranarray(A0))
            ctx.upload(1, np.asfortranarray(A1))
            ctx.set_state(u.U, u.S, u.V)
            host = ctx.step_bug_adjoint(0, 1, *cot)
            ctx.step_adjoint_ptr("bug", 0, 1, ptrs[0].value, n, ptrs[1].value, r, ptrs[2].value, m)
            ctx.synchronize()
        for p, x, h in zip(ptrs, cot, host):
            got = np.empty_like(x)
            assert hip.hipMemcpy(got.ctypes.data, p, x.nbytes, D2H) == 0
            assert np.array_equal(got, h)
    finally:
        for p in ptrs:
            hip.hipFree(p)
