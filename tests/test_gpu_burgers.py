"""Parity of the matrix-ODE steps with a quadratic term -(D Y) o Y (viscous Burgers)
against the CPU oracle (needs an MI355X).

The reference's own DE test (test/data_agnostic_approximation.jl:3-58) integrates
Burgers' equation, one parameter sample per column, with BUG, the three KSL
orders and RA(1e-4, rmax=10); it asserts nothing (a smoke test).  Here the same
problem runs on the engine and in the oracle (oracle/dlra_de.py, the dense
right-hand side with sparse stencils) with identical fixed-step RK4 sub-steps.

Tolerance: FP64, relative Frobenius error on Y <= 1e-10 after one step,
<= 1e-9 after ten; projectors <= 1e-8 (RA: on the kept directions only).
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import dlra, dlra_de
from oracle.data import random_factors
from oracle.lra import truncated_svd
from oracle.metrics import proj_dist, rel_err

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_quadratic(lri, n, m, q, periodic, seed):
    rng = np.random.default_rng(seed)
    A = lri.SymTridiagonal(-rng.uniform(20, 40, n), rng.uniform(5, 15, n if periodic else n - 1), periodic)
    B = lri.SymTridiagonal(-rng.uniform(10, 20, m), rng.uniform(2, 6, m - 1))
    D = lri.Tridiagonal(rng.uniform(-3, 3, n), rng.uniform(-1, 1, n), rng.uniform(-3, 3, n), periodic)
    G, H = rng.standard_normal((n, q)), rng.standard_normal((m, q))
    dev = lri.QuadraticOperator(A, B, D, G, H)
    host = dlra_de.QuadraticOperator(A.dense(), B.dense(), G, H, D.dense())
    return dev, host


def run_one(lri, name, dev, u, dt, N, tol=1e-3, rmax=16):
    n, m, r = u.U.shape[0], u.V.shape[0], u.S.shape[0]
    rcap = min(rmax, n, m) if name == "ra" else r
    with lri.native.Context(n, m, rcap) as ctx:
        ctx.set_state(u.U, u.S, u.V)
        q = dev.G.shape[1]
        ctx.set_operator_lyapunov(dev.A.d, dev.A.e, dev.A.periodic, dev.B.d, dev.B.e, dev.B.periodic,
                                  dev.G if q else None, dev.H if q else None)
        ctx.set_operator_quadratic(dev.D.dl, dev.D.dd, dev.D.du, dev.D.periodic)
        nat = lri.native
        if name == "primal":
            ctx.step_de_ksl(nat.PRIMAL, dt, N)
        elif name == "dual":
            ctx.step_de_ksl(nat.DUAL, dt, N)
        elif name == "strang":
            ctx.step_de_ksl(nat.STRANG, dt, N)
        elif name == "bug":
            ctx.step_de_bug(dt, N)
        else:
            ctx.step_de_bug_ra(dt, N, tol, rmax)
        return ctx.get_state()


def oracle_one(name, host, u, dt, N, tol=1e-3, rmax=16):
    if name == "primal":
        return dlra_de.ksl_primal_de_step(host, u.U, u.S, u.V, 0.0, dt, N)
    if name == "dual":
        return dlra_de.ksl_dual_de_step(host, u.U, u.S, u.V, 0.0, dt, N)
    if name == "strang":
        a = dlra_de.ksl_primal_de_step(host, u.U, u.S, u.V, 0.0, dt / 2, N)
        return dlra_de.ksl_dual_de_step(host, *a, dt / 2, dt / 2, N)
    if name == "bug":
        return dlra_de.bug_de_step(host, u.U, u.S, u.V, 0.0, dt, N)
    return dlra_de.rank_adaptive_bug_de_step(host, u.U, u.S, u.V, 0.0, dt, N, tol, rmax)[:3]


ALGS = ["primal", "dual", "strang", "bug", "ra"]


@pytest.mark.parametrize("name", ALGS)
@pytest.mark.parametrize("periodic", [True, False])
def test_one_step_vs_oracle(lri, name, periodic):
    n, m, r = 50, 37, 4
    say(f"case: one quadratic-operator step {name}, periodic={periodic}, n={n} m={m} r={r}")
    dev, host = random_quadratic(lri, n, m, 1, periodic, seed=1)
    u = random_factors(np.random.default_rng(2), n, m, r)
    dt = 0.02
    ymax = np.abs(u.to_matrix()).max()
    N = lri.rk4_substeps(dev.spectral_bound(2 * ymax), dt / 2 if name == "strang" else dt)
    got = run_one(lri, name, dev, u, dt, N)
    ref = oracle_one(name, host, u, dt, N)
    assert got[1].shape == ref[1].shape
    e = rel_err(*got, *ref)
    say(f"  N={N}: rel err {e:.2e}, proj U {proj_dist(got[0], ref[0]):.2e}")
    assert e <= 1e-10
    if name != "ra":
        assert proj_dist(got[0], ref[0]) <= 1e-8 and proj_dist(got[2], ref[2]) <= 1e-8


def burgers_case(lri, n=1000, msq=20, r=5):
    """test/data_agnostic_approximation.jl:4-46: n = 1000 cells of [0, pi), nu = 0.005,
    m = 20^2 samples, rank-5 truncated SVD of the initial samples."""
    Lap, Grad = dlra_de.burgers_matrices(n, 0.005)
    m = msq * msq
    host = dlra_de.QuadraticOperator(sp.csr_matrix(Lap), sp.csr_matrix((m, m)), np.zeros((n, 1)),
                                     np.zeros((m, 1)), sp.csr_matrix(Grad))
    dev = lri.burgers_operator(n, m, 0.005)
    u0 = truncated_svd(dlra_de.burgers_initial(n, msq), r)
    return dev, host, u0


def test_burgers_operator_matches_reference_stencils(lri):
    dev, host, u0 = burgers_case(lri, n=200, msq=4)
    Y = u0.to_matrix()
    assert np.allclose(dev(Y), host(Y), rtol=0, atol=1e-9 * np.abs(host(Y)).max())


@pytest.mark.parametrize("alg", ["bug", "primal", "dual", "strang", "ra"])
def test_reference_burgers_problem(lri, alg):
    """The reference's Burgers DE test on t in [0, 0.1] (10 of its 100 steps of
    dt = 1e-2) for each of its five solvers, engine vs oracle."""
    dev, host, u0 = burgers_case(lri)
    dt, tf = 1e-2, 0.1
    ymax = np.abs(u0.to_matrix()).max()
    N = lri.rk4_substeps(dev.spectral_bound(2 * ymax), dt / 2 if alg == "strang" else dt)
    say(f"case: reference Burgers problem, {alg}, n=1000 m=400 r=5, {N} RK4 sub-steps")
    A = {"bug": lri.UnconventionalAlgorithm(), "primal": lri.ProjectorSplitting(lri.PrimalLieTrotter()),
         "dual": lri.ProjectorSplitting(lri.DualLieTrotter()), "strang": lri.ProjectorSplitting(lri.Strang()),
         "ra": lri.RankAdaptiveUnconventionalAlgorithm(1e-4, rmax=10)}[alg]
    O = {"bug": dlra.UnconventionalAlgorithm(), "primal": dlra.ProjectorSplitting(dlra.PrimalLieTrotter()),
         "dual": dlra.ProjectorSplitting(dlra.DualLieTrotter()), "strang": dlra.ProjectorSplitting(dlra.Strang()),
         "ra": dlra.RankAdaptiveUnconventionalAlgorithm(1e-4, rmax=10)}[alg]
    y0 = lri.SVDLikeRepresentation(u0.U, u0.S, u0.V)
    sol = lri.solve(lri.MatrixDEProblem(dev, y0, (0.0, tf)), A, dt, substeps=N)
    ref = dlra_de.solve_de(dlra_de.MatrixDEProblem(host, u0, (0.0, tf)), O, dt, N)
    assert len(sol.Y) == len(ref.Y) == 11
    for k in (1, 10):
        a, b = sol.Y[k], ref.Y[k]
        assert a.S.shape == b.S.shape
        e = rel_err(a.U, a.S, a.V, b.U, b.S, b.V)
        say(f"  step {k}: rank {a.S.shape[0]}, rel err {e:.2e}")
        assert e <= (1e-10 if k == 1 else 1e-9)


def test_quadratic_rank_limits(lri):
    n, m = 40, 40
    dev, _ = random_quadratic(lri, n, m, 0, True, seed=3)
    u = random_factors(np.random.default_rng(4), n, m, 3)
    with lri.native.Context(n, m, 40) as ctx:
        ctx.set_state(np.linalg.qr(np.random.default_rng(5).standard_normal((n, 33)))[0], np.eye(33),
                      np.linalg.qr(np.random.default_rng(6).standard_normal((m, 33)))[0])
        ctx.set_operator_lyapunov(dev.A.d, dev.A.e, True, dev.B.d, dev.B.e, False)
        ctx.set_operator_quadratic(dev.D.dl, dev.D.dd, dev.D.du, True)
        with pytest.raises(ValueError):
            ctx.step_de_bug(0.01, 2)  # rank 33 > 32
        ctx.set_state(u.U, u.S, u.V)
        ctx.step_de_bug(0.01, 2)
        ctx.set_state(np.linalg.qr(np.random.default_rng(7).standard_normal((n, 17)))[0], np.eye(17),
                      np.linalg.qr(np.random.default_rng(8).standard_normal((m, 17)))[0])
        with pytest.raises(ValueError):
            ctx.step_de_bug_ra(0.01, 2, 1e-3, 20)  # 2 r = 34 > 32
