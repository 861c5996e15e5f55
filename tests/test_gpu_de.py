"""Parity of the matrix-ODE (MatrixDEProblem) steps with the CPU oracle (needs an MI355X).

The right-hand side is F(Y) = A Y + Y B' + G H' with A, B symmetric tridiagonal
(Dirichlet or periodic) and a rank-q forcing; every sub-problem runs N classical
RK4 steps, on the GPU in the sub-problem's eigenbasis, in the oracle
(oracle/dlra_de.py) directly on the dense sub-problem.

Tolerance: FP64 throughout, relative Frobenius error on Y <= 1e-10 and
projector distance <= 1e-9.  The first test needs no oracle: with A = a I and
B = b I each integrator's result is Y0 times a product of RK4 stability
polynomials (see scalar_factor).
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import dlra, dlra_de
from oracle.data import random_factors
from oracle.metrics import proj_dist, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class SparseLyapunov:
    """F(Y) = A Y + Y B' + G H' with sparse A, B (host side of the oracle at larger n)."""

    def __init__(self, A, B, G, H):
        self.A, self.B, self.G, self.H = sp.csr_matrix(A), sp.csr_matrix(B), G, H

    def __call__(self, Y, t=0.0):
        return self.A @ Y + (self.B @ Y.T).T + self.G @ self.H.T


def tri(lri, n, scale, periodic, rng=None):
    if rng is None:
        return lri.laplacian_1d(n, scale, periodic)
    # variable coefficients
    d = -scale * (2.0 + rng.uniform(0.0, 1.0, n))
    e = scale * rng.uniform(0.5, 1.0, n if periodic else n - 1)
    return lri.SymTridiagonal(d, e, periodic)


def operator_pair(lri, n, m, q, periodic=False, variable=False, seed=0, sa=60.0, sb=40.0):
    rng = np.random.default_rng(seed)
    A = tri(lri, n, sa, periodic, rng if variable else None)
    B = tri(lri, m, sb, periodic, rng if variable else None)
    G = rng.standard_normal((n, q)) if q else None
    H = rng.standard_normal((m, q)) if q else None
    dev = lri.LyapunovOperator(A, B, G, H)
    Gd = G if q else np.zeros((n, 1))
    Hd = H if q else np.zeros((m, 1))
    if n * m <= 200 * 200:
        host = dlra_de.LyapunovOperator(A.dense(), B.dense(), Gd, Hd)
    else:
        host = SparseLyapunov(A.dense(), B.dense(), Gd, Hd)
    return dev, host


def run_one(lri, name, dev, u, dt, N, tol=1e-3, rmax=64, resident=False):
    n, m, r = u.U.shape[0], u.V.shape[0], u.S.shape[0]
    rcap = min(2 * r, 64, n, m) if name == "ra" else r
    ctx = lri.native.Context(n, m, rcap)
    try:
        ctx.set_state(u.U, u.S, u.V)
        q = dev.G.shape[1]
        ctx.set_operator_lyapunov(dev.A.d, dev.A.e, dev.A.periodic, dev.B.d, dev.B.e, dev.B.periodic,
                                  dev.G if q else None, dev.H if q else None)
        if resident:
            ctx.set_resident_eigenbasis(True)
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
    finally:
        ctx.close()


def oracle_one(name, host, u, dt, N, tol=1e-3, rmax=64):
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


def assert_close(got, ref, tol=TOL, proj_tol=1e-9):
    e = rel_err(*got, *ref)
    pu = proj_dist(got[0], ref[0])
    pv = proj_dist(got[2], ref[2])
    say(f"  rel err {e:.3e}, proj U {pu:.2e}, proj V {pv:.2e}")
    assert e <= tol
    assert pu <= proj_tol and pv <= proj_tol


ALGS = ["primal", "dual", "strang", "bug", "ra"]


def rk4_poly(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def scalar_factor(name, z, N):
    """Y1 / Y0 for F(Y) = c Y, z = h c: BUG runs the K, L and S flows forward
    (R(z)^N); KSL runs K and L forward and S backward (R(z)^2 R(-z) per sub-step),
    Strang does that twice with half the step."""
    if name == "bug":
        return rk4_poly(z) ** N
    if name in ("primal", "dual"):
        return (rk4_poly(z) ** 2 * rk4_poly(-z)) ** N
    return (rk4_poly(z / 2) ** 2 * rk4_poly(-z / 2)) ** (2 * N)


@pytest.mark.parametrize("name", ALGS[:4])
def test_scalar_operator_closed_form(lri, name):
    """A = a I, B = b I, no forcing: Y1 = f Y0 with the RK4 stability polynomial products above."""
    n, m, r = 40, 36, 3
    say(f"case: scalar operator closed form, {name}, n={n} m={m} r={r}")
    rng = np.random.default_rng(5)
    u = random_factors(rng, n, m, r)
    a, b, dt, N = -3.0, -1.5, 0.2, 7
    dev = lri.LyapunovOperator(lri.SymTridiagonal(a * np.ones(n), np.zeros(n - 1)),
                               lri.SymTridiagonal(b * np.ones(m), np.zeros(m - 1)))
    got = run_one(lri, name, dev, u, dt, N)
    ref = (u.U, u.S * scalar_factor(name, dt / N * (a + b), N), u.V)
    assert_close(got, ref, 1e-12)


@pytest.mark.parametrize("name", ALGS)
@pytest.mark.parametrize("periodic,q,variable,resident", [(False, 1, False, False), (True, 2, True, False),
                                                          (False, 0, True, False), (False, 1, False, True),
                                                          (True, 2, False, True)])
def test_one_step_vs_oracle(lri, name, periodic, q, variable, resident):
    """resident: the state kept in the operators' analytic eigenbasis (lri_set_resident_eigenbasis)."""
    n, m, r = 50, 37, 4
    say(f"case: one DE step {name}, periodic={periodic} q={q} variable={variable} resident={resident}, "
        f"n={n} m={m} r={r}")
    dev, host = operator_pair(lri, n, m, q, periodic, variable, seed=1)
    u = random_factors(np.random.default_rng(2), n, m, r)
    dt = 0.05
    N = lri.rk4_substeps(dev.spectral_bound(), dt / 2 if name == "strang" else dt)
    got = run_one(lri, name, dev, u, dt, N, resident=resident)
    ref = oracle_one(name, host, u, dt, N)
    assert got[1].shape == ref[1].shape
    assert_close(got, ref)


@pytest.mark.parametrize("n,m", [(5000, 24), (9000, 20), (64, 4100)])
def test_tall_sizes(lri, n, m):
    """Column lengths served by the 8- and 16-rows-per-thread RK4 kernels."""
    r = 3
    say(f"case: tall BUG step n={n} m={m} r={r}")
    dev, host = operator_pair(lri, n, m, 1, False, True, seed=3, sa=0.5, sb=0.4)
    u = random_factors(np.random.default_rng(4), n, m, r)
    dt = 1.0
    N = lri.rk4_substeps(dev.spectral_bound(), dt)
    got = run_one(lri, "bug", dev, u, dt, N)
    ref = oracle_one("bug", host, u, dt, N)
    assert_close(got, ref)


def test_resident_eigenbasis_solve(lri):
    """Ten BUG and three rank-adaptive steps of the heat equation with forcing, n = m = 300, with the
    state resident in the eigenbasis against the same steps in the original basis (1e-11), the oracle
    (1e-10), a trajectory saved while resident, and leaving the eigenbasis for a data-path call.
    The operator scale 0.05 n^2 keeps the problem well conditioned: with the full n^2 scale and
    forcing, two evaluations of the oracle whose operator products differ only in summation order
    (sparse and dense) already disagree at 5e-5 after one step, so no comparison at 1e-12 means
    anything there."""
    n = m = 300
    r, dt = 8, 2e-4
    dev, host = operator_pair(lri, n, m, 1, False, False, seed=8, sa=0.05 * n * n, sb=0.05 * m * m)
    u = random_factors(np.random.default_rng(9), n, m, r, decay=0.5)
    N = lri.rk4_substeps(dev.spectral_bound(), dt)
    say(f"case: resident eigenbasis, BUG and RA, n={n} r={r}, {N} RK4 sub-steps per step")
    q = dev.G.shape[1]
    # rank-adaptive: three steps; the kept singular values sit at tol, where later steps let two
    # evaluations that differ only in rounding drift apart (the oracle against itself as well)
    for name, steps in (("bug", 10), ("ra", 3)):
        got = {}
        for resident in (False, True):
            with lri.native.Context(n, m, 32) as ctx:
                ctx.set_operator_lyapunov(dev.A.d, dev.A.e, dev.A.periodic, dev.B.d, dev.B.e, dev.B.periodic,
                                          dev.G, dev.H)
                if resident:
                    ctx.set_resident_eigenbasis(True)
                ctx.set_state(u.U, u.S, u.V)
                ctx.trajectory_reserve(steps + 1)
                for k in range(steps):
                    if name == "bug":
                        ctx.step_de_bug(dt, N)
                    else:
                        ctx.step_de_bug_ra(dt, N, 1e-6, 32)
                    ctx.trajectory_save(k)
                got[resident] = (ctx.get_state(), ctx.trajectory_get(steps // 2))
                if resident:  # the one-copy range refuses slots saved in the eigenbasis
                    with pytest.raises(ValueError):
                        ctx.trajectory_get_range(0, 2)
                else:
                    rng = ctx.trajectory_get_range(steps // 2, 1)[0]
                    assert all(np.array_equal(a, b) for a, b in zip(rng, got[False][1]))
                if resident:  # a data-path call leaves the eigenbasis; the state is unchanged
                    ctx.set_resident_eigenbasis(False)
                    assert rel_err(*ctx.get_state(), *got[True][0]) <= 1e-13
        e = rel_err(*got[True][0], *got[False][0])
        et = rel_err(*got[True][1], *got[False][1])
        say(f"  {name}: resident vs original basis {e:.2e} (trajectory slot {et:.2e})")
        assert e <= 1e-11 and et <= 1e-11
        U, S, V = u.U, u.S, u.V
        for k in range(steps):
            if name == "bug":
                U, S, V = dlra_de.bug_de_step(host, U, S, V, k * dt, dt, N)
            else:
                U, S, V, _ = dlra_de.rank_adaptive_bug_de_step(host, U, S, V, k * dt, dt, N, 1e-6, 32)
        assert_close(got[True][0], (U, S, V))


@pytest.mark.parametrize("name", ["bug", "primal"])
def test_stiff_heat_equation(lri, name):
    """Dirichlet Laplacian with the n^2 scale (rho ~ 7e4) and 80 RK4 sub-steps.

    dt keeps dt * rho ~ 7: with much larger dt the K step damps the random
    initial directions by many orders of magnitude, K1 becomes ill-conditioned
    and orth(K1) itself is only determined to ~eps * cond(K1) (2e-10 at dt = 1e-3
    here), in the oracle as much as on the GPU."""
    n = m = 96
    r = 6
    say(f"case: stiff heat equation {name}, n={n} r={r}")
    dev, host = operator_pair(lri, n, m, 1, False, False, seed=6, sa=n * n, sb=m * m)
    u = random_factors(np.random.default_rng(7), n, m, r)
    dt = 1e-4
    assert lri.rk4_substeps(dev.spectral_bound(), dt) >= 3
    N = 80
    got = run_one(lri, name, dev, u, dt, N)
    ref = oracle_one(name, host, u, dt, N)
    assert_close(got, ref)


def test_rank_adaptive_changes_rank(lri):
    n, m, r = 48, 48, 6
    say(f"case: RA DE rank change, n={n} r={r}")
    dev, host = operator_pair(lri, n, m, 1, False, False, seed=8)
    u = random_factors(np.random.default_rng(9), n, m, r, decay=1.0)
    dt = 0.02
    N = lri.rk4_substeps(dev.spectral_bound(), dt)
    for tol, rmax in ((1e-3, 64), (1e-12, 9), (1e-12, 64)):
        got = run_one(lri, "ra", dev, u, dt, N, tol, rmax)
        ref = dlra_de.rank_adaptive_bug_de_step(host, u.U, u.S, u.V, 0.0, dt, N, tol, rmax)
        say(f"  tol={tol} rmax={rmax}: rank {got[1].shape[0]} (oracle {ref[3]})")
        assert got[1].shape[0] == ref[3]
        assert_close(got, ref[:3])


@pytest.mark.parametrize("alg", ["bug", "strang", "ra"])
def test_solve_matches_oracle(lri, alg):
    """solve(MatrixDEProblem, alg, dt) over 8 steps against the oracle's driver."""
    n, m, r = 40, 32, 4
    say(f"case: solve DE {alg}, n={n} m={m} r={r}")
    dev, host = operator_pair(lri, n, m, 1, False, False, seed=10)
    u = random_factors(np.random.default_rng(11), n, m, r)
    dt, tf = 0.01, 0.08
    N = 5
    u0 = lri.SVDLikeRepresentation(u.U, u.S, u.V)
    A = {"bug": lri.UnconventionalAlgorithm(), "strang": lri.ProjectorSplitting(lri.Strang()),
         "ra": lri.RankAdaptiveUnconventionalAlgorithm(1e-4, rmax=12)}[alg]
    O = {"bug": dlra.UnconventionalAlgorithm(), "strang": dlra.ProjectorSplitting(dlra.Strang()),
         "ra": dlra.RankAdaptiveUnconventionalAlgorithm(1e-4, rmax=12)}[alg]
    sol = lri.solve(lri.MatrixDEProblem(dev, u0, (0.0, tf)), A, dt, substeps=N)
    ref = dlra_de.solve_de(dlra_de.MatrixDEProblem(host, u, (0.0, tf)), O, dt, N)
    assert len(sol.Y) == len(ref.Y) == 9
    assert np.allclose(sol.t, ref.t, rtol=0, atol=1e-15)
    for k in (1, 4, 8):
        a, b = sol.Y[k], ref.Y[k]
        assert a.S.shape == b.S.shape
        # after rank changes the last kept singular value sits near tol: its
        # direction is only determined to ~(error / gap), while Y itself is not
        # sensitive to it, so the subspaces get the looser bound there
        assert_close((a.U, a.S, a.V), (b.U, b.S, b.V), 1e-9, 1e-9 if alg != "ra" or k == 1 else 1e-6)


def test_operator_validation(lri):
    u = random_factors(np.random.default_rng(0), 16, 16, 2)
    ctx = lri.native.Context(16, 16, 2)
    try:
        ctx.set_state(u.U, u.S, u.V)
        with pytest.raises(ValueError):
            ctx.step_de_bug(0.1, 2)  # no operator yet
        ctx.set_operator_lyapunov(np.ones(16), np.ones(15), False, np.ones(16), np.ones(15), False)
        with pytest.raises(ValueError):
            ctx.step_de_bug(0.1, 0)
        with pytest.raises(ValueError):
            ctx.step_de_bug(-0.1, 2)
        with pytest.raises(ValueError):
            ctx.step_de_ksl(7, 0.1, 2)
    finally:
        ctx.close()


@pytest.mark.parametrize("name", ["bug", "primal", "dual"])
def test_graded_state_needs_shifted_cholqr(lri, name):
    """Singular values 10^(-i/4), i < 32 (cond ~ 6e7, the initial data of the 4096 x 4096
    heat-equation workload): CholeskyQR2's Gram matrix would be numerically singular;
    the shifted CholeskyQR3 of the DE steps matches the oracle's Householder QR.  Y is
    compared to 1e-10; the directions of the smallest singular values are determined
    only to ~eps / sigma_min (~1e-8), so the subspaces get a 1e-6 bound."""
    n, m, r = 200, 160, 32
    say(f"case: graded state {name}, n={n} m={m} r={r}")
    dev, host = operator_pair(lri, n, m, 1, False, False, seed=12, sa=float(n * n), sb=float(m * m))
    u = random_factors(np.random.default_rng(13), n, m, r, decay=0.25)
    dt = 1e-5
    N = lri.rk4_substeps(dev.spectral_bound(), dt)
    got = run_one(lri, name, dev, u, dt, N)
    ref = oracle_one(name, host, u, dt, N)
    assert_close(got, ref, 1e-10, 1e-6)


def test_stiff_collapse_is_not_an_error(lri):
    """dt * rho ~ 50: the K and L sub-steps damp every direction but the forcing's
    steady state, so K1 is numerically rank-deficient.  The engine keeps the
    dependent directions as zero columns (Householder QR would complete them
    arbitrarily), so the step must not fail, and Y must still agree with the
    oracle up to the size of what the oracle's arbitrary directions carry."""
    n = m = 128
    r = 8
    say(f"case: stiff collapse BUG, n={n} r={r}")
    dev, host = operator_pair(lri, n, m, 1, False, False, seed=14, sa=n * n, sb=m * m)
    u = random_factors(np.random.default_rng(15), n, m, r, decay=1.0)
    dt = 2e-4
    N = lri.rk4_substeps(dev.spectral_bound(), dt)
    got = run_one(lri, "bug", dev, u, dt, N)
    assert np.all(np.isfinite(got[0])) and np.all(np.isfinite(got[1])) and np.all(np.isfinite(got[2]))
    ref = oracle_one("bug", host, u, dt, N)
    e = rel_err(*got, *ref)
    say(f"  rel err vs oracle {e:.2e}")
    assert e < 1e-4


@pytest.mark.parametrize("name", ["bug", "primal"])
def test_mixed_spectral_and_stencil_sides(lri, name):
    """A constant (closed-form sub-steps in its analytic eigenbasis), B variable
    (RK4 stencil kernel): both integrators in one step."""
    n, m, r = 44, 39, 5
    say(f"case: mixed sides {name}, n={n} m={m} r={r}")
    rng = np.random.default_rng(16)
    A = lri.laplacian_1d(n, 70.0, periodic=True)
    B = tri(lri, m, 30.0, False, rng)
    G, H = rng.standard_normal((n, 1)), rng.standard_normal((m, 1))
    dev = lri.LyapunovOperator(A, B, G, H)
    host = dlra_de.LyapunovOperator(A.dense(), B.dense(), G, H)
    u = random_factors(np.random.default_rng(17), n, m, r)
    dt = 0.03
    N = lri.rk4_substeps(dev.spectral_bound(), dt)
    got = run_one(lri, name, dev, u, dt, N)
    ref = oracle_one(name, host, u, dt, N)
    assert_close(got, ref)


def test_bug_steps_reuse_galerkin_data(lri):
    """One operator on both sides (n = m): the K and L sub-steps share one pair of
    eigenbasis transforms, and a BUG step reuses the Galerkin matrices and
    eigenpairs its predecessor's S sub-step left for the new bases.  A set_state
    or a KSL step in between must make the next BUG step recompute them."""
    n = m = 64
    r = 5
    say(f"case: BUG steps with reused Galerkin data, n=m={n} r={r}")
    dev, host = operator_pair(lri, n, m, 1, False, False, seed=18, sa=30.0, sb=30.0)
    u = random_factors(np.random.default_rng(19), n, m, r)
    u2 = random_factors(np.random.default_rng(20), n, m, r)
    dt = 0.02
    N = lri.rk4_substeps(dev.spectral_bound(), dt)
    nat = lri.native
    ctx = nat.Context(n, m, r)
    try:
        ctx.set_state(u.U, u.S, u.V)
        ctx.set_operator_lyapunov(dev.A.d, dev.A.e, False, dev.B.d, dev.B.e, False, dev.G, dev.H)
        ref = (u.U, u.S, u.V)
        for step in range(3):
            ctx.step_de_bug(dt, N)
            ref = dlra_de.bug_de_step(host, *ref, 0.0, dt, N)
            say(f"  BUG step {step + 1}")
            assert_close(ctx.get_state(), ref)
        ctx.set_state(u2.U, u2.S, u2.V)
        ctx.step_de_bug(dt, N)
        ref = dlra_de.bug_de_step(host, u2.U, u2.S, u2.V, 0.0, dt, N)
        say("  BUG step after set_state")
        assert_close(ctx.get_state(), ref)
        ctx.step_de_ksl(nat.PRIMAL, dt, N)
        ref = dlra_de.ksl_primal_de_step(host, *ref, 0.0, dt, N)
        say("  KSL primal step")
        assert_close(ctx.get_state(), ref)
        ctx.step_de_bug(dt, N)
        ref = dlra_de.bug_de_step(host, *ref, 0.0, dt, N)
        say("  BUG step after KSL")
        assert_close(ctx.get_state(), ref)
    finally:
        ctx.close()
