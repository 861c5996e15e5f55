"""Parity of the Greedy integrator on data problems with the CPU oracle (needs an MI355X).

greedy_integrator.jl:84-111: an SVD-like state takes U1 = orth(X V0),
V1 = orth(X' U0), S1 = U1' X V1 with X = y(t+dt); a two-factor state U Z'
takes Z1 = X' U0 and U1 = the orthogonal polar factor of X Z1.

Tolerances: FP64 relative Frobenius error on Y <= 1e-10 against the oracle
on identical inputs, subspaces through projectors (<= 1e-9); the polar factor
is unique for a full-rank X Z1, so the two-factor U is compared directly
(<= 1e-10).  Mixed precision (BF16 / FP32 snapshots on matrix cores) against
the FP64 engine on the same decoded snapshots: <= 1e-6.
"""
import sys

import numpy as np
import pytest

from oracle import dlra
from oracle.data import bf16_round, random_factors, rotating_data
from oracle.lra import TwoFactor
from oracle.metrics import orth_defect, proj_dist, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def case(seed, n, m, r):
    rng = np.random.default_rng(seed)
    u = random_factors(rng, n, m, r, decay=0.2)
    X = u.to_matrix() + 1e-2 * rng.standard_normal((n, m))
    return u, X


def run_greedy(lri, U, S, V, snaps, twofactor=False, dtype="f64", precision="f64"):
    n, m, r = U.shape[0], V.shape[0], U.shape[1]
    with lri.native.Context(n, m, r) as ctx:
        ctx.set_precision(lri.native.PREC_MIXED if precision == "mixed" else lri.native.PREC_F64)
        ctx.set_state(U, S, V)
        ctx.reserve(len(snaps), lri.native.DTYPES[dtype])
        for k, a in enumerate(snaps):
            ctx.upload(k, lri.api._encode(a, dtype))
        for k in range(len(snaps)):
            if twofactor:
                ctx.step_greedy_twofactor(k)
            else:
                ctx.step_greedy(k)
        return ctx.get_state()


@pytest.mark.parametrize("shape", [(256, 256, 8), (200, 136, 5), (37, 53, 3), (1000, 700, 32), (300, 300, 64)])
def test_greedy_one_step_fp64(lri, shape):
    n, m, r = shape
    say(f"case: one Greedy step (SVD-like) n={n} m={m} r={r} FP64")
    u, X = case(1, n, m, r)
    U, S, V = run_greedy(lri, u.U, u.S, u.V, [X])
    Uo, So, Vo = dlra.greedy_step(u.U, u.S, u.V, X)
    e = rel_err(U, S, V, Uo, So, Vo)
    say(f"  rel err {e:.2e}")
    assert e <= FP64_TOL
    assert proj_dist(U, Uo) <= 1e-9 and proj_dist(V, Vo) <= 1e-9
    assert orth_defect(U) < 1e-13 and orth_defect(V) < 1e-13


@pytest.mark.parametrize("shape", [(256, 256, 8), (200, 136, 5), (1000, 700, 32)])
def test_greedy_twofactor_one_step_fp64(lri, shape):
    n, m, r = shape
    say(f"case: one Greedy step (two-factor) n={n} m={m} r={r} FP64")
    u, X = case(2, n, m, r)
    Z0 = u.V @ u.S.T
    U, _, Z = run_greedy(lri, u.U, np.eye(r), Z0, [X], twofactor=True)
    Uo, Zo = dlra.greedy_step_twofactor(u.U, Z0, X)
    eu = np.linalg.norm(U - Uo) / np.linalg.norm(Uo)
    ez = np.linalg.norm(Z - Zo) / np.linalg.norm(Zo)
    say(f"  U err {eu:.2e}, Z err {ez:.2e}")
    assert eu <= FP64_TOL and ez <= 1e-13
    assert orth_defect(U) < 1e-13


def test_greedy_several_steps_fp64(lri):
    """Ten steps over rotating data (the driver's snapshot sequence), both state kinds."""
    d = rotating_data(300, 240, 6, noise=1e-3, seed=3)
    u0 = d.initial()
    snaps = [d(0.1 * k) for k in range(1, 11)]
    U, S, V = run_greedy(lri, u0.U, u0.S, u0.V, snaps)
    Uo, So, Vo = u0.U, u0.S, u0.V
    for X in snaps:
        Uo, So, Vo = dlra.greedy_step(Uo, So, Vo, X)
    assert rel_err(U, S, V, Uo, So, Vo) <= FP64_TOL
    Z0 = u0.V @ u0.S.T
    U, _, Z = run_greedy(lri, u0.U, np.eye(6), Z0, snaps, twofactor=True)
    Uo, Zo = u0.U, Z0
    for X in snaps:
        Uo, Zo = dlra.greedy_step_twofactor(Uo, Zo, X)
    assert np.linalg.norm(U @ Z.T - Uo @ Zo.T) / np.linalg.norm(Uo @ Zo.T) <= FP64_TOL


@pytest.mark.parametrize("dtype,n,m", [("bf16", 1024, 1024), ("f32", 1024, 1024), ("bf16", 1000, 700),
                                       ("bf16", 900, 520)])
def test_greedy_mixed_precision(lri, dtype, n, m):
    """Ragged n (1000, 900: row blocks that end inside a 64- or 128-row step of
    pass T, and rows that are not a multiple of 8) next to the aligned case."""
    r = 32
    say(f"case: Greedy mixed precision {dtype} n={n} m={m} r={r} 10 steps")
    d = rotating_data(n, m, r, noise=1e-3, seed=4)
    u0 = d.initial()
    snaps = [d(k / 199.0) for k in range(1, 11)]
    snaps = [bf16_round(a) for a in snaps] if dtype == "bf16" else [a.astype(np.float32).astype(np.float64)
                                                                     for a in snaps]
    ref = run_greedy(lri, u0.U, u0.S, u0.V, snaps, dtype=dtype, precision="f64")
    mix = run_greedy(lri, u0.U, u0.S, u0.V, snaps, dtype=dtype, precision="mixed")
    e = rel_err(*mix, *ref)
    say(f"  rel err mixed vs FP64 = {e:.3e}")
    assert e <= 1e-6


def test_greedy_function_vs_array_data(lri):
    """The reference's consistency test (test/data_driven_approximation.jl:22-31) for
    GreedyIntegrator on the engine, SVD-like and two-factor, against the oracle too."""
    d = rotating_data(128, 96, 5, noise=1e-3, seed=5)
    u0 = d.initial()
    snaps = [d(0.01 * k) for k in range(101)]
    y0 = lri.SVDLikeRepresentation(u0.U, u0.S, u0.V)
    tf0 = lri.TwoFactorRepresentation(u0.U.copy(), u0.V @ u0.S.T)
    for start, ostart in ((y0, u0), (tf0, TwoFactor(tf0.U, tf0.Z))):
        a = lri.solve(lri.MatrixDataProblem(d, start, (0.0, 1.0)), lri.GreedyIntegrator(), 1e-2)
        b = lri.solve(lri.MatrixDataProblem(snaps, start), lri.GreedyIntegrator())
        assert len(a.Y) == 101 and len(b.Y) == 101
        assert type(a.Y[-1]) is type(start)
        A, B = a.Y[-1].matrix(), b.Y[-1].matrix()
        assert np.all(np.abs(A - B) <= np.sqrt(np.finfo(float).eps) * np.maximum(np.abs(A), np.abs(B)))
        o = dlra.solve(dlra.MatrixDataProblem(snaps, ostart), dlra.GreedyIntegrator()).Y[-1].to_matrix()
        e = np.linalg.norm(B - o) / np.linalg.norm(o)
        say(f"  {type(start).__name__}: engine vs oracle after 100 steps {e:.2e}")
        assert e <= FP64_TOL

