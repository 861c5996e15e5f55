"""Parity of the HIP engine with the CPU oracle (needs an MI355X).

Tolerances:
* FP64 products: relative Frobenius error on Y <= 1e-10 against the oracle on
  identical inputs; subspaces compared through projectors (<= 1e-9).
* Mixed precision: BF16 snapshots (hi/lo-split BF16 matrix cores, FP32
  accumulation) and F32 snapshots (FP32 matrix cores) against the FP64 engine
  on the same decoded snapshots: <= 1e-6.
"""
import functools
import sys

import numpy as np
import pytest

from oracle import dlra
from oracle.data import bf16_bits, bf16_round, random_factors, rotating_data
from oracle.metrics import orth_defect, proj_dist, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@functools.lru_cache(maxsize=None)
def cfg0_data():
    d = rotating_data(256, 256, 8, noise=1e-3, seed=0)
    return d, d.initial()


# ------------------------------------------------------------------ lane layouts


def test_mfma_layouts(lri):
    rng = np.random.default_rng(0)
    # FP64 16x16x4: A[i][k] at lane 16k+i, B[k][j] at lane 16k+j, D[4*v+l/16][l%16]
    A = rng.integers(-4, 5, (16, 4)).astype(np.float64)
    B = rng.integers(-4, 5, (4, 16)).astype(np.float64)
    a = np.array([A[l % 16, l // 16] for l in range(64)])
    b = np.array([B[l // 16, l % 16] for l in range(64)])
    out = lri.native.selftest_mfma(0, a, b).reshape(64, 4)
    D = A @ B
    got = np.zeros((16, 16))
    for l in range(64):
        for v in range(4):
            got[4 * v + l // 16, l % 16] = out[l, v]
    assert np.array_equal(got, D)
    # BF16 32x32x16: A[i][8*(l/32)+e] at lane l (i = l%32), B[8*(l/32)+e][l%32];
    # D[8*(v/4)+4*(l/32)+(v%4)][l%32]
    A = rng.integers(-4, 5, (32, 16)).astype(np.float64)
    B = rng.integers(-4, 5, (16, 32)).astype(np.float64)
    a = np.zeros((64, 8), np.uint16)
    b = np.zeros((64, 8), np.uint16)
    for l in range(64):
        for e in range(8):
            a[l, e] = bf16_bits(np.array([A[l % 32, 8 * (l // 32) + e]]))[0]
            b[l, e] = bf16_bits(np.array([B[8 * (l // 32) + e, l % 32]]))[0]
    out = lri.native.selftest_mfma(1, a, b).reshape(64, 16)
    D = A @ B
    got = np.zeros((32, 32))
    for l in range(64):
        for v in range(16):
            got[8 * (v // 4) + 4 * (l // 32) + (v % 4), l % 32] = out[l, v]
    assert np.array_equal(got, D)
    # FP32 32x32x2
    A = rng.integers(-4, 5, (32, 2)).astype(np.float32)
    B = rng.integers(-4, 5, (2, 32)).astype(np.float32)
    a = np.array([A[l % 32, l // 32] for l in range(64)], np.float32)
    b = np.array([B[l // 32, l % 32] for l in range(64)], np.float32)
    out = lri.native.selftest_mfma(2, a, b).reshape(64, 16)
    got = np.zeros((32, 32))
    for l in range(64):
        for v in range(16):
            got[8 * (v // 4) + 4 * (l // 32) + (v % 4), l % 32] = out[l, v]
    assert np.array_equal(got, A.astype(np.float64) @ B.astype(np.float64))


def test_mfma_layout_decode(lri):
    """Decode where each (row, col) of D lands: A[i][0] = 1, A[i][1] = i, B[0][j] = 64 j, B[1][j] = 1,
    so D[i][j] = 64 j + i.  Prints the observed map of the first lanes (diagnostic for the test above)."""
    res = {}
    # FP64 16x16x4
    A = np.zeros((16, 4)); A[:, 0] = 1; A[:, 1] = np.arange(16)
    B = np.zeros((4, 16)); B[0, :] = 64 * np.arange(16); B[1, :] = 1
    a = np.array([A[l % 16, l // 16] for l in range(64)])
    b = np.array([B[l // 16, l % 16] for l in range(64)])
    out = lri.native.selftest_mfma(0, a, b).reshape(64, 4)
    bad = [(l, v, int(out[l, v]) % 64, int(out[l, v]) // 64) for l in range(64) for v in range(4)
           if (int(out[l, v]) % 64, int(out[l, v]) // 64) != (4 * v + l // 16, l % 16)]
    res["f64"] = bad[:8]
    # BF16 32x32x16
    A = np.zeros((32, 16)); A[:, 0] = 1; A[:, 1] = np.arange(32)
    B = np.zeros((16, 32)); B[0, :] = 64 * np.arange(32); B[1, :] = 1
    a = np.zeros((64, 8), np.uint16); b = np.zeros((64, 8), np.uint16)
    for l in range(64):
        for e in range(8):
            a[l, e] = bf16_bits(np.array([A[l % 32, 8 * (l // 32) + e]]))[0]
            b[l, e] = bf16_bits(np.array([B[8 * (l // 32) + e, l % 32]]))[0]
    out = lri.native.selftest_mfma(1, a, b).reshape(64, 16)
    bad = [(l, v, int(out[l, v]) % 64, int(out[l, v]) // 64) for l in range(64) for v in range(16)
           if (int(out[l, v]) % 64, int(out[l, v]) // 64) != (8 * (v // 4) + 4 * (l // 32) + (v % 4), l % 32)]
    res["bf16"] = bad[:8]
    say(f"layout mismatches (lane, reg, row, col): {res}")
    assert not res["f64"] and not res["bf16"]


# ------------------------------------------------------------------ the product passes in isolation


@pytest.mark.parametrize("dtype,precision,tol", [("f64", "f64", 1e-13), ("bf16", "f64", 1e-13),
                                                 ("f32", "f64", 1e-13), ("bf16", "mixed", 2e-5),
                                                 ("f32", "mixed", 2e-6)])
@pytest.mark.parametrize("shape", [(256, 256, 8), (1000, 700, 32), (37, 53, 3), (512, 384, 64)])
def test_products(lri, dtype, precision, tol, shape):
    n, m, r = shape
    say(f"case: products n={n} m={m} r={r} {dtype} {precision}")
    u, A0, A1 = one_step_case(9, n, m, r)
    if dtype == "bf16":
        A0, A1 = bf16_round(A0), bf16_round(A1)
    elif dtype == "f32":
        A0, A1 = A0.astype(np.float32).astype(np.float64), A1.astype(np.float32).astype(np.float64)
    with lri.native.Context(n, m, r) as ctx:
        ctx.set_precision(lri.native.PREC_MIXED if precision == "mixed" else lri.native.PREC_F64)
        ctx.set_state(u.U, u.S, u.V)
        ctx.reserve(2, lri.native.DTYPES[dtype])
        ctx.upload(0, lri.api._encode(A0, dtype))
        ctx.upload(1, lri.api._encode(A1, dtype))
        P = ctx.product(0, 0, 1)
        Q = ctx.product(1, 0, 1)
    dA = A1 - A0
    Pr, Qr = dA @ u.V, dA.T @ u.U
    eP = np.linalg.norm(P - Pr) / np.linalg.norm(Pr)
    eQ = np.linalg.norm(Q - Qr) / np.linalg.norm(Qr)
    say(f"  rel err dA V = {eP:.2e}, dA' U = {eQ:.2e}")
    assert eP <= tol and eQ <= tol


@pytest.mark.parametrize("r", [6, 24])
def test_zero_increment_steps(lri, r):
    u, A0, _ = one_step_case(8, 96, 80, r)
    for kind in ("primal", "dual", "bug"):
        U, S, V = run_native_step(lri, kind, u, [A0, A0.copy()])
        assert rel_err(U, S, V, u.U, u.S, u.V) <= 1e-13, kind


def rank_deficient_case(r, seed=11, n=120, m=100):
    """A state whose S has two zero singular values (K = U S has rank r - 2)."""
    rng = np.random.default_rng(seed)
    u = random_factors(rng, n, m, r, decay=0.05)
    sig = np.linalg.svd(u.S, compute_uv=False)
    sig[2] = sig[r - 3] = 0.0
    Us, _ = np.linalg.qr(rng.standard_normal((r, r)))
    Vs, _ = np.linalg.qr(rng.standard_normal((r, r)))
    return type(u)(u.U, Us @ np.diag(sig) @ Vs.T, u.V), rng


@pytest.mark.parametrize("r", [6, 24, 40])
def test_rank_deficient_zero_increment_returns_y0(lri, r):
    """A rank-deficient state and a zero increment: K = U S drops two Cholesky pivots.
    The reference's Householder qr! (projector_splitting.jl:137,149,174,186) returns an
    orthonormal basis and the step returns Y0; so must the engine, whose dropped columns
    are completed (r = 6 and 24 run the factorisations inside the row kernels, r = 40 as
    separate epilogue launches)."""
    say(f"case: rank-deficient state, zero increment, r={r}")
    u, rng = rank_deficient_case(r)
    A0 = rng.standard_normal((120, 100))
    for kind in ("primal", "dual", "strang"):
        snaps = [A0, A0.copy(), A0.copy()] if kind == "strang" else [A0, A0.copy()]
        U, S, V = run_native_step(lri, kind, u, snaps)
        e = rel_err(U, S, V, u.U, u.S, u.V)
        say(f"  {kind}: rel err vs Y0 {e:.2e}, orthogonality {orth_defect(U):.1e} {orth_defect(V):.1e}")
        assert e <= 1e-13, kind
        assert orth_defect(U) <= 1e-13 and orth_defect(V) <= 1e-13, kind


@pytest.mark.parametrize("r", [6, 24, 40])
def test_rank_deficient_step_identities(lri, r):
    """A rank-deficient state and an increment dA = U0 C V0' + 1e-2 (I - U0 U0') D (I - V0 V0')
    whose C keeps K = U0 S0 + dA V0 (and L = V0 S0' + dA' U0) of rank r - 2, while its
    other part reaches the completed directions.  The completion is a free choice
    (Householder makes another), but every orthonormal completion gives a primal step with
    U1 orthonormal, range(K) inside span(U1) and Y1 = K V0' + U1 U1' dA (I - V0 V0')
    (projector_splitting.jl:129-152: S~ = R - U1' dA V0, L = V0 S~' + dA' U1, Y1 = U1 L');
    the dual step (projector_splitting.jl:166-189) the mirror identity
    Y1 = U0 L' + (I - U0 U0') dA V1 V1'."""
    say(f"case: rank-deficient state, rank-deficient increment, r={r}")
    rng = np.random.default_rng(12)
    u = random_factors(rng, 120, 100, r, decay=0.05)
    sig = np.linalg.svd(u.S, compute_uv=False)
    sig[2] = sig[r - 3] = 0.0
    Us, _ = np.linalg.qr(rng.standard_normal((r, r)))
    Vs, _ = np.linalg.qr(rng.standard_normal((r, r)))
    U0, S0, V0 = u.U, Us @ np.diag(sig) @ Vs.T, u.V
    c = 1e-2 * rng.uniform(0.5, 1.0, r)
    c[2] = c[r - 3] = 0.0
    PU, PV = np.eye(120) - U0 @ U0.T, np.eye(100) - V0 @ V0.T
    dA = U0 @ (Us @ np.diag(c) @ Vs.T) @ V0.T + 1e-2 * PU @ rng.standard_normal((120, 100)) @ PV
    A0 = rng.standard_normal((120, 100))
    A1 = A0 + dA
    dA = A1 - A0  # the increment the engine sees
    u = type(u)(U0, S0, V0)
    K = U0 @ S0 + dA @ V0
    assert np.linalg.matrix_rank(K) == r - 2
    U, S, V = run_native_step(lri, "primal", u, [A0, A1])
    Y1 = K @ V0.T + U @ (U.T @ dA) @ PV
    e = np.linalg.norm(U @ S @ V.T - Y1) / np.linalg.norm(Y1)
    say(f"  primal: identity residual {e:.2e}, orthogonality {orth_defect(U):.1e} {orth_defect(V):.1e}")
    assert e <= 1e-12 and orth_defect(U) <= 1e-13 and orth_defect(V) <= 1e-13
    assert np.linalg.norm(K - U @ (U.T @ K)) <= 1e-12 * np.linalg.norm(K)
    U, S, V = run_native_step(lri, "dual", u, [A0, A1])
    L = V0 @ S0.T + dA.T @ U0
    Y1 = U0 @ L.T + PU @ dA @ V @ V.T
    e = np.linalg.norm(U @ S @ V.T - Y1) / np.linalg.norm(Y1)
    say(f"  dual: identity residual {e:.2e}, orthogonality {orth_defect(U):.1e} {orth_defect(V):.1e}")
    assert e <= 1e-12 and orth_defect(U) <= 1e-13 and orth_defect(V) <= 1e-13
    assert np.linalg.norm(L - V @ (V.T @ L)) <= 1e-12 * np.linalg.norm(L)


# ------------------------------------------------------------------ FP64 parity, one step


def one_step_case(seed, n, m, r, inc=1e-2):
    rng = np.random.default_rng(seed)
    u = random_factors(rng, n, m, r, decay=0.2)
    A0 = rng.standard_normal((n, m))
    A1 = A0 + inc * rng.standard_normal((n, m))
    return u, A0, A1


def run_native_step(lri, kind, u, snaps, dtype="f64", precision="f64", tol=1e-8, rmax=64, rcap=None):
    n, m = u.U.shape[0], u.V.shape[0]
    rcap = rcap or (min(64, n, m) if kind == "ra" else u.rank)
    with lri.native.Context(n, m, rcap) as ctx:
        ctx.set_precision(lri.native.PREC_MIXED if precision == "mixed" else lri.native.PREC_F64)
        ctx.set_state(u.U, u.S, u.V)
        ctx.reserve(len(snaps), lri.native.DTYPES[dtype])
        for k, a in enumerate(snaps):
            ctx.upload(k, lri.api._encode(a, dtype))
        if kind == "primal":
            ctx.step_ksl(lri.native.PRIMAL, 0, 1)
        elif kind == "dual":
            ctx.step_ksl(lri.native.DUAL, 0, 1)
        elif kind == "strang":
            ctx.step_ksl_strang(0, 1, 2)
        elif kind == "bug":
            ctx.step_bug(0, 1)
        elif kind == "ra":
            ctx.step_bug_ra(0, 1, tol, rmax)
        return ctx.get_state()


ORACLE = {
    "primal": dlra.ksl_primal_step,
    "dual": dlra.ksl_dual_step,
    "bug": dlra.bug_step,
}


@pytest.mark.parametrize("kind", ["primal", "dual", "bug"])
@pytest.mark.parametrize("shape", [(256, 256, 8), (200, 136, 5), (37, 53, 3), (1000, 700, 32), (300, 300, 64)])
def test_one_step_fp64(lri, kind, shape):
    n, m, r = shape
    say(f"case: one {kind} step n={n} m={m} r={r} FP64")
    u, A0, A1 = one_step_case(1, n, m, r)
    U, S, V = run_native_step(lri, kind, u, [A0, A1])
    Uo, So, Vo = ORACLE[kind](u.U, u.S, u.V, A1 - A0)
    assert rel_err(U, S, V, Uo, So, Vo) <= FP64_TOL
    assert proj_dist(U, Uo) <= 1e-9 and proj_dist(V, Vo) <= 1e-9
    assert orth_defect(U) < 1e-13 and orth_defect(V) < 1e-13
    if kind == "primal":
        assert np.abs(np.triu(S, 1)).max() == 0.0  # S = R' lower triangular
    if kind == "dual":
        assert np.abs(np.tril(S, -1)).max() == 0.0


def test_one_step_strang_fp64(lri):
    u, A0, A1 = one_step_case(2, 300, 240, 12)
    Am = 0.5 * (A0 + A1) + 1e-3 * np.random.default_rng(5).standard_normal(A0.shape)
    U, S, V = run_native_step(lri, "strang", u, [A0, Am, A1])
    Uo, So, Vo = dlra.ksl_primal_step(u.U, u.S, u.V, Am - A0)
    Uo, So, Vo = dlra.ksl_dual_step(Uo, So, Vo, A1 - Am)
    assert rel_err(U, S, V, Uo, So, Vo) <= FP64_TOL


@pytest.mark.parametrize("shape,tol", [((256, 256, 8), 1e-3), ((200, 150, 12), 2e-2), ((64, 96, 32), 1e-6)])
def test_one_step_rank_adaptive_fp64(lri, shape, tol):
    n, m, r = shape
    say(f"case: one RA step n={n} m={m} r={r} tol={tol}")
    u, A0, A1 = one_step_case(3, n, m, r, inc=1e-2)
    U, S, V = run_native_step(lri, "ra", u, [A0, A1], tol=tol, rmax=64)
    Uo, So, Vo, r1 = dlra.rank_adaptive_bug_step(u.U, u.S, u.V, A1 - A0, tol, 64)
    assert S.shape[0] == r1
    assert rel_err(U, S, V, Uo, So, Vo) <= 1e-9
    assert np.allclose(np.diag(S), np.diag(So), rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("shape,tol,rmax", [((300, 260, 48), 1e-14, 64), ((260, 300, 48), 3e-2, 64),
                                           ((400, 330, 64), 1e-14, 64)])
def test_one_step_rank_adaptive_wide_fp64(lri, shape, tol, rmax):
    """Ranks 48 and 64: the augmented core is 96 x 96 or 128 x 128, so the SVD runs the
    256-thread Jacobi kernel (2 r > 64, V in LDS or in global scratch); tol = 1e-14 keeps
    the rank at rmax, tol = 3e-2 truncates below it."""
    n, m, r = shape
    say(f"case: one RA step n={n} m={m} r={r} tol={tol} rmax={rmax}")
    u, A0, A1 = one_step_case(13, n, m, r, inc=1e-2)
    Uo, So, Vo, r1 = dlra.rank_adaptive_bug_step(u.U, u.S, u.V, A1 - A0, tol, rmax)
    U, S, V = run_native_step(lri, "ra", u, [A0, A1], tol=tol, rmax=rmax)
    say(f"  new rank {S.shape[0]} (oracle {r1})")
    assert S.shape[0] == r1
    assert rel_err(U, S, V, Uo, So, Vo) <= 1e-9
    assert np.allclose(np.diag(S), np.diag(So), rtol=1e-10, atol=1e-14)
    assert orth_defect(U) < 1e-12 and orth_defect(V) < 1e-12


def test_rank_adaptive_capacity_overflow(lri):
    """rmax above the context capacity and a tolerance that keeps every direction: the
    step needs rank 2 r = 24 > rcap = 16, so it reports LRI_ENOTSUP (no silent clip) and
    leaves the state at the step's start; the same context then steps with rmax = rcap."""
    n, m, r, rcap = 128, 112, 12, 16
    say(f"case: RA capacity overflow n={n} m={m} r={r} rcap={rcap}")
    u, A0, A1 = one_step_case(14, n, m, r, inc=1e-1)
    with lri.native.Context(n, m, rcap) as ctx:
        ctx.set_state(u.U, u.S, u.V)
        ctx.reserve(2, lri.native.F64)
        ctx.upload(0, lri.api._encode(A0, "f64"))
        ctx.upload(1, lri.api._encode(A1, "f64"))
        with pytest.raises(lri.native.NotSupportedError):
            ctx.step_bug_ra(0, 1, 1e-14, 64)
        U, S, V = ctx.get_state()
        unchanged = S.shape[0] == r and np.array_equal(U, u.U) and np.array_equal(S, u.S) and np.array_equal(V, u.V)
        assert unchanged
        assert ctx.step_bug_ra(0, 1, 1e-14, rcap) == rcap
        U, S, V = ctx.get_state()
    Uo, So, Vo, r1 = dlra.rank_adaptive_bug_step(u.U, u.S, u.V, A1 - A0, 1e-14, rcap)
    assert r1 == rcap and rel_err(U, S, V, Uo, So, Vo) <= 1e-9


def test_solve_capacity_overflow_raises(lri):
    """Through the API: the reference's default rmax (2^62) and a tolerance of 0 on data
    whose increments add rank every step: the rank doubles to 64 (the capacity), then the
    next step needs more and raises instead of clipping."""
    n = 200
    rng = np.random.default_rng(15)
    u = random_factors(rng, n, n, 16, decay=0.5)
    snaps = [rng.standard_normal((n, n)) for _ in range(4)]
    prob = lri.MatrixDataProblem(snaps, lri.SVDLikeRepresentation(u.U, u.S, u.V))
    with pytest.raises(lri.native.NotSupportedError):
        lri.solve(prob, lri.RankAdaptiveUnconventionalAlgorithm(0.0))


def test_rank_adaptive_zero_increment_keeps_rank(lri):
    u, A0, _ = one_step_case(4, 128, 96, 6)
    U, S, V = run_native_step(lri, "ra", u, [A0, A0.copy()], tol=1e-10, rmax=64)
    assert S.shape[0] == 6
    assert rel_err(U, S, V, u.U, u.S, u.V) <= 1e-12


def test_rank_adaptive_rmax(lri):
    u, A0, A1 = one_step_case(5, 128, 96, 6, inc=1e-1)
    U, S, V = run_native_step(lri, "ra", u, [A0, A1], tol=1e-14, rmax=9)
    assert S.shape[0] == 9


# ------------------------------------------------------------------ full solves vs fixture


GOLDEN_ALGS = {
    "ksl_primal": lambda lri: lri.ProjectorSplitting(lri.PrimalLieTrotter()),
    "ksl_dual": lambda lri: lri.ProjectorSplitting(lri.DualLieTrotter()),
    "ksl_strang": lambda lri: lri.ProjectorSplitting(lri.Strang()),
    "bug": lambda lri: lri.UnconventionalAlgorithm(),
    "ra_bug": lambda lri: lri.RankAdaptiveUnconventionalAlgorithm(2e-2, rmax=16),
}


@pytest.mark.parametrize("name", list(GOLDEN_ALGS))
def test_cfg0_solve_matches_fixture(lri, name):
    """Workload: n=m=256, r=8, t in [0,1], dt=1e-2 (100 steps), function-form data, FP64."""
    import os

    say(f"case: cfg0 solve {name}")
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg0_golden.npz"))
    d, u0 = cfg0_data()
    prob = lri.MatrixDataProblem(d, lri.SVDLikeRepresentation(u0.U, u0.S, u0.V), (0.0, 1.0))
    sol = lri.solve(prob, GOLDEN_ALGS[name](lri), 1e-2)
    assert len(sol.Y) == 101 and sol.t[-1] == z[f"{name}_t"][-1]
    yf = sol.Y[-1]
    assert rel_err(yf.U, yf.S, yf.V, z[f"{name}_U"], z[f"{name}_S"], z[f"{name}_V"]) <= FP64_TOL
    norms = np.array([np.linalg.norm(y.S) for y in sol.Y])
    assert np.allclose(norms, z[f"{name}_norms"], rtol=1e-11, atol=0)
    assert [y.rank for y in sol.Y] == list(z[f"{name}_ranks"])


def test_array_data_matches_function_data(lri):
    """The reference's own consistency test (test/data_driven_approximation.jl:27-31) on the engine."""
    d, u0 = cfg0_data()
    y0 = lri.SVDLikeRepresentation(u0.U, u0.S, u0.V)
    snaps = [d(0.01 * k) for k in range(101)]
    for alg in (lri.ProjectorSplitting(lri.PrimalLieTrotter()), lri.UnconventionalAlgorithm()):
        a = lri.solve(lri.MatrixDataProblem(d, y0, (0.0, 1.0)), alg, 1e-2).Y[-1]
        b = lri.solve(lri.MatrixDataProblem(snaps, y0), alg).Y[-1]
        A, B = a.matrix(), b.matrix()
        assert np.all(np.abs(A - B) <= np.sqrt(np.finfo(float).eps) * np.maximum(np.abs(A), np.abs(B)))


def test_array_data_late_start(lri):
    """tspan = (3, 9) on array data: the cache's yprev is y[1] whatever t0 is
    (projector_splitting.jl:91, unconventional.jl:91), so the first increment is
    y[4] - y[1], the later ones y[k+1] - y[k]; the engine's solve against the oracle's."""
    d, u0 = cfg0_data()
    y0 = lri.SVDLikeRepresentation(u0.U, u0.S, u0.V)
    snaps = [d(0.01 * k) for k in range(10)]
    from oracle import lra as olra

    for name in ("ksl_primal", "bug", "ra_bug"):
        alg = GOLDEN_ALGS[name](lri)
        got = lri.solve(lri.MatrixDataProblem(snaps, y0, (3, 9)), alg)
        oalg = {"ksl_primal": dlra.ProjectorSplitting(dlra.PrimalLieTrotter()),
                "bug": dlra.UnconventionalAlgorithm(),
                "ra_bug": dlra.RankAdaptiveUnconventionalAlgorithm(2e-2, rmax=16)}[name]
        ref = dlra.solve(dlra.MatrixDataProblem(snaps, olra.SVDLike(u0.U, u0.S, u0.V), (3, 9)), oalg)
        assert len(got.Y) == len(ref.Y) == 7 and list(got.t) == list(ref.t)
        a, b = got.Y[-1], ref.Y[-1]
        assert rel_err(a.U, a.S, a.V, b.U, b.S, b.V) <= FP64_TOL, name


def test_device_trajectory_roundtrip(lri):
    """lri_trajectory_save / _get: each slot returns the state it was saved from, also
    across rank changes of the rank-adaptive step, and an unsaved slot is refused."""
    u, A0, A1 = one_step_case(16, 150, 120, 6, inc=1e-1)
    A2 = A1 + 1e-1 * np.random.default_rng(17).standard_normal(A1.shape)
    with lri.native.Context(150, 120, 24) as ctx:
        ctx.set_state(u.U, u.S, u.V)
        ctx.reserve(3, lri.native.F64)
        for k, a in enumerate((A0, A1, A2)):
            ctx.upload(k, lri.api._encode(a, "f64"))
        ctx.trajectory_reserve(4)
        ctx.trajectory_save(0)
        want = [ctx.get_state()]
        for k in range(2):
            ctx.step_bug_ra(k, k + 1, 1e-12, 24)
            ctx.trajectory_save(k + 1)
            want.append(ctx.get_state())
        assert [w[1].shape[0] for w in want] == [6, 12, 24]
        for k, w in enumerate(want):
            got = ctx.trajectory_get(k)
            assert all(np.array_equal(a, b) for a, b in zip(got, w)), k
        # lri_trajectory_get_range: one copy of a run of slots, the same bits
        for slot0, count in ((0, 3), (1, 2), (2, 1), (0, 0)):
            got = ctx.trajectory_get_range(slot0, count)
            assert len(got) == count
            for k, g in enumerate(got):
                assert all(np.array_equal(a, b) for a, b in zip(g, want[slot0 + k])), (slot0, k)
        with pytest.raises(ValueError):
            ctx.trajectory_get(3)
        with pytest.raises(ValueError):
            ctx.trajectory_get_range(2, 2)  # slot 3 never saved
        with pytest.raises(ValueError):
            ctx.trajectory_get_range(3, 2)  # past the reservation


# ------------------------------------------------------------------ mixed precision


def rotating_snaps(n, r, steps, seed, dtype):
    d = rotating_data(n, n, r, noise=1e-3, seed=seed)
    u0 = d.initial()
    ts = np.arange(steps + 1) / 199.0
    snaps = [d(t) for t in ts]
    if dtype == "bf16":
        snaps = [bf16_round(a) for a in snaps]
    elif dtype == "f32":
        snaps = [a.astype(np.float32).astype(np.float64) for a in snaps]
    return u0, snaps


def run_solve(lri, u0, snaps, kind, precision, dtype):
    n, m = u0.U.shape[0], u0.V.shape[0]
    with lri.native.Context(n, m, u0.rank) as ctx:
        ctx.set_precision(lri.native.PREC_MIXED if precision == "mixed" else lri.native.PREC_F64)
        ctx.set_state(u0.U, u0.S, u0.V)
        ctx.reserve(len(snaps), lri.native.DTYPES[dtype])
        for k, a in enumerate(snaps):
            ctx.upload(k, lri.api._encode(a, dtype))
        for k in range(len(snaps) - 1):
            if kind == "primal":
                ctx.step_ksl(lri.native.PRIMAL, k, k + 1)
            elif kind == "dual":
                ctx.step_ksl(lri.native.DUAL, k, k + 1)
            else:
                ctx.step_bug(k, k + 1)
        return ctx.get_state()


@pytest.mark.parametrize("dtype", ["bf16", "f32"])
@pytest.mark.parametrize("kind", ["primal", "dual", "bug"])
def test_mixed_precision_error(lri, dtype, kind):
    say(f"case: mixed {kind} {dtype} n=1024 r=32 20 steps")
    u0, snaps = rotating_snaps(1024, 32, 20, seed=1, dtype=dtype)
    ref = run_solve(lri, u0, snaps, kind, "f64", dtype)
    mix = run_solve(lri, u0, snaps, kind, "mixed", dtype)
    err = rel_err(*mix, *ref)
    say(f"  rel err mixed vs FP64 = {err:.3e}")
    assert err <= 1e-6
    # the FP64 engine on decoded snapshots agrees with the CPU oracle
    if kind == "primal" and dtype == "bf16":
        U, S, V = u0.U, u0.S, u0.V
        for k in range(len(snaps) - 1):
            U, S, V = dlra.ksl_primal_step(U, S, V, snaps[k + 1] - snaps[k])
        assert rel_err(*ref, U, S, V) <= FP64_TOL


def test_mixed_precision_ragged(lri):
    """Sizes that are not multiples of the tiles (BF16 vector and scalar load paths): the products to
    2e-5 relative (three BF16 products per term drop terms below 2^-17), and one BUG step on data whose
    increment is 1e-3 of the data to 1e-6."""
    rng = np.random.default_rng(7)
    for (n, m, r) in [(200, 136, 5), (517, 389, 40), (64, 1000, 33)]:
        say(f"case: ragged mixed n={n} m={m} r={r}")
        X = rng.standard_normal((n, r)) @ np.diag(2.0 ** -np.arange(r)) @ rng.standard_normal((r, m))
        A0 = bf16_round(X + 1e-3 * rng.standard_normal((n, m)))
        A1 = bf16_round(A0 + 1e-3 * rng.standard_normal((n, m)))
        U, s, Vt = np.linalg.svd(A0, full_matrices=False)
        from oracle.lra import SVDLike

        u = SVDLike(U[:, :r].copy(), np.diag(s[:r]), Vt[:r].T.copy())
        with lri.native.Context(n, m, r) as ctx:
            ctx.set_precision(lri.native.PREC_MIXED)
            ctx.set_state(u.U, u.S, u.V)
            ctx.reserve(2, lri.native.BF16)
            ctx.upload(0, lri.api._encode(A0, "bf16"))
            ctx.upload(1, lri.api._encode(A1, "bf16"))
            P, Q = ctx.product(0, 0, 1), ctx.product(1, 0, 1)
        dA = A1 - A0
        assert np.linalg.norm(P - dA @ u.V) <= 2e-5 * np.linalg.norm(dA @ u.V)
        assert np.linalg.norm(Q - dA.T @ u.U) <= 2e-5 * np.linalg.norm(dA.T @ u.U)
        ref = run_solve(lri, u, [A0, A1], "bug", "f64", "bf16")
        mix = run_solve(lri, u, [A0, A1], "bug", "mixed", "bf16")
        assert rel_err(*mix, *ref) <= 1e-6


# ------------------------------------------------------------------ errors, determinism


def test_nan_snapshot_reports_nonfinite(lri):
    u, A0, A1 = one_step_case(6, 64, 64, 4)
    A1[3, 5] = np.nan
    with pytest.raises(lri.native.NonFiniteError):
        run_native_step(lri, "primal", u, [A0, A1])


def test_bad_indices_and_missing_state(lri):
    with lri.native.Context(32, 32, 4) as ctx:
        with pytest.raises(ValueError):
            ctx.step_ksl(lri.native.PRIMAL, 0, 1)  # no state
        rng = np.random.default_rng(0)
        u = random_factors(rng, 32, 32, 4)
        ctx.set_state(u.U, u.S, u.V)
        with pytest.raises(ValueError):
            ctx.step_ksl(lri.native.PRIMAL, 0, 1)  # no snapshots
        ctx.reserve(2, lri.native.F64)
        with pytest.raises(ValueError):
            ctx.step_ksl(lri.native.PRIMAL, 0, 2)
        with pytest.raises(ValueError):
            ctx.step_ksl(lri.native.STRANG, 0, 1)
        with pytest.raises(ValueError):
            ctx.set_state(u.U[:, :3], u.S[:3, :3], u.V)


def test_bitwise_deterministic(lri):
    u0, snaps = rotating_snaps(512, 16, 5, seed=2, dtype="bf16")
    a = run_solve(lri, u0, snaps, "primal", "mixed", "bf16")
    b = run_solve(lri, u0, snaps, "primal", "mixed", "bf16")
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_snapshot_pitch_is_transparent(lri):
    """BF16 snapshots in the library's slots (their pitch padded to n + 64 where n
    elements are a multiple of 4 KB), bound as device arrays at pitch n, and bound at
    pitch n + 64: the same bits after three mixed-precision KSL steps.  The bound
    arrays come from the HIP runtime the library itself uses (hipMalloc, hipMemcpy2D)."""
    import ctypes

    n, m, r = 2048, 600, 8
    d = rotating_data(n, m, r, noise=1e-3, seed=41)
    u0 = d.initial()
    snaps = [lri.api._encode(bf16_round(d(t)), "bf16") for t in np.arange(4) / 199.0]  # uint16, n x m
    try:
        hip = ctypes.CDLL("libamdhip64.so.7")
    except OSError:
        hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipMemcpy2D.argtypes = [vp, ctypes.c_size_t, vp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                ctypes.c_int]
    hip.hipFree.argtypes = [vp]
    host_to_device = 1
    allocs = []

    def run(setup):
        with lri.native.Context(n, m, r) as ctx:
            ctx.set_precision(lri.native.PREC_MIXED)
            ctx.set_state(u0.U, u0.S, u0.V)
            setup(ctx)
            for k in range(3):
                ctx.step_ksl(lri.native.PRIMAL, k, k + 1)
            return ctx.get_state()

    def owned(ctx):
        ctx.reserve(len(snaps), lri.native.BF16)
        for k, a in enumerate(snaps):
            ctx.upload(k, a)

    def bound(pad):
        def setup(ctx):
            ld = n + pad
            ptrs = []
            for a in snaps:
                p = vp()
                assert hip.hipMalloc(ctypes.byref(p), 2 * ld * m) == 0
                allocs.append(p.value)
                assert hip.hipMemcpy2D(p, 2 * ld, a.ctypes.data, 2 * n, 2 * n, m, host_to_device) == 0
                ptrs.append(p.value)
            ctx.bind(ptrs, lri.native.BF16, ld)
        return setup

    say(f"case: BF16 snapshot pitch n={n} m={m} r={r}: owned slots, bound at pitch {n} and {n + 64}")
    try:
        a, b, c = run(owned), run(bound(0)), run(bound(64))
    finally:
        for p in allocs:
            hip.hipFree(vp(p))
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, y) and np.array_equal(x, z)
