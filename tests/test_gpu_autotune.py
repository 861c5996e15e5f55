"""The pass autotuner (lri_autotune_passes; needs an MI355X).

The tuner times every candidate reduction chunk of both passes and keeps the
fastest for the context.  A chunk only changes how the pass's partial results
are split and summed, so a tuned context must agree with the FP64 CPU oracle
to the FP64 bound (relative Frobenius error <= 1e-10) in FP64, and with the
FP64 engine to the mixed-precision bound (<= 1e-6) with BF16 snapshots.
"""
import sys

import numpy as np
import pytest

from oracle import dlra
from oracle.data import bf16_round, rotating_data
from oracle.metrics import rel_err

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def solve(lri, u0, snaps, precision, dtype, tune):
    n, m = u0.U.shape[0], u0.V.shape[0]
    with lri.native.Context(n, m, u0.rank) as ctx:
        ctx.set_precision(precision)
        ctx.set_state(u0.U, u0.S, u0.V)
        ctx.reserve(len(snaps), lri.native.DTYPES[dtype])
        for k, a in enumerate(snaps):
            ctx.upload(k, lri.api._encode(a, dtype))
        report = ctx.autotune_passes(0, 1, reps=3) if tune else None
        for k in range(len(snaps) - 1):
            ctx.step_ksl(lri.native.PRIMAL, k, k + 1)
        return ctx.get_state(), report


@pytest.mark.parametrize("n,r", [(1500, 16), (3000, 32)])
def test_tuned_fp64_matches_oracle(lri, n, r):
    say(f"case: autotuned FP64 primal KSL n={n} r={r}, 4 steps")
    d = rotating_data(n, n, r, noise=1e-3, seed=11)
    u0 = d.initial()
    snaps = [d(t) for t in np.arange(5) / 199.0]
    got, rep = solve(lri, u0, snaps, lri.native.PREC_F64, "f64", True)
    say(f"  report {rep}")
    for k in ("pass_n", "pass_t"):
        assert rep[k]["chunk"] >= 256 and rep[k]["chunk"] % 256 == 0 and rep[k]["splits"] >= 1
        assert 0 < rep[k]["us"] <= rep[k]["default_us"] * (1 + 1e-9)
    U, S, V = u0.U, u0.S, u0.V
    for k in range(len(snaps) - 1):
        U, S, V = dlra.ksl_primal_step(U, S, V, snaps[k + 1] - snaps[k])
    e = rel_err(*got, U, S, V)
    say(f"  tuned engine vs oracle: {e:.2e}")
    assert e <= 1e-10


def test_tuned_mixed_matches_fp64(lri):
    n, r = 4096, 32
    say(f"case: autotuned mixed primal KSL, BF16 snapshots n={n} r={r}, 6 steps")
    d = rotating_data(n, n, r, noise=1e-3, seed=12)
    u0 = d.initial()
    snaps = [bf16_round(d(t)) for t in np.arange(7) / 199.0]
    ref, _ = solve(lri, u0, snaps, lri.native.PREC_F64, "bf16", False)
    mix, rep = solve(lri, u0, snaps, lri.native.PREC_MIXED, "bf16", True)
    plain, _ = solve(lri, u0, snaps, lri.native.PREC_MIXED, "bf16", False)
    say(f"  report {rep}")
    e, e0 = rel_err(*mix, *ref), rel_err(*plain, *ref)
    say(f"  tuned mixed vs FP64: {e:.2e} (default plan: {e0:.2e})")
    assert e <= 1e-6
