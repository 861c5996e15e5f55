"""Write tests/golden/cfg0_golden.npz from the CPU oracle.

Workload: rotating low-rank-plus-noise data, n = m = 256, rank 8, t in [0, 1],
dt = 1e-2 (100 steps), seed 0, function-form MatrixDataProblem; Y0 is the
rank-8 truncated SVD of A(0).  For each integrator the fixture holds the final
factors, the accumulated t of every step and ||Y_k||_F of every step.

The reference (Julia) is not runnable here, so these vectors come from the
oracle, whose steps are pinned by tests/test_oracle.py; the fixture freezes
them so the GPU tests and later oracle edits are checked against fixed data.

Run from the repo root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import dlra  # noqa: E402
from oracle.data import rotating_data  # noqa: E402
from oracle.metrics import frob_norm  # noqa: E402

N, R, SEED, DT = 256, 8, 0, 1e-2
ALGS = {
    "ksl_primal": dlra.ProjectorSplitting(dlra.PrimalLieTrotter()),
    "ksl_dual": dlra.ProjectorSplitting(dlra.DualLieTrotter()),
    "ksl_strang": dlra.ProjectorSplitting(dlra.Strang()),
    "bug": dlra.UnconventionalAlgorithm(),
    "ra_bug": dlra.RankAdaptiveUnconventionalAlgorithm(2e-2, rmax=16),
}


def compute():
    d = rotating_data(N, N, R, noise=1e-3, seed=SEED)
    u0 = d.initial()
    out = {"meta": np.array([N, N, R, SEED]), "dt": np.array(DT)}
    for name, alg in ALGS.items():
        sol = dlra.solve(dlra.MatrixDataProblem(d, u0, (0.0, 1.0)), alg, DT)
        yf = sol.Y[-1]
        out[f"{name}_U"] = yf.U
        out[f"{name}_S"] = yf.S
        out[f"{name}_V"] = yf.V
        out[f"{name}_t"] = np.array(sol.t, dtype=np.float64)
        out[f"{name}_norms"] = np.array([frob_norm(y.U, y.S, y.V) for y in sol.Y])
        out[f"{name}_ranks"] = np.array([y.rank for y in sol.Y])
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg0_golden.npz")
    np.savez_compressed(path, **compute())
    print("wrote", path, os.path.getsize(path), "bytes")
