"""The oracle reproduces the committed cfg0 fixture (n=m=256, r=8, 100 steps, seed 0).

The fixture (tests/golden/cfg0_golden.npz, written by tests/golden/make_golden.py)
freezes the oracle's output so that the GPU tests and any later change to the
oracle are checked against fixed vectors.
"""
import os

import numpy as np
import pytest

from oracle import dlra
from oracle.data import rotating_data
from oracle.metrics import frob_norm, rel_err

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "cfg0_golden.npz"))


@pytest.fixture(scope="module")
def data():
    d = rotating_data(256, 256, 8, noise=1e-3, seed=0)
    return d, d.initial()


@pytest.mark.parametrize("name,alg", [
    ("ksl_primal", dlra.ProjectorSplitting(dlra.PrimalLieTrotter())),
    ("bug", dlra.UnconventionalAlgorithm()),
])
def test_oracle_matches_fixture(golden, data, name, alg):
    d, u0 = data
    sol = dlra.solve(dlra.MatrixDataProblem(d, u0, (0.0, 1.0)), alg, 1e-2)
    yf = sol.Y[-1]
    assert rel_err(yf.U, yf.S, yf.V, golden[f"{name}_U"], golden[f"{name}_S"], golden[f"{name}_V"]) < 1e-12
    norms = np.array([frob_norm(y.U, y.S, y.V) for y in sol.Y])
    assert np.allclose(norms, golden[f"{name}_norms"], rtol=1e-12, atol=0)
    assert np.array_equal(np.array(sol.t), golden[f"{name}_t"])


def test_fixture_is_self_consistent(golden):
    assert list(golden["meta"]) == [256, 256, 8, 0]
    for name in ("ksl_primal", "ksl_dual", "ksl_strang", "bug", "ra_bug"):
        U, S, V = golden[f"{name}_U"], golden[f"{name}_S"], golden[f"{name}_V"]
        assert np.abs(U.T @ U - np.eye(U.shape[1])).max() < 1e-13
        assert np.abs(V.T @ V - np.eye(V.shape[1])).max() < 1e-13
        assert len(golden[f"{name}_t"]) == 101
        assert abs(golden[f"{name}_norms"][-1] - frob_norm(U, S, V)) < 1e-13
