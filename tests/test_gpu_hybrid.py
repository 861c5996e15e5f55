"""The Greedy integrator on a MatrixHybridProblem (needs an MI355X).

The engine's step (``lri_step_hybrid_greedy``) against the CPU oracle
(``oracle.dlra_de.greedy_hybrid_step``: greedy_integrator.jl:73-82 with the Z
dynamics f(Z, U, t) = F(U Z')' U integrated by fixed-step RK4), FP64, relative
Frobenius error of Y = U Z' <= 1e-10; and the reference's own hybrid test
(test/data_informed_approximation.jl): data-informed Burgers with n = 1000 cells,
10 x 10 parameter samples, r = 15, dt = 1e-2 over [0, 1], final error against
the true solution <= 0.1.
"""
import math
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import dlra_de

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_y(U1, Z1, U2, Z2):
    Y1, Y2 = U1 @ Z1.T, U2 @ Z2.T
    return np.linalg.norm(Y1 - Y2) / np.linalg.norm(Y2)


def test_hybrid_linear_operator_matches_oracle(lri):
    """Variable-coefficient A and B with a rank-1 forcing (the RK4 column kernel),
    function-form data, 6 steps."""
    rng = np.random.default_rng(41)
    n, m, r, dt, N, steps = 200, 150, 6, 0.05, 8, 6
    A = lri.SymTridiagonal(-rng.uniform(4, 8, n), rng.uniform(1, 2, n - 1))
    B = lri.SymTridiagonal(-rng.uniform(4, 8, m), rng.uniform(1, 2, m - 1))
    g, h = rng.standard_normal((n, 1)), rng.standard_normal((m, 1))
    F = lri.LyapunovOperator(A, B, g, h)
    X0, X1 = rng.standard_normal((n, m)), rng.standard_normal((n, m))
    y = lambda t: X0 + math.sin(t) * X1  # noqa: E731
    U0, _ = np.linalg.qr(rng.standard_normal((n, r)))
    Z0 = y(0.0).T @ U0
    say(f"case: hybrid Greedy, linear operator, n={n} m={m} r={r}, {steps} steps")
    prob = lri.MatrixHybridProblem(y, F, lri.TwoFactorRepresentation(U0, Z0), (0.0, steps * dt))
    sol = lri.solve(prob, lri.GreedyIntegrator(), dt, substeps=N)
    host = dlra_de.LyapunovOperator(A.dense(), B.dense(), g, h)
    Ys, ts = dlra_de.solve_hybrid(dlra_de.MatrixHybridProblem(y, host, U0, Z0, (0.0, steps * dt)), dt, N)
    assert len(sol.Y) == len(Ys) == steps + 1
    for k in (1, steps):
        e = rel_y(sol.Y[k].U, sol.Y[k].Z, *Ys[k])
        say(f"  step {k}: rel err {e:.2e}")
        assert e <= 1e-10
    assert np.linalg.norm(sol.Y[-1].U.T @ sol.Y[-1].U - np.eye(r)) <= 1e-12


def test_hybrid_constant_operator_array_data(lri):
    """Constant-coefficient A (the analytic eigenbasis path), array-form data."""
    rng = np.random.default_rng(42)
    n, m, r, N = 128, 96, 5, 4
    A = lri.laplacian_1d(n, 0.3)  # h rho = 0.25 (4 (0.3 + 0.2)) = 0.5: RK4-stable with dt = 1, N = 4
    B = lri.laplacian_1d(m, 0.2, periodic=True)
    F = lri.LyapunovOperator(A, B)
    snaps = [rng.standard_normal((n, m)) for _ in range(5)]
    U0, _ = np.linalg.qr(rng.standard_normal((n, r)))
    Z0 = snaps[0].T @ U0
    prob = lri.MatrixHybridProblem(snaps, F, lri.TwoFactorRepresentation(U0, Z0))
    sol = lri.solve(prob, lri.GreedyIntegrator(), 1, substeps=N)
    host = dlra_de.LyapunovOperator(A.dense(), B.dense(), np.zeros((n, 0)), np.zeros((m, 0)))
    U, Z = U0, Z0
    for k in range(4):  # t = 1..5, dt = 1, X = y[t + dt]
        U, Z = dlra_de.greedy_hybrid_step(host, U, Z, snaps[k + 1], k + 1, 1, N)
    e = rel_y(sol.Y[-1].U, sol.Y[-1].Z, U, Z)
    say(f"case: hybrid Greedy, constant operator, array data: rel err {e:.2e}")
    assert e <= 1e-10


def burgers_truth(n, msq, nu, dt, steps, length=math.pi):
    """The true solution of the reference's hybrid test (test/data_informed_approximation.jl:40-62):
    each sample column integrated on the periodic grid (here by RK4 with h * rho <= 1, in place of
    DifferentialEquations' adaptive solve), saved at t = 0, dt, ..., steps * dt."""
    dx = length / n
    left = np.roll(np.arange(n), 1)
    right = np.roll(np.arange(n), -1)
    lap = sp.csr_matrix((np.concatenate([np.full(n, nu / dx ** 2), np.full(n, -2 * nu / dx ** 2),
                                         np.full(n, nu / dx ** 2)]),
                         (np.tile(np.arange(n), 3), np.concatenate([left, np.arange(n), right]))), shape=(n, n))
    grad = sp.csr_matrix((np.concatenate([np.full(n, -0.5 / dx), np.full(n, 0.5 / dx)]),
                          (np.tile(np.arange(n), 2), np.concatenate([left, right]))), shape=(n, n))
    x = dx / 2 + dx * np.arange(n)
    ub = 0.5 * (np.exp(np.cos(x)) - 1.5) * np.sin(x + 2 * math.pi * 0.37)
    xs = np.linspace(-1, 1, msq)
    cols = [ub + 0.5 * x1 * np.sin(2 * math.pi * x) + 0.5 * x2 * np.sin(3 * math.pi * x) for x2 in xs for x1 in xs]
    rho = np.stack(cols, axis=1)
    f = lambda R: lap @ R - (grad @ R) * R  # noqa: E731
    out = [rho.copy()]
    sub = int(math.ceil(dt * (4 * nu / dx ** 2 + 2 * np.abs(rho).max() / dx)))
    h = dt / sub
    for _ in range(steps):
        for _ in range(sub):
            k1 = f(rho)
            k2 = f(rho + 0.5 * h * k1)
            k3 = f(rho + 0.5 * h * k2)
            k4 = f(rho + h * k3)
            rho = rho + (h / 6) * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(rho.copy())
    return out, lap, grad


def test_reference_data_informed_burgers(lri):
    """test/data_informed_approximation.jl on the engine: U0 from the rank-15 truncated SVD of
    [Y(0) Y(dt)], Z0 = Y(0)' U0, f(Z, U, t) = burgers(U Z')' U, data interpolated from the true
    solution; the reference asserts a final relative error <= 0.1.  The engine also matches the
    oracle's hybrid solve (the same RK4 sub-steps) to 1e-9."""
    n, msq, r, nu, dt = 1000, 10, 15, 0.005, 1e-2
    steps = 100
    say(f"case: data-informed Burgers n={n} m={msq * msq} r={r}, {steps} steps (true solution on the CPU)")
    truth, lap, grad = burgers_truth(n, msq, nu, dt, steps)
    t_range_len = len(truth)

    def data(t):  # interpolate_data (test/data_informed_approximation.jl:70-73), 0-based
        idx = min(int(math.floor(t / dt)), t_range_len - 1)
        return truth[idx]

    Ub, _, _ = np.linalg.svd(np.hstack([truth[0], truth[1]]), full_matrices=False)
    U0 = Ub[:, :r].copy()
    Z0 = truth[0].T @ U0
    F = lri.burgers_operator(n, msq * msq, nu)
    prob = lri.MatrixHybridProblem(data, F, lri.TwoFactorRepresentation(U0, Z0), (0.0, 1.0))
    N = lri.rk4_substeps(F.spectral_bound(2.0 * float(np.abs(truth[0]).max())), dt)
    sol = lri.solve(prob, lri.GreedyIntegrator(), dt, substeps=N)
    yT = sol.Y[-1]
    err = np.linalg.norm(truth[-1] - yT.U @ yT.Z.T) / np.linalg.norm(truth[-1])
    say(f"  {len(sol.Y) - 1} steps, {N} RK4 sub-steps each: error vs the true solution {err:.3e}")
    assert err <= 0.1
    host = dlra_de.QuadraticOperator(lap, np.zeros((msq * msq, msq * msq)), np.zeros((n, 0)),
                                     np.zeros((msq * msq, 0)), grad)
    Ys, _ = dlra_de.solve_hybrid(dlra_de.MatrixHybridProblem(data, host, U0, Z0, (0.0, 1.0)), dt, N)
    e = rel_y(yT.U, yT.Z, *Ys[-1])
    say(f"  engine vs oracle hybrid solve: {e:.2e}")
    assert e <= 1e-9
