"""FP64 CPU restatement of the reference's matrix-ODE (MatrixDEProblem) step path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Reference files followed:

* DE sub-problem wiring, KSL ......... ``src/integrators/projector_splitting.jl:43-85,129-189``
* DE sub-problem wiring, BUG ......... ``src/integrators/unconventional.jl:39-84,133-157``
* DE sub-problem wiring, RA-BUG ...... ``src/integrators/rank_adaptive_unconventional.jl:47-92,194-233``
* Greedy on a MatrixHybridProblem .... ``src/integrators/greedy_integrator.jl:41-46,61-82,106-111``
* driver ............................. ``src/primitives.jl:68-80``

The reference integrates each sub-problem (K, S, L) with an OrdinaryDiffEq
integrator (default Tsit5, adaptive) stepped by exactly dt
(``step!(integ, dt, true)``).  This restatement uses the classical fixed-step
RK4 with N sub-steps of h = dt/N for every sub-problem (the ``RK4`` sub-step
algorithm of the engine); Tsit5 itself is not restated ("parity unpinned" for
Tsit5-driven results: that needs DifferentialEquations.jl).

The right-hand sides are the reference's defaults:
  K' = F(K V0', t) V0                     (projector_splitting.jl:53-54)
  L' = F(U0 L', t)' U0                    (:73-74)
  S' = -U1' F(U1 S V0', t) V0  (KSL, :63-64);  S' = U1' F(U1 S V1', t) V1 (BUG, unconventional.jl:72-73)
F acts on dense n x m matrices here (small sizes only).

Pinning (``tests/test_oracle_de.py``): the RK4 tableau on scalar linear ODEs
(local error order 5, against the exact exponential); the sub-step flows
against the matrix exponential of the linear sub-problem when F is linear; the
closed forms of one step in the small-dt limit; exactness for F whose flow
stays in a fixed rank-r manifold; agreement of the K/L/S sub-problems with the
data path when F(Y, t) = dA/dt is constant in Y.
"""
from __future__ import annotations

import math

from dataclasses import dataclass, field
from typing import Any, Callable, List

import numpy as np

from .dlra import DLRSolution, PrimalLieTrotter, DualLieTrotter, Strang, ProjectorSplitting, \
    UnconventionalAlgorithm, RankAdaptiveUnconventionalAlgorithm, _init_sol_len, _qr
from .lra import SVDLike, truncate_to_tolerance


@dataclass
class MatrixDEProblem:
    """``MatrixDEProblem{f,u0,tspan}`` (primitives.jl:13-17): dY/dt = f(Y, t)."""

    f: Callable[[np.ndarray, float], np.ndarray]
    u0: SVDLike
    tspan: tuple


@dataclass
class LyapunovOperator:
    """F(Y, t) = A Y + Y B' + sum_k g_k h_k' with dense A (n x n), B (m x m), G = g h'."""

    A: np.ndarray
    B: np.ndarray
    g: np.ndarray  # n x q
    h: np.ndarray  # m x q

    def __call__(self, Y, t=0.0):
        return self.A @ Y + Y @ self.B.T + self.g @ self.h.T


@dataclass
class QuadraticOperator:
    """F(Y, t) = A Y + Y B' + g h' - (D Y) o Y (o: elementwise), dense A, B, D.

    With A = nu * Laplacian and D = central gradient (both periodic) and no B,
    g, h this is the viscous Burgers right-hand side of the reference's
    ``test/data_agnostic_approximation.jl:31-33`` (Delta*rho - (grad*rho) .* rho),
    each column one parameter sample."""

    A: np.ndarray
    B: np.ndarray
    g: np.ndarray  # n x q
    h: np.ndarray  # m x q
    D: np.ndarray

    def __call__(self, Y, t=0.0):
        return self.A @ Y + Y @ self.B.T + self.g @ self.h.T - (self.D @ Y) * Y


def burgers_matrices(n: int, nu: float, length: float = math.pi):
    """Delta (times nu) and the central gradient on a periodic grid of n cells,
    entry by entry as ``test/data_agnostic_approximation.jl:6-29`` builds them."""
    dx = length / n
    left = lambda i: i - 1 if i > 0 else n - 1  # noqa: E731
    right = lambda i: i + 1 if i < n - 1 else 0  # noqa: E731
    Lap = np.zeros((n, n))
    Grad = np.zeros((n, n))
    for i in range(n):
        Lap[i, left(i)] = nu / dx ** 2
        Lap[i, i] = -2 * nu / dx ** 2
        Lap[i, right(i)] = nu / dx ** 2
    for i in range(n):
        Grad[i, left(i)] = -1 / 2 / dx
        Grad[i, right(i)] = 1 / 2 / dx
    return Lap, Grad


def burgers_initial(n: int, m: int, sigma=(0.5, 0.5), length: float = math.pi):
    """rho0_mat of ``test/data_agnostic_approximation.jl:35-44``: the deterministic
    profile plus sigma_1 xi_1 sin(2 pi x) + sigma_2 xi_2 sin(3 pi x) over an m x m
    grid of (xi_1, xi_2) in [-1, 1]^2 (columns in Julia's column-major order of the
    comprehension), n x m^2."""
    dx = length / n
    x = dx / 2 + dx * np.arange(n)
    ub = 0.5 * (np.exp(np.cos(x)) - 1.5) * np.sin(x + 2 * math.pi * 0.37)
    xs = np.linspace(-1.0, 1.0, m)
    cols = []
    for x2 in xs:  # the comprehension's first index (xi_1) runs fastest
        for x1 in xs:
            cols.append(ub + sigma[0] * x1 * np.sin(2 * math.pi * x) + sigma[1] * x2 * np.sin(3 * math.pi * x))
    return np.stack(cols, axis=1)


def dirichlet_laplacian(n: int, scale: float | None = None) -> np.ndarray:
    """Tridiagonal [1, -2, 1] * scale (default n^2), homogeneous Dirichlet boundaries."""
    s = float(n * n) if scale is None else scale
    return s * (np.diag(-2.0 * np.ones(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1))


def rk4_substeps(rho: float, dt: float, cap: float = 2.5) -> int:
    """Fixed RK4 sub-steps with h*rho <= cap (RK4 is stable on [-2.78, 0] of the real axis)."""
    return max(1, int(np.ceil(dt * rho / cap)))


def rk4(f, y, t, dt, N):
    """N classical RK4 steps of size dt/N for y' = f(y, t)."""
    h = dt / N
    for s in range(N):
        ts = t + s * h
        k1 = f(y, ts)
        k2 = f(y + 0.5 * h * k1, ts + 0.5 * h)
        k3 = f(y + 0.5 * h * k2, ts + 0.5 * h)
        k4 = f(y + h * k3, ts + h)
        y = y + (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    return y


# ----------------------------------------------------------------------------
# steps
# ----------------------------------------------------------------------------


def _K_rhs(F, V):
    return lambda K, t: F(K @ V.T, t) @ V


def _L_rhs(F, U):
    return lambda L, t: F(U @ L.T, t).T @ U


def ksl_primal_de_step(F, U, S, V, t, dt, N):
    """projector_splitting.jl:129-152 with the DE sub-problems (:52-80)."""
    K = rk4(_K_rhs(F, V), U @ S, t, dt, N)
    U1, R = _qr(K)
    St = rk4(lambda S_, tt: -(U1.T @ F(U1 @ S_ @ V.T, tt) @ V), R, t, dt, N)
    L = rk4(_L_rhs(F, U1), V @ St.T, t, dt, N)
    V1, R2 = _qr(L)
    return U1, R2.T.copy(), V1


def ksl_dual_de_step(F, U, S, V, t, dt, N):
    """projector_splitting.jl:166-189 with the DE sub-problems."""
    L = rk4(_L_rhs(F, U), V @ S.T, t, dt, N)
    V1, R = _qr(L)
    St = rk4(lambda S_, tt: -(U.T @ F(U @ S_ @ V1.T, tt) @ V1), R.T.copy(), t, dt, N)
    K = rk4(_K_rhs(F, V1), U @ St, t, dt, N)
    U1, R2 = _qr(K)
    return U1, R2, V1


def bug_de_step(F, U, S, V, t, dt, N):
    """unconventional.jl:133-157 with the DE sub-problems (:51-79)."""
    K = rk4(_K_rhs(F, V), U @ S, t, dt, N)
    U1, _ = _qr(K)
    M = U1.T @ U
    L = rk4(_L_rhs(F, U), V @ S.T, t, dt, N)
    V1, _ = _qr(L)
    Nn = V1.T @ V
    S1 = rk4(lambda S_, tt: U1.T @ F(U1 @ S_ @ V1.T, tt) @ V1, M @ S @ Nn.T, t, dt, N)
    return U1, S1, V1


def rank_adaptive_bug_de_step(F, U, S, V, t, dt, N, tol, rmax):
    """rank_adaptive_unconventional.jl:194-233 with the DE sub-problems (:59-86)."""
    K = rk4(_K_rhs(F, V), U @ S, t, dt, N)
    Uh, _ = _qr(np.hstack([K, U]))
    M = Uh.T @ U
    L = rk4(_L_rhs(F, U), V @ S.T, t, dt, N)
    Vh, _ = _qr(np.hstack([L, V]))
    Nn = Vh.T @ V
    Sh = rk4(lambda S_, tt: Uh.T @ F(Uh @ S_ @ Vh.T, tt) @ Vh, M @ S @ Nn.T, t, dt, N)
    Us, sig, Vst = np.linalg.svd(Sh)
    r1 = int(min(rmax, truncate_to_tolerance(sig, tol)))
    return Uh @ Us[:, :r1], np.diag(sig[:r1]), Vh @ Vst[:r1].T, r1


class LowRankLyapunov:
    """F(Y) = A Y + Y B' + G H' applied to low-rank arguments without forming n x m
    matrices, as LowRankArithmetic propagates the factors of a TwoFactor /
    SVDLike argument through such an F (the algebra below is exact).  A and B
    may be scipy sparse matrices; G is n x q and H is m x q."""

    def __init__(self, A, B, G, H):
        self.A, self.B = A, B
        self.G = np.asarray(G, dtype=np.float64).reshape(A.shape[0], -1)
        self.H = np.asarray(H, dtype=np.float64).reshape(B.shape[0], -1)

    def dense(self) -> LyapunovOperator:
        d = lambda X: X.toarray() if hasattr(X, "toarray") else np.asarray(X)
        return LyapunovOperator(d(self.A), d(self.B), self.G, self.H)

    def K_rhs(self, V):
        """K' = F(K V') V = A K + K (V'B'V) + G (H'V)  (V'V = I)."""
        Bt = V.T @ (self.B.T @ V)
        c = self.G @ (self.H.T @ V)
        return lambda K, t: self.A @ K + K @ Bt + c

    def L_rhs(self, U):
        """L' = F(U L')' U = B L + L (U'AU)' + H (G'U)  (U'U = I)."""
        At = U.T @ (self.A @ U)
        c = self.H @ (self.G.T @ U)
        return lambda L, t: self.B @ L + L @ At.T + c

    def S_rhs(self, U, V, sign=1.0):
        """S' = sign U'F(U S V')V = sign ((U'AU) S + S (V'B'V) + (U'G)(H'V))."""
        At = U.T @ (self.A @ U)
        Bt = V.T @ (self.B.T @ V)
        c = (U.T @ self.G) @ (self.H.T @ V)
        return lambda S_, t: sign * (At @ S_ + S_ @ Bt + c)


def bug_de_step_lowrank(op: LowRankLyapunov, U, S, V, t, dt, N):
    """unconventional.jl:133-157 with the DE sub-problems evaluated in factored form
    (the cost profile of the reference with LowRankArithmetic: O((n + m) r^2) per
    right-hand side instead of O(n m r))."""
    K = rk4(op.K_rhs(V), U @ S, t, dt, N)
    U1, _ = _qr(K)
    M = U1.T @ U
    L = rk4(op.L_rhs(U), V @ S.T, t, dt, N)
    V1, _ = _qr(L)
    Nn = V1.T @ V
    S1 = rk4(op.S_rhs(U1, V1), M @ S @ Nn.T, t, dt, N)
    return U1, S1, V1


def rank_adaptive_bug_de_step_lowrank(op: LowRankLyapunov, U, S, V, t, dt, N, tol, rmax):
    """rank_adaptive_unconventional.jl:194-233 with the DE sub-problems in factored form
    (as ``bug_de_step_lowrank``); the S sub-problem in the augmented bases is
    S' = Uh' F(Uh S Vh') Vh = (Uh'AUh) S + S (Vh'B'Vh) + (Uh'G)(H'Vh)."""
    K = rk4(op.K_rhs(V), U @ S, t, dt, N)
    Uh, _ = _qr(np.hstack([K, U]))
    M = Uh.T @ U
    L = rk4(op.L_rhs(U), V @ S.T, t, dt, N)
    Vh, _ = _qr(np.hstack([L, V]))
    Nn = Vh.T @ V
    Sh = rk4(op.S_rhs(Uh, Vh), M @ S @ Nn.T, t, dt, N)
    Us, sig, Vst = np.linalg.svd(Sh)
    r1 = int(min(rmax, truncate_to_tolerance(sig, tol)))
    return Uh @ Us[:, :r1], np.diag(sig[:r1]), Vh @ Vst[:r1].T, r1


# ----------------------------------------------------------------------------
# the Greedy integrator on a MatrixHybridProblem (greedy_integrator.jl:41-82)
# ----------------------------------------------------------------------------


@dataclass
class MatrixHybridProblem:
    """``MatrixHybridProblem{y,f,u0,tspan}`` (primitives.jl:36-41): data y(t), and the
    Z dynamics dZ/dt = f(Z, U, t) of a two-factor state U Z'.  Here f is given by
    an operator F on n x m matrices, f(Z, U, t) = F(U Z', t)' U, the form of the
    reference's own hybrid test (test/data_informed_approximation.jl:75,
    FZ(Z, U, t) = burgers(U Z')' U)."""

    y: Callable[[float], np.ndarray]
    F: Callable[[np.ndarray, float], np.ndarray]
    U0: np.ndarray
    Z0: np.ndarray
    tspan: tuple


def greedy_hybrid_step(F, U, Z, X, t, dt, N):
    """greedy_step!(::TwoFactorRepresentation, ..., ::Type{<:MatrixHybridProblem})
    (greedy_integrator.jl:73-82): the Z integrator (an ODEProblem in Z with the
    parameter U, :44) advances Z over [t, t + dt] with U held (here N classical
    RK4 steps instead of Tsit5), then X = y(t + dt), XZ = X Z and U = Q P' from the
    SVD XZ = Q Sigma P' (U is updated in place after Z, so the Z step sees the old U).
    Returns (U1, Z1)."""
    Z1 = rk4(_L_rhs(F, U), Z, t, dt, N)
    Q, _, Pt = np.linalg.svd(X @ Z1, full_matrices=False)
    return Q @ Pt, Z1


def solve_hybrid(prob: MatrixHybridProblem, dt, substeps: int):
    """``solve(prob::MatrixHybridProblem, GreedyIntegrator(), dt)`` (primitives.jl:68-76,
    greedy_integrator.jl:61-71,106-111); returns the list of (U, Z) and of t."""
    t0, tf = prob.tspan
    if not tf > t0:
        raise AssertionError("Integration in reverse time direction is not supported")
    U, Z = prob.U0.copy(), prob.Z0.copy()
    Ys: List[Any] = [None] * _init_sol_len(dt, t0, tf)
    ts: List[Any] = list(np.linspace(t0, tf, len(Ys)))
    Ys[0] = (U.copy(), Z.copy())
    t, it, T = t0, 0, tf - t0
    while (tf - t) / T > 1e-8:
        U, Z = greedy_hybrid_step(prob.F, U, Z, np.asarray(prob.y(t + dt), dtype=np.float64), t, dt, substeps)
        t = t + dt
        it += 1
        if it <= len(Ys) - 1:
            Ys[it] = (U.copy(), Z.copy())
            ts[it] = t
        else:
            Ys.append((U.copy(), Z.copy()))
            ts.append(t)
    return Ys, ts


def solve_de(prob: MatrixDEProblem, alg, dt, substeps: int) -> DLRSolution:
    """``solve(prob::MatrixDEProblem, alg, dt)`` (primitives.jl:68-76) with RK4 sub-steps."""
    t0, tf = prob.tspan
    if not tf > t0:
        raise AssertionError("Integration in reverse time direction is not supported")
    u = prob.u0.copy()
    Ys: List[Any] = [None] * _init_sol_len(dt, t0, tf)
    ts: List[Any] = list(np.linspace(t0, tf, len(Ys)))
    Ys[0] = prob.u0.copy()
    t, it, T = t0, 0, tf - t0
    F = prob.f
    while (tf - t) / T > 1e-8:
        U, S, V = u.U, u.S, u.V
        if isinstance(alg, ProjectorSplitting):
            if isinstance(alg.order, PrimalLieTrotter):
                U, S, V = ksl_primal_de_step(F, U, S, V, t, dt, substeps)
            elif isinstance(alg.order, DualLieTrotter):
                U, S, V = ksl_dual_de_step(F, U, S, V, t, dt, substeps)
            elif isinstance(alg.order, Strang):
                U, S, V = ksl_primal_de_step(F, U, S, V, t, dt / 2, substeps)
                U, S, V = ksl_dual_de_step(F, U, S, V, t + dt / 2, dt / 2, substeps)
        elif isinstance(alg, UnconventionalAlgorithm):
            U, S, V = bug_de_step(F, U, S, V, t, dt, substeps)
        elif isinstance(alg, RankAdaptiveUnconventionalAlgorithm):
            U, S, V, _ = rank_adaptive_bug_de_step(F, U, S, V, t, dt, substeps, alg.tol, alg.rmax)
        else:
            raise TypeError(f"unsupported algorithm {alg!r}")
        u = SVDLike(U, S, V)
        t = t + dt
        it += 1
        if it <= len(Ys) - 1:
            Ys[it] = u.copy()
            ts[it] = t
        else:
            Ys.append(u.copy())
            ts.append(t)
    return DLRSolution(Ys, ts)
