"""FP64 CPU restatement of the reference's data-driven DLRA step path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Reference files followed (paths relative to the reference checkout):

* driver, problem and solution types ...... ``src/primitives.jl:13-104``
* data sub-integrator, snapshot access ..... ``src/integrators/data_integrator.jl:1-28``
* projector splitting (KSL) ................ ``src/integrators/projector_splitting.jl:87-211``
* unconventional (BUG) ..................... ``src/integrators/unconventional.jl:86-164``
* rank-adaptive BUG ........................ ``src/integrators/rank_adaptive_unconventional.jl:106-233``
* greedy (data problems) ................... ``src/integrators/greedy_integrator.jl:24-39,84-111``

Each step exists in two modes:

``faithful``  the reference's operation order: the increment is materialised
              as ``ycurr - yprev``, ``I'*dy`` / ``dy'`` copies are made, and the
              S sub-step makes a third pass ``(sign*left')*dy*right`` over dy.
``efficient`` the same mathematics with two passes over dy (the S sub-step of
              KSL reuses ``dy*V0``), the form the HIP engine implements.

QR is LAPACK Householder (``numpy.linalg.qr`` = dgeqrf/dorgqr, the same
reflector convention as Julia's ``qr!`` = geqrt), SVD is LAPACK gesdd (Julia's
default ``svd``), so the factors carry the reference's sign conventions.

Pinning (``tests/test_oracle.py``)
---------------------------------
* every step: the basis-free closed forms, brute-forced with explicit n x n
  projectors built from SVD-based (not QR-based) orthonormal bases;
* KSL and BUG: the exactness theorems (rank-r data is reproduced exactly);
* invariants: orthonormal factors, triangular/diagonal S shapes, dy = 0 leaves
  Y unchanged, RA rank bounds;
* faithful vs efficient modes agree;
* the reference's own test (``test/data_driven_approximation.jl:27-31``):
  function-form and snapshot-form data give the same final Y;
* driver semantics (step count, accumulated t, 1-based snapshot indexing).
* ``rank_adaptive_bug_step``'s truncation rule inherits "parity unpinned" from
  ``lra.truncate_to_tolerance``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Callable, List, Sequence, Union

import numpy as np

from .lra import SVDLike, TwoFactor, truncate_to_tolerance

# ----------------------------------------------------------------------------
# problem / algorithm / solution types  (primitives.jl:13-63, *_integrator.jl)
# ----------------------------------------------------------------------------


@dataclass
class MatrixDataProblem:
    """``MatrixDataProblem{y,u0,tspan}`` (primitives.jl:23-30).

    ``y`` is a callable t -> (n x m) array or a sequence of snapshots.  For a
    sequence, ``tspan`` defaults to ``(1, len(y))`` (1-based Int indices,
    primitives.jl:28-30).
    """

    y: Union[Callable[[float], np.ndarray], Sequence[np.ndarray]]
    u0: SVDLike
    tspan: tuple = None

    def __post_init__(self):
        if self.tspan is None:
            if callable(self.y):
                raise ValueError("a function-form MatrixDataProblem needs tspan")
            self.tspan = (1, len(self.y))


class PrimalLieTrotter:  # projector_splitting.jl:1
    pass


class DualLieTrotter:  # projector_splitting.jl:2
    pass


class Strang:  # projector_splitting.jl:3
    pass


@dataclass
class ProjectorSplitting:  # projector_splitting.jl:31-41
    order: Any = field(default_factory=PrimalLieTrotter)


@dataclass
class UnconventionalAlgorithm:  # unconventional.jl:13-21
    pass


@dataclass
class RankAdaptiveUnconventionalAlgorithm:  # rank_adaptive_unconventional.jl:16-23
    tol: float
    rmax: int = 2 ** 62


@dataclass
class GreedyIntegrator:  # greedy_integrator.jl:16-22 (Z_alg / Z_kwargs: hybrid problems only)
    pass


@dataclass
class DLRSolution:  # primitives.jl:46-49
    Y: List[SVDLike]
    t: List[Any]


# ----------------------------------------------------------------------------
# data access  (data_integrator.jl:22-28)
# ----------------------------------------------------------------------------


def update_data(y, t, dt) -> np.ndarray:
    """``update_data!``: y(t+dt) for a function, a copy of y[t+dt] (1-based) for arrays."""
    if callable(y):
        return np.array(y(t + dt), dtype=np.float64)
    if not (isinstance(t, (int, np.integer)) and isinstance(dt, (int, np.integer))):
        # data_integrator.jl:26 only defines the array method for Int t and dt
        raise TypeError("array-form data needs integer t and dt (no method for a fractional index)")
    return np.array(y[t + dt - 1], dtype=np.float64)


def initial_data(y, t0) -> np.ndarray:
    """``yprev`` at cache construction (projector_splitting.jl:91)."""
    if callable(y):
        return np.array(y(t0), dtype=np.float64)
    return np.array(y[0], dtype=np.float64)


# ----------------------------------------------------------------------------
# helpers
# ----------------------------------------------------------------------------


def _qr(A):
    """LAPACK Householder thin QR (Julia ``qr!`` + ``Matrix(F.Q)``)."""
    Q, R = np.linalg.qr(A, mode="reduced")
    return Q, R


def _data_integrate(u, sign, left, dy, right, faithful):
    """``MatrixDataIntegrator`` step: ``u .+= sign*left'*dy*right`` (data_integrator.jl:14-16).

    ``left is None`` stands for ``I`` (UniformScaling).  In faithful mode the
    reference's ``I'*dy`` copy and left-to-right evaluation are reproduced.
    """
    if left is None:
        prod = (np.array(dy, copy=True) if faithful else dy) @ right
        return u + sign * prod
    prod = ((sign * left.T) @ dy) @ right
    return u + prod


# ----------------------------------------------------------------------------
# the steps
# ----------------------------------------------------------------------------


def ksl_primal_step(U, S, V, dA, mode="efficient"):
    """Primal Lie-Trotter KSL step for data (projector_splitting.jl:129-152).

    K: US = U*S; US += dA*V; U1, R = qr(US)            (:133-138)
    S: S~ = R - U1'*dA*V                               (:141-142, sign -1 at :99)
    L: VS = V*S~'; VS += dA'*U1; V1, R2 = qr(VS); S1 = R2'   (:145-151)
    """
    faithful = mode == "faithful"
    US = U @ S
    if faithful:
        US = _data_integrate(US, 1, None, dA, V, True)
    else:
        P = dA @ V
        US = US + P
    U1, R = _qr(US)
    if faithful:
        St = _data_integrate(R, -1, U1, dA, V, True)
        VS = V @ St.T
        VS = _data_integrate(VS, 1, None, np.array(dA.T, copy=True), U1, True)
    else:
        St = R - U1.T @ P
        VS = V @ St.T + dA.T @ U1
    V1, R2 = _qr(VS)
    return U1, R2.T.copy(), V1


def ksl_dual_step(U, S, V, dA, mode="efficient"):
    """Dual Lie-Trotter KSL step for data (projector_splitting.jl:166-189).

    L: VS = V*S'; VS += dA'*U; V1, R = qr(VS)          (:170-175)
    S: S~ = R' - U'*dA*V1                              (:178-179)
    K: US = U*S~; US += dA*V1; U1, R2 = qr(US); S1 = R2   (:182-188)
    """
    faithful = mode == "faithful"
    VS = V @ S.T
    if faithful:
        VS = _data_integrate(VS, 1, None, np.array(dA.T, copy=True), U, True)
    else:
        Q = dA.T @ U
        VS = VS + Q
    V1, R = _qr(VS)
    if faithful:
        St = _data_integrate(R.T.copy(), -1, U, dA, V1, True)
        US = U @ St
        US = _data_integrate(US, 1, None, dA, V1, True)
    else:
        St = R.T - Q.T @ V1
        US = U @ St + dA @ V1
    U1, R2 = _qr(US)
    return U1, R2, V1


def bug_step(U, S, V, dA, mode="efficient"):
    """Unconventional (BUG) step for data (unconventional.jl:133-157).

    K from old V: U1 = orth(U*S + dA*V), M = U1'*U       (:137-142)
    L from old U: V1 = orth(V*S' + dA'*U), N = V1'*V     (:145-150)
    S1 = M*S*N' + U1'*dA*V1                              (:154-156, sign +1 at :102)
    """
    faithful = mode == "faithful"
    if faithful:
        US = _data_integrate(U @ S, 1, None, dA, V, True)
        VS = _data_integrate(V @ S.T, 1, None, np.array(dA.T, copy=True), U, True)
    else:
        US = U @ S + dA @ V
        VS = V @ S.T + dA.T @ U
    U1, _ = _qr(US)
    M = U1.T @ U
    V1, _ = _qr(VS)
    N = V1.T @ V
    S0 = M @ S @ N.T
    if faithful:
        S1 = _data_integrate(S0, 1, U1, dA, V1, True)
    else:
        S1 = S0 + U1.T @ (dA @ V1)
    return U1, S1, V1


def rank_adaptive_bug_step(U, S, V, dA, tol, rmax, mode="efficient"):
    """Rank-adaptive BUG step for data (rank_adaptive_unconventional.jl:194-233).

    Uhat = orth([U*S + dA*V | U]) (n x 2r), M = Uhat'*U      (:198-206)
    Vhat = orth([V*S' + dA'*U | V]) (m x 2r), N = Vhat'*V    (:209-217)
    Shat = M*S*N' + Uhat'*dA*Vhat (2r x 2r)                  (:219-220)
    Uh, sig, Vh = svd(Shat); r1 = min(rmax, truncate_to_tolerance(sig, tol))
    U1 = Uhat*Uh[:, :r1], S1 = diag(sig[:r1]), V1 = Vhat*Vh[:, :r1]  (:222-231)
    Returns (U1, S1, V1, r1).
    """
    faithful = mode == "faithful"
    if faithful:
        US = _data_integrate(U @ S, 1, None, dA, V, True)
        VS = _data_integrate(V @ S.T, 1, None, np.array(dA.T, copy=True), U, True)
    else:
        US = U @ S + dA @ V
        VS = V @ S.T + dA.T @ U
    Uhat, _ = _qr(np.hstack([US, U]))
    M = Uhat.T @ U
    Vhat, _ = _qr(np.hstack([VS, V]))
    N = Vhat.T @ V
    S0 = M @ S @ N.T
    if faithful:
        Shat = _data_integrate(S0, 1, Uhat, dA, Vhat, True)
    else:
        Shat = S0 + Uhat.T @ (dA @ Vhat)
    Uh, sig, Vht = np.linalg.svd(Shat)
    r1 = int(min(rmax, truncate_to_tolerance(sig, tol)))
    U1 = Uhat @ Uh[:, :r1]
    V1 = Vhat @ Vht[:r1, :].T
    return U1, np.diag(sig[:r1]), V1, r1


def greedy_step(U, S, V, X):
    """``greedy_step!`` for an SVD-like state (greedy_integrator.jl:94-104); X = y(t+dt).

    XV = X*V; XU = X'*U (old bases)      (:97-99)
    U1 = Q of qr(XV); V1 = Q of qr(XU)   (:100-101)
    S1 = U1' * (X*V1)                    (:102-103)
    """
    XV = X @ V
    XU = X.T @ U
    U1, _ = _qr(XV)
    V1, _ = _qr(XU)
    S1 = U1.T @ (X @ V1)
    return U1, S1, V1


def greedy_step_twofactor(U, Z, X):
    """``greedy_step!`` for Y = U*Z' with data (greedy_integrator.jl:84-92); X = y(t+dt).

    Z1 = X'*U (old U)                    (:88)
    Q, _, P = svd(X*Z1); U1 = Q*P'       (:89-91)
    """
    Z1 = X.T @ U
    Q, _, Pt = np.linalg.svd(X @ Z1, full_matrices=False)
    return Q @ Pt, Z1


# ----------------------------------------------------------------------------
# the driver  (primitives.jl:68-104)
# ----------------------------------------------------------------------------


def _init_sol_len(dt, t0, tf):
    """``init_sol`` slot count: floor((tf-t0)/dt)+1 for Float dt, length(t0:dt:tf) for Int."""
    if isinstance(dt, (int, np.integer)) and isinstance(t0, (int, np.integer)):
        return len(range(t0, tf + 1, dt))
    return int(math.floor((tf - t0) / dt)) + 1


def solve(prob: MatrixDataProblem, alg, dt=None, mode="efficient", save_everystep=True) -> DLRSolution:
    """``solve(prob, alg, dt)`` / ``solve(prob::MatrixDataProblem, alg)`` (primitives.jl:68-80)."""
    if dt is None:
        if callable(prob.y):
            raise AssertionError("If the data is not provided as array, integration stepsize needs to be specified")
        dt = 1
    t0, tf = prob.tspan
    if not tf > t0:
        raise AssertionError("Integration in reverse time direction is not supported")
    u = prob.u0.copy()
    nslots = _init_sol_len(dt, t0, tf)
    Ys: List[Any] = [None] * nslots
    if isinstance(dt, (int, np.integer)) and isinstance(t0, (int, np.integer)):
        ts: List[Any] = list(range(t0, tf + 1, dt))
    else:
        ts = list(np.linspace(t0, tf, nslots))
    Ys[0] = prob.u0.copy()
    yprev = initial_data(prob.y, t0)
    t = t0
    it = 0
    T = tf - t0
    while (tf - t) / T > 1e-8:
        if isinstance(alg, GreedyIntegrator):
            # greedy_integrator.jl:106-111: only X = y(t+dt) enters the step
            ycur = update_data(prob.y, t, dt)
            if isinstance(u, TwoFactor):
                u = TwoFactor(*greedy_step_twofactor(u.U, u.Z, ycur))
            else:
                u = SVDLike(*greedy_step(u.U, u.S, u.V, ycur))
            t = t + dt
            it += 1
            if save_everystep or (tf - t) / T <= 1e-8:
                if it <= len(Ys) - 1:
                    Ys[it] = u.copy()
                    ts[it] = t
                else:
                    Ys.append(u.copy())
                    ts.append(t)
            continue
        if isinstance(u, TwoFactor):
            raise TypeError("a TwoFactor state needs the GreedyIntegrator (greedy_integrator.jl:24-30)")
        U, S, V = u.U, u.S, u.V
        if isinstance(alg, ProjectorSplitting) and isinstance(alg.order, Strang):
            # projector_splitting.jl:205-211: primal(t, dt/2) then dual(t+dt/2, dt/2)
            ymid = update_data(prob.y, t, dt / 2)
            U, S, V = ksl_primal_step(U, S, V, ymid - yprev, mode)
            ycur = update_data(prob.y, t + dt / 2, dt / 2)
            U, S, V = ksl_dual_step(U, S, V, ycur - ymid, mode)
        else:
            ycur = update_data(prob.y, t, dt)
            dA = ycur - yprev
            if isinstance(alg, ProjectorSplitting):
                if isinstance(alg.order, PrimalLieTrotter):
                    U, S, V = ksl_primal_step(U, S, V, dA, mode)
                elif isinstance(alg.order, DualLieTrotter):
                    U, S, V = ksl_dual_step(U, S, V, dA, mode)
                else:
                    raise TypeError(f"unknown projector-splitting order {alg.order!r}")
            elif isinstance(alg, UnconventionalAlgorithm):
                U, S, V = bug_step(U, S, V, dA, mode)
            elif isinstance(alg, RankAdaptiveUnconventionalAlgorithm):
                U, S, V, _ = rank_adaptive_bug_step(U, S, V, dA, alg.tol, alg.rmax, mode)
            else:
                raise TypeError(f"unsupported algorithm {alg!r}")
        yprev = ycur
        u = SVDLike(U, S, V)
        t = t + dt
        it += 1
        if save_everystep or (tf - t) / T <= 1e-8:
            if it <= len(Ys) - 1:
                Ys[it] = u.copy()
                ts[it] = t
            else:
                Ys.append(u.copy())
                ts.append(t)
    return DLRSolution(Ys, ts)
