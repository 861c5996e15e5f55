"""Host-side mirror of the reference's operator API for the data-driven DLRA path.

Same names, argument meaning and error behaviour as LowRankIntegrators.jl
(src/primitives.jl, src/integrators/*.jl); every step runs in liblri_hip.so on
the GPU.  Julia's ``step!`` is spelled ``step`` here.

    prob = MatrixDataProblem(y, u0, tspan)           # y: t -> matrix, or a list of snapshots
    sol  = solve(prob, ProjectorSplitting(PrimalLieTrotter()), dt)
    sol.Y[-1].U, sol.Y[-1].S, sol.Y[-1].V, sol.t

Matrix ODEs run too when F is given as an operator descriptor (a Python
closure cannot run on the device):

    F    = LyapunovOperator(laplacian_1d(n), laplacian_1d(m), G, H)   # A Y + Y B' + G H'
    prob = MatrixDEProblem(F, u0, (0.0, 0.05))
    sol  = solve(prob, UnconventionalAlgorithm(), 1e-4)          # RK4 sub-steps

Extensions beyond the reference (keyword-only): ``precision`` ("f64" or
"mixed"), ``snapshot_dtype`` ("f64", "f32", "bf16"), ``save_everystep``,
``device``, and for matrix ODEs ``substeps`` (RK4 steps per sub-problem; by
default the smallest count with h * rho <= 2.5 for the Gershgorin bound rho of
the operator).  The custom K/L/S right-hand-side hooks raise
NotImplementedError.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Callable, List, Optional, Sequence, Union

import numpy as np

from . import native

# ----------------------------------------------------------------------------
# low-rank types (LowRankArithmetic.jl names)
# ----------------------------------------------------------------------------


@dataclass
class SVDLikeRepresentation:
    """Y = U * S * V' (LowRankArithmetic ``SVDLikeRepresentation``), host FP64 arrays."""

    U: np.ndarray
    S: np.ndarray
    V: np.ndarray

    @property
    def rank(self) -> int:
        return int(self.S.shape[0])

    @property
    def shape(self):
        return (self.U.shape[0], self.V.shape[0])

    def matrix(self) -> np.ndarray:
        return self.U @ self.S @ self.V.T

    def copy(self) -> "SVDLikeRepresentation":
        return SVDLikeRepresentation(self.U.copy(), self.S.copy(), self.V.copy())


@dataclass
class TwoFactorRepresentation:
    """Y = U * Z' (LowRankArithmetic ``TwoFactorRepresentation``)."""

    U: np.ndarray
    Z: np.ndarray

    @property
    def rank(self) -> int:
        return int(self.U.shape[1])

    @property
    def shape(self):
        return (self.U.shape[0], self.Z.shape[0])

    def matrix(self) -> np.ndarray:
        return self.U @ self.Z.T

    def copy(self) -> "TwoFactorRepresentation":
        return TwoFactorRepresentation(self.U.copy(), self.Z.copy())


def Matrix(u) -> np.ndarray:
    """Dense matrix of a low-rank representation (Julia ``Matrix(u)``)."""
    return u.matrix()


def rank(u) -> int:
    return u.rank


def truncated_svd(A: np.ndarray, r: Optional[int] = None, tol: Optional[float] = None) -> SVDLikeRepresentation:
    """Rank-r (or tolerance) truncated SVD of a dense host matrix (initial data, LAPACK)."""
    U, s, Vt = np.linalg.svd(np.asarray(A, dtype=np.float64), full_matrices=False)
    if r is None:
        if tol is None:
            raise ValueError("truncated_svd needs r or tol")
        tail = np.sqrt(np.concatenate([np.cumsum(s[::-1] ** 2)[::-1], [0.0]]))
        r = next(rho for rho in range(1, len(s) + 1) if tail[rho] <= tol) if len(s) else 0
    r = int(min(r, len(s)))
    return SVDLikeRepresentation(U[:, :r].copy(), np.diag(s[:r]), Vt[:r].T.copy())


# ----------------------------------------------------------------------------
# problems (primitives.jl:13-41)
# ----------------------------------------------------------------------------


class AbstractDLRProblem:
    pass


@dataclass
class MatrixDataProblem(AbstractDLRProblem):
    """Track a low-rank factorisation of data y(t) (primitives.jl:23-30).

    ``y`` is a callable t -> (n x m) array or a sequence of snapshots (host
    arrays, or device tensors exposing ``data_ptr()``, all of one dtype and
    column-major / Fortran-ordered).  For a sequence the time span defaults to
    ``(1, len(y))`` and the step is the integer index step 1.
    """

    y: Union[Callable[[float], Any], Sequence[Any]]
    u0: SVDLikeRepresentation
    tspan: Optional[tuple] = None

    def __post_init__(self):
        if self.tspan is None:
            if callable(self.y):
                raise TypeError("MatrixDataProblem(y::Function, u0) needs a tspan")
            self.tspan = (1, len(self.y))


class SymTridiagonal:
    """Symmetric tridiagonal n x n matrix: ``d`` on the diagonal, ``e[i]`` coupling rows
    i and i+1 (n - 1 entries), plus ``e[n-1]`` coupling rows n-1 and 0 when periodic."""

    def __init__(self, d, e, periodic: bool = False):
        self.d = np.asarray(d, dtype=np.float64).reshape(-1).copy()
        self.e = np.asarray(e, dtype=np.float64).reshape(-1).copy()
        self.periodic = bool(periodic)
        n = self.d.shape[0]
        if self.e.shape[0] != (n if self.periodic else n - 1):
            raise ValueError(f"a {'periodic ' if periodic else ''}{n} x {n} tridiagonal needs "
                             f"{n if periodic else n - 1} off-diagonal entries")
        if self.periodic and n < 3:
            raise ValueError("a periodic tridiagonal needs n >= 3")

    @property
    def n(self) -> int:
        return self.d.shape[0]

    def dense(self) -> np.ndarray:
        n = self.n
        A = np.diag(self.d)
        A[np.arange(n - 1), np.arange(1, n)] = self.e[: n - 1]
        A[np.arange(1, n), np.arange(n - 1)] = self.e[: n - 1]
        if self.periodic:
            A[n - 1, 0] += self.e[n - 1]
            A[0, n - 1] += self.e[n - 1]
        return A

    def is_constant(self) -> bool:
        """d I + e (shift + shift') (the off-diagonal's wrap entry too when periodic)?"""
        ne = self.n if self.periodic else self.n - 1
        return bool(np.all(self.d == self.d[0]) and (ne < 1 or np.all(self.e[:ne] == self.e[0])))

    def gershgorin(self) -> float:
        """Upper bound of the spectral radius."""
        n = self.n
        a = np.abs(self.e)
        up = np.zeros(n)
        dn = np.zeros(n)
        dn[: n - 1] = a[: n - 1]
        up[1:] = a[: n - 1]
        if self.periodic:
            dn[n - 1] += a[n - 1]
            up[0] += a[n - 1]
        return float(np.max(np.abs(self.d) + up + dn))


def laplacian_1d(n: int, scale: Optional[float] = None, periodic: bool = False) -> SymTridiagonal:
    """scale * tridiag(1, -2, 1) (default scale n^2), Dirichlet or periodic boundaries."""
    s = float(n * n) if scale is None else float(scale)
    return SymTridiagonal(-2.0 * s * np.ones(n), s * np.ones(n if periodic else n - 1), periodic)


class LyapunovOperator:
    """F(Y, t) = A Y + Y B' + G H' with A, B ``SymTridiagonal`` and G (n x q), H (m x q), q <= 4.

    The device form of a matrix-ODE right-hand side.  Calling it evaluates F on a
    dense host matrix (for small checks)."""

    def __init__(self, A: SymTridiagonal, B: SymTridiagonal, G=None, H=None):
        self.A, self.B = A, B
        if (G is None) != (H is None):
            raise ValueError("give both G and H, or neither")
        if G is None:
            self.G = np.zeros((A.n, 0))
            self.H = np.zeros((B.n, 0))
        else:
            self.G = np.asarray(G, dtype=np.float64).reshape(A.n, -1)
            self.H = np.asarray(H, dtype=np.float64).reshape(B.n, -1)
            if self.G.shape[1] != self.H.shape[1] or self.G.shape[1] > native.MAX_FORCE:
                raise ValueError(f"G and H need the same number q <= {native.MAX_FORCE} of columns")

    def __call__(self, Y, t=0.0):
        return self.A.dense() @ Y + Y @ self.B.dense().T + self.G @ self.H.T

    def spectral_bound(self) -> float:
        return self.A.gershgorin() + self.B.gershgorin()


class Tridiagonal:
    """General tridiagonal n x n matrix acting on rows: row i is
    ``dl[i] Y[i-1] + dd[i] Y[i] + du[i] Y[i+1]`` (n entries each); with ``periodic``
    row 0 takes ``dl[0] Y[n-1]`` and row n-1 ``du[n-1] Y[0]``, else those are unused."""

    def __init__(self, dl, dd, du, periodic: bool = False):
        self.dl, self.dd, self.du = (np.asarray(x, dtype=np.float64).reshape(-1).copy() for x in (dl, dd, du))
        self.periodic = bool(periodic)
        n = self.dd.shape[0]
        if self.dl.shape[0] != n or self.du.shape[0] != n:
            raise ValueError("dl, dd and du need n entries each")
        if self.periodic and n < 3:
            raise ValueError("a periodic tridiagonal needs n >= 3")

    @property
    def n(self) -> int:
        return self.dd.shape[0]

    def dense(self) -> np.ndarray:
        n = self.n
        D = np.diag(self.dd)
        D[np.arange(1, n), np.arange(n - 1)] = self.dl[1:]
        D[np.arange(n - 1), np.arange(1, n)] = self.du[: n - 1]
        if self.periodic:
            D[0, n - 1] += self.dl[0]
            D[n - 1, 0] += self.du[n - 1]
        return D

    def row_bound(self) -> float:
        """max_i sum_j |D_ij| (infinity norm)."""
        lo = np.abs(self.dl).copy()
        hi = np.abs(self.du).copy()
        if not self.periodic:
            lo[0] = 0.0
            hi[-1] = 0.0
        return float(np.max(lo + np.abs(self.dd) + hi))


class QuadraticOperator(LyapunovOperator):
    """F(Y, t) = A Y + Y B' + G H' - (D Y) o Y (o: elementwise), D a ``Tridiagonal`` on the
    n rows.  Viscous Burgers with one parameter sample per column
    (test/data_agnostic_approximation.jl:16-33) is ``burgers_operator``."""

    def __init__(self, A: SymTridiagonal, B: SymTridiagonal, D: Tridiagonal, G=None, H=None):
        super().__init__(A, B, G, H)
        if D.n != A.n:
            raise ValueError(f"D is {D.n} x {D.n}, A is {A.n} x {A.n}")
        self.D = D

    def __call__(self, Y, t=0.0):
        return super().__call__(Y, t) - (self.D.dense() @ Y) * Y

    def spectral_bound(self, ymax: float = 0.0) -> float:
        """Bound of the Jacobian's spectral radius near states with max |Y_ij| <= ymax:
        rho(A) + rho(B) + 2 ||D||_inf ymax."""
        return super().spectral_bound() + 2.0 * self.D.row_bound() * ymax


def burgers_operator(n: int, m: int, nu: float, length: float = math.pi) -> QuadraticOperator:
    """Delta rho - (grad rho) .* rho on a periodic grid of n cells of [0, length)
    (test/data_agnostic_approximation.jl:4-33): Delta = nu/dx^2 tridiag(1, -2, 1),
    grad = 1/(2 dx) tridiag(-1, 0, 1), both periodic; m columns (samples)."""
    dx = length / n
    A = SymTridiagonal(-2.0 * nu / dx ** 2 * np.ones(n), nu / dx ** 2 * np.ones(n), periodic=True)
    B = SymTridiagonal(np.zeros(m), np.zeros(m - 1))
    g = 0.5 / dx
    D = Tridiagonal(-g * np.ones(n), np.zeros(n), g * np.ones(n), periodic=True)
    return QuadraticOperator(A, B, D)


def burgers_samples(n: int, msq: int, sigma=(0.5, 0.5), length: float = math.pi) -> np.ndarray:
    """The initial samples of the reference's Burgers problem
    (test/data_agnostic_approximation.jl:35-44), n x msq^2: the profile
    0.5 (exp(cos x) - 1.5) sin(x + 0.74 pi) on the cell centres of [0, length) plus
    sigma_1 xi_1 sin(2 pi x) + sigma_2 xi_2 sin(3 pi x) for (xi_1, xi_2) on an
    msq x msq grid of [-1, 1]^2, xi_1 varying fastest."""
    dx = length / n
    x = dx / 2 + dx * np.arange(n)
    ub = 0.5 * (np.exp(np.cos(x)) - 1.5) * np.sin(x + 2 * math.pi * 0.37)
    xi = np.linspace(-1.0, 1.0, msq)
    x1 = np.tile(xi, msq)
    x2 = np.repeat(xi, msq)
    return (ub[:, None] + sigma[0] * np.sin(2 * math.pi * x)[:, None] * x1[None, :]
            + sigma[1] * np.sin(3 * math.pi * x)[:, None] * x2[None, :])


def rk4_substeps(rho: float, dt: float, cap: float = 2.5) -> int:
    """Fewest RK4 steps with h * rho <= cap (RK4's real stability interval is [-2.78, 0])."""
    return max(1, int(math.ceil(abs(dt) * rho / cap)))


class MatrixDEProblem(AbstractDLRProblem):
    """Matrix ODE dY/dt = f(Y, t) (primitives.jl:13-17).  On this engine ``f`` is a
    ``LyapunovOperator``."""

    def __init__(self, f, u0, tspan):
        self.f, self.u0, self.tspan = f, u0, tspan


class MatrixHybridProblem(AbstractDLRProblem):
    """``MatrixHybridProblem{y,f,u0,tspan}`` (primitives.jl:36-41): data y(t) (a function,
    or a sequence of snapshots with tspan = (1, len(y))) and the Z dynamics
    dZ/dt = f(Z, U, t) of a two-factor state U Z' (Greedy integrator only).  On this
    engine ``f`` is an operator descriptor F (``LyapunovOperator`` /
    ``QuadraticOperator``) with f(Z, U, t) = F(U Z')' U, the form of the reference's
    hybrid test (test/data_informed_approximation.jl:75, FZ = burgers(U Z')' U)."""

    def __init__(self, y, f, u0, tspan=None):
        if tspan is None:
            if callable(y):
                raise ValueError("a function-form MatrixHybridProblem needs tspan")
            tspan = (1, len(y))
        self.y, self.f, self.u0, self.tspan = y, f, u0, tspan


# ----------------------------------------------------------------------------
# algorithms (projector_splitting.jl:1-41, unconventional.jl:1-21,
#             rank_adaptive_unconventional.jl:1-23)
# ----------------------------------------------------------------------------


class PrimalLieTrotter:
    pass


class DualLieTrotter:
    pass


class Strang:
    pass


class AbstractDLRAlgorithm:
    pass


def _no_hooks(kw):
    hooks = [k for k in ("S_rhs", "L_rhs", "K_rhs") if kw.get(k) is not None]
    if hooks:
        raise NotImplementedError(f"custom {', '.join(hooks)} hooks belong to the DE path, which this engine does not run")


class ProjectorSplitting(AbstractDLRAlgorithm):
    """KSL projector-splitting integrator (Lubich-Oseledets 2014)."""

    def __init__(self, order=None, *, S_rhs=None, L_rhs=None, K_rhs=None, S_kwargs=None, L_kwargs=None,
                 K_kwargs=None, S_alg=None, L_alg=None, K_alg=None):
        _no_hooks(dict(S_rhs=S_rhs, L_rhs=L_rhs, K_rhs=K_rhs))
        self.order = PrimalLieTrotter() if order is None else order
        if not isinstance(self.order, (PrimalLieTrotter, DualLieTrotter, Strang)):
            raise TypeError(f"unknown projector-splitting order {self.order!r}")


class UnconventionalAlgorithm(AbstractDLRAlgorithm):
    """Unconventional (BUG) integrator (Ceruti-Lubich 2021)."""

    def __init__(self, *, S_rhs=None, L_rhs=None, K_rhs=None, S_kwargs=None, L_kwargs=None, K_kwargs=None,
                 S_alg=None, L_alg=None, K_alg=None):
        _no_hooks(dict(S_rhs=S_rhs, L_rhs=L_rhs, K_rhs=K_rhs))


class RankAdaptiveUnconventionalAlgorithm(AbstractDLRAlgorithm):
    """Rank-adaptive BUG integrator (Ceruti-Kusch-Lubich); rank capped at ``native.MAX_RANK``."""

    def __init__(self, tol, *, S_rhs=None, L_rhs=None, K_rhs=None, S_kwargs=None, L_kwargs=None,
                 K_kwargs=None, S_alg=None, L_alg=None, K_alg=None, rmax=2 ** 62):
        _no_hooks(dict(S_rhs=S_rhs, L_rhs=L_rhs, K_rhs=K_rhs))
        self.tol = float(tol)
        self.rmax = int(rmax)


class GreedyIntegrator(AbstractDLRAlgorithm):
    """``GreedyIntegrator(; Z_alg, Z_kwargs)`` (greedy_integrator.jl:16-22).  ``Z_alg`` and
    ``Z_kwargs`` drive the Z sub-problem of a ``MatrixHybridProblem``, which this
    engine does not run; on a ``MatrixDataProblem`` they are unused, as in the
    reference."""

    def __init__(self, *, Z_alg=None, Z_kwargs=None):
        self.Z_alg = Z_alg
        self.Z_kwargs = dict(Z_kwargs or {})


# ----------------------------------------------------------------------------
# solution / integrator (primitives.jl:46-104)
# ----------------------------------------------------------------------------


@dataclass
class DLRSolution:
    Y: List[Any]
    t: List[Any]


def _is_int(x) -> bool:
    return isinstance(x, (int, np.integer)) and not isinstance(x, bool)


def init_sol(dt, t0, tf, u0) -> DLRSolution:
    """``init_sol`` (primitives.jl:92-104)."""
    if _is_int(dt) and _is_int(t0) and _is_int(tf):
        steps = list(range(t0, tf + 1, dt))
        return DLRSolution([None] * len(steps), steps)
    n = int(math.floor((tf - t0) / dt)) + 1
    return DLRSolution([None] * n, list(np.linspace(t0, tf, n)))


def _encode(a, dtype: str) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    if dtype == "f64":
        return np.asfortranarray(a)
    if dtype == "f32":
        return np.asfortranarray(a, dtype=np.float32)
    f = np.asfortranarray(a, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) & 0xFFFF
    return np.asfortranarray(b.astype(np.uint16))


class DLRIntegrator:
    """``DLRIntegrator`` (primitives.jl:54-63) holding a device context."""

    def __init__(self, prob, alg, dt, ctx, u, sol, *, snapshot_dtype, save_everystep):
        self.prob = prob
        self.alg = alg
        self.dt = dt
        self.t = prob.tspan[0]
        self.iter = 0
        self.ctx = ctx
        self.u = u
        self.sol = sol
        self.probType = type(prob)
        self.snapshot_dtype = snapshot_dtype
        self.save_everystep = save_everystep
        self._array_data = isinstance(prob, MatrixDataProblem) and not callable(prob.y)
        self._slot_of_t = 0  # function data: ring slot (of 3) holding y(t)
        # array data: the snapshot index of yprev; alg_cache seeds it with y[1] whatever t0 is
        # (projector_splitting.jl:91, unconventional.jl:91, rank_adaptive_unconventional.jl:113)
        self._kprev = 0
        self.rank = u.rank
        self.twofactor = isinstance(u, TwoFactorRepresentation)

    def refresh_u(self):
        U, S, V = self.ctx.get_state()
        # a two-factor state keeps Z in the V slot of the device state
        self.u = TwoFactorRepresentation(U, V) if self.twofactor else SVDLikeRepresentation(U, S, V)
        return self.u


def _device_ptr(x):
    dp = getattr(x, "data_ptr", None)
    return dp() if callable(dp) else None


def init(prob, alg, dt, *, precision: str = "f64", snapshot_dtype: Optional[str] = None,
         save_everystep: bool = True, device: int = 0, substeps: Optional[int] = None,
         eigenbasis: bool = True) -> DLRIntegrator:
    """``init(prob, alg, dt)`` (projector_splitting.jl:107-115 and the BUG/RA counterparts)."""
    if isinstance(prob, MatrixHybridProblem):
        return _init_hybrid(prob, alg, dt, precision=precision, snapshot_dtype=snapshot_dtype,
                            save_everystep=save_everystep, device=device, substeps=substeps)
    if isinstance(prob, MatrixDEProblem):
        return _init_de(prob, alg, dt, save_everystep=save_everystep, device=device, substeps=substeps,
                        eigenbasis=eigenbasis)
    if not isinstance(alg, (ProjectorSplitting, UnconventionalAlgorithm, RankAdaptiveUnconventionalAlgorithm,
                            GreedyIntegrator)):
        raise TypeError(f"unsupported algorithm {alg!r}")
    t0, tf = prob.tspan
    if not tf > t0:
        raise AssertionError("Integration in reverse time direction is not supported")
    u0 = prob.u0
    if isinstance(u0, TwoFactorRepresentation) and not isinstance(alg, GreedyIntegrator):
        # only greedy_integrator.jl has TwoFactorRepresentation methods (:24-30,84-92)
        raise TypeError("a TwoFactorRepresentation state needs the GreedyIntegrator")
    twofactor = isinstance(u0, TwoFactorRepresentation)
    n = u0.U.shape[0]
    m = (u0.Z if twofactor else u0.V).shape[0]
    r = u0.rank
    if isinstance(alg, RankAdaptiveUnconventionalAlgorithm):
        rcap = int(min(alg.rmax, native.MAX_RANK, n, m))
        if r > rcap:
            raise ValueError(f"initial rank {r} exceeds the capacity {rcap}")
    else:
        rcap = r
    if r > native.MAX_RANK:
        raise ValueError(f"rank {r} exceeds LRI_MAX_RANK = {native.MAX_RANK}")
    ctx = native.Context(n, m, rcap, device)
    ctx.set_precision(native.PREC_MIXED if precision == "mixed" else native.PREC_F64)
    if twofactor:
        ctx.set_state(u0.U, np.eye(r), u0.Z)  # Z in the V slot, S unused
    else:
        ctx.set_state(u0.U, u0.S, u0.V)

    # data: device tensors are bound as borrowed snapshots, host data is uploaded
    y = prob.y
    if callable(y):
        sdt = snapshot_dtype or "f64"
        ctx.reserve(3, native.DTYPES[sdt])
        ctx.upload(0, _encode(y(t0), sdt))
    else:
        ptr0 = _device_ptr(y[0])
        if ptr0 is not None:
            sdt = snapshot_dtype or {"torch.float64": "f64", "torch.float32": "f32",
                                     "torch.bfloat16": "bf16"}[str(y[0].dtype)]
            ld = int(y[0].stride(1)) if y[0].dim() == 2 else n
            ctx.bind([_device_ptr(a) for a in y], native.DTYPES[sdt], ld, owners=y)
        else:
            sdt = snapshot_dtype or "f64"
            ctx.reserve(len(y), native.DTYPES[sdt])
            for k, a in enumerate(y):
                ctx.upload(k, _encode(a, sdt))
    sol = init_sol(dt, t0, tf, u0)
    sol.Y[0] = u0.copy()
    return DLRIntegrator(prob, alg, dt, ctx, u0.copy(), sol, snapshot_dtype=sdt, save_everystep=save_everystep)


def _init_hybrid(prob, alg, dt, *, precision, snapshot_dtype, save_everystep, device, substeps) -> DLRIntegrator:
    """``init(prob::MatrixHybridProblem, ::GreedyIntegrator, dt)`` (greedy_integrator.jl:41-46,61-71):
    the operator descriptor of the Z dynamics and the data ring go to the device."""
    F = prob.f
    if not isinstance(F, LyapunovOperator):
        raise NotImplementedError("a Python callable cannot run on the device: give the Z dynamics as a "
                                  "LyapunovOperator or QuadraticOperator F, f(Z, U, t) = F(U Z')' U")
    if not isinstance(alg, GreedyIntegrator):
        raise TypeError("a MatrixHybridProblem is solved by the GreedyIntegrator (greedy_integrator.jl:41-82)")
    u0 = prob.u0
    if not isinstance(u0, TwoFactorRepresentation):
        raise TypeError("a MatrixHybridProblem has a two-factor state U Z' (greedy_integrator.jl:73-82)")
    t0, tf = prob.tspan
    if not tf > t0:
        raise AssertionError("Integration in reverse time direction is not supported")
    n, m, r = u0.U.shape[0], u0.Z.shape[0], u0.rank
    if (F.A.n, F.B.n) != (n, m):
        raise ValueError(f"operator is {F.A.n} x {F.B.n}, state is {n} x {m}")
    if r > native.MAX_RANK:
        raise ValueError(f"rank {r} exceeds LRI_MAX_RANK = {native.MAX_RANK}")
    quad = isinstance(F, QuadraticOperator)
    if quad and r > native.QUAD_MAX_RANK:
        raise ValueError(f"the quadratic operator needs rank <= {native.QUAD_MAX_RANK}")
    ctx = native.Context(n, m, r, device)
    ctx.set_precision(native.PREC_MIXED if precision == "mixed" else native.PREC_F64)
    ctx.set_state(u0.U, np.eye(r), u0.Z)  # Z in the V slot
    q = F.G.shape[1]
    ctx.set_operator_lyapunov(F.A.d, F.A.e, F.A.periodic, F.B.d, F.B.e, F.B.periodic,
                              F.G if q else None, F.H if q else None)
    if quad:
        ctx.set_operator_quadratic(F.D.dl, F.D.dd, F.D.du, F.D.periodic)
    y = prob.y
    sdt = snapshot_dtype or "f64"
    if callable(y):
        ctx.reserve(3, native.DTYPES[sdt])
        ctx.upload(0, _encode(y(t0), sdt))
    else:
        ctx.reserve(len(y), native.DTYPES[sdt])
        for k, a in enumerate(y):
            ctx.upload(k, _encode(a, sdt))
    sol = init_sol(dt, t0, tf, u0)
    sol.Y[0] = u0.copy()
    integ = DLRIntegrator(prob, alg, dt, ctx, u0.copy(), sol, snapshot_dtype=sdt, save_everystep=save_everystep)
    integ._array_data = not callable(y)
    integ.substeps = int(substeps) if substeps is not None else None
    if quad and integ.substeps is None:
        # |Y_ij| <= ||U_i|| ||Z_j|| bounds the Jacobian of the quadratic term (factor 2 of headroom)
        integ.quad_ymax = 2.0 * float(np.max(np.linalg.norm(u0.U, axis=1)) * np.max(np.linalg.norm(u0.Z, axis=1)))
    return integ


def _step_hybrid(integrator, dt) -> None:
    """greedy_step!(::TwoFactorRepresentation, ..., ::Type{<:MatrixHybridProblem}) (greedy_integrator.jl:73-82)."""
    ctx, t = integrator.ctx, integrator.t
    N = integrator.substeps
    if N is None:
        F = integrator.prob.f
        rho = F.spectral_bound(integrator.quad_ymax) if isinstance(F, QuadraticOperator) else F.spectral_bound()
        N = rk4_substeps(rho, dt)
    y = integrator.prob.y
    if integrator._array_data:
        if not (_is_int(t) and _is_int(dt)):
            raise TypeError("array-form data needs integer t and dt")
        kn = t + dt - 1
        if not 0 <= kn < len(y):
            raise IndexError(f"snapshot index {kn + 1} outside the data (length {len(y)})")
    else:
        kn = [s_ for s_ in range(3) if s_ != integrator._slot_of_t][0]
        ctx.upload(kn, _encode(y(t + dt), integrator.snapshot_dtype))
        integrator._slot_of_t = kn
    ctx.step_hybrid_greedy(kn, dt, N)


def _init_de(prob, alg, dt, *, save_everystep, device, substeps, eigenbasis=True) -> DLRIntegrator:
    """``init`` for a matrix ODE: the operator descriptor is copied to the device."""
    F = prob.f
    if not isinstance(F, LyapunovOperator):
        raise NotImplementedError("a Python callable cannot run on the device: give the right-hand side as a "
                                  "LyapunovOperator (A Y + Y B' + G H') or a QuadraticOperator "
                                  "(... - (D Y) o Y)")
    if not isinstance(alg, (ProjectorSplitting, UnconventionalAlgorithm, RankAdaptiveUnconventionalAlgorithm)):
        raise TypeError(f"unsupported algorithm {alg!r}")
    t0, tf = prob.tspan
    if not tf > t0:
        raise AssertionError("Integration in reverse time direction is not supported")
    u0 = prob.u0
    n, m = u0.U.shape[0], u0.V.shape[0]
    if (F.A.n, F.B.n) != (n, m):
        raise ValueError(f"operator is {F.A.n} x {F.B.n}, state is {n} x {m}")
    r = u0.rank
    if isinstance(alg, RankAdaptiveUnconventionalAlgorithm):
        rcap = int(min(alg.rmax, native.MAX_RANK, n, m))
        if r > rcap:
            raise ValueError(f"initial rank {r} exceeds the capacity {rcap}")
    else:
        rcap = r
    if r > native.MAX_RANK:
        raise ValueError(f"rank {r} exceeds LRI_MAX_RANK = {native.MAX_RANK}")
    ctx = native.Context(n, m, rcap, device)
    ctx.set_state(u0.U, u0.S, u0.V)
    q = F.G.shape[1]
    ctx.set_operator_lyapunov(F.A.d, F.A.e, F.A.periodic, F.B.d, F.B.e, F.B.periodic,
                              F.G if q else None, F.H if q else None)
    quad = isinstance(F, QuadraticOperator)
    if eigenbasis and not quad and F.A.is_constant() and F.B.is_constant():
        # constant coefficients: the state stays in the operator's eigenbasis (no n x n transforms per step)
        ctx.set_resident_eigenbasis(True)
    if quad:
        # the rank-adaptive S sub-step runs in the augmented 2 r bases
        kmax = 2 * rcap if isinstance(alg, RankAdaptiveUnconventionalAlgorithm) else r
        if kmax > native.QUAD_MAX_RANK:
            raise ValueError(f"the quadratic operator needs rank <= {native.QUAD_MAX_RANK} "
                             f"(rank-adaptive: 2 rmax <= {native.QUAD_MAX_RANK})")
        ctx.set_operator_quadratic(F.D.dl, F.D.dd, F.D.du, F.D.periodic)
    sol = init_sol(dt, t0, tf, u0)
    sol.Y[0] = u0.copy()
    integ = DLRIntegrator(prob, alg, dt, ctx, u0.copy(), sol, snapshot_dtype=None, save_everystep=save_everystep)
    integ.substeps = int(substeps) if substeps is not None else None
    if quad and integ.substeps is None:
        # the Jacobian of -(D Y) o Y scales with max |Y|: bounded from the factors of Y0
        # (|Y_ij| <= ||(U S)_i|| ||V_j||), with a factor 2 of headroom for growth
        ymax = float(np.max(np.linalg.norm(u0.U @ u0.S, axis=1)) * np.max(np.linalg.norm(u0.V, axis=1)))
        integ.quad_ymax = 2.0 * ymax
    return integ


def _step_de(integrator, alg, dt) -> None:
    ctx = integrator.ctx
    N = integrator.substeps
    if N is None:
        half = isinstance(alg, ProjectorSplitting) and isinstance(alg.order, Strang)
        F = integrator.prob.f
        rho = F.spectral_bound(integrator.quad_ymax) if isinstance(F, QuadraticOperator) else F.spectral_bound()
        N = rk4_substeps(rho, dt / 2 if half else dt)
    if isinstance(alg, ProjectorSplitting):
        order = {PrimalLieTrotter: native.PRIMAL, DualLieTrotter: native.DUAL, Strang: native.STRANG}[type(alg.order)]
        ctx.step_de_ksl(order, dt, N)
    elif isinstance(alg, UnconventionalAlgorithm):
        ctx.step_de_bug(dt, N)
    else:
        r_new = ctx.step_de_bug_ra(dt, N, alg.tol, alg.rmax)
        if r_new != integrator.rank:
            print(f"rank adjusted: new rank = {r_new}")  # rank_adaptive_unconventional.jl:230
            integrator.rank = r_new


def step(integrator: DLRIntegrator, alg=None, dt=None) -> None:
    """``step!(integrator, alg, dt)``: one step on the device, then t += dt, iter += 1."""
    alg = integrator.alg if alg is None else alg
    dt = integrator.dt if dt is None else dt
    ctx = integrator.ctx
    t = integrator.t
    if isinstance(integrator.prob, MatrixHybridProblem):
        _step_hybrid(integrator, dt)
    elif isinstance(integrator.prob, MatrixDEProblem):
        _step_de(integrator, alg, dt)
    elif integrator._array_data:
        if isinstance(alg, ProjectorSplitting) and isinstance(alg.order, Strang):
            # update_data!(x, y::AbstractArray, t::Int, dt::Int) has no method for dt/2 (data_integrator.jl:26)
            raise TypeError("array-form data has no half-step snapshot: Strang needs function-form data")
        if not (_is_int(t) and _is_int(dt)):
            raise TypeError("array-form data needs integer t and dt")
        # update_data! reads y[t + dt]; dy = y[t + dt] - yprev, then yprev = y[t + dt]
        # (projector_splitting.jl:117-121, data_integrator.jl:26-28)
        kp, kn = integrator._kprev, t + dt - 1
        if not 0 <= kn < len(integrator.prob.y):
            raise IndexError(f"snapshot index {kn + 1} outside the data (length {len(integrator.prob.y)})")
        _dispatch(integrator, alg, kp, kn, None)
        integrator._kprev = kn
    else:
        cur = integrator._slot_of_t
        free = [s for s in range(3) if s != cur]
        if isinstance(alg, GreedyIntegrator):
            # greedy_step! reads only y(t + dt) (update_data!, data_integrator.jl:22-24)
            ctx.upload(free[0], _encode(integrator.prob.y(t + dt), integrator.snapshot_dtype))
            _dispatch(integrator, alg, cur, free[0], None)
            integrator._slot_of_t = free[0]
        elif isinstance(alg, ProjectorSplitting) and isinstance(alg.order, Strang):
            ctx.upload(free[0], _encode(integrator.prob.y(t + dt / 2), integrator.snapshot_dtype))
            ctx.upload(free[1], _encode(integrator.prob.y(t + dt / 2 + dt / 2), integrator.snapshot_dtype))
            ctx.step_ksl_strang(cur, free[0], free[1])
            integrator._slot_of_t = free[1]
        else:
            ctx.upload(free[0], _encode(integrator.prob.y(t + dt), integrator.snapshot_dtype))
            _dispatch(integrator, alg, cur, free[0], None)
            integrator._slot_of_t = free[0]
    integrator.t = t + dt
    integrator.iter += 1


def _dispatch(integrator, alg, kp, kn, _):
    ctx = integrator.ctx
    if isinstance(alg, GreedyIntegrator):
        if integrator.twofactor:
            ctx.step_greedy_twofactor(kn)
        else:
            ctx.step_greedy(kn)
    elif isinstance(alg, ProjectorSplitting):
        ctx.step_ksl(native.PRIMAL if isinstance(alg.order, PrimalLieTrotter) else native.DUAL, kp, kn)
    elif isinstance(alg, UnconventionalAlgorithm):
        ctx.step_bug(kp, kn)
    else:
        r_new = ctx.step_bug_ra(kp, kn, alg.tol, alg.rmax)
        if r_new != integrator.rank:
            print(f"rank adjusted: new rank = {r_new}")  # rank_adaptive_unconventional.jl:230
            integrator.rank = r_new


def update_sol(integrator: DLRIntegrator, force: bool = False) -> None:
    """``update_sol!`` (primitives.jl:82-90): store a copy of u and t.  With a device
    trajectory (``solve`` reserves one) the copy stays on the device
    (``lri_trajectory_save``, no synchronisation) until ``fetch_sol``."""
    if not (integrator.save_everystep or force):
        return
    k = integrator.iter
    traj = getattr(integrator, "_traj", None)
    if traj is not None and k < integrator._traj_slots:
        integrator.ctx.trajectory_save(k)
        traj[k] = k
        if k <= len(integrator.sol.Y) - 1:
            integrator.sol.Y[k] = None
            integrator.sol.t[k] = integrator.t
        else:
            integrator.sol.Y.append(None)
            integrator.sol.t.append(integrator.t)
        return
    u = integrator.refresh_u().copy()
    if k <= len(integrator.sol.Y) - 1:
        integrator.sol.Y[k] = u
        integrator.sol.t[k] = integrator.t
    else:
        integrator.sol.Y.append(u)
        integrator.sol.t.append(integrator.t)


def fetch_sol(integrator: DLRIntegrator) -> DLRSolution:
    """Copy the states that ``update_sol`` left in the device trajectory into ``sol.Y``."""
    traj = getattr(integrator, "_traj", None) or {}
    ctx = integrator.ctx

    def put(k, U, S, V):
        integrator.sol.Y[k] = TwoFactorRepresentation(U, V) if integrator.twofactor else SVDLikeRepresentation(U, S, V)

    items = sorted(traj.items())
    if getattr(ctx, "resident", False):
        # slots saved in the eigenbasis are transformed back one at a time
        for k, slot in items:
            put(k, *ctx.trajectory_get(slot))
    else:
        i = 0
        while i < len(items):  # runs of consecutive slots, one copy each
            j = i + 1
            while j < len(items) and items[j][1] == items[j - 1][1] + 1:
                j += 1
            for (k, _), usv in zip(items[i:j], ctx.trajectory_get_range(items[i][1], j - i)):
                put(k, *usv)
            i = j
    traj.clear()
    return integrator.sol


def solve(prob, alg, dt=None, **kw) -> DLRSolution:
    """``solve(prob, alg, dt)`` / ``solve(prob::MatrixDataProblem, alg)`` (primitives.jl:68-80).

    The per-step copies of ``update_sol!`` go to a device trajectory (one slot per
    saved step, reserved here) and are copied out once, after the last step."""
    if dt is None:
        if not isinstance(prob, MatrixDataProblem) or callable(prob.y):
            raise AssertionError("If the data is not provided as array, integration stepsize needs to be specified")
        dt = 1
    integ = init(prob, alg, dt, **kw)
    try:
        tf, t0 = prob.tspan[1], prob.tspan[0]
        T = tf - t0
        if integ.save_everystep:
            integ._traj_slots = len(integ.sol.Y) + 2  # the accumulated t can add a step past init_sol's slots
            integ.ctx.trajectory_reserve(integ._traj_slots)
            integ._traj = {}
        while (tf - integ.t) / T > 1e-8:
            step(integ, alg, dt)
            last = not ((tf - integ.t) / T > 1e-8)
            update_sol(integ, force=last)
        integ.ctx.synchronize()
        return fetch_sol(integ)
    finally:
        integ.ctx.close()


def solve_batch(probs, alg, dt, *, substeps=None, save_everystep: bool = False, device: int = 0,
                eigenbasis: bool = True):
    """``solve`` for many independent ``MatrixDEProblem``s of uneven sizes and ranks at once
    (the reference solves them one ``solve`` call at a time, primitives.jl:68-80): one
    context per problem in an ``lri_batch``, whose steps queue every member's work
    before waiting for any.  ``alg``: ``UnconventionalAlgorithm`` or
    ``RankAdaptiveUnconventionalAlgorithm``; all problems share ``dt`` and ``tspan``.
    Returns one ``DLRSolution`` per problem (with ``save_everystep=False``: the initial
    and the final state).  With ``eigenbasis`` (default), members with constant-coefficient
    linear operators keep their state in the operator's eigenbasis; when every member does,
    a rank-adaptive batch step runs each of its stages as one launch over all members."""
    if not probs:
        return []
    ra = isinstance(alg, RankAdaptiveUnconventionalAlgorithm)
    if not (ra or isinstance(alg, UnconventionalAlgorithm)):
        raise TypeError("solve_batch runs the unconventional and rank-adaptive integrators")
    t0, tf = probs[0].tspan
    if any(p.tspan != (t0, tf) for p in probs):
        raise ValueError("all problems of a batch share tspan")
    if not tf > t0:
        raise AssertionError("Integration in reverse time direction is not supported")
    ns, ms, rcaps = [], [], []
    for p in probs:
        if not isinstance(p, MatrixDEProblem) or not isinstance(p.f, LyapunovOperator):
            raise TypeError("solve_batch takes MatrixDEProblems with a LyapunovOperator or QuadraticOperator")
        n, m, r = p.u0.U.shape[0], p.u0.V.shape[0], p.u0.rank
        ns.append(n)
        ms.append(m)
        rcaps.append(int(min(alg.rmax, native.MAX_RANK, n, m)) if ra else r)
        if r > rcaps[-1]:
            raise ValueError(f"initial rank {r} exceeds the capacity {rcaps[-1]}")
    b = native.Batch(ns, ms, rcaps, device)
    try:
        Ns = []
        for p, c in zip(probs, b.members):
            F = p.f
            c.set_state(p.u0.U, p.u0.S, p.u0.V)
            q = F.G.shape[1]
            c.set_operator_lyapunov(F.A.d, F.A.e, F.A.periodic, F.B.d, F.B.e, F.B.periodic,
                                    F.G if q else None, F.H if q else None)
            if eigenbasis and not isinstance(F, QuadraticOperator) and F.A.is_constant() and F.B.is_constant():
                c.set_resident_eigenbasis(True)
            if isinstance(F, QuadraticOperator):
                c.set_operator_quadratic(F.D.dl, F.D.dd, F.D.du, F.D.periodic)
                ymax = float(np.max(np.linalg.norm(p.u0.U @ p.u0.S, axis=1)) * np.max(np.linalg.norm(p.u0.V, axis=1)))
                rho = F.spectral_bound(2.0 * ymax)
            else:
                rho = F.spectral_bound()
            Ns.append(int(substeps) if substeps is not None else rk4_substeps(rho, dt))
        sols = [init_sol(dt, t0, tf, p.u0) for p in probs]
        for s_, p in zip(sols, probs):
            s_.Y[0] = p.u0.copy()
        t, it, T = t0, 0, tf - t0
        while (tf - t) / T > 1e-8:
            if ra:
                b.step_de_bug_ra(dt, Ns, alg.tol, alg.rmax)
            else:
                b.step_de_bug(dt, Ns)
            t = t + dt
            it += 1
            last = not ((tf - t) / T > 1e-8)
            if save_everystep or last:
                for s_, c in zip(sols, b.members):
                    U, S, V = c.get_state()
                    u = SVDLikeRepresentation(U, S, V)
                    if it <= len(s_.Y) - 1:
                        s_.Y[it] = u
                        s_.t[it] = t
                    else:
                        s_.Y.append(u)
                        s_.t.append(t)
        b.synchronize()
        if not save_everystep:
            for s_ in sols:
                keep = [k for k, y in enumerate(s_.Y) if y is not None]
                s_.Y = [s_.Y[k] for k in keep]
                s_.t = [s_.t[k] for k in keep]
        return sols
    finally:
        b.close()


def solve_gradient(prob, alg, loss_grad, dt=None, *, checkpoint: int = 20, precision: str = "f64",
                   snapshot_dtype: Optional[str] = None, substeps: Optional[int] = None, device: int = 0):
    """Gradient of a loss of the final state of ``solve(prob, alg[, dt])`` with respect to
    the initial factors (U0, S0, V0), by the engine's hand-written reverse steps with
    checkpointed recomputation: states are kept every ``checkpoint`` steps on the way
    forward and recomputed segment by segment on the way back.

    - array-form data (``MatrixDataProblem``), ``alg`` ``UnconventionalAlgorithm()``,
      ``RankAdaptiveUnconventionalAlgorithm(tol; rmax)`` (the ranks of the forward solve held
      fixed), ``ProjectorSplitting(PrimalLieTrotter())`` or ``ProjectorSplitting(DualLieTrotter())``:
      ``lri_step_bug_adjoint`` / ``lri_step_bug_ra_adjoint`` / ``lri_step_ksl_adjoint`` /
      ``lri_step_ksl_dual_adjoint``; returns ``(value, (Ub0, Sb0, Vb0))``;
    - a matrix ODE (``MatrixDEProblem`` with a ``QuadraticOperator``; D = 0 for a linear
      one) and ``UnconventionalAlgorithm()``, ``RankAdaptiveUnconventionalAlgorithm(tol; rmax)``
      (2 rmax <= LRI_QUAD_MAX_RANK; the ranks of the forward solve held fixed) or
      ``ProjectorSplitting(PrimalLieTrotter())``:
      ``lri_step_de_bug_adjoint`` / ``lri_step_de_bug_ra_adjoint`` / ``lri_step_de_ksl_adjoint``, with ``dt`` and
      the RK4 ``substeps`` as ``solve``; returns ``(value, (Ub0, Sb0, Vb0), (abar, bbar))``,
      abar and bbar the derivatives along scales of the operator's A and B
      (A = nu Laplacian: dL/dnu = abar / nu).

    ``loss_grad(U, S, V) -> (value, Ub, Sb, Vb)`` gives the loss of the final factors
    and its cotangents.  (The reference has no adjoint; in Julia this is the body of a
    ChainRulesCore ``rrule`` for ``solve``.)"""
    if checkpoint < 1:
        raise ValueError("checkpoint must be >= 1")
    if isinstance(prob, MatrixDEProblem):
        return _solve_gradient_de(prob, alg, loss_grad, dt, checkpoint, substeps, device)
    ksl = isinstance(alg, ProjectorSplitting) and isinstance(alg.order, PrimalLieTrotter)
    dual = isinstance(alg, ProjectorSplitting) and isinstance(alg.order, DualLieTrotter)
    ra = isinstance(alg, RankAdaptiveUnconventionalAlgorithm)
    if isinstance(alg, ProjectorSplitting) and isinstance(alg.order, Strang):
        raise TypeError("array-form data has no half-step snapshot: the Strang reverse step is "
                        "native.Context.step_ksl_strang_adjoint on function-form snapshots")
    if not (ksl or dual or ra or isinstance(alg, UnconventionalAlgorithm)):
        raise TypeError("the adjoint is built for the unconventional (BUG), rank-adaptive BUG and primal "
                        "and dual KSL integrators")
    if ksl:
        fwd = lambda c, a, b: c.step_ksl(native.PRIMAL, a, b)  # noqa: E731
        rev = lambda c, a, b, *x: c.step_ksl_adjoint(a, b, *x)  # noqa: E731
    elif dual:
        fwd = lambda c, a, b: c.step_ksl(native.DUAL, a, b)  # noqa: E731
        rev = lambda c, a, b, *x: c.step_ksl_dual_adjoint(a, b, *x)  # noqa: E731
    elif ra:
        fwd = lambda c, a, b: c.step_bug_ra(a, b, alg.tol, alg.rmax)  # noqa: E731
        rev = lambda c, a, b, *x: c.step_bug_ra_adjoint(a, b, alg.tol, alg.rmax, *x)  # noqa: E731
    else:
        fwd = lambda c, a, b: c.step_bug(a, b)  # noqa: E731
        rev = lambda c, a, b, *x: c.step_bug_adjoint(a, b, *x)  # noqa: E731
    if not isinstance(prob, MatrixDataProblem) or callable(prob.y):
        raise TypeError("solve_gradient takes a MatrixDataProblem with array-form data")
    integ = init(prob, alg, 1, precision=precision, snapshot_dtype=snapshot_dtype, save_everystep=False,
                 device=device)
    ctx = integ.ctx
    try:
        t0, tf = prob.tspan
        T = tf - t0  # steps over snapshot indices t0 .. tf - 1 (0-based)
        k0 = t0 - 1

        def kp(s_):  # yprev starts at y[1] whatever t0 is (as in step)
            return 0 if s_ == 0 else k0 + s_

        cps = {0: ctx.get_state()}
        for s_ in range(T):
            fwd(ctx, kp(s_), k0 + s_ + 1)
            if (s_ + 1) % checkpoint == 0 and s_ + 1 < T:
                cps[s_ + 1] = ctx.get_state()
        value, Ub, Sb, Vb = loss_grad(*ctx.get_state())
        for c0 in sorted(cps, reverse=True):
            c1 = min(c0 + checkpoint, T)
            states = [cps[c0]]
            ctx.set_state(*cps[c0])
            for s_ in range(c0, c1 - 1):
                fwd(ctx, kp(s_), k0 + s_ + 1)
                states.append(ctx.get_state())
            for s_ in range(c1 - 1, c0 - 1, -1):
                ctx.set_state(*states[s_ - c0])
                Ub, Sb, Vb = rev(ctx, kp(s_), k0 + s_ + 1, Ub, Sb, Vb)
        return value, (Ub, Sb, Vb)
    finally:
        ctx.close()


def _de_step_count(tspan, dt) -> int:
    """The number of steps ``solve`` takes (t accumulated, primitives.jl:71-74)."""
    t0, tf = tspan
    T = tf - t0
    t, k = t0, 0
    while (tf - t) / T > 1e-8:
        t = t + dt
        k += 1
    return k


def de_substeps(prob, dt, substeps: Optional[int] = None) -> int:
    """The RK4 sub-steps ``solve`` uses for a matrix ODE with step ``dt``."""
    if substeps is not None:
        return int(substeps)
    F, u0 = prob.f, prob.u0
    if isinstance(F, QuadraticOperator):
        ymax = float(np.max(np.linalg.norm(u0.U @ u0.S, axis=1)) * np.max(np.linalg.norm(u0.V, axis=1)))
        return rk4_substeps(F.spectral_bound(2.0 * ymax), dt)
    return rk4_substeps(F.spectral_bound(), dt)


def _solve_gradient_de(prob, alg, loss_grad, dt, checkpoint, substeps, device):
    ksl = isinstance(alg, ProjectorSplitting) and isinstance(alg.order, PrimalLieTrotter)
    ra = isinstance(alg, RankAdaptiveUnconventionalAlgorithm)
    if not (ksl or ra or isinstance(alg, UnconventionalAlgorithm)):
        raise TypeError("the matrix-ODE adjoint is built for the unconventional (BUG), rank-adaptive BUG and "
                        "primal KSL integrators")
    if ksl:
        fwd = lambda c: c.step_de_ksl(native.PRIMAL, dt, N)  # noqa: E731
        rev = lambda c, *x: c.step_de_ksl_adjoint(dt, N, *x)  # noqa: E731
    elif ra:  # the ranks of the forward solve held fixed
        fwd = lambda c: c.step_de_bug_ra(dt, N, alg.tol, alg.rmax)  # noqa: E731
        rev = lambda c, *x: c.step_de_bug_ra_adjoint(dt, N, alg.tol, alg.rmax, *x)  # noqa: E731
    else:
        fwd = lambda c: c.step_de_bug(dt, N)  # noqa: E731
        rev = lambda c, *x: c.step_de_bug_adjoint(dt, N, *x)  # noqa: E731
    if not isinstance(prob.f, QuadraticOperator):
        raise TypeError("the matrix-ODE adjoint takes a QuadraticOperator (D = 0 for a linear right-hand side)")
    if dt is None:
        raise AssertionError("If the data is not provided as array, integration stepsize needs to be specified")
    N = de_substeps(prob, dt, substeps)
    steps = _de_step_count(prob.tspan, dt)
    integ = init(prob, alg, dt, save_everystep=False, device=device, substeps=N)
    ctx = integ.ctx
    try:
        cps = {0: ctx.get_state()}
        for s_ in range(steps):
            fwd(ctx)
            if (s_ + 1) % checkpoint == 0 and s_ + 1 < steps:
                cps[s_ + 1] = ctx.get_state()
        value, Ub, Sb, Vb = loss_grad(*ctx.get_state())
        abar = bbar = 0.0
        for c0 in sorted(cps, reverse=True):
            c1 = min(c0 + checkpoint, steps)
            states = [cps[c0]]
            ctx.set_state(*cps[c0])
            for _ in range(c0, c1 - 1):
                fwd(ctx)
                states.append(ctx.get_state())
            for s_ in range(c1 - 1, c0 - 1, -1):
                ctx.set_state(*states[s_ - c0])
                Ub, Sb, Vb, (a, b) = rev(ctx, Ub, Sb, Vb)
                abar += a
                bbar += b
        return value, (Ub, Sb, Vb), (abar, bbar)
    finally:
        ctx.close()


# ----------------------------------------------------------------------------
# the parametric Burgers problem of the differentiation workload
# ----------------------------------------------------------------------------


def _qr_pos(A):
    Q, R = np.linalg.qr(A, mode="reduced")
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return Q * s, R * s[:, None]


def periodic_burgers_operator(n: int, m: int, nu: float) -> QuadraticOperator:
    """nu u_xx - u u_x on n points x_i = i / n of the periodic [0, 1), central differences:
    A = nu / dx^2 tridiag(1, -2, 1), D = 1 / (2 dx) tridiag(-1, 0, 1), both periodic;
    m columns (samples), no coupling between them (B = 0)."""
    dx = 1.0 / n
    A = SymTridiagonal(-2.0 * nu / dx ** 2 * np.ones(n), nu / dx ** 2 * np.ones(n), periodic=True)
    B = SymTridiagonal(np.zeros(m), np.zeros(m - 1))
    g = 0.5 / dx
    D = Tridiagonal(-g * np.ones(n), np.zeros(n), g * np.ones(n), periodic=True)
    return QuadraticOperator(A, B, D)


class SineBurgers:
    """Parametric viscous Burgers with sine initial data (one sample per column):
    u0(x; xi) = sum_{k=1..K} xi_k sin(2 pi k x) / k on n points of [0, 1), xi K x m,
    viscosity nu, rank r >= K.  The rank-K initial matrix Y0 = Phi xi is represented
    as U0 S0 V0' with U0 the normalised sines k = 1..r (orthonormal on the grid, and
    independent of xi), V0 = Q([xi' | omega]) (the QR with a positive diagonal;
    omega a fixed m x (r - K) Gaussian of ``seed``) and S0 = U0' Phi xi V0.

    ``problem(tspan)`` is the ``MatrixDEProblem``; ``pullback(Ub0, Sb0, Vb0)`` maps
    cotangents of (U0, S0, V0) to that of xi (U0 does not depend on xi)."""

    def __init__(self, n: int, xi: np.ndarray, nu: float, r: int, seed: int = 0):
        xi = np.asarray(xi, dtype=np.float64)
        K, m = xi.shape
        if not (K <= r <= m):
            raise ValueError(f"need K = {K} <= r = {r} <= m = {m}")
        self.n, self.m, self.r, self.nu, self.xi = n, m, r, float(nu), xi
        x = np.arange(n) / n
        self.U0 = np.stack([np.sin(2 * np.pi * k * x) for k in range(1, r + 1)], axis=1) * np.sqrt(2.0 / n)
        Phi = np.stack([np.sin(2 * np.pi * k * x) / k for k in range(1, K + 1)], axis=1)
        self.C = self.U0.T @ Phi  # r x K
        self.omega = np.random.default_rng(seed).standard_normal((m, r - K))
        self.V0, self.R = _qr_pos(np.hstack([xi.T, self.omega]))
        self.S0 = self.C @ (xi @ self.V0)
        self.op = periodic_burgers_operator(n, m, nu)

    def initial(self) -> SVDLikeRepresentation:
        return SVDLikeRepresentation(self.U0.copy(), self.S0.copy(), self.V0.copy())

    def problem(self, tspan=(0.0, 0.02)) -> "MatrixDEProblem":
        return MatrixDEProblem(self.op, self.initial(), tspan)

    def pullback(self, Ub0, Sb0, Vb0) -> np.ndarray:
        K = self.xi.shape[0]
        xib = self.C.T @ Sb0 @ self.V0.T                  # S0 = C xi V0
        Vb = Vb0 + self.xi.T @ (self.C.T @ Sb0)
        Q, R = self.V0, self.R                            # V0 = Q([xi' | omega]), R not used
        M = -Vb.T @ Q
        Mlt = np.tril(M) + np.tril(M, -1).T
        Ab = np.linalg.solve(R, (Vb + Q @ Mlt).T).T
        return xib + Ab[:, :K].T


def mean_column_loss(U, S, V):
    """L = || mean over columns of U S V' ||^2 and its cotangents (Ub, Sb, Vb)."""
    m = V.shape[0]
    w = V.sum(axis=0) / m
    y = S @ w
    z = U @ y
    Uz = U.T @ z
    return float(z @ z), 2.0 * np.outer(z, y), 2.0 * np.outer(Uz, w), np.outer(np.ones(m), 2.0 * S.T @ Uz) / m
