"""MI355X (gfx950) engine for the per-step DLRA hot path of LowRankIntegrators.jl.

The steps of the projector-splitting (KSL), unconventional (BUG) and
rank-adaptive BUG integrators (and the Greedy integrator) on data problems, and on matrix ODEs with a
tridiagonal Lyapunov-type right-hand side, run in ``liblri_hip.so``
(hand-written HIP kernels behind the C ABI of ``include/lri.h``); this package
is the host side: a ctypes binding (``native``) and a mirror of the
reference's Julia API (``api``).
"""
from . import metrics, native  # noqa: F401
from .api import (  # noqa: F401
    DLRIntegrator,
    DLRSolution,
    DualLieTrotter,
    GreedyIntegrator,
    LyapunovOperator,
    Matrix,
    QuadraticOperator,
    SineBurgers,
    Tridiagonal,
    burgers_operator,
    burgers_samples,
    de_substeps,
    mean_column_loss,
    periodic_burgers_operator,
    MatrixDataProblem,
    MatrixDEProblem,
    MatrixHybridProblem,
    PrimalLieTrotter,
    ProjectorSplitting,
    RankAdaptiveUnconventionalAlgorithm,
    Strang,
    SVDLikeRepresentation,
    SymTridiagonal,
    TwoFactorRepresentation,
    UnconventionalAlgorithm,
    init,
    laplacian_1d,
    rank,
    rk4_substeps,
    solve,
    solve_batch,
    solve_gradient,
    step,
    truncated_svd,
    update_sol,
    fetch_sol,
)

step_ = step  # Julia's step!
