"""MI355X (gfx950) engine for the per-step DLRA hot path of LowRankIntegrators.jl.

The steps of the projector-splitting (KSL), unconventional (BUG) and
rank-adaptive BUG integrators on data problems run in ``liblri_hip.so``
(hand-written HIP kernels behind the C ABI of ``include/lri.h``); this package
is the host side: a ctypes binding (``native``) and a mirror of the
reference's Julia API (``api``).
"""
from . import native  # noqa: F401
from .api import (  # noqa: F401
    DLRIntegrator,
    DLRSolution,
    DualLieTrotter,
    Matrix,
    MatrixDataProblem,
    MatrixDEProblem,
    MatrixHybridProblem,
    PrimalLieTrotter,
    ProjectorSplitting,
    RankAdaptiveUnconventionalAlgorithm,
    Strang,
    SVDLikeRepresentation,
    TwoFactorRepresentation,
    UnconventionalAlgorithm,
    init,
    rank,
    solve,
    step,
    truncated_svd,
    update_sol,
)

step_ = step  # Julia's step!
