"""Low-rank representation types and truncation rules (oracle side).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference takes these from the external package LowRankArithmetic.jl
(compat "0.1.2", ``Project.toml:15``), which is not vendored with the
reference and is not installed here.  Call sites in the reference:

* ``SVDLikeRepresentation(U, S, V)`` -- Y = U*S*V'
  (``src/integrators/rank_adaptive_unconventional.jl:231``)
* ``TwoFactorRepresentation(U, Z)`` -- Y = U*Z'
  (``src/integrators/projector_splitting.jl:54,74``)
* ``rank(u)``, ``size(u)`` (``rank_adaptive_unconventional.jl:50-51``)
* ``truncate_to_tolerance(S, tol)`` (``rank_adaptive_unconventional.jl:223``)
* ``truncated_svd(Y0; tol)`` (``test/data_driven_approximation.jl:18``)

Pinning
-------
* ``SVDLike`` / ``TwoFactor``: pinned by ``to_matrix`` against the dense
  definition in the tests (trivial algebra).
* ``truncated_svd(A, r)``: pinned by the Eckart-Young property (error equals
  the norm of the discarded singular values) in ``tests/test_oracle.py``.
* ``truncate_to_tolerance``: **parity unpinned** for its boundary rule.  The
  rule adopted is the Ceruti-Kusch-Lubich one, r1 = min{rho >= 1 :
  sqrt(sum_{j>rho} sigma_j^2) <= tol}; whether LowRankArithmetic 0.1.2 uses an
  absolute or relative tolerance, ``<`` or ``<=``, and which minimum rank, can
  only be settled by importing that package.  The tests pin the mathematical
  content (monotonicity, the discarded tail really is <= tol, minimality).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SVDLike:
    """Y = U @ S @ V.T with U (n x r), S (r x r), V (m x r), all float64."""

    U: np.ndarray
    S: np.ndarray
    V: np.ndarray

    @property
    def rank(self) -> int:
        return self.S.shape[0]

    @property
    def shape(self):
        return (self.U.shape[0], self.V.shape[0])

    def to_matrix(self) -> np.ndarray:
        return self.U @ self.S @ self.V.T

    def copy(self) -> "SVDLike":
        return SVDLike(self.U.copy(), self.S.copy(), self.V.copy())


@dataclass
class TwoFactor:
    """Y = U @ Z.T (LowRankArithmetic ``TwoFactorRepresentation``)."""

    U: np.ndarray
    Z: np.ndarray

    @property
    def rank(self) -> int:
        return int(self.U.shape[1])

    def to_matrix(self) -> np.ndarray:
        return self.U @ self.Z.T

    def copy(self) -> "TwoFactor":
        return TwoFactor(self.U.copy(), self.Z.copy())


def truncate_to_tolerance(sigma: np.ndarray, tol: float) -> int:
    """Smallest rank rho >= 1 whose discarded tail has 2-norm <= tol.

    ``sigma`` is sorted non-increasing (LAPACK gesdd order, as returned by
    ``svd`` at ``rank_adaptive_unconventional.jl:222``).  Parity unpinned for
    the boundary rule (see module header).
    """
    s = np.asarray(sigma, dtype=np.float64)
    k = s.shape[0]
    if k == 0:
        return 0
    # tail[rho] = sqrt(sum_{j >= rho} s_j^2) for 0-based rho; keep rho values.
    tail2 = np.concatenate([np.cumsum((s[::-1]) ** 2)[::-1], [0.0]])
    for rho in range(1, k + 1):
        if np.sqrt(tail2[rho]) <= tol:
            return rho
    return k


def truncated_svd(A: np.ndarray, r: int | None = None, tol: float | None = None) -> SVDLike:
    """Rank-r (or tolerance-tol) truncated SVD of a dense matrix, FP64 LAPACK gesdd."""
    U, s, Vt = np.linalg.svd(np.asarray(A, dtype=np.float64), full_matrices=False)
    if r is None:
        if tol is None:
            raise ValueError("truncated_svd needs a rank r or a tolerance tol")
        r = truncate_to_tolerance(s, tol)
    r = int(min(r, s.shape[0]))
    return SVDLike(U[:, :r].copy(), np.diag(s[:r]), Vt[:r, :].T.copy())
