"""Seeded synthetic data for the tests and the CPU leg of the bench.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The rotating low-rank-plus-noise family used by the workloads:

    A(t) = Q1(t) (A1 + e^t A2) Q2(t)',   Qi(t) = expm(t Ti)

* Ti: n x n (m x m) skew-symmetric, upper triangle N(0, 1/n), mirrored with
  a zero diagonal (the construction of ``test/data_driven_approximation.jl:4-13``
  with the N(0,1/n) scaling of the workload definition);
* A1 = Ua diag(2^-i, i = 0..r-1) Va' with Ua, Va from the QR of Gaussian
  matrices;
* A2 = noise * N(0, 1)  (noise = 1e-3; 0 gives exactly rank-r data).

All draws come from one ``numpy.random.Generator(PCG64(seed))`` in a fixed
order, so a seed identifies the data.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

from .lra import SVDLike, truncated_svd


def skew(rng, n):
    """Skew-symmetric matrix, upper triangle N(0, 1/n), zero diagonal."""
    W = rng.standard_normal((n, n)) / np.sqrt(n)
    W = np.triu(W, 1)
    return W - W.T


@dataclass
class RotatingData:
    n: int
    m: int
    r: int
    T1: np.ndarray
    T2: np.ndarray
    A1: np.ndarray
    A2: np.ndarray
    cache: dict = field(default_factory=dict, repr=False)

    def __call__(self, t: float) -> np.ndarray:
        """A(t); memoised per t value (the solvers ask for the same grid repeatedly)."""
        key = float(t)
        hit = self.cache.get(key)
        if hit is None:
            Q1 = sla.expm(key * self.T1)
            Q2 = sla.expm(key * self.T2)
            hit = Q1 @ (self.A1 + np.exp(key) * self.A2) @ Q2.T
            hit.setflags(write=False)
            self.cache[key] = hit
        return hit.copy()

    def snapshots(self, ts) -> list:
        """A(t_k) for each t in ts (each snapshot evaluated from its own expm)."""
        return [self(float(t)) for t in ts]

    def initial(self, r: int | None = None) -> SVDLike:
        """Rank-r truncated SVD of A(t0 = 0)."""
        return truncated_svd(self(0.0), r=self.r if r is None else r)


def rotating_data(n: int, m: int, r: int, noise: float = 1e-3, seed: int = 0) -> RotatingData:
    rng = np.random.Generator(np.random.PCG64(seed))
    T1 = skew(rng, n)
    T2 = skew(rng, m)
    Ua, _ = np.linalg.qr(rng.standard_normal((n, r)))
    Va, _ = np.linalg.qr(rng.standard_normal((m, r)))
    A1 = Ua @ np.diag(2.0 ** -np.arange(r)) @ Va.T
    A2 = noise * rng.standard_normal((n, m))
    return RotatingData(n, m, r, T1, T2, A1, A2)


def random_factors(rng, n, m, r, decay=0.25):
    """Orthonormal U (n x r), V (m x r) and S = diag(10^(-decay*i))."""
    U, _ = np.linalg.qr(rng.standard_normal((n, r)))
    V, _ = np.linalg.qr(rng.standard_normal((m, r)))
    S = np.diag(10.0 ** (-decay * np.arange(r)))
    return SVDLike(U, S, V)


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float64 values to bfloat16 (round-to-nearest-even via float32), returned as float64.

    The float64 -> float32 step rounds first; the bench and tests use it only
    for values that are already float32, for which it is exact.
    """
    f = np.ascontiguousarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    lsb = (b >> 16) & 1
    b = ((b + 0x7FFF + lsb) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """bfloat16 bit patterns (uint16) of x, round-to-nearest-even via float32."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    lsb = (b >> 16) & 1
    return (((b + 0x7FFF + lsb) >> 16) & 0xFFFF).astype(np.uint16)
