"""Factored-form comparison of low-rank results (no n x m matrix is formed).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Bases from different QR implementations differ by rotations (LAPACK
Householder R has mixed-sign diagonals, CholeskyQR2 a positive one), so
results are compared as matrices Y = U S V' and as subspaces (projectors).
Both measures avoid the cancellation of ``|S1|^2 + |S2|^2 - 2 tr(...)``, which
would limit them to ~1e-8: the difference is formed as a 2r x 2r core between
the QR factors of the stacked bases.

Pinning: ``tests/test_oracle.py`` checks both against dense NumPy
computations on small matrices, including differences near 1e-13.
"""
from __future__ import annotations

import numpy as np


def frob_diff(U1, S1, V1, U2, S2, V2) -> float:
    """||U1 S1 V1' - U2 S2 V2'||_F via QR of [U1 U2] and [V1 V2]."""
    Qa, Ra = np.linalg.qr(np.hstack([U1, U2]), mode="reduced")
    Qb, Rb = np.linalg.qr(np.hstack([V1, V2]), mode="reduced")
    r1, r2 = S1.shape[0], S2.shape[0]
    D = np.zeros((r1 + r2, r1 + r2))
    D[:r1, :r1] = S1
    D[r1:, r1:] = -S2
    return float(np.linalg.norm(Ra @ D @ Rb.T))


def frob_norm(U, S, V) -> float:
    """||U S V'||_F (exact for any U, V via their R factors)."""
    _, Ra = np.linalg.qr(U, mode="reduced")
    _, Rb = np.linalg.qr(V, mode="reduced")
    return float(np.linalg.norm(Ra @ S @ Rb.T))


def rel_err(Ua, Sa, Va, Ub, Sb, Vb) -> float:
    """Relative Frobenius error ||Ya - Yb||_F / ||Yb||_F, Y = U S V'."""
    return frob_diff(Ua, Sa, Va, Ub, Sb, Vb) / frob_norm(Ub, Sb, Vb)


def proj_dist(Q1, Q2) -> float:
    """||Q1 Q1' - Q2 Q2'||_F for orthonormal Q1, Q2 of equal width.

    Uses ||P1 - P2||_F = sqrt(2) ||(I - P2) Q1||_F (equal ranks), which has no
    cancellation of squares.
    """
    if Q1.shape[1] != Q2.shape[1]:
        raise ValueError("projector distance needs equal ranks")
    E = Q1 - Q2 @ (Q2.T @ Q1)
    return float(np.sqrt(2.0) * np.linalg.norm(E))


def orth_defect(Q) -> float:
    """||Q'Q - I||_F."""
    k = Q.shape[1]
    return float(np.linalg.norm(Q.T @ Q - np.eye(k)))
