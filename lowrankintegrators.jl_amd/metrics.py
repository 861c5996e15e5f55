"""Comparison of low-rank results in factored form (no n x m matrix is formed).

Bases from different QR methods differ by rotations and signs, so results are
compared as matrices Y = U S V'.  The difference is formed as a small core
between the R factors of the stacked bases, which avoids the cancellation of
||S1||^2 + ||S2||^2 - 2 tr(...) (that form stops near 1e-8).
"""
import numpy as np


def rel_err(Ua, Sa, Va, Ub, Sb, Vb) -> float:
    """||Ua Sa Va' - Ub Sb Vb'||_F / ||Ub Sb Vb'||_F."""
    _, Ru = np.linalg.qr(np.hstack([Ua, Ub]), mode="reduced")
    _, Rv = np.linalg.qr(np.hstack([Va, Vb]), mode="reduced")
    ra, rb = Sa.shape[0], Sb.shape[0]
    core = np.zeros((ra + rb, ra + rb))
    core[:ra, :ra] = Sa
    core[ra:, ra:] = -Sb
    diff = np.linalg.norm(Ru @ core @ Rv.T)
    _, R1 = np.linalg.qr(Ub, mode="reduced")
    _, R2 = np.linalg.qr(Vb, mode="reduced")
    return float(diff / np.linalg.norm(R1 @ Sb @ R2.T))
