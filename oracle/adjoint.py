"""Reverse-mode (adjoint) restatement of the data-driven BUG step.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference has no adjoint: its integrators are differentiated, if at all,
by Julia AD through ``unconventional_step!`` (``src/integrators/unconventional.jl:133-157``).
This module is the CPU reference gradient for the engine's hand-written
reverse step: the vector-Jacobian product of

    K1 = U0 S0 + dA V0,   U1 = Q(K1)           (unconventional.jl:137-142)
    L1 = V0 S0' + dA' U0, V1 = Q(L1)           (:145-150)
    M = U1' U0, N = V1' V0
    S1 = M S0 N' + U1' dA V1                   (:154-156)

with the thin-QR rule for Q = Q(K), K = Q R, R unused:
    Kbar = (Qbar + Q copyltu(-Qbar' Q)) R^{-T},  copyltu(X) = tril(X) + tril(X, -1)'.

Pinning (``tests/test_oracle_adjoint.py``): central finite differences of the
forward step (``dlra.bug_step``) along random directions, one step and
several steps; the QR rules separately against finite differences of
``numpy.linalg.qr``; and a forward-mode (tangent) restatement of the step,
``bug_step_jvp``, derived independently, against the reverse one through
<ybar, J xdot> = <J' ybar, xdot>.
"""
from __future__ import annotations

import numpy as np

from .dlra import _qr


def copyltu(X):
    return np.tril(X) + np.tril(X, -1).T


def qr_vjp(Q, R, Qbar):
    """Cotangent of K for Q = Q(K) (thin QR, R not used downstream)."""
    M = -Qbar.T @ Q
    return np.linalg.solve(R, (Qbar + Q @ copyltu(M)).T).T  # (.) R^{-T}


def qr_vjp_full(Q, R, Qbar, Rbar):
    """Cotangent of K for (Q, R) = QR(K) when both factors are used downstream:
    Kbar = (Qbar + Q copyltu(R Rbar' - Qbar' Q)) R^{-T}."""
    M = R @ Rbar.T - Qbar.T @ Q
    return np.linalg.solve(R, (Qbar + Q @ copyltu(M)).T).T


def ksl_primal_step_vjp(U, S, V, dA, Ub1, Sb1, Vb1):
    """Cotangents (Ub0, Sb0, Vb0, dAb) of the primal KSL data step
    (projector_splitting.jl:129-152):
        P = dA V, K = U S + P -> (U1, R1);  St = R1 - U1' P;
        L = V St' + dA' U1 -> (V1, R2);     S1 = R2'."""
    P = dA @ V
    U1, R1 = _qr(U @ S + P)
    St = R1 - U1.T @ P
    V1, R2 = _qr(V @ St.T + dA.T @ U1)
    Lb = qr_vjp_full(V1, R2, Vb1, Sb1.T)
    Vb0 = Lb @ St
    Stb = Lb.T @ V
    Ub1 = Ub1 + dA @ Lb - P @ Stb.T
    dAb = U1 @ Lb.T
    Pb = -U1 @ Stb
    Kb = qr_vjp_full(U1, R1, Ub1, Stb)
    Ub0 = Kb @ S.T
    Sb0 = U.T @ Kb
    Pb = Pb + Kb
    Vb0 = Vb0 + dA.T @ Pb
    dAb = dAb + Pb @ V.T
    return Ub0, Sb0, Vb0, dAb


def ksl_dual_step_vjp(U, S, V, dA, Ub1, Sb1, Vb1):
    """Cotangents (Ub0, Sb0, Vb0, dAb) of the dual KSL data step
    (projector_splitting.jl:166-189):
        Qd = dA' U, L = V S' + Qd -> (V1, R);  St = R' - Qd' V1;
        K = U St + dA V1 -> (U1, R2);          S1 = R2."""
    Qd = dA.T @ U
    V1, R = _qr(V @ S.T + Qd)
    St = R.T - Qd.T @ V1
    U1, R2 = _qr(U @ St + dA @ V1)
    Kb = qr_vjp_full(U1, R2, Ub1, Sb1)
    Ub0 = Kb @ St.T
    Stb = U.T @ Kb
    dAb = Kb @ V1.T
    V1b = Vb1 + dA.T @ Kb - Qd @ Stb
    Qdb = -V1 @ Stb.T
    Lb = qr_vjp_full(V1, R, V1b, Stb.T)
    Vb0 = Lb @ S
    Sb0 = Lb.T @ V
    Qdb = Qdb + Lb
    dAb = dAb + U @ Qdb.T
    Ub0 = Ub0 + dA @ Qdb
    return Ub0, Sb0, Vb0, dAb


def ksl_strang_step_vjp(U, S, V, dA1, dA2, Ub1, Sb1, Vb1):
    """Cotangents (Ub0, Sb0, Vb0, dAb1, dAb2) of the Strang KSL data step
    (projector_splitting.jl:205-211): the primal half step on dA1, then the
    dual half step on dA2 from its result."""
    from .dlra import ksl_primal_step

    Um, Sm, Vm = ksl_primal_step(U, S, V, dA1)
    Ubm, Sbm, Vbm, dAb2 = ksl_dual_step_vjp(Um, Sm, Vm, dA2, Ub1, Sb1, Vb1)
    Ub0, Sb0, Vb0, dAb1 = ksl_primal_step_vjp(U, S, V, dA1, Ubm, Sbm, Vbm)
    return Ub0, Sb0, Vb0, dAb1, dAb2


def ksl_primal_step_jvp(U, S, V, dA, Ud, Sd, Vd, dAd):
    """Forward-mode tangents (U1d, S1d, V1d) of the primal KSL data step."""
    P = dA @ V
    Pd = dAd @ V + dA @ Vd
    K = U @ S + P
    Kd = Ud @ S + U @ Sd + Pd
    U1, R1 = _qr(K)
    U1d, R1d = qr_jvp(U1, R1, Kd)
    St = R1 - U1.T @ P
    Std = R1d - U1d.T @ P - U1.T @ Pd
    L = V @ St.T + dA.T @ U1
    Ld = Vd @ St.T + V @ Std.T + dAd.T @ U1 + dA.T @ U1d
    V1, R2 = _qr(L)
    V1d, R2d = qr_jvp(V1, R2, Ld)
    return U1d, R2d.T, V1d


def bug_step_vjp(U, S, V, dA, Ub1, Sb1, Vb1):
    """Cotangents (Ub0, Sb0, Vb0, dAb) of the BUG data step's inputs, given those
    of its outputs (U1, S1, V1); recomputes the forward step (same QR as dlra)."""
    K = U @ S + dA @ V
    U1, RK = _qr(K)
    L = V @ S.T + dA.T @ U
    V1, RL = _qr(L)
    M = U1.T @ U
    N = V1.T @ V
    # S1 = M S N' + U1' dA V1
    Mb = Sb1 @ N @ S.T
    Nb = Sb1.T @ M @ S
    Sb0 = M.T @ Sb1 @ N
    Ub1 = Ub1 + dA @ V1 @ Sb1.T + U @ Mb.T
    Vb1 = Vb1 + dA.T @ U1 @ Sb1 + V @ Nb.T
    dAb = U1 @ Sb1 @ V1.T
    Ub0 = U1 @ Mb
    Vb0 = V1 @ Nb
    # QR backward
    Kb = qr_vjp(U1, RK, Ub1)
    Lb = qr_vjp(V1, RL, Vb1)
    # K = U S + dA V ; L = V S' + dA' U
    Ub0 = Ub0 + Kb @ S.T + dA @ Lb
    Sb0 = Sb0 + U.T @ Kb + Lb.T @ V
    Vb0 = Vb0 + dA.T @ Kb + Lb @ S
    dAb = dAb + Kb @ V.T + U @ Lb.T
    return Ub0, Sb0, Vb0, dAb


def qr_jvp(Q, R, Kd):
    """Tangents (Qd, Rd) of the thin QR K = Q R along Kd (forward mode):
    C = Q' Kd R^{-1}, X = Rd R^{-1} upper triangular with Q'Qd = C - X antisymmetric."""
    Ri = np.linalg.inv(R)
    C = Q.T @ Kd @ Ri
    X = np.triu(C) + np.triu(C.T, 1)
    Qd = Q @ (C - X) + (Kd @ Ri - Q @ (Q.T @ Kd @ Ri))
    return Qd, X @ R


def bug_step_jvp(U, S, V, dA, Ud, Sd, Vd, dAd):
    """Forward-mode tangents (U1d, S1d, V1d) of the BUG data step along (Ud, Sd, Vd, dAd):
    an independent derivation of the same derivative as bug_step_vjp (the two agree
    through <ybar, J xdot> = <J' ybar, xdot>)."""
    K = U @ S + dA @ V
    Kd = Ud @ S + U @ Sd + dAd @ V + dA @ Vd
    U1, RK = _qr(K)
    U1d, _ = qr_jvp(U1, RK, Kd)
    L = V @ S.T + dA.T @ U
    Ld = Vd @ S.T + V @ Sd.T + dAd.T @ U + dA.T @ Ud
    V1, RL = _qr(L)
    V1d, _ = qr_jvp(V1, RL, Ld)
    M, N = U1.T @ U, V1.T @ V
    Md = U1d.T @ U + U1.T @ Ud
    Nd = V1d.T @ V + V1.T @ Vd
    S1d = (Md @ S @ N.T + M @ Sd @ N.T + M @ S @ Nd.T
           + U1d.T @ dA @ V1 + U1.T @ dAd @ V1 + U1.T @ dA @ V1d)
    return U1d, S1d, V1d


def bug_solve_gradient(U0, S0, V0, snaps, loss_grad, checkpoint=20, ksl=False):
    """Gradient of loss(U_T, S_T, V_T) after the BUG (ksl True or "primal": primal
    KSL; "dual": dual KSL) steps over snaps[0..T] with respect to (U0, S0, V0):
    checkpoints every `checkpoint` steps on the way forward, recomputation inside
    each segment on the way back.  loss_grad(U, S, V) -> (value, Ub, Sb, Vb)."""
    from .dlra import bug_step, ksl_dual_step, ksl_primal_step

    if ksl == "dual":
        step, vjp = ksl_dual_step, ksl_dual_step_vjp
    elif ksl:
        step, vjp = ksl_primal_step, ksl_primal_step_vjp
    else:
        step, vjp = bug_step, bug_step_vjp

    T = len(snaps) - 1
    cps = {0: (U0, S0, V0)}
    U, S, V = U0, S0, V0
    for k in range(T):
        U, S, V = step(U, S, V, snaps[k + 1] - snaps[k])
        if (k + 1) % checkpoint == 0 and k + 1 < T:
            cps[k + 1] = (U, S, V)
    value, Ub, Sb, Vb = loss_grad(U, S, V)
    for c0 in sorted(cps, reverse=True):
        c1 = min(c0 + checkpoint, T)
        states = [cps[c0]]
        for k in range(c0, c1 - 1):
            states.append(step(*states[-1], snaps[k + 1] - snaps[k]))
        for k in range(c1 - 1, c0 - 1, -1):
            Ub, Sb, Vb, _ = vjp(*states[k - c0], snaps[k + 1] - snaps[k], Ub, Sb, Vb)
    return value, Ub, Sb, Vb


# ----------------------------------------------------------------------------
# rank-adaptive BUG step: SVD backward
# ----------------------------------------------------------------------------


def svd_vjp(Us, sig, Vs, Ub, sb, Vb):
    """Cotangent of a square A = Us diag(sig) Vs' (full SVD, singular values distinct)
    given cotangents of its leading factors Us[:, :k], sig[:k], Vs[:, :k] (Ub 2r x k,
    sb k, Vb 2r x k; the discarded factors carry none):
        Abar = Us [ (F o (Us'Ubf - Ubf'Us)) Sig + diag(sbf) + Sig (F o (Vs'Vbf - Vbf'Vs)) ] Vs',
    F_ij = 1 / (sig_j^2 - sig_i^2) (i != j; 0 where both i, j >= k), Ubf, Vbf, sbf the
    cotangents padded with zeros."""
    q = Us.shape[0]
    k = Ub.shape[1]
    Ubf = np.zeros((q, q))
    Vbf = np.zeros((q, q))
    Ubf[:, :k] = Ub
    Vbf[:, :k] = Vb
    sbf = np.zeros(q)
    sbf[:k] = sb
    s2 = sig ** 2
    D = s2[None, :] - s2[:, None]
    F = np.zeros((q, q))
    keep = np.zeros((q, q), dtype=bool)
    keep[:k, :] = True
    keep[:, :k] = True
    np.fill_diagonal(keep, False)
    F[keep] = 1.0 / D[keep]
    J = Us.T @ Ubf
    K = Vs.T @ Vbf
    Sg = np.diag(sig)
    core = (F * (J - J.T)) @ Sg + np.diag(sbf) + Sg @ (F * (K - K.T))
    return Us @ core @ Vs.T


def _complete(B, X):
    """Q with [B | Q] orthonormal spanning [B | X] (B orthonormal): the QR of
    (I - B B') X, and its R."""
    Z = X - B @ (B.T @ X)
    Q, R = np.linalg.qr(Z, mode="reduced")
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return Q * s, R * s[:, None], Z


def ra_bug_step_aug(U, S, V, dA, tol, rmax):
    """The rank-adaptive BUG data step in the form its reverse differentiates:
    Uhat = [U0 | Q((I - U0 U0') K1)], Vhat likewise, then as the reference.  The
    result U1 S1 V1' equals the reference's (same spans); the factors may differ in
    the signs of singular-vector pairs, so cotangents for ``ra_bug_step_vjp`` are
    formed from these factors.  Returns (U1, S1, V1, r1)."""
    K1 = U @ S + dA @ V
    L1 = V @ S.T + dA.T @ U
    Uh = np.hstack([U, _complete(U, K1)[0]])
    Vh = np.hstack([V, _complete(V, L1)[0]])
    Us, sig, Vst = np.linalg.svd(Uh.T @ (U @ S @ V.T + dA) @ Vh)
    from .lra import truncate_to_tolerance

    r1 = int(min(rmax, truncate_to_tolerance(sig, tol)))
    return Uh @ Us[:, :r1], np.diag(sig[:r1]), Vh @ Vst[:r1].T, r1


def ra_bug_step_vjp(U, S, V, dA, tol, rmax, Ub1, Sb1, Vb1):
    """Cotangents (Ub0, Sb0, Vb0, dAb) of the rank-adaptive BUG data step
    (rank_adaptive_unconventional.jl:194-233) given those of its result
    (U1 n x r1, S1 = diag(sigma[:r1]), V1 m x r1), the rank r1 held fixed.  The
    augmented bases are taken as Uhat = [U0 | Q((I - U0 U0') K1)], which spans the same
    space as the reference's Q([K1 U0]) (the result U1 S1 V1' does not depend on the
    choice); only the diagonal of Sb1 enters (S1 is diagonal).  Returns also r1."""
    K1 = U @ S + dA @ V
    L1 = V @ S.T + dA.T @ U
    Qu, Ru, _ = _complete(U, K1)
    Qv, Rv, _ = _complete(V, L1)
    Uh = np.hstack([U, Qu])
    Vh = np.hstack([V, Qv])
    Y = U @ S @ V.T + dA
    Sh = Uh.T @ Y @ Vh
    Us, sig, Vst = np.linalg.svd(Sh)
    Vs = Vst.T
    from .lra import truncate_to_tolerance

    r1 = int(min(rmax, truncate_to_tolerance(sig, tol)))
    assert Ub1.shape[1] == r1, (Ub1.shape, r1)
    # U1 = Uh Us[:, :r1], V1 = Vh Vs[:, :r1], S1 = diag(sig[:r1])
    Uhb = Ub1 @ Us[:, :r1].T
    Vhb = Vb1 @ Vs[:, :r1].T
    Shb = svd_vjp(Us, sig, Vs, Uh.T @ Ub1, np.diag(Sb1).copy(), Vh.T @ Vb1)
    # Sh = Uh' Y Vh, Y = U S V' + dA
    Uhb = Uhb + Y @ Vh @ Shb.T
    Vhb = Vhb + Y.T @ Uh @ Shb
    Yb = Uh @ Shb @ Vh.T
    r = U.shape[1]
    Ub0 = Uhb[:, :r] + Yb @ V @ S.T
    Vb0 = Vhb[:, :r] + Yb.T @ U @ S
    Sb0 = U.T @ Yb @ V
    dAb = Yb
    # Q = Q((I - U U') K1): Zbar by the thin-QR rule, then Z = K1 - U (U'K1)
    for side in (0, 1):
        B, X, Qc, Rc, Qb = (U, K1, Qu, Ru, Uhb[:, r:]) if side == 0 else (V, L1, Qv, Rv, Vhb[:, r:])
        Zb = qr_vjp(Qc, Rc, Qb)
        Xb = Zb - B @ (B.T @ Zb)
        Bb = -Zb @ (X.T @ B) - X @ (Zb.T @ B)
        if side == 0:
            Ub0 = Ub0 + Bb + Xb @ S.T  # K1 = U S + dA V
            Sb0 = Sb0 + U.T @ Xb
            Vb0 = Vb0 + dA.T @ Xb
            dAb = dAb + Xb @ V.T
        else:
            Vb0 = Vb0 + Bb + Xb @ S  # L1 = V S' + dA' U
            Sb0 = Sb0 + Xb.T @ V
            Ub0 = Ub0 + dA @ Xb
            dAb = dAb + U @ Xb.T
    return Ub0, Sb0, Vb0, dAb, r1
