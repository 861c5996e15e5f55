"""Reverse- and forward-mode restatement of the matrix-ODE BUG and primal KSL steps.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The step is ``unconventional_step!`` (``src/integrators/unconventional.jl:133-157``)
with the default sub-problem right-hand sides (``unconventional.jl:51-79``),

    K' = F(K V0') V0,            K(0) = U0 S0,   U1 = Q(K(dt)),  M = U1'U0
    L' = F(U0 L')' U0,           L(0) = V0 S0',  V1 = Q(L(dt)),  N = V1'V0
    S' = U1' F(U1 S V1') V1,     S(0) = M S0 N', S1 = S(dt),

each sub-problem integrated by N classical RK4 steps of size dt / N, for

    F(Y) = a A Y + b Y B' + G H' - (D Y) o Y          (o: elementwise product)

at a = b = 1 (``a``, ``b``: scales of the two linear parts, whose derivatives
give e.g. the viscosity derivative of Burgers' equation, A = nu Laplacian:
dL/dnu = abar / nu).  The reference has no adjoint (Julia AD would
differentiate the step); this is the derivative of the discrete map the engine
computes, with the thin QR taken with a positive diagonal of R (the CholeskyQR
convention), so the bases here and the engine's agree column by column.

Everything is dense in (n, m): with Y (n x m) the sub-problem's matrix,
    J_F(Y)[d]    = A d + d B' - (D d) o Y - (D Y) o d,
    J_F(Y)'[Yb]  = A' Yb + Yb B - D'(Yb o Y) - Yb o (D Y),
an independent route from the engine's r x r x r tensor form.

The primal KSL step (``projector_splitting.jl:129-152`` with the sub-problems of
``:52-80``) is restated the same way (``ksl_primal_de_step*``): K -> (U1, R1), the S
sub-step backwards in time from R1, L from V0 S~' -> (V1, R2), S1 = R2'.

Pinning (``tests/test_oracle_adjoint_de.py``): central finite differences of
the forward step in every input (U0, S0, V0, a, b) along random directions,
one step and a checkpointed solve; the forward-mode restatement
``bug_de_step_jvp`` (tangent RK4 and the tangent QR rule, derived separately)
against the reverse one through <ybar, J xdot> = <J' ybar, xdot>; the RK4
reverse alone against the closed form of a linear ODE (the RK4 stability
polynomial).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .adjoint import qr_jvp, qr_vjp


def qr_pos(K):
    """Thin QR with a non-negative diagonal of R (the QR that CholeskyQR computes)."""
    Q, R = np.linalg.qr(K, mode="reduced")
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return Q * s, R * s[:, None]


@dataclass
class ScaledQuadratic:
    """F(Y) = a A Y + b Y B' + g h' - (D Y) o Y; A, B, D dense or scipy sparse."""

    A: object
    B: object
    g: np.ndarray
    h: np.ndarray
    D: object
    a: float = 1.0
    b: float = 1.0

    def __call__(self, Y):
        return self.a * (self.A @ Y) + self.b * (self.B @ Y.T).T + self.g @ self.h.T - (self.D @ Y) * Y

    def jvp(self, Y, d, ad=0.0, bd=0.0):
        """J_F(Y)[d] plus the derivative along the scales (ad, bd)."""
        out = self.a * (self.A @ d) + self.b * (self.B @ d.T).T - (self.D @ d) * Y - (self.D @ Y) * d
        if ad:
            out = out + ad * (self.A @ Y)
        if bd:
            out = out + bd * (self.B @ Y.T).T
        return out

    def vjp(self, Y, Yb):
        """(J_F(Y)' Yb, <Yb, A Y>, <Yb, Y B'>)."""
        AY = self.A @ Y
        YB = (self.B @ Y.T).T
        Xb = (self.a * (self.A.T @ Yb) + self.b * (self.B.T @ Yb.T).T - self.D.T @ (Yb * Y)
              - Yb * (self.D @ Y))
        return Xb, float(np.sum(Yb * AY)), float(np.sum(Yb * YB))

    def scaled(self, a, b):
        return ScaledQuadratic(self.A, self.B, self.g, self.h, self.D, a, b)


# ----------------------------------------------------------------------------
# RK4: forward, reverse, tangent
# ----------------------------------------------------------------------------


def rk4_states(f, y, h, N):
    """Sub-step starts y_0 .. y_N of N classical RK4 steps of y' = f(y)."""
    ys = [y]
    for _ in range(N):
        k1 = f(y)
        k2 = f(y + 0.5 * h * k1)
        k3 = f(y + 0.5 * h * k2)
        k4 = f(y + h * k3)
        y = y + (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
        ys.append(y)
    return ys


def rk4_vjp(f, fvjp, ys, h, ybar):
    """Cotangent of y_0 given that of y_N.  fvjp(x, kbar) returns J_f(x)' kbar and
    accumulates the parameter cotangents of the stage itself."""
    for s in range(len(ys) - 2, -1, -1):
        y = ys[s]
        k1 = f(y)
        X2 = y + 0.5 * h * k1
        k2 = f(X2)
        X3 = y + 0.5 * h * k2
        k3 = f(X3)
        X4 = y + h * k3
        yb = ybar.copy()
        xb = fvjp(X4, (h / 6.0) * ybar)
        yb += xb
        xb = fvjp(X3, (h / 3.0) * ybar + h * xb)
        yb += xb
        xb = fvjp(X2, (h / 3.0) * ybar + 0.5 * h * xb)
        yb += xb
        xb = fvjp(y, (h / 6.0) * ybar + 0.5 * h * xb)
        yb += xb
        ybar = yb
    return ybar


def rk4_jvp(f, fjvp, y, yd, h, N):
    """(y_N, its tangent) for y_0 = y with tangent yd; fjvp(x, xd) = J_f(x) xd plus the
    parameters' tangents."""
    for _ in range(N):
        k1, k1d = f(y), fjvp(y, yd)
        X2, X2d = y + 0.5 * h * k1, yd + 0.5 * h * k1d
        k2, k2d = f(X2), fjvp(X2, X2d)
        X3, X3d = y + 0.5 * h * k2, yd + 0.5 * h * k2d
        k3, k3d = f(X3), fjvp(X3, X3d)
        X4, X4d = y + h * k3, yd + h * k3d
        k4, k4d = f(X4), fjvp(X4, X4d)
        y = y + (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
        yd = yd + (h / 6.0) * (k1d + 2.0 * k2d + 2.0 * k3d + k4d)
    return y, yd


# ----------------------------------------------------------------------------
# the BUG step
# ----------------------------------------------------------------------------


def _rhs(F, U0, V0):
    fK = lambda K: F(K @ V0.T) @ V0  # noqa: E731
    fL = lambda L: F(U0 @ L.T).T @ U0  # noqa: E731
    return fK, fL


def bug_de_step(F: ScaledQuadratic, U0, S0, V0, dt, N):
    """The forward step (positive-diagonal QR)."""
    h = dt / N
    fK, fL = _rhs(F, U0, V0)
    U1, _ = qr_pos(rk4_states(fK, U0 @ S0, h, N)[-1])
    V1, _ = qr_pos(rk4_states(fL, V0 @ S0.T, h, N)[-1])
    M, Nn = U1.T @ U0, V1.T @ V0
    fS = lambda S: U1.T @ F(U1 @ S @ V1.T) @ V1  # noqa: E731
    return U1, rk4_states(fS, M @ S0 @ Nn.T, h, N)[-1], V1


def bug_de_step_vjp(F: ScaledQuadratic, U0, S0, V0, dt, N, Ub1, Sb1, Vb1):
    """Cotangents (Ub0, Sb0, Vb0, abar, bbar) of the step's inputs and of the scales
    a, b of F, given the cotangents of (U1, S1, V1).  Recomputes the forward step."""
    h = dt / N
    fK, fL = _rhs(F, U0, V0)
    Ks = rk4_states(fK, U0 @ S0, h, N)
    Ls = rk4_states(fL, V0 @ S0.T, h, N)
    U1, RK = qr_pos(Ks[-1])
    V1, RL = qr_pos(Ls[-1])
    M, Nn = U1.T @ U0, V1.T @ V0
    fS = lambda S: U1.T @ F(U1 @ S @ V1.T) @ V1  # noqa: E731
    Ss = rk4_states(fS, M @ S0 @ Nn.T, h, N)
    acc = {"U1": Ub1.copy(), "V1": Vb1.copy(), "U0": np.zeros_like(U0), "V0": np.zeros_like(V0), "a": 0.0,
           "b": 0.0}

    def vjp_S(S, sb):
        Y = U1 @ S @ V1.T
        Gb, da, db = F.vjp(Y, U1 @ sb @ V1.T)
        FY = F(Y)
        acc["U1"] += FY @ V1 @ sb.T + Gb @ V1 @ S.T
        acc["V1"] += FY.T @ U1 @ sb + Gb.T @ U1 @ S
        acc["a"] += da
        acc["b"] += db
        return U1.T @ Gb @ V1

    def vjp_K(K, kb):
        Y = K @ V0.T
        Gb, da, db = F.vjp(Y, kb @ V0.T)
        acc["V0"] += F(Y).T @ kb + Gb.T @ K
        acc["a"] += da
        acc["b"] += db
        return Gb @ V0

    def vjp_L(L, lb):
        Y = U0 @ L.T
        Gb, da, db = F.vjp(Y, U0 @ lb.T)
        acc["U0"] += F(Y) @ lb + Gb @ L
        acc["a"] += da
        acc["b"] += db
        return Gb.T @ U0

    Sib = rk4_vjp(fS, vjp_S, Ss, h, Sb1)
    Mb = Sib @ Nn @ S0.T
    Nb = Sib.T @ M @ S0
    Sb0 = M.T @ Sib @ Nn
    Kb = qr_vjp(U1, RK, acc["U1"] + U0 @ Mb.T)
    Lb = qr_vjp(V1, RL, acc["V1"] + V0 @ Nb.T)
    K0b = rk4_vjp(fK, vjp_K, Ks, h, Kb)
    L0b = rk4_vjp(fL, vjp_L, Ls, h, Lb)
    Ub0 = U1 @ Mb + K0b @ S0.T + acc["U0"]
    Vb0 = V1 @ Nb + L0b @ S0 + acc["V0"]
    Sb0 = Sb0 + U0.T @ K0b + L0b.T @ V0
    return Ub0, Sb0, Vb0, acc["a"], acc["b"]


def ra_bug_de_step(F: ScaledQuadratic, U0, S0, V0, dt, N, tol, rmax):
    """The rank-adaptive matrix-ODE step (rank_adaptive_unconventional.jl:194-233 with
    the sub-problems of :59-86) in the form its reverse differentiates: the augmented
    bases Uhat = [U0 | Q((I - U0 U0') K1)], Vhat likewise (the same spans as the
    reference's Q([K1 U0]), so U1 S1 V1' is the same), the S sub-step from
    M S0 N' (M = Uhat'U0, N = Vhat'V0), then the SVD truncated to
    min(rmax, the tolerance rank).  Returns (U1, S1, V1, r1)."""
    from .adjoint import _complete
    from .lra import truncate_to_tolerance

    h = dt / N
    fK, fL = _rhs(F, U0, V0)
    K1 = rk4_states(fK, U0 @ S0, h, N)[-1]
    L1 = rk4_states(fL, V0 @ S0.T, h, N)[-1]
    Uh = np.hstack([U0, _complete(U0, K1)[0]])
    Vh = np.hstack([V0, _complete(V0, L1)[0]])
    M, Nn = Uh.T @ U0, Vh.T @ V0
    fS = lambda S: Uh.T @ F(Uh @ S @ Vh.T) @ Vh  # noqa: E731
    Us, sig, Vst = np.linalg.svd(rk4_states(fS, M @ S0 @ Nn.T, h, N)[-1])
    r1 = int(min(rmax, truncate_to_tolerance(sig, tol)))
    return Uh @ Us[:, :r1], np.diag(sig[:r1]), Vh @ Vst[:r1].T, r1


def ra_bug_de_step_vjp(F: ScaledQuadratic, U0, S0, V0, dt, N, tol, rmax, Ub1, Sb1, Vb1):
    """Cotangents (Ub0, Sb0, Vb0, abar, bbar) of ``ra_bug_de_step``'s inputs and of the
    scales a, b of F, given those of its result (U1 n x r1, S1 = diag(sigma[:r1]),
    V1 m x r1; only the diagonal of Sb1 enters), the rank r1 held fixed; also r1.
    The reverse of the data step's truncation and completions (``adjoint.ra_bug_step_vjp``)
    around the RK4 reverses of ``bug_de_step_vjp``, with the S sub-step in 2r:
      Uhbar = Ub1 Us[:, :r1]' + (S sub-step cotangent of Uhat) + U0 Mb',
      Sibar from the SVD rule through the S reverse, Mb = Sibar N S0', Nb = Sibar' M S0,
      Ub0 = Uhbar[:, :r] + Uhat Mb + (completion reverse) + Kbar0 S0' + (L sub-step cotangent of U0),
      Sb0 = M' Sibar N + U0'Kbar0 + Lbar0'V0."""
    from .adjoint import _complete, svd_vjp
    from .lra import truncate_to_tolerance

    h = dt / N
    r = U0.shape[1]
    fK, fL = _rhs(F, U0, V0)
    Ks = rk4_states(fK, U0 @ S0, h, N)
    Ls = rk4_states(fL, V0 @ S0.T, h, N)
    K1, L1 = Ks[-1], Ls[-1]
    Qu, Ru, _ = _complete(U0, K1)
    Qv, Rv, _ = _complete(V0, L1)
    Uh = np.hstack([U0, Qu])
    Vh = np.hstack([V0, Qv])
    M, Nn = Uh.T @ U0, Vh.T @ V0
    fS = lambda S: Uh.T @ F(Uh @ S @ Vh.T) @ Vh  # noqa: E731
    Ss = rk4_states(fS, M @ S0 @ Nn.T, h, N)
    Us, sig, Vst = np.linalg.svd(Ss[-1])
    Vs = Vst.T
    r1 = int(min(rmax, truncate_to_tolerance(sig, tol)))
    assert Ub1.shape[1] == r1, (Ub1.shape, r1)
    acc = {"Uh": Ub1 @ Us[:, :r1].T, "Vh": Vb1 @ Vs[:, :r1].T, "U0": np.zeros_like(U0), "V0": np.zeros_like(V0),
           "a": 0.0, "b": 0.0}
    Shb = svd_vjp(Us, sig, Vs, Uh.T @ Ub1, np.diag(Sb1).copy(), Vh.T @ Vb1)

    def vjp_S(S, sb):
        Y = Uh @ S @ Vh.T
        Gb, da, db = F.vjp(Y, Uh @ sb @ Vh.T)
        FY = F(Y)
        acc["Uh"] += FY @ Vh @ sb.T + Gb @ Vh @ S.T
        acc["Vh"] += FY.T @ Uh @ sb + Gb.T @ Uh @ S
        acc["a"] += da
        acc["b"] += db
        return Uh.T @ Gb @ Vh

    def vjp_K(K, kb):
        Y = K @ V0.T
        Gb, da, db = F.vjp(Y, kb @ V0.T)
        acc["V0"] += F(Y).T @ kb + Gb.T @ K
        acc["a"] += da
        acc["b"] += db
        return Gb @ V0

    def vjp_L(L, lb):
        Y = U0 @ L.T
        Gb, da, db = F.vjp(Y, U0 @ lb.T)
        acc["U0"] += F(Y) @ lb + Gb @ L
        acc["a"] += da
        acc["b"] += db
        return Gb.T @ U0

    Sib = rk4_vjp(fS, vjp_S, Ss, h, Shb)
    Mb = Sib @ Nn @ S0.T
    Nb = Sib.T @ M @ S0
    Sb0 = M.T @ Sib @ Nn
    Uhb = acc["Uh"] + U0 @ Mb.T
    Vhb = acc["Vh"] + V0 @ Nb.T
    Ub0 = Uhb[:, :r] + Uh @ Mb
    Vb0 = Vhb[:, :r] + Vh @ Nb
    # Q = Q((I - B B') X1): Zbar by the thin-QR rule, then Z = X1 - B (B'X1)
    Xb = []
    for B, X1, Qc, Rc, Qb in ((U0, K1, Qu, Ru, Uhb[:, r:]), (V0, L1, Qv, Rv, Vhb[:, r:])):
        Zb = qr_vjp(Qc, Rc, Qb)
        Xb.append(Zb - B @ (B.T @ Zb))
        Bb = -Zb @ (X1.T @ B) - X1 @ (Zb.T @ B)
        if B is U0:
            Ub0 = Ub0 + Bb
        else:
            Vb0 = Vb0 + Bb
    K0b = rk4_vjp(fK, vjp_K, Ks, h, Xb[0])
    L0b = rk4_vjp(fL, vjp_L, Ls, h, Xb[1])
    Ub0 = Ub0 + K0b @ S0.T + acc["U0"]
    Vb0 = Vb0 + L0b @ S0 + acc["V0"]
    Sb0 = Sb0 + U0.T @ K0b + L0b.T @ V0
    return Ub0, Sb0, Vb0, acc["a"], acc["b"], r1


def bug_de_step_jvp(F: ScaledQuadratic, U0, S0, V0, dt, N, Ud, Sd, Vd, ad=0.0, bd=0.0):
    """Forward-mode tangents (U1d, S1d, V1d) of the step along (Ud, Sd, Vd, ad, bd)."""
    h = dt / N
    fK, fL = _rhs(F, U0, V0)

    def jK(K, Kd):
        Y = K @ V0.T
        return F.jvp(Y, Kd @ V0.T + K @ Vd.T, ad, bd) @ V0 + F(Y) @ Vd

    def jL(L, Ld):
        Y = U0 @ L.T
        return F.jvp(Y, Ud @ L.T + U0 @ Ld.T, ad, bd).T @ U0 + F(Y).T @ Ud

    K, Kd = rk4_jvp(fK, jK, U0 @ S0, Ud @ S0 + U0 @ Sd, h, N)
    L, Ld = rk4_jvp(fL, jL, V0 @ S0.T, Vd @ S0.T + V0 @ Sd.T, h, N)
    U1, RK = qr_pos(K)
    V1, RL = qr_pos(L)
    U1d, _ = qr_jvp(U1, RK, Kd)
    V1d, _ = qr_jvp(V1, RL, Ld)
    M, Nn = U1.T @ U0, V1.T @ V0
    Md = U1d.T @ U0 + U1.T @ Ud
    Nd = V1d.T @ V0 + V1.T @ Vd
    fS = lambda S: U1.T @ F(U1 @ S @ V1.T) @ V1  # noqa: E731

    def jS(S, Sd_):
        Y = U1 @ S @ V1.T
        Yd = U1d @ S @ V1.T + U1 @ Sd_ @ V1.T + U1 @ S @ V1d.T
        FY = F(Y)
        return U1d.T @ FY @ V1 + U1.T @ F.jvp(Y, Yd, ad, bd) @ V1 + U1.T @ FY @ V1d

    S0s = M @ S0 @ Nn.T
    S0sd = Md @ S0 @ Nn.T + M @ Sd @ Nn.T + M @ S0 @ Nd.T
    _, S1d = rk4_jvp(fS, jS, S0s, S0sd, h, N)
    return U1d, S1d, V1d


def ksl_primal_de_step(F: ScaledQuadratic, U0, S0, V0, dt, N):
    """primal_LT_step! with the DE sub-problems (projector_splitting.jl:52-80,129-152),
    positive-diagonal QR:  K -> (U1, R1);  S~' = -U1'F(U1 S V0')V0 from R1;
    L -> (V1, R2) from V0 S~';  S1 = R2'."""
    h = dt / N
    fK = lambda K: F(K @ V0.T) @ V0  # noqa: E731
    U1, R1 = qr_pos(rk4_states(fK, U0 @ S0, h, N)[-1])
    fS = lambda S: -(U1.T @ F(U1 @ S @ V0.T) @ V0)  # noqa: E731
    St = rk4_states(fS, R1, h, N)[-1]
    fL = lambda L: F(U1 @ L.T).T @ U1  # noqa: E731
    V1, R2 = qr_pos(rk4_states(fL, V0 @ St.T, h, N)[-1])
    return U1, R2.T.copy(), V1


def ksl_primal_de_step_vjp(F: ScaledQuadratic, U0, S0, V0, dt, N, Ub1, Sb1, Vb1):
    """Cotangents (Ub0, Sb0, Vb0, abar, bbar) of the primal KSL matrix-ODE step; both QR
    factors carry cotangents: Kbar = (Qbar + Q copyltu(R Rbar' - Qbar'Q)) R^{-T}."""
    from .adjoint import qr_vjp_full

    h = dt / N
    fK = lambda K: F(K @ V0.T) @ V0  # noqa: E731
    Ks = rk4_states(fK, U0 @ S0, h, N)
    U1, R1 = qr_pos(Ks[-1])
    fS = lambda S: -(U1.T @ F(U1 @ S @ V0.T) @ V0)  # noqa: E731
    Ss = rk4_states(fS, R1, h, N)
    St = Ss[-1]
    fL = lambda L: F(U1 @ L.T).T @ U1  # noqa: E731
    Ls = rk4_states(fL, V0 @ St.T, h, N)
    V1, R2 = qr_pos(Ls[-1])
    acc = {"U1": Ub1.copy(), "V0": np.zeros_like(V0), "a": 0.0, "b": 0.0}

    def vjp_L(L, lb):
        Y = U1 @ L.T
        Gb, da, db = F.vjp(Y, U1 @ lb.T)
        acc["U1"] += F(Y) @ lb + Gb @ L
        acc["a"] += da
        acc["b"] += db
        return Gb.T @ U1

    def vjp_S(S, sb):  # f = -U1' F(U1 S V0') V0
        Y = U1 @ S @ V0.T
        Gb, da, db = F.vjp(Y, -(U1 @ sb @ V0.T))
        FY = F(Y)
        acc["U1"] += -(FY @ V0 @ sb.T) + Gb @ V0 @ S.T
        acc["V0"] += -(FY.T @ U1 @ sb) + Gb.T @ U1 @ S
        acc["a"] += da
        acc["b"] += db
        return U1.T @ Gb @ V0

    def vjp_K(K, kb):
        Y = K @ V0.T
        Gb, da, db = F.vjp(Y, kb @ V0.T)
        acc["V0"] += F(Y).T @ kb + Gb.T @ K
        acc["a"] += da
        acc["b"] += db
        return Gb @ V0

    Lb = qr_vjp_full(V1, R2, Vb1, Sb1.T)           # S1 = R2'
    L0b = rk4_vjp(fL, vjp_L, Ls, h, Lb)
    Vb0 = L0b @ St                                  # L(0) = V0 St'
    Stb = L0b.T @ V0
    R1b = rk4_vjp(fS, vjp_S, Ss, h, Stb)
    Kb = qr_vjp_full(U1, R1, acc["U1"], R1b)
    K0b = rk4_vjp(fK, vjp_K, Ks, h, Kb)
    return K0b @ S0.T, U0.T @ K0b, Vb0 + acc["V0"], acc["a"], acc["b"]


def ksl_primal_de_step_jvp(F: ScaledQuadratic, U0, S0, V0, dt, N, Ud, Sd, Vd, ad=0.0, bd=0.0):
    """Forward-mode tangents (U1d, S1d, V1d) of the primal KSL matrix-ODE step."""
    h = dt / N
    fK = lambda K: F(K @ V0.T) @ V0  # noqa: E731

    def jK(K, Kd):
        Y = K @ V0.T
        return F.jvp(Y, Kd @ V0.T + K @ Vd.T, ad, bd) @ V0 + F(Y) @ Vd

    K, Kd = rk4_jvp(fK, jK, U0 @ S0, Ud @ S0 + U0 @ Sd, h, N)
    U1, R1 = qr_pos(K)
    U1d, R1d = qr_jvp(U1, R1, Kd)
    fS = lambda S: -(U1.T @ F(U1 @ S @ V0.T) @ V0)  # noqa: E731

    def jS(S, Sd_):
        Y = U1 @ S @ V0.T
        Yd = U1d @ S @ V0.T + U1 @ Sd_ @ V0.T + U1 @ S @ Vd.T
        FY = F(Y)
        return -(U1d.T @ FY @ V0 + U1.T @ F.jvp(Y, Yd, ad, bd) @ V0 + U1.T @ FY @ Vd)

    St, Std = rk4_jvp(fS, jS, R1, R1d, h, N)
    fL = lambda L: F(U1 @ L.T).T @ U1  # noqa: E731

    def jL(L, Ld):
        Y = U1 @ L.T
        return F.jvp(Y, U1d @ L.T + U1 @ Ld.T, ad, bd).T @ U1 + F(Y).T @ U1d

    L, Ld = rk4_jvp(fL, jL, V0 @ St.T, Vd @ St.T + V0 @ Std.T, h, N)
    V1, R2 = qr_pos(L)
    V1d, R2d = qr_jvp(V1, R2, Ld)
    return U1d, R2d.T.copy(), V1d


def bug_de_solve(F, U, S, V, dt, N, steps, ksl=False):
    step = ksl_primal_de_step if ksl else bug_de_step
    for _ in range(steps):
        U, S, V = step(F, U, S, V, dt, N)
    return U, S, V


def bug_de_solve_gradient(F: ScaledQuadratic, U0, S0, V0, dt, N, steps, loss_grad, checkpoint=20, ksl=False,
                          ra=None):
    """Gradient of loss(U_T, S_T, V_T) after `steps` BUG (ksl: primal KSL; ra = (tol, rmax):
    rank-adaptive BUG, the ranks of the forward solve held) steps with respect to
    (U0, S0, V0) and the scales (a, b), with checkpoints every `checkpoint` steps and
    recomputation inside each segment.  Returns (value, Ub0, Sb0, Vb0, abar, bbar)."""
    if ra is not None:
        tol, rmax = ra
        bug_de_step_ = lambda F_, U, S, V, dt_, N_: ra_bug_de_step(F_, U, S, V, dt_, N_, tol, rmax)[:3]  # noqa: E731
        bug_de_step_vjp_ = lambda F_, U, S, V, dt_, N_, *c: ra_bug_de_step_vjp(  # noqa: E731
            F_, U, S, V, dt_, N_, tol, rmax, *c)[:5]
    else:
        bug_de_step_ = ksl_primal_de_step if ksl else bug_de_step
        bug_de_step_vjp_ = ksl_primal_de_step_vjp if ksl else bug_de_step_vjp
    cps = {0: (U0, S0, V0)}
    U, S, V = U0, S0, V0
    for k in range(steps):
        U, S, V = bug_de_step_(F, U, S, V, dt, N)
        if (k + 1) % checkpoint == 0 and k + 1 < steps:
            cps[k + 1] = (U, S, V)
    value, Ub, Sb, Vb = loss_grad(U, S, V)
    abar = bbar = 0.0
    for c0 in sorted(cps, reverse=True):
        c1 = min(c0 + checkpoint, steps)
        states = [cps[c0]]
        for _ in range(c0, c1 - 1):
            states.append(bug_de_step_(F, *states[-1], dt, N))
        for k in range(c1 - 1, c0 - 1, -1):
            Ub, Sb, Vb, da, db = bug_de_step_vjp_(F, *states[k - c0], dt, N, Ub, Sb, Vb)
            abar += da
            bbar += db
    return value, Ub, Sb, Vb, abar, bbar


# ----------------------------------------------------------------------------
# the parametric Burgers problem of the differentiation workload
# ----------------------------------------------------------------------------


def periodic_burgers(n: int, nu: float):
    """nu * Laplacian and the central gradient on n points x_i = i / n of [0, 1), periodic
    (dense)."""
    dx = 1.0 / n
    I = np.eye(n)
    right = np.roll(I, 1, axis=1)  # (right @ y)_i = y_{i+1}
    left = right.T
    return nu / dx ** 2 * (left - 2 * I + right), (right - left) / (2 * dx)


def sine_state(n: int, xi: np.ndarray, r: int, omega: np.ndarray):
    """Y0 = sum_k xi_k sin(2 pi k x) / k (k = 1..K, one column per sample) as
    U0 S0 V0' with U0 the normalised sines k = 1..r, V0 = Q([xi' | omega]) (positive
    diagonal), S0 = U0' Y0 V0.  Computed densely (the restatement of the setup)."""
    K, m = xi.shape
    x = np.arange(n) / n
    Phi = np.stack([np.sin(2 * np.pi * k * x) / k for k in range(1, K + 1)], axis=1)
    U0 = np.stack([np.sin(2 * np.pi * k * x) for k in range(1, r + 1)], axis=1) * np.sqrt(2.0 / n)
    V0, _ = qr_pos(np.hstack([xi.T, omega]))
    Y0 = Phi @ xi
    return U0, U0.T @ Y0 @ V0, V0, Y0


def mean_column_loss(U, S, V):
    """L = || mean over columns of U S V' ||^2 and its cotangents."""
    m = V.shape[0]
    w = V.sum(axis=0) / m
    y = S @ w
    z = U @ y
    Uz = U.T @ z
    return float(z @ z), 2.0 * np.outer(z, y), 2.0 * np.outer(Uz, w), np.outer(np.ones(m), 2.0 * S.T @ Uz) / m
