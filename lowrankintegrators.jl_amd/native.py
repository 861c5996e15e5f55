"""ctypes binding of liblri_hip.so (declarations in include/lri.h).

The library is loaded from this package directory; there is no fallback: if it
is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblri_hip.so")

# status codes (include/lri.h)
OK, EINVAL, EHIP, EOOM, EBREAKDOWN, ENONFINITE, ENOTSUP = 0, -1, -2, -3, -4, -5, -6
F64, F32, BF16 = 0, 1, 2
PREC_F64, PREC_MIXED = 0, 1
PRIMAL, DUAL, STRANG = 0, 1, 2
KERNEL_PASS_N, KERNEL_PASS_T, KERNEL_SMALL, KERNEL_DE = 0, 1, 2, 3
MAX_RANK = 64
MAX_FORCE = 4
DE_MAX_ROWS = 16384
QUAD_MAX_RANK = 32

DTYPES = {"f64": F64, "f32": F32, "bf16": BF16}

_i64 = ctypes.c_int64
_int = ctypes.c_int
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "lri_version": (_int, []),
    "lri_status_string": (ctypes.c_char_p, [_int]),
    "lri_last_error": (ctypes.c_char_p, [_vp]),
    "lri_create": (_int, [ctypes.POINTER(_vp), _int, _i64, _i64, _i64]),
    "lri_destroy": (_int, [_vp]),
    "lri_set_precision": (_int, [_vp, _int]),
    "lri_set_state": (_int, [_vp, _i64, _dp, _i64, _dp, _i64, _dp, _i64]),
    "lri_get_state": (_int, [_vp, ctypes.POINTER(_i64), _dp, _i64, _dp, _i64, _dp, _i64]),
    "lri_get_rank": (_int, [_vp, ctypes.POINTER(_i64)]),
    "lri_snapshots_reserve": (_int, [_vp, _i64, _int]),
    "lri_snapshot_upload": (_int, [_vp, _i64, _vp, _i64]),
    "lri_snapshots_bind": (_int, [_vp, ctypes.POINTER(_vp), _i64, _int, _i64]),
    "lri_step_ksl": (_int, [_vp, _int, _i64, _i64]),
    "lri_step_ksl_strang": (_int, [_vp, _i64, _i64, _i64]),
    "lri_step_bug": (_int, [_vp, _i64, _i64]),
    "lri_step_bug_ra": (_int, [_vp, _i64, _i64, ctypes.c_double, _i64, ctypes.POINTER(_i64)]),
    "lri_step_bug_adjoint": (_int, [_vp, _i64, _i64, _dp, _i64, _dp, _i64, _dp, _i64]),
    "lri_step_ksl_adjoint": (_int, [_vp, _i64, _i64, _dp, _i64, _dp, _i64, _dp, _i64]),
    "lri_step_ksl_dual_adjoint": (_int, [_vp, _i64, _i64, _dp, _i64, _dp, _i64, _dp, _i64]),
    "lri_step_ksl_strang_adjoint": (_int, [_vp, _i64, _i64, _i64, _dp, _i64, _dp, _i64, _dp, _i64]),
    "lri_step_bug_ra_adjoint": (_int, [_vp, _i64, _i64, ctypes.c_double, _i64, _i64, _dp, _i64, _dp, _i64, _dp,
                                       _i64]),
    "lri_step_greedy": (_int, [_vp, _i64]),
    "lri_step_greedy_twofactor": (_int, [_vp, _i64]),
    "lri_step_hybrid_greedy": (_int, [_vp, _i64, ctypes.c_double, _i64]),
    "lri_set_operator_lyapunov": (_int, [_vp, _dp, _dp, _int, _dp, _dp, _int, _i64, _dp, _i64, _dp, _i64]),
    "lri_set_operator_quadratic": (_int, [_vp, _dp, _dp, _dp, _int]),
    "lri_set_resident_eigenbasis": (_int, [_vp, _int]),
    "lri_step_de_ksl": (_int, [_vp, _int, ctypes.c_double, _i64]),
    "lri_step_de_bug": (_int, [_vp, ctypes.c_double, _i64]),
    "lri_step_de_bug_ra": (_int, [_vp, ctypes.c_double, _i64, ctypes.c_double, _i64, ctypes.POINTER(_i64)]),
    "lri_step_de_bug_adjoint": (_int, [_vp, ctypes.c_double, _i64, _dp, _i64, _dp, _i64, _dp, _i64, _dp]),
    "lri_step_de_bug_ra_adjoint": (_int, [_vp, ctypes.c_double, _i64, ctypes.c_double, _i64, _i64, _dp, _i64, _dp,
                                          _i64, _dp, _i64, _dp]),
    "lri_step_de_ksl_adjoint": (_int, [_vp, ctypes.c_double, _i64, _dp, _i64, _dp, _i64, _dp, _i64, _dp]),
    "lri_batch_create": (_int, [ctypes.POINTER(_vp), _int, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                ctypes.POINTER(_i64)]),
    "lri_batch_destroy": (_int, [_vp]),
    "lri_batch_last_error": (ctypes.c_char_p, [_vp]),
    "lri_batch_context": (_int, [_vp, _i64, ctypes.POINTER(_vp)]),
    "lri_batch_step_de_bug": (_int, [_vp, ctypes.c_double, ctypes.POINTER(_i64)]),
    "lri_batch_step_de_bug_ra": (_int, [_vp, ctypes.c_double, ctypes.POINTER(_i64), ctypes.c_double, _i64,
                                        ctypes.POINTER(_i64)]),
    "lri_batch_use_graphs": (_int, [_vp, _int]),
    "lri_batch_use_merged_launches": (_int, [_vp, _int]),
    "lri_batch_launch_stats": (_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "lri_batch_member_status": (_int, [_vp, _i64, ctypes.POINTER(_int)]),
    "lri_batch_synchronize": (_int, [_vp]),
    "lri_synchronize": (_int, [_vp]),
    "lri_trajectory_reserve": (_int, [_vp, _i64]),
    "lri_trajectory_save": (_int, [_vp, _i64]),
    "lri_trajectory_get": (_int, [_vp, _i64, ctypes.POINTER(_i64), _dp, _i64, _dp, _i64, _dp, _i64]),
    "lri_trajectory_get_range": (_int, [_vp, _i64, _i64, ctypes.POINTER(_i64), _dp]),
    "lri_get_stream": (_int, [_vp, ctypes.POINTER(_vp)]),
    "lri_kernel_timing": (_int, [_vp, _int]),
    "lri_kernel_timing_read": (_int, [_vp, _int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i64)]),
    "lri_kernel_timing_sample": (_int, [_vp, _int]),
    "lri_kernel_timing_samples": (_int, [_vp, _int, ctypes.POINTER(ctypes.c_float), _i64, ctypes.POINTER(_i64)]),
    "lri_autotune_passes": (_int, [_vp, _i64, _i64, _i64, ctypes.POINTER(ctypes.c_double)]),
    "lri_selftest_mfma": (_int, [_int, _vp, _vp, _vp]),
    "lri_stream_probe": (_int, [_int, _dp]),
    "lri_debug_product": (_int, [_vp, _int, _i64, _i64, _dp, _i64]),
    "lri_product_noalloc": (_int, [_int, _int, _int, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _vp]),
    "lri_noalloc_last_error": (ctypes.c_char_p, []),
}

_LIB = None


class LRIError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class BreakdownError(LRIError):
    pass


class NonFiniteError(LRIError):
    pass


class NotSupportedError(LRIError):
    """LRI_ENOTSUP: e.g. a rank-adaptive step that needs a rank above the context capacity."""


def lib():
    """Load (once) and return the library with typed entry points."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(status, ctx=None):
    if status == OK:
        return
    L = lib()
    msg = L.lri_last_error(ctx).decode(errors="replace") if L.lri_last_error(ctx) else ""
    if not msg:
        msg = L.lri_status_string(status).decode()
    if status == EBREAKDOWN:
        raise BreakdownError(status, msg)
    if status == ENONFINITE:
        raise NonFiniteError(status, msg)
    if status == ENOTSUP:
        raise NotSupportedError(status, msg)
    if status == EINVAL:
        raise ValueError(msg)
    raise LRIError(status, msg)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class Context:
    """Owning wrapper of an lri_ctx (one problem size on one device)."""

    def __init__(self, n: int, m: int, rcap: int, device: int = 0):
        L = lib()
        h = _vp()
        st = L.lri_create(ctypes.byref(h), int(device), int(n), int(m), int(rcap))
        check(st, None)
        self.h = h
        self.n, self.m, self.rcap = int(n), int(m), int(rcap)
        self._bound = None  # keeps borrowed snapshot owners alive

    def close(self):
        if getattr(self, "h", None):
            lib().lri_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state
    def set_state(self, U, S, V):
        U = np.asfortranarray(U, dtype=np.float64)
        S = np.asfortranarray(S, dtype=np.float64)
        V = np.asfortranarray(V, dtype=np.float64)
        r = S.shape[0]
        if U.shape != (self.n, r) or V.shape != (self.m, r) or S.shape != (r, r):
            raise ValueError(f"shape mismatch: U {U.shape}, S {S.shape}, V {V.shape} for n={self.n}, m={self.m}")
        check(lib().lri_set_state(self.h, r, _dptr(U), self.n, _dptr(S), r, _dptr(V), self.m), self.h)

    def rank(self) -> int:
        r = _i64()
        check(lib().lri_get_rank(self.h, ctypes.byref(r)), self.h)
        return int(r.value)

    def get_state(self):
        r = _i64()
        check(lib().lri_get_rank(self.h, ctypes.byref(r)), self.h)
        r = int(r.value)
        U = np.zeros((self.n, r), order="F")
        S = np.zeros((r, r), order="F")
        V = np.zeros((self.m, r), order="F")
        r2 = _i64()
        check(lib().lri_get_state(self.h, ctypes.byref(r2), _dptr(U), self.n, _dptr(S), r, _dptr(V), self.m), self.h)
        return U, S, V

    def set_precision(self, prec: int):
        check(lib().lri_set_precision(self.h, int(prec)), self.h)

    # -- data
    def reserve(self, count: int, dtype: int):
        check(lib().lri_snapshots_reserve(self.h, int(count), int(dtype)), self.h)
        self.dtype = dtype

    def upload(self, k: int, host: np.ndarray):
        """Upload a host snapshot (already encoded in the reserved dtype)."""
        a = np.asfortranarray(host)
        check(lib().lri_snapshot_upload(self.h, int(k), a.ctypes.data_as(_vp), a.shape[0]), self.h)

    def bind(self, ptrs, dtype: int, ld: int, owners=None):
        arr = (_vp * len(ptrs))(*[_vp(int(p)) for p in ptrs])
        check(lib().lri_snapshots_bind(self.h, arr, len(ptrs), int(dtype), int(ld)), self.h)
        self._bound = owners
        self.dtype = dtype

    # -- steps
    def step_ksl(self, order: int, kp: int, kn: int):
        check(lib().lri_step_ksl(self.h, int(order), int(kp), int(kn)), self.h)

    def step_ksl_strang(self, kp: int, km: int, kn: int):
        check(lib().lri_step_ksl_strang(self.h, int(kp), int(km), int(kn)), self.h)

    def step_bug(self, kp: int, kn: int):
        check(lib().lri_step_bug(self.h, int(kp), int(kn)), self.h)

    def step_bug_ra(self, kp: int, kn: int, tol: float, rmax: int) -> int:
        r = _i64()
        check(lib().lri_step_bug_ra(self.h, int(kp), int(kn), float(tol), int(rmax), ctypes.byref(r)), self.h)
        return int(r.value)

    def step_bug_adjoint(self, kp: int, kn: int, Ub, Sb, Vb):
        """Cotangents of the BUG step's inputs from those of its result (host arrays)."""
        return self._adjoint(lib().lri_step_bug_adjoint, kp, kn, Ub, Sb, Vb)

    def step_ksl_adjoint(self, kp: int, kn: int, Ub, Sb, Vb):
        """Cotangents of the primal KSL step's inputs from those of its result (host arrays)."""
        return self._adjoint(lib().lri_step_ksl_adjoint, kp, kn, Ub, Sb, Vb)

    def step_ksl_dual_adjoint(self, kp: int, kn: int, Ub, Sb, Vb):
        """Cotangents of the dual KSL step's inputs from those of its result (host arrays)."""
        return self._adjoint(lib().lri_step_ksl_dual_adjoint, kp, kn, Ub, Sb, Vb)

    def step_ksl_strang_adjoint(self, kp: int, km: int, kn: int, Ub, Sb, Vb):
        """Cotangents of the Strang KSL step's inputs (A[kp] -> A[km] -> A[kn]) from
        those of its result (host arrays)."""
        fn = lib().lri_step_ksl_strang_adjoint
        return self._adjoint(lambda h, a, b, *x: fn(h, a, int(km), b, *x), kp, kn, Ub, Sb, Vb)

    def step_adjoint_ptr(self, kind: str, kp: int, kn: int, ub: int, ldub: int, sb: int, ldsb: int, vb: int,
                         ldvb: int):
        """The BUG (kind "bug"), primal KSL ("ksl") or dual KSL ("ksl_dual") reverse step on cotangents at raw
        addresses -- device memory (e.g. torch tensors' data_ptr()) or host -- column-major
        with the given leading dimensions, overwritten in place."""
        fn = {"bug": lib().lri_step_bug_adjoint, "ksl": lib().lri_step_ksl_adjoint,
              "ksl_dual": lib().lri_step_ksl_dual_adjoint}[kind]
        P = ctypes.POINTER(ctypes.c_double)
        check(fn(self.h, int(kp), int(kn), ctypes.cast(ub, P), int(ldub), ctypes.cast(sb, P), int(ldsb),
                 ctypes.cast(vb, P), int(ldvb)), self.h)

    def step_bug_ra_adjoint(self, kp: int, kn: int, tol: float, rmax: int, Ub, Sb, Vb):
        """Cotangents of the rank-adaptive BUG step's inputs (rank r) from those of its
        result (rank r1 = Ub.shape[1], the rank the forward step returned)."""
        n, r1 = Ub.shape
        m = Vb.shape[0]
        r = self.rank()
        w = max(r, r1)
        U = np.zeros((n, w), order="F")
        V = np.zeros((m, w), order="F")
        S = np.zeros((w, w), order="F")
        U[:, :r1] = Ub
        V[:, :r1] = Vb
        S[:r1, :r1] = Sb
        check(lib().lri_step_bug_ra_adjoint(self.h, int(kp), int(kn), float(tol), int(rmax), int(r1), _dptr(U), n,
                                            _dptr(S), w, _dptr(V), m), self.h)
        return U[:, :r].copy(), S[:r, :r].copy(), V[:, :r].copy()

    def _adjoint(self, fn, kp, kn, Ub, Sb, Vb):
        Ub = np.asfortranarray(Ub, dtype=np.float64).copy(order="F")
        Sb = np.asfortranarray(Sb, dtype=np.float64).copy(order="F")
        Vb = np.asfortranarray(Vb, dtype=np.float64).copy(order="F")
        check(fn(self.h, int(kp), int(kn), _dptr(Ub), Ub.shape[0], _dptr(Sb), Sb.shape[0], _dptr(Vb), Vb.shape[0]),
              self.h)
        return Ub, Sb, Vb

    def step_greedy(self, kn: int):
        check(lib().lri_step_greedy(self.h, int(kn)), self.h)

    def step_greedy_twofactor(self, kn: int):
        check(lib().lri_step_greedy_twofactor(self.h, int(kn)), self.h)

    def set_resident_eigenbasis(self, on: bool):
        """Keep the state in the operator's analytic eigenbasis (constant coefficients, linear)."""
        check(lib().lri_set_resident_eigenbasis(self.h, int(bool(on))), self.h)
        self.resident = bool(on)

    def step_hybrid_greedy(self, kn: int, dt: float, substeps: int):
        check(lib().lri_step_hybrid_greedy(self.h, int(kn), float(dt), int(substeps)), self.h)

    # -- matrix ODE
    def set_operator_lyapunov(self, a_diag, a_off, a_periodic, b_diag, b_off, b_periodic, G=None, H=None):
        """F(Y) = A Y + Y B' + G H' with A, B symmetric tridiagonal (diagonal, off-diagonal)."""
        def vec(x, n, name):
            x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
            if x.shape[0] != n:
                raise ValueError(f"{name} must have {n} entries, got {x.shape[0]}")
            return x

        ad = vec(a_diag, self.n, "a_diag")
        bd = vec(b_diag, self.m, "b_diag")
        # off-diagonals: n - 1 entries, or n when periodic (padded to n either way)
        ao = np.zeros(self.n)
        bo = np.zeros(self.m)
        a_off = np.asarray(a_off, dtype=np.float64).reshape(-1)
        b_off = np.asarray(b_off, dtype=np.float64).reshape(-1)
        na, nb = (self.n if a_periodic else self.n - 1), (self.m if b_periodic else self.m - 1)
        if a_off.shape[0] != na or b_off.shape[0] != nb:
            raise ValueError(f"off-diagonals need {na} and {nb} entries, got {a_off.shape[0]} and {b_off.shape[0]}")
        ao[:na] = a_off
        bo[:nb] = b_off
        if G is None or H is None:
            q, Gf, Hf = 0, np.zeros((self.n, 1), order="F"), np.zeros((self.m, 1), order="F")
        else:
            Gf = np.asfortranarray(np.asarray(G, dtype=np.float64).reshape(self.n, -1))
            Hf = np.asfortranarray(np.asarray(H, dtype=np.float64).reshape(self.m, -1))
            if Gf.shape[1] != Hf.shape[1]:
                raise ValueError("G and H need the same number of columns")
            q = Gf.shape[1]
        check(lib().lri_set_operator_lyapunov(self.h, _dptr(ad), _dptr(ao), int(bool(a_periodic)), _dptr(bd), _dptr(bo),
                                              int(bool(b_periodic)), int(q), _dptr(Gf), self.n, _dptr(Hf), self.m),
              self.h)

    def set_operator_quadratic(self, dl, dd, du, periodic):
        """Adds -(D Y) o Y, D = tridiag(dl, dd, du) on the n rows (n entries each)."""
        vs = []
        for x, name in ((dl, "dl"), (dd, "dd"), (du, "du")):
            x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
            if x.shape[0] != self.n:
                raise ValueError(f"{name} must have {self.n} entries, got {x.shape[0]}")
            vs.append(x)
        check(lib().lri_set_operator_quadratic(self.h, _dptr(vs[0]), _dptr(vs[1]), _dptr(vs[2]), int(bool(periodic))),
              self.h)

    def step_de_ksl(self, order: int, dt: float, substeps: int):
        check(lib().lri_step_de_ksl(self.h, int(order), float(dt), int(substeps)), self.h)

    def step_de_bug(self, dt: float, substeps: int):
        check(lib().lri_step_de_bug(self.h, float(dt), int(substeps)), self.h)

    def step_de_bug_ra(self, dt: float, substeps: int, tol: float, rmax: int) -> int:
        r = _i64()
        check(lib().lri_step_de_bug_ra(self.h, float(dt), int(substeps), float(tol), int(rmax), ctypes.byref(r)),
              self.h)
        return int(r.value)

    def step_de_bug_adjoint(self, dt: float, substeps: int, Ub, Sb, Vb):
        """Cotangents of the matrix-ODE BUG step's inputs from those of its result (host
        arrays), and the derivatives (a, b) along the scales of the operator's linear parts."""
        return self._de_adjoint(lib().lri_step_de_bug_adjoint, dt, substeps, Ub, Sb, Vb)

    def step_de_ksl_adjoint(self, dt: float, substeps: int, Ub, Sb, Vb):
        """The same for the primal KSL matrix-ODE step."""
        return self._de_adjoint(lib().lri_step_de_ksl_adjoint, dt, substeps, Ub, Sb, Vb)

    def step_de_bug_ra_adjoint(self, dt: float, substeps: int, tol: float, rmax: int, Ub, Sb, Vb):
        """Cotangents of the rank-adaptive matrix-ODE step's inputs (rank r) from those of its
        result (rank r1 = Ub.shape[1], the rank the forward step returned), and the
        derivatives (a, b) along the scales of the operator's linear parts."""
        n, r1 = Ub.shape
        m = Vb.shape[0]
        r = self.rank()
        w = max(r, r1)
        U = np.zeros((n, w), order="F")
        V = np.zeros((m, w), order="F")
        S = np.zeros((w, w), order="F")
        U[:, :r1] = Ub
        V[:, :r1] = Vb
        S[:r1, :r1] = Sb
        sc = np.zeros(2)
        check(lib().lri_step_de_bug_ra_adjoint(self.h, float(dt), int(substeps), float(tol), int(rmax), int(r1),
                                               _dptr(U), n, _dptr(S), w, _dptr(V), m, _dptr(sc)), self.h)
        return U[:, :r].copy(), S[:r, :r].copy(), V[:, :r].copy(), (float(sc[0]), float(sc[1]))

    def _de_adjoint(self, fn, dt, substeps, Ub, Sb, Vb):
        Ub = np.asfortranarray(Ub, dtype=np.float64).copy(order="F")
        Sb = np.asfortranarray(Sb, dtype=np.float64).copy(order="F")
        Vb = np.asfortranarray(Vb, dtype=np.float64).copy(order="F")
        sc = np.zeros(2)
        check(fn(self.h, float(dt), int(substeps), _dptr(Ub), Ub.shape[0], _dptr(Sb), Sb.shape[0], _dptr(Vb),
                 Vb.shape[0], _dptr(sc)), self.h)
        return Ub, Sb, Vb, (float(sc[0]), float(sc[1]))

    def synchronize(self):
        check(lib().lri_synchronize(self.h), self.h)

    def stream(self) -> int:
        s = _vp()
        check(lib().lri_get_stream(self.h, ctypes.byref(s)), self.h)
        return int(s.value or 0)

    def product(self, which: int, kp: int, kn: int) -> np.ndarray:
        """dA @ V (which=0, n x r) or dA.T @ U (which=1, m x r) with the current state."""
        r = self.rank()
        rows = self.n if which == 0 else self.m
        out = np.zeros((rows, r), order="F")
        check(lib().lri_debug_product(self.h, int(which), int(kp), int(kn), _dptr(out), rows), self.h)
        return out

    # -- timing
    def timing(self, mode):
        """0/False: off; 1/True: every kernel family; 2: the two product passes only."""
        check(lib().lri_kernel_timing(self.h, int(mode)), self.h)

    def timing_read(self, family: int):
        ms = ctypes.c_double()
        cnt = _i64()
        check(lib().lri_kernel_timing_read(self.h, int(family), ctypes.byref(ms), ctypes.byref(cnt)), self.h)
        return float(ms.value), int(cnt.value)

    def trajectory_reserve(self, slots: int):
        """Device storage for ``slots`` saved states (replaces any earlier reservation)."""
        check(lib().lri_trajectory_reserve(self.h, int(slots)), self.h)

    def trajectory_save(self, slot: int):
        """Queue a device copy of the current state into ``slot`` (no synchronisation)."""
        check(lib().lri_trajectory_save(self.h, int(slot)), self.h)

    def trajectory_get(self, slot: int):
        """(U, S, V) saved in ``slot`` (synchronising)."""
        r = _i64()
        check(lib().lri_trajectory_get(self.h, int(slot), ctypes.byref(r), None, self.n, None, 1, None, self.m), self.h)
        rr = int(r.value)
        U = np.zeros((self.n, rr), order="F")
        S = np.zeros((rr, rr), order="F")
        V = np.zeros((self.m, rr), order="F")
        check(lib().lri_trajectory_get(self.h, int(slot), ctypes.byref(r), _dptr(U), self.n, _dptr(S), rr, _dptr(V),
                                       self.m), self.h)
        return U, S, V

    def trajectory_get_range(self, slot0: int, count: int):
        """The states of slots slot0 .. slot0+count-1 in one copy: a list of (U, S, V),
        U and V views into one host buffer (their slots' leading columns), S a copy of the
        leading block."""
        n, m = self.n, self.m
        rc = self.rcap
        stride = n * rc + rc * rc + m * rc
        buf = np.empty(max(count, 1) * stride)
        ranks = np.zeros(max(count, 1), dtype=np.int64)
        check(lib().lri_trajectory_get_range(self.h, int(slot0), int(count),
                                             ranks.ctypes.data_as(ctypes.POINTER(_i64)), _dptr(buf)), self.h)
        out = []
        for i in range(count):
            o, r = i * stride, int(ranks[i])
            U = buf[o:o + n * r].reshape((n, r), order="F")
            S = np.array(buf[o + n * rc:o + n * rc + rc * rc].reshape((rc, rc), order="F")[:r, :r], order="F")
            V = buf[o + n * rc + rc * rc:o + n * rc + rc * rc + m * r].reshape((m, r), order="F")
            out.append((U, S, V))
        return out

    def timing_sample(self, every: int):
        """Time one launch in ``every`` of each family."""
        check(lib().lri_kernel_timing_sample(self.h, int(every)), self.h)

    def timing_samples(self, family: int) -> np.ndarray:
        """Per-launch durations (ms) of a family since timing was enabled."""
        cnt = _i64()
        check(lib().lri_kernel_timing_samples(self.h, int(family), None, 0, ctypes.byref(cnt)), self.h)
        out = np.zeros(int(cnt.value), np.float32)
        if out.size:
            check(lib().lri_kernel_timing_samples(self.h, int(family), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                  out.size, ctypes.byref(cnt)), self.h)
        return out.astype(np.float64)

    def autotune_passes(self, kp: int, kn: int, reps: int = 20) -> dict:
        """Time the candidate reduction chunks of pass N and pass T over A[kn] - A[kp] at the
        current rank and precision (three interleaved rounds, the median per candidate) and
        keep the fastest, the default unless beaten by 1 % (lri_autotune_passes)."""
        rep = np.zeros(8, np.float64)
        check(lib().lri_autotune_passes(self.h, int(kp), int(kn), int(reps),
                                        rep.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), self.h)
        out = {}
        for d, name in enumerate(("pass_n", "pass_t")):
            c, s_, us, dus = rep[4 * d:4 * d + 4]
            out[name] = {"chunk": int(c), "splits": int(s_), "us": float(us), "default_us": float(dus)}
        return out


def selftest_mfma(which: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Raw per-lane results of one matrix-core instruction (see lri_selftest_mfma)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    out = np.zeros(256, np.float64) if which == 0 else np.zeros(1024, np.float32)
    check(lib().lri_selftest_mfma(int(which), a.ctypes.data_as(_vp), b.ctypes.data_as(_vp),
                                  out.ctypes.data_as(_vp)), None)
    return out


def _check_batch(status, b):
    if status == OK:
        return
    L = lib()
    raw = L.lri_batch_last_error(b)
    msg = raw.decode(errors="replace") if raw else L.lri_status_string(status).decode()
    if status == EBREAKDOWN:
        raise BreakdownError(status, msg)
    if status == ENONFINITE:
        raise NonFiniteError(status, msg)
    if status == ENOTSUP:
        raise NotSupportedError(status, msg)
    if status == EINVAL:
        raise ValueError(msg)
    raise LRIError(status, msg)


class _Member(Context):
    """A batch member: the context API on a context the batch owns."""

    def __init__(self, h, n, m, rcap):  # noqa: super().__init__ would create a context
        self.h = h
        self.n, self.m, self.rcap = int(n), int(m), int(rcap)
        self._bound = None

    def close(self):
        self.h = None


class Batch:
    """Owning wrapper of an lri_batch: independent problems of uneven sizes, one
    context each; ``members[i]`` is member i's context (state, operator)."""

    def __init__(self, ns, ms, rcaps, device: int = 0):
        L = lib()
        ns, ms, rcaps = (np.ascontiguousarray(x, dtype=np.int64) for x in (ns, ms, rcaps))
        if not (ns.shape == ms.shape == rcaps.shape) or ns.ndim != 1:
            raise ValueError("ns, ms and rcaps need one entry per problem")
        h = _vp()
        ip = ctypes.POINTER(_i64)
        st = L.lri_batch_create(ctypes.byref(h), int(device), ns.shape[0], ns.ctypes.data_as(ip),
                                ms.ctypes.data_as(ip), rcaps.ctypes.data_as(ip))
        if st != OK:
            _check_batch(st, None)
        self.h = h
        self.members = []
        for i in range(ns.shape[0]):
            c = _vp()
            _check_batch(L.lri_batch_context(h, i, ctypes.byref(c)), h)
            self.members.append(_Member(c, ns[i], ms[i], rcaps[i]))

    def __len__(self):
        return len(self.members)

    def close(self):
        if getattr(self, "h", None):
            for c in self.members:
                c.close()
            lib().lri_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _sub(self, substeps):
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(substeps, dtype=np.int64), (len(self),)))
        return s, s.ctypes.data_as(ctypes.POINTER(_i64))

    def step_de_bug(self, dt: float, substeps):
        s, p = self._sub(substeps)
        _check_batch(lib().lri_batch_step_de_bug(self.h, float(dt), p), self.h)

    def step_de_bug_ra(self, dt: float, substeps, tol: float, rmax: int) -> np.ndarray:
        s, p = self._sub(substeps)
        r = np.zeros(len(self), dtype=np.int64)
        _check_batch(lib().lri_batch_step_de_bug_ra(self.h, float(dt), p, float(tol), int(rmax),
                                                     r.ctypes.data_as(ctypes.POINTER(_i64))), self.h)
        return r

    def member_status(self) -> np.ndarray:
        """Each member's status from the last batch call (0 = OK, else its LRI_E* code)."""
        out = np.zeros(len(self), dtype=np.int64)
        st = _int()
        for i in range(len(self)):
            _check_batch(lib().lri_batch_member_status(self.h, i, ctypes.byref(st)), self.h)
            out[i] = st.value
        return out

    def use_graphs(self, on: bool):
        _check_batch(lib().lri_batch_use_graphs(self.h, int(bool(on))), self.h)

    def use_merged_launches(self, on: bool):
        """Merged launches (default on): the members share one stream, and a rank-adaptive
        step of members in their resident eigenbasis runs each stage as one launch over all."""
        _check_batch(lib().lri_batch_use_merged_launches(self.h, int(bool(on))), self.h)

    def launch_stats(self):
        """(batch steps run as merged launches, kernel launches of the last of them)."""
        a, b = _i64(), _i64()
        _check_batch(lib().lri_batch_launch_stats(self.h, ctypes.byref(a), ctypes.byref(b)), self.h)
        return a.value, b.value

    def synchronize(self):
        _check_batch(lib().lri_batch_synchronize(self.h), self.h)


def product_noalloc(which: int, dtype: int, precision: int, n: int, m: int, a_prev: int, a_next: int, lda: int,
                    B: int, ldb: int, r: int, out: int, ldo: int, stream: int = 0) -> None:
    """lri_product_noalloc on device pointers (asynchronous on ``stream``; allocates nothing)."""
    L = lib()
    st = L.lri_product_noalloc(int(which), int(dtype), int(precision), int(n), int(m), _vp(a_prev), _vp(a_next),
                               int(lda), _vp(B), int(ldb), int(r), _vp(out), int(ldo), _vp(stream))
    if st != OK:
        msg = (L.lri_noalloc_last_error() or b"").decode(errors="replace") or L.lri_status_string(st).decode()
        if st == EINVAL:
            raise ValueError(msg)
        raise LRIError(st, msg)


def stream_probe(pattern: int) -> float:
    """Bare streaming-read bandwidth (TB/s) of the device (see lri_stream_probe)."""
    v = np.zeros(1)
    check(lib().lri_stream_probe(int(pattern), _dptr(v)))
    return float(v[0])
