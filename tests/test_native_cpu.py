"""C-ABI checks that need no GPU: the library loads, exports every symbol that
include/lri.h declares, and validates arguments before touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lri.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lri_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("lri_create", "lri_destroy", "lri_set_state", "lri_get_state", "lri_snapshots_bind",
                     "lri_snapshots_reserve", "lri_snapshot_upload", "lri_step_ksl", "lri_step_ksl_strang",
                     "lri_step_bug", "lri_step_bug_ra", "lri_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol(native):
    for name in declared_functions():
        assert hasattr(native, name), f"{name} declared in lri.h but not exported"


def test_binding_covers_the_header(lri):
    assert set(declared_functions()) == set(lri.native.SIGNATURES)


def test_status_strings(native):
    for s in (0, -1, -2, -3, -4, -5, -6):
        assert native.lri_status_string(s)
    assert native.lri_version() >= 100


def test_create_validates_before_the_device(native):
    h = ctypes.c_void_p()
    assert native.lri_create(ctypes.byref(h), 0, 0, 16, 4) == -1
    assert b"rcap" in native.lri_last_error(None)
    assert native.lri_create(ctypes.byref(h), 0, 16, 16, 65) == -1
    assert native.lri_create(ctypes.byref(h), 0, 8, 16, 9) == -1
    assert native.lri_create(None, 0, 8, 16, 4) == -1


def test_null_context_is_rejected(native):
    assert native.lri_destroy(None) == 0
    assert native.lri_synchronize(None) == -1
    assert native.lri_step_ksl(None, 0, 0, 1) == -1


def test_api_rejects_out_of_scope_problems(lri):
    u0 = lri.SVDLikeRepresentation(np.eye(4, 2), np.eye(2), np.eye(4, 2))
    with pytest.raises(NotImplementedError):
        lri.init(lri.MatrixDEProblem(lambda y, t: y, u0, (0.0, 1.0)), lri.UnconventionalAlgorithm(), 0.1)
    with pytest.raises(NotImplementedError):
        lri.ProjectorSplitting(lri.PrimalLieTrotter(), K_rhs=lambda *a: None)
    with pytest.raises(AssertionError):
        lri.solve(lri.MatrixDataProblem(lambda t: np.zeros((4, 4)), u0, (0.0, 1.0)), lri.UnconventionalAlgorithm())
    with pytest.raises(AssertionError):
        lri.init(lri.MatrixDataProblem(lambda t: np.zeros((4, 4)), u0, (1.0, 0.0)), lri.UnconventionalAlgorithm(), 0.1)
    # only the Greedy integrator has two-factor methods (greedy_integrator.jl:24-30)
    tf0 = lri.TwoFactorRepresentation(np.eye(4, 2), np.ones((4, 2)))
    with pytest.raises(TypeError):
        lri.init(lri.MatrixDataProblem(lambda t: np.zeros((4, 4)), tf0, (0.0, 1.0)), lri.UnconventionalAlgorithm(), 0.1)
    with pytest.raises(NotImplementedError):  # the hybrid problem's Z dynamics are a closure
        lri.init(lri.MatrixHybridProblem(None, lambda z, u, t: z, tf0, (0.0, 1.0)), lri.GreedyIntegrator(), 0.1)


def test_init_sol_slots(lri):
    s = lri.api.init_sol(1e-2, 0.0, 1.0, None)
    assert len(s.Y) == 101
    s = lri.api.init_sol(1, 1, 11, None)
    assert s.t == list(range(1, 12))


def test_bf16_encoding_round_to_nearest_even(lri):
    x = np.array([[1.0, 1.00390625, 1.01171875, -3.14159, 0.0]])
    bits = lri.api._encode(x, "bf16")
    back = (bits.astype(np.uint32) << 16).view(np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa -> 1.0
    assert back[0, 0] == 1.0 and back[0, 1] == 1.0
    assert abs(back[0, 3] - np.float32(-3.140625)) == 0
    assert back[0, 4] == 0.0


def test_operator_descriptor_matches_the_oracle(lri):
    """Host side of the matrix-ODE descriptor: the tridiagonal, its bound and F agree with the oracle's."""
    from oracle import dlra_de

    n, m = 12, 9
    A = lri.laplacian_1d(n)
    assert np.array_equal(A.dense(), dlra_de.dirichlet_laplacian(n))
    P = lri.laplacian_1d(m, 2.0, periodic=True)
    Pd = P.dense()
    assert Pd[0, m - 1] == 2.0 and Pd[m - 1, 0] == 2.0 and np.allclose(Pd, Pd.T)
    for T in (A, P, lri.SymTridiagonal(np.arange(5.0), [1.0, -2.0, 0.5, 3.0])):
        assert T.gershgorin() >= np.max(np.abs(np.linalg.eigvalsh(T.dense()))) - 1e-12
    rng = np.random.default_rng(0)
    G, H = rng.standard_normal((n, 2)), rng.standard_normal((m, 2))
    F = lri.LyapunovOperator(A, P, G, H)
    Fo = dlra_de.LyapunovOperator(A.dense(), Pd, G, H)
    Y = rng.standard_normal((n, m))
    assert np.allclose(F(Y), Fo(Y))
    assert lri.rk4_substeps(F.spectral_bound(), 1e-3) == dlra_de.rk4_substeps(F.spectral_bound(), 1e-3)


def test_operator_descriptor_validation(lri):
    with pytest.raises(ValueError):
        lri.SymTridiagonal(np.ones(4), np.ones(4))  # Dirichlet needs n - 1 off-diagonals
    with pytest.raises(ValueError):
        lri.SymTridiagonal(np.ones(2), np.ones(2), periodic=True)
    with pytest.raises(ValueError):
        lri.LyapunovOperator(lri.laplacian_1d(4), lri.laplacian_1d(4), np.ones((4, 5)), np.ones((4, 5)))
    with pytest.raises(ValueError):
        lri.LyapunovOperator(lri.laplacian_1d(4), lri.laplacian_1d(4), np.ones((4, 1)), None)
    u0 = lri.SVDLikeRepresentation(np.eye(4, 2), np.eye(2), np.eye(4, 2))
    F = lri.LyapunovOperator(lri.laplacian_1d(4), lri.laplacian_1d(5))
    with pytest.raises(ValueError):  # operator 4 x 5 vs state 4 x 4
        lri.init(lri.MatrixDEProblem(F, u0, (0.0, 1.0)), lri.UnconventionalAlgorithm(), 0.1)


def test_de_entry_points_reject_a_null_context(native):
    assert native.lri_step_de_bug(None, 0.1, 1) == -1
    assert native.lri_step_de_ksl(None, 0, 0.1, 1) == -1
    assert native.lri_set_operator_lyapunov(None, None, None, 0, None, None, 0, 0, None, 0, None, 0) == -1
