"""lri_product_noalloc (the allocation-free product) against lri_debug_product, the
entry point whose kernels it shares, through the C ABI (needs an MI355X).

Two calls in a row on the same GPU (the second with a smaller input, so it runs
over what the first left in the shared static scratch), each checked against a
context's product on the same snapshots and operand: the same kernels in the same
order, so the results agree bit for bit.  A shape beyond the scratch is refused.
"""
import ctypes
import sys

import numpy as np
import pytest

from oracle.data import bf16_round, random_factors

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def context_product(lri, which, u, A0, A1, dtype, precision):
    n, m = A0.shape
    with lri.native.Context(n, m, u.rank) as ctx:
        ctx.set_precision(precision)
        ctx.set_state(u.U, u.S, u.V)
        ctx.reserve(2, lri.native.DTYPES[dtype])
        ctx.upload(0, lri.api._encode(A0, dtype))
        ctx.upload(1, lri.api._encode(A1, dtype))
        return ctx.product(which, 0, 1)


class _Hip:
    """The HIP runtime the library itself runs on (libamdhip64.so.7, already loaded by
    liblri_hip.so): device buffers for the caller-owned arrays of the call."""

    def __init__(self):
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMalloc.restype = ctypes.c_int
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipFree.restype = ctypes.c_int
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpy.restype = ctypes.c_int
        h.hipDeviceSynchronize.restype = ctypes.c_int
        self.h = h
        self.bufs = []

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 1)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, ptr, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def bf16_bits(a):
    """BF16 bits (round to nearest even) of values that are already BF16-representable."""
    return (np.ascontiguousarray(a, dtype=np.float32).view(np.uint32) >> 16).astype(np.uint16)


def noalloc_product(lri, hip, which, u, A0, A1, dtype, precision):
    n, m = A0.shape
    enc = {"f64": lambda a: np.asfortranarray(a, dtype=np.float64),
           "f32": lambda a: np.asfortranarray(a, dtype=np.float32),
           "bf16": lambda a: np.asfortranarray(bf16_bits(a))}[dtype]
    a0 = hip.upload(enc(A0).ravel(order="F"))
    a1 = hip.upload(enc(A1).ravel(order="F"))
    Bh = u.V if which == 0 else u.U
    B = hip.upload(np.asfortranarray(Bh).ravel(order="F"))
    rows = n if which == 0 else m
    out = hip.upload(np.full(rows * u.rank, np.nan))
    lri.native.product_noalloc(which, lri.native.DTYPES[dtype], precision, n, m, a0, a1, n, B, Bh.shape[0], u.rank,
                               out, rows, 0)
    assert hip.h.hipDeviceSynchronize() == 0
    return hip.download(out, (u.rank, rows), np.float64).T.copy()


@pytest.mark.parametrize("dtype,precision", [("bf16", 1), ("f32", 1), ("f64", 0), ("bf16", 0)])
def test_noalloc_matches_context_product(lri, dtype, precision):
    lri.native.lib()  # the library first, so that the runtime below is the one it loaded
    hip = _Hip()
    rng = np.random.default_rng(21)
    for (n, m, r) in [(700, 530, 32), (300, 181, 8)]:  # the second call is smaller
        say(f"case: noalloc product n={n} m={m} r={r} {dtype} precision={precision}")
        u = random_factors(rng, n, m, r)
        A0 = rng.standard_normal((n, m))
        A1 = A0 + 1e-2 * rng.standard_normal((n, m))
        if dtype == "bf16":
            A0, A1 = bf16_round(A0), bf16_round(A1)
        elif dtype == "f32":
            A0, A1 = A0.astype(np.float32).astype(np.float64), A1.astype(np.float32).astype(np.float64)
        for which in (0, 1):
            ref = context_product(lri, which, u, A0, A1, dtype, precision)
            got = noalloc_product(lri, hip, which, u, A0, A1, dtype, precision)
            assert got.shape == ref.shape and np.array_equal(got, ref), (which, n, m, r)
    hip.free()


def test_noalloc_rejects_oversize(lri):
    with pytest.raises(ValueError, match="LRI_NOALLOC_SCRATCH_BYTES"):
        lri.native.product_noalloc(0, lri.native.F64, 0, 65536, 65536, 1, 1, 65536, 1, 65536, 64, 1, 65536)
