"""CPU FP64 oracle for the per-step DLRA path of LowRankIntegrators.jl.

TEST INFRASTRUCTURE ONLY.  This package restates, in NumPy/SciPy FP64
(OpenBLAS/LAPACK), the algorithms of the reference's data path so that the HIP
engine can be checked against it.  It may be imported only by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``, and
only as the checker: the product (the ``lowrankintegrators.jl_amd`` package
and ``liblri_hip.so``) never imports, links or calls anything in here.

Modules
-------
lra      low-rank representation types and truncation rules (the parts of the
         external LowRankArithmetic.jl the reference calls)
dlra     problems, driver loop and the KSL / BUG / rank-adaptive BUG steps
data     seeded synthetic data generators used by the tests and the bench
metrics  factored-form error measures (no n x m matrix is formed)

Pinning (how each function is checked by something other than itself) is
listed in the header of each module; functions that nothing but an import of
the reference could pin are marked "parity unpinned" there.
"""
