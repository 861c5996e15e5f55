"""Build liblri_hip.so (gfx950) in-tree with hipcc.

Each csrc/*.hip file is compiled to an object, then linked into
``lowrankintegrators.jl_amd/liblri_hip.so``.  All device code is built with
packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblri_hip.so")
SOURCES = ["lri_tsmm.hip", "lri_rowops.hip", "lri_de.hip", "lri_quad.hip", "lri_quad_adj.hip", "lri_engine.hip", "lri_noalloc.hip"]

ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found (ROCm is required to build liblri_hip.so)")
    return path


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature and says so; that line is noise
    out = "\n".join(line for line in p.stdout.splitlines() if "packed-fp32-ops" not in line)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


# The host code again with AddressSanitizer and UBSan (host pass only; the
# device code is unchanged), as liblri_hip_san.so: tests/test_sanitize_host.py
# links a C program against it and drives the C ABI's argument validation and
# error paths on the CPU.  The sanitizer runtime comes from that program.
LIB_SAN = os.path.join(HERE, "liblri_hip_san.so")
SAN_FLAGS = [
    "-Xarch_host", "-fsanitize=address",
    "-Xarch_host", "-fsanitize=undefined",
    "-Xarch_host", "-fno-sanitize-recover=all",
    "-Xarch_host", "-fno-omit-frame-pointer",
]


def _build_variant(cc, flags, objdir, lib, verbose):
    os.makedirs(objdir, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        out = _run([cc, *flags, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj, out

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(compile_one, SOURCES))
    for _, out in results:
        if out.strip() and verbose:
            print(out)
    tmp = lib + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *[o for o, _ in results]])
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


def build(verbose: bool = True, sanitized: bool = True) -> str:
    cc = hipcc()
    _build_variant(cc, FLAGS, OBJDIR, LIB, verbose)
    if sanitized:
        _build_variant(cc, FLAGS + SAN_FLAGS, os.path.join(OBJDIR, "san"), LIB_SAN, verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
