"""The C ABI's host code under AddressSanitizer and UBSan (no GPU needed).

`build()` also compiles the library's host pass with -fsanitize=address and
-fsanitize=undefined (liblri_hip_san.so).  This test links
tests/host/abi_host_check.c against it with the same sanitizers (the runtime
comes from the program) and runs it: argument validation, null handles, a
context request that fails cleanly without a device, batch and allocation-free
entry points.  A memory error or undefined behaviour aborts the program."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lowrankintegrators.jl_amd")
LIB_SAN = os.path.join(PKG, "liblri_hip_san.so")
SRC = os.path.join(ROOT, "tests", "host", "abi_host_check.c")
CLANG = "/opt/rocm/llvm/bin/clang"


@pytest.mark.skipif(not os.path.exists(LIB_SAN), reason="liblri_hip_san.so not built (run __graft_entry__.build())")
@pytest.mark.skipif(not os.path.exists(CLANG), reason="the ROCm clang is not installed")
def test_abi_host_paths_under_sanitizers(tmp_path):
    exe = tmp_path / "abi_host_check"
    cmd = [CLANG, "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"), SRC, "-o", str(exe),
           "-L", PKG, "-l:liblri_hip_san.so", f"-Wl,-rpath,{PKG}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ)
    # the HIP runtime's own allocations are not ours to judge; memory errors and UB still abort
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    p = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    report = p.stdout + p.stderr
    assert "AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert p.returncode == 0, report[-4000:]
    assert "abi host checks passed" in p.stdout
