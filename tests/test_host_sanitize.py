"""The C++ host API's CPU-only utilities (validation digits, Frobenius error,
timer, CSV logging, op helpers) built with AddressSanitizer and
UndefinedBehaviorSanitizer and run on the CPU: any invalid access or
undefined operation aborts the check."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_host_utilities_under_sanitizers(tmp_path):
    exe = str(tmp_path / "hostsan")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROCM, "include"),
           os.path.join(ROOT, "distributed-gemm-thesis_amd", "host", "utilities.cpp"),
           os.path.join(ROOT, "tests", "host_sanitize", "check.cpp"),
           "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-3000:])
    assert "host utilities: ok" in r.stdout
