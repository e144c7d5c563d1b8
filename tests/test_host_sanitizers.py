"""The library's host C++ code (imps_small.cpp: Jacobi and Hessenberg eigen routines of the
Krylov drivers; imps_io.cpp: state and MPO files) built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++, host only, no GPU) and run through tests/host/sanitize_host.cpp,
which checks each routine over a range of sizes and every error path of the file readers
(truncated, flipped, oversized, mismatched and missing files).  Any sanitizer report fails the
run (-fno-sanitize-recover=all; leaks are reported at exit)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "infinitempsalgorithms_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the ROCm headers")
def test_host_code_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           os.path.join(ROOT, "tests", "host", "sanitize_host.cpp"),
           os.path.join(CSRC, "imps_small.cpp"), os.path.join(CSRC, "imps_io.cpp"), "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    scratch = tmp_path / "files"
    scratch.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(scratch)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and "host checks passed" in r.stdout, (r.stdout + r.stderr)[-6000:]
