"""The library's host set-up code (csrc/fdtd_setup.cpp) and the CPU oracle built with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer (host code only), driven by
tests/sanitize/sanitize_driver.cpp through every plane-wave mode in float and double: any
out-of-bounds access, use after free, leak or undefined operation fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_and_oracle_are_sanitizer_clean(tmp_path):
    exe = tmp_path / "sanitize_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fopenmp",
           "-I", os.path.join(ROOT, "maxwell_solver_amd", "csrc"),
           os.path.join(ROOT, "tests", "sanitize", "sanitize_driver.cpp"),
           os.path.join(ROOT, "maxwell_solver_amd", "csrc", "fdtd_setup.cpp"),
           os.path.join(ROOT, "oracle", "fdtd_oracle.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "SANITIZE OK" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]
