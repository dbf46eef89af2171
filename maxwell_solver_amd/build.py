"""Builds libfdtd_hip.so in-tree with hipcc for gfx950.

Flags:
  -ffp-contract=off            every multiply and add rounded on its own (bit parity with the
                               CPU oracle, which is built the same way)
  -packed-fp32-ops disabled    no v_pk_*_f32 generation
Correctly rounded FP32 division and IEEE denormals are HIP's defaults and are kept.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfdtd_hip.so")
SOURCES = ["fdtd_kernels.hip", "fdtd_setup.cpp", "fdtd_api.cpp"]
ARCH = os.environ.get("FDTD_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [
        hipcc(), "-O3", "-std=c++17", "-fPIC", "-shared",
        f"--offload-arch={ARCH}",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I", os.path.join(HERE, "..", "include"),
        "-o", tmp,
    ] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
