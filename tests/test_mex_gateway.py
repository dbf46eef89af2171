"""The C MEX gateway (maxwell_solver_amd/mex/simFDTD_gateway.c) driven through a test-only fake
mx layer: error identifiers on the CPU, the highfrequency.m op sequence on the GPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "mex_host")
EXE = os.path.join(HOST, "replay_highfrequency")


def _build():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", HOST], check=True)
    return EXE


def test_gateway_error_identifiers():
    r = subprocess.run([_build(), "errors"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERRORS OK" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [3, 1, 2, 4, 6, 7])
def test_gateway_replays_highfrequency(tmp_path, mode):
    from maxwell_solver_amd import scenarios as S
    from oracle.oracle import OracleSim
    N, T = 40, 70
    dt = np.float32(0.99 * 0.01 / (1.7320508075688772 * 299792458.0))
    src = S.gauss_sine(T, float(dt), 3e9)
    src.tofile(tmp_path / "src.bin")
    r = subprocess.run([_build(), "run", str(mode), str(N), str(T), str(tmp_path / "src.bin"), str(tmp_path / "ez.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ez = np.fromfile(tmp_path / "ez.bin", dtype=np.float32)
    o = OracleSim(np.float32)
    o.set_const([3e9, 0.01, float(dt)], [T, N, N, N, 0])
    o.set_source(src)
    o.ready()
    o.init_1d(1.0, src)
    o.update_coeffs()
    o.reset()
    o.run(0, T, mode)
    want = o.field({1: "Ey", 2: "Ex", 3: "Ez", 4: "Ey", 6: "Ex", 7: "Ez"}[mode]).reshape(-1)
    assert np.abs(want).max() > 1e-3
    np.testing.assert_array_equal(ez, want)
