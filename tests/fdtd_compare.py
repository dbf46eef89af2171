"""Comparison helpers shared by the parity tests."""
from __future__ import annotations

import numpy as np

FIELDS = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]
REL_L2_TOL = 1e-5  # FP32 GPU vs FP32 CPU oracle (relative L2); bit equality is also reported


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def compare_state(gpu, orc, psi=True):
    """Per-array (rel-L2, max-abs, number of unequal elements)."""
    out = {}
    for n in FIELDS:
        a, b = gpu.field(n), orc.field(n)
        out[n] = (rel_l2(a, b), float(np.max(np.abs(a.astype(np.float64) - b))), int(np.count_nonzero(a != b)))
    if psi:
        for q in range(12):
            a, b = gpu.psi(q), orc.psi(q)
            out[f"psi{q}"] = (rel_l2(a, b), float(np.max(np.abs(a.astype(np.float64) - b))), int(np.count_nonzero(a != b)))
    return out


def assert_close(report, tol=REL_L2_TOL, exact=False):
    bad = {k: v for k, v in report.items() if v[0] > tol or (exact and v[2] != 0)}
    assert not bad, f"parity failure: {bad}"
