"""Pins of the oracle's source and output functions: AddCurrents (K17), the probe recorder and the
snapshot cadence.

* AddCurrents (src/gpu.h:55-82 via src/simulation.h:256-286): one call on a field of random values
  changes exactly the cells of the box, of the one component the polarisation names, by the
  reference's expression E - (((Ch s) dx) / dx) / dx in float32, with Ch the per-cell coefficient
  (itself pinned by its closed form in test_oracle_pins.py) and s = source[n] of SetSource; the
  upper ends are clipped to the grid, and an empty box, a negative start or a polarisation outside
  1..3 changes nothing (the reference's size_t box).  In vacuum the change equals the current
  density form -dt s / (eps0 dx^2) of Ampere's law to float rounding.
* Probe recorder: after step n the record row n of component c and probe a holds that component at
  the cell (P + x, P + y, P + z) of the probe's offsets (src/simulation.h:1047-1067), read from the
  field itself after the step -- in fused runs, per-op sequences (a step closes at UpdateE_PML) and
  with a CPML width other than 10.
* Snapshots: snapshot q of a run with cadence s is the field after step (q + 1) s counted since
  ResetFieldAll, also for a run that starts mid-way.
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim

C0 = 299792458.0
MU0 = 4 * 3.1415926535897 * 1e-7
EPS0 = 1 / C0 / C0 / MU0
DX = 0.01
DT = 0.99 * DX / (np.sqrt(3) * C0)
FIELDS = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]


def _k17_sim(I, J, K, T=8, vacuum=False, seed=1):
    s = OracleSim(np.float32)
    s.set_const([3e9, DX, DT], [T, I, J, K, 0])
    s.set_tfsf(3)
    rng = np.random.default_rng(seed)
    s.set_source((rng.standard_normal(T) * 1e3).astype(np.float32))
    if not vacuum:
        for f in range(1, 4):
            s.update_material(f, rng.uniform(0, 0.3, I * J * K).astype(np.float32))
        for f in range(4, 7):
            s.update_material(f, rng.uniform(1, 9, I * J * K).astype(np.float32))
    s.ready()
    for n in FIELDS:
        s.set_field(n, rng.standard_normal((K, J, I)).astype(np.float32))
    return s


def _expected_k17(s, old, pol, box, n, src):
    """The reference expression of gpu.h:76 over the clipped box, in float32 arithmetic."""
    I, J, K = s.dims
    ch = s.coeffs()[1][pol - 1].reshape(K, J, I)
    out = old.copy()
    i0, i1, j0, j1, k0, k1 = box
    if i1 < i0 or j1 < j0 or k1 < k0 or min(i0, j0, k0) < 0:
        return out
    sl = (slice(k0, min(k1, K - 1) + 1), slice(j0, min(j1, J - 1) + 1), slice(i0, min(i1, I - 1) + 1))
    sv = np.float32(src[n])
    dx = np.float32(DX)
    v = (((ch[sl] * sv) * dx) / dx) / dx
    out[sl] = out[sl] - v
    return out


BOXES = [
    [3, 17, 5, 5, 9, 9],        # a line along x
    [2, 4, 3, 6, 1, 20],        # a block
    [20, 40, 15, 30, 25, 50],   # upper ends past the grid: clipped
    [0, 0, 0, 0, 0, 0],         # the first cell
    [23, 23, 22, 22, 28, 28],   # the last cell (den_ex = 0 there: Ch = 0)
]


@pytest.mark.parametrize("pol", [1, 2, 3])
@pytest.mark.parametrize("box", BOXES)
def test_add_currents_reference_expression(pol, box):
    I, J, K = 24, 23, 29
    s = _k17_sim(I, J, K)
    before = {n: s.field(n).copy() for n in FIELDS}
    n = 5
    src = np.random.default_rng(1).standard_normal(8).astype(np.float32) * np.float32(1e3)
    s.add_currents(n, pol, box)
    for q, name in enumerate(FIELDS):
        got = s.field(name)
        if q == pol - 1:
            want = _expected_k17(s, before[name], pol, box, n, src)
            np.testing.assert_array_equal(got, want)
            changed = np.count_nonzero(got != before[name])
            assert changed > 0 or box == [23, 23, 22, 22, 28, 28]
        else:
            np.testing.assert_array_equal(got, before[name])


@pytest.mark.parametrize("box,pol", [
    ([-1, 5, 2, 4, 2, 4], 1),     # negative start along x: the reference's size_t box writes nothing
    ([2, 5, -3, 4, 2, 4], 2),
    ([2, 5, 2, 4, -1, 4], 3),
    ([-5, -1, 2, 4, 2, 4], 3),    # both ends negative: every index past the grid
    ([5, 2, 2, 4, 2, 4], 1),      # empty box (i1 < i0)
    ([2, 5, 2, 4, 4, 2], 2),
    ([2, 5, 2, 4, 2, 4], 0),      # polarisation outside 1..3
    ([2, 5, 2, 4, 2, 4], 4),
])
def test_add_currents_writes_nothing(box, pol):
    s = _k17_sim(24, 23, 29)
    before = {n: s.field(n).copy() for n in FIELDS}
    s.add_currents(3, pol, box)
    for name in FIELDS:
        np.testing.assert_array_equal(s.field(name), before[name])


def test_add_currents_is_amperes_law_in_vacuum():
    """In vacuum Ch = 2 dt / (2 eps0) / dx, so the change is -dt s / (eps0 dx^2): the current s
    through one cell face of area dx^2 (an independent float64 form of the update)."""
    I, J, K = 24, 25, 26
    s = _k17_sim(I, J, K, vacuum=True, seed=4)
    src = np.random.default_rng(4).standard_normal(8).astype(np.float32) * np.float32(1e3)
    before = s.field("Ey").astype(np.float64)
    box = [4, 12, 6, 6, 7, 9]
    s.add_currents(2, 2, box)
    d = s.field("Ey").astype(np.float64) - before
    dxf, dtf = float(np.float32(DX)), float(np.float32(DT))
    want = -dtf * float(src[2]) / (EPS0 * dxf * dxf)
    cells = d[7:10, 6:7, 4:13]
    np.testing.assert_allclose(cells, want, rtol=2e-6, atol=2e-6 * abs(want) + 1e-6 * np.abs(before).max())
    mask = np.ones_like(d, bool)
    mask[7:10, 6:7, 4:13] = False
    assert not np.any(d[mask])


def _probe_sim(P, comps, mode, T=24):
    I, J, K = 40, 42, 44
    s = OracleSim(np.float32)
    s.set_const([3e9, DX, DT], [T, I, J, K, 5])
    s.set_cpml([P, 3, 5, 0.05, 1, 0.8])
    s.set_tfsf(P + 5)
    t = np.arange(T) * DT
    src = (np.exp(-((t - 8 * DT) / (3 * DT)) ** 2)).astype(np.float32)
    s.set_source(src)
    s.set_probe_components(comps)
    rng = np.random.default_rng(9)
    loc = np.stack([rng.integers(0, I - 2 * P, 5), rng.integers(0, J - 2 * P, 5), rng.integers(0, K - 2 * P, 5)])
    loc[:, 0] = [I - 2 * P - 1, J - 2 * P - 1, K - 2 * P - 1]  # the last cell a probe may take
    loc[:, 1] = [0, 0, 0]
    s.set_antenna_locations(loc.reshape(-1))
    eps = np.ones(I * J * K, np.float32)
    eps[(np.arange(I * J * K) // I) % 3 == 0] = 4.0
    s.update_material(4, eps)
    s.update_material(6, eps)
    s.ready()
    if mode:
        s.init_1d(1.0, src)
    s.reset()
    return s, loc


@pytest.mark.parametrize("P,mode", [(10, 3), (8, 1), (12, 0)])
def test_probe_record_is_the_field_after_each_step(P, mode):
    comps = [1, 3, 5, 6]
    s, loc = _probe_sim(P, comps, mode)
    T = s.T
    seen = {}
    for n in range(T - 4):  # the last rows are never reached
        if mode == 0:
            s.add_currents(n, 3, [20, 20, 18, 18, 22, 22])  # an excitation without a plane wave
        s.run(n, 1, mode)
        for c in comps:
            f = s.field(FIELDS[c - 1])
            seen[(c, n)] = f[P + loc[2], P + loc[1], P + loc[0]]
    rec = s.probes()
    assert rec.shape == (len(comps), T, 5)
    for q, c in enumerate(comps):
        for n in range(T - 4):
            np.testing.assert_array_equal(rec[q, n], seen[(c, n)])
        assert not np.any(rec[q, T - 4:])
    assert np.any(rec != 0)


def test_probe_record_of_per_op_steps():
    """The reference op sequence (highfrequency.m:85-91): a step's record is taken when
    UpdateE_PML closes it, so plane-wave and PML ops between steps land in the right row."""
    comps = [3, 5]
    s, loc = _probe_sim(10, comps, 3)
    P = 10
    for n in range(12):
        s.update_h_all(); s.update_h_planewave(n, 3); s.update_h_pml()
        s.update_e_all(); s.update_e_planewave(n, 3); s.update_e_pml()
        rec = s.probes()
        for q, c in enumerate(comps):
            f = s.field(FIELDS[c - 1])
            np.testing.assert_array_equal(rec[q, n], f[P + loc[2], P + loc[1], P + loc[0]])
            if n + 1 < s.T:
                assert not np.any(rec[q, n + 1:])


@pytest.mark.parametrize("every,comp", [(3, "Ez"), (5, "Hy"), (1, "Ex")])
def test_snapshots_are_fields_after_every_kth_step(every, comp):
    s, _ = _probe_sim(10, [3], 3, T=24)
    states = {}
    for n in range(17):
        s.run(n, 1, 3)
        states[n + 1] = s.field(comp).reshape(-1)
    t, _ = _probe_sim(10, [3], 3, T=24)
    first = t.run(0, 7, 3, snap_every=every, snap_comp=comp)
    rest = t.run(7, 10, 3, snap_every=every, snap_comp=comp)  # the cadence counts from ResetFieldAll
    got = [first[q] for q in range(7 // every)]
    n_rest = 17 // every - 7 // every
    got += [rest[q] for q in range(n_rest)]
    want_steps = [every * (q + 1) for q in range(17 // every)]
    assert len(got) == len(want_steps)
    for g, n in zip(got, want_steps):
        np.testing.assert_array_equal(g.astype(np.float32), states[n])
    np.testing.assert_array_equal(t.field(comp), s.field(comp))
