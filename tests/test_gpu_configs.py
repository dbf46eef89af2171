"""GPU parity at the sizes of the five workloads, against the CPU oracle (FP32, same seeded inputs).

* 256^3 (40 spheres, oblique plane wave, 64 probes Ez/Hy, Ez snapshot every 250 steps): 400 steps,
  past the pulse peak (step ~70) crossing the spheres and reaching the CPML slabs; fields, all
  twelve psi arrays, every probe record and the step-250 snapshot compared elementwise.
* 384x384x256 lossy layered ground (12-cell CPML, vertical pulse on the layered 1D line, 16x16
  surface probes Ex/Hy): 400 steps, past the pulse peak (step ~258) with the wave ~80 cells into
  the box and over the probe plane; fields, psi and probes compared elementwise.
* The 96-grid batched workload: the whole generator's batch through the batched engine; every
  grid's probe record equals its run alone, and a seeded stratified sample of 10 grids -- per class
  of largest edge ([32,64), [64,96), [96,128), [128,192), [192,..)) the grid with the fewest cell
  updates and the one with the most steps, at their full step counts (14.7 G cell updates) --
  equals the oracle in fields, psi and probes.
* The 2000-request served workload: a stratified sample of 8 requests -- per class of largest edge
  ([48,80), [80,128), [128,200), [200,320]) and material kind (spheres / layered ground) the
  request with the fewest cell updates, at its full step count (11.7 G cell updates) -- submitted
  at once to the request server with 4 slots (so requests queue and join running ones), each
  equal to the oracle in fields, psi and probes.
Equality is the expectation (IEEE float arithmetic with no contraction on both sides); the
acceptance bar is relative L2 <= 1e-5 with the count of unequal elements reported.
"""
import numpy as np
import pytest

from tests.fdtd_compare import assert_close, compare_state, rel_l2

pytestmark = pytest.mark.gpu


def _pair(sc, **kw):
    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    from oracle.oracle import OracleSim
    g, o = Simulation(), OracleSim(np.float32)
    S.setup(g, sc, **kw)
    S.setup(o, sc, **kw)
    return g, o


def _check(g, o, exact=True):
    rep = compare_state(g, o)
    print({k: v for k, v in rep.items() if v[2]})
    assert_close(rep, exact=exact)
    np.testing.assert_array_equal(g.probes(), o.probes())


def test_workload2_full_size_400_steps():
    from maxwell_solver_amd import scenarios as S
    sc = S.config2(T=400)
    assert sc.snap_every == 250
    g, o = _pair(sc)
    g.run(0, sc.T, sc.mode, snap_every=sc.snap_every, snap_comp="Ez")
    snap_o = o.run(0, sc.T, sc.mode, snap_every=sc.snap_every, snap_comp="Ez")
    _check(g, o)
    snaps = g.snapshots()
    assert snaps.shape[0] == 1
    np.testing.assert_array_equal(snaps[0], snap_o[0].astype(np.float32))
    # the test reaches what it is meant to: spheres lit, probes and the CPML slabs carrying field
    assert np.abs(g.probes()).max() > 1e-2
    assert max(np.abs(g.psi(q)).max() for q in range(12)) > 0
    g.close()
    o.close()


def test_workload3_full_size_400_steps():
    from maxwell_solver_amd import scenarios as S
    sc = S.config3(T=400)
    assert (sc.I, sc.J, sc.K, sc.P, len(sc.probes)) == (384, 384, 256, 12, 256)
    g, o = _pair(sc)
    g.run(0, sc.T, sc.mode)
    o.run(0, sc.T, sc.mode)
    _check(g, o)
    assert np.abs(g.probes()).max() > 0 and np.abs(g.field("Ex")).max() > 0
    g.close()
    o.close()


# ---- the 96-grid batched workload ----
C4_CLASSES = [(32, 64), (64, 96), (96, 128), (128, 192), (192, 10 ** 6)]


def c4_sample(work):
    """Per class of largest edge: the grid with the fewest cell updates and the grid with the most
    steps (indices into the generator's list)."""
    pick = []
    for lo, hi in C4_CLASSES:
        m = [(q, sc, st) for q, (sc, st) in enumerate(work) if lo <= max(sc.I, sc.J, sc.K) < hi]
        a = min(m, key=lambda x: x[1].cells * x[2])
        b = max(m, key=lambda x: x[2])
        pick += [a[0]] if a[0] == b[0] else [a[0], b[0]]
    return pick


@pytest.fixture(scope="module")
def batch96():
    from maxwell_solver_amd import Batch, Simulation
    from maxwell_solver_amd import scenarios as S
    work = S.config4()
    sims = []
    b = Batch()
    for sc, steps in work:
        g = Simulation()
        S.setup(g, sc)
        sims.append(g)
        b.add(g, 0, steps, sc.mode)
    b.run()
    b.close()
    yield work, sims
    for g in sims:
        g.close()


def test_workload4_batch_equals_each_grid_alone(batch96):
    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    work, sims = batch96
    assert len(work) == 96
    for (sc, steps), g in zip(work, sims):
        a = Simulation()
        S.setup(a, sc)
        a.run(0, steps, sc.mode)
        np.testing.assert_array_equal(g.probes(), a.probes(), err_msg=sc.name)
        np.testing.assert_array_equal(g.field("Ez"), a.field("Ez"), err_msg=sc.name)
        a.close()


def test_workload4_device_queue_equals_default(batch96):
    """The whole 96-grid batch again through the device work queue (persistent launches pulling
    (grid, tile) items from an atomic counter): every grid equals the default engine's result."""
    from maxwell_solver_amd import Batch, Simulation
    from maxwell_solver_amd import scenarios as S
    work, sims = batch96
    b = Batch()
    b.set_strategy(3)
    mine = []
    for sc, steps in work:
        g = Simulation()
        S.setup(g, sc)
        mine.append(g)
        b.add(g, 0, steps, sc.mode)
    b.run()
    b.close()
    for (sc, steps), g, ref in zip(work, mine, sims):
        np.testing.assert_array_equal(g.probes(), ref.probes(), err_msg=sc.name)
        np.testing.assert_array_equal(g.field("Ez"), ref.field("Ez"), err_msg=sc.name)
        np.testing.assert_array_equal(g.psi(2), ref.psi(2), err_msg=sc.name)
        g.close()


def test_workload4_sample_is_stratified():
    from maxwell_solver_amd import scenarios as S
    work = S.config4()
    pick = c4_sample(work)
    assert len(pick) == 10
    cu = sum(work[q][0].cells * work[q][1] for q in pick)
    assert 14e9 < cu < 15e9


@pytest.mark.parametrize("slot", range(10))
def test_workload4_sampled_grid_matches_oracle(batch96, slot):
    from maxwell_solver_amd import scenarios as S
    from oracle.oracle import OracleSim
    work, sims = batch96
    q = c4_sample(work)[slot]
    sc, steps = work[q]
    o = OracleSim(np.float32)
    S.setup(o, sc)
    o.run(0, steps, sc.mode)
    print(sc.name, sc.I, sc.J, sc.K, steps)
    _check(sims[q], o)
    assert max(np.abs(o.field(n)).max() for n in ("Ex", "Ey", "Ez")) > 0
    o.close()


# ---- the served workload ----
C5_CLASSES = [(48, 80), (80, 128), (128, 200), (200, 10 ** 6)]


def c5_sample(reqs):
    pick = []
    for lo, hi in C5_CLASSES:
        for mode in (5, 2):  # spheres (oblique plane wave) / layered ground (vertical pulse)
            m = [(q, sc, st) for q, (sc, st, _) in enumerate(reqs) if lo <= max(sc.I, sc.J, sc.K) < hi and sc.mode == mode]
            pick.append(min(m, key=lambda x: x[1].cells * x[2])[0])
    return pick


def test_workload5_sample_served_matches_oracle():
    from maxwell_solver_amd import Server, Simulation
    from maxwell_solver_amd import scenarios as S
    from oracle.oracle import OracleSim
    reqs = S.config5(2000)
    pick = c5_sample(reqs)
    assert len(set(pick)) == 8
    srv = Server(hbm_budget=288 << 30, max_active=4, steps_per_round=64)
    ids = {}
    for q in pick:
        sc, steps, _ = reqs[q]
        sim = Simulation()
        S.setup_host(sim, sc)
        ids[srv.submit(sim, steps, sc.mode, source=sc.source)] = q
    done = []
    while len(done) < len(pick):
        active, queued = srv.round()
        done += srv.collect()
        assert active <= 4
    for rid, ta, tadm, tdone in done:
        assert srv.request_error(rid) == (0, "")
        q = ids[rid]
        sc, steps, _ = reqs[q]
        o = OracleSim(np.float32)
        S.setup(o, sc)
        o.run(0, steps, sc.mode)
        g = srv.simulation(rid)
        print(sc.name, sc.I, sc.J, sc.K, steps, "admitted after", round(tadm - ta, 3), "s")
        _check(g, o)
        assert max(np.abs(o.field(n)).max() for n in ("Ex", "Ey", "Ez")) > 0
        o.close()
        srv.release(rid)
    srv.close()
