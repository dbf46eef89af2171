"""Request server (served form): continuous batching of solves that arrive one at a time.

GPU tests: every request's probe record, fields and psi equal those of the same simulation run
alone (bit for bit: the batched engine runs the same kernels on the same state), requests are
admitted in arrival order within the HBM budget, and a served request matches the CPU oracle.
CPU tests: argument and state checks of the C ABI that make no device call.
"""
import numpy as np
import pytest

from tests.fdtd_compare import FIELDS, assert_close, compare_state, rel_l2


def _requests(n=6, seed_steps=(30, 90)):
    from maxwell_solver_amd import scenarios as S
    return S.config5(n, edge_lo=48, edge_hi=64, steps_lo=seed_steps[0], steps_hi=seed_steps[1])


def _alone(sc, steps):
    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    g = Simulation()
    S.setup(g, sc)
    g.run(0, steps, sc.mode)
    g.synchronize()
    return g


@pytest.mark.gpu
def test_served_results_equal_alone_and_fit_the_budget():
    from maxwell_solver_amd import Server, Simulation
    from maxwell_solver_amd import scenarios as S
    reqs = _requests(6)
    # a budget that holds about two of these requests at once (bounds taken from a probe server)
    probe = Server(hbm_budget=1 << 40, max_active=1, steps_per_round=1)
    sizes = []
    for sc, steps, _ in reqs:
        sim = Simulation()
        S.setup_host(sim, sc)
        sizes.append(probe.request_bytes(probe.submit(sim, steps, sc.mode, source=sc.source)))
    probe.close()
    budget = int(2.5 * max(sizes))
    srv = Server(hbm_budget=budget, max_active=3, steps_per_round=16)
    ids = {}
    # arrivals: three at once, then one per round
    pending = list(range(len(reqs)))
    order = []
    while pending or len(order) < len(reqs):
        burst = 3 if not ids else 1
        for _ in range(min(burst, len(pending))):
            q = pending.pop(0)
            sc, steps, _ = reqs[q]
            sim = Simulation()
            S.setup_host(sim, sc)
            ids[srv.submit(sim, steps, sc.mode, source=sc.source)] = q
        srv.round()
        for rid, ta, tadm, tdone in srv.collect():
            assert ta <= tadm <= tdone
            order.append((rid, ta, tadm, tdone))
            q = ids[rid]
            sc, steps, _ = reqs[q]
            g = srv.simulation(rid)
            ref = _alone(sc, steps)
            np.testing.assert_array_equal(g.probes(), ref.probes())
            for n in FIELDS:
                np.testing.assert_array_equal(g.field(n), ref.field(n))
            for k in range(12):
                np.testing.assert_array_equal(g.psi(k), ref.psi(k))
            ref.close()
            srv.release(rid)
    in_use, peak, cells = srv.stats()
    assert in_use == 0
    assert 0 < peak <= budget
    assert cells == sum(sc.cells * steps for sc, steps, _ in reqs)
    # admission in arrival order
    by_id = sorted(order)
    adm = [t[2] for t in by_id]
    assert adm == sorted(adm)
    srv.close()


@pytest.mark.gpu
def test_served_request_matches_oracle():
    from maxwell_solver_amd import Server, Simulation
    from maxwell_solver_amd import scenarios as S
    from oracle.oracle import OracleSim
    reqs = _requests(4, (40, 60))
    srv = Server(hbm_budget=64 << 30, max_active=4, steps_per_round=25)
    sims = {}
    for sc, steps, _ in reqs:
        sim = Simulation()
        S.setup_host(sim, sc)
        sims[srv.submit(sim, steps, sc.mode, source=sc.source)] = (sc, steps)
    done = []
    while len(done) < len(reqs):
        srv.round()
        done += srv.collect()
    for rid, *_ in done:
        sc, steps = sims[rid]
        o = OracleSim(np.float32)
        S.setup(o, sc)
        o.run(0, steps, sc.mode)
        g = srv.simulation(rid)
        assert_close(compare_state(g, o))
        assert rel_l2(g.probes(), o.probes()) <= 1e-5
        o.close()
        srv.release(rid)
    srv.close()


def test_server_argument_and_state_checks():
    """Checks the server makes before any device call (runs without a GPU)."""
    from maxwell_solver_amd import Server, SimFDTDError, Simulation
    from maxwell_solver_amd import scenarios as S
    with pytest.raises(SimFDTDError):
        Server(hbm_budget=0)
    with pytest.raises(SimFDTDError):
        Server(hbm_budget=1 << 30, max_active=0)
    srv = Server(hbm_budget=1 << 30, max_active=2, steps_per_round=8)
    sc, steps, _ = _requests(1)[0]
    s0 = Simulation()
    with pytest.raises(SimFDTDError) as e:  # no SetConst
        srv.submit(s0, 10, 0)
    assert e.value.identifier == "simFDTD:InvalidInput"
    s0.close()
    sim = Simulation()
    S.setup_host(sim, sc)
    with pytest.raises(SimFDTDError):
        srv.submit(sim, 0, sc.mode, source=sc.source)  # no steps
    with pytest.raises(SimFDTDError):
        srv.submit(sim, steps, sc.mode)  # plane wave without its source samples
    with pytest.raises(SimFDTDError):
        srv.submit(sim, steps + 1, sc.mode, source=sc.source)  # beyond the source length
    tiny = Server(hbm_budget=1 << 20, max_active=2, steps_per_round=8)
    with pytest.raises(SimFDTDError) as e:
        tiny.submit(sim, steps, sc.mode, source=sc.source)
    assert e.value.identifier == "simFDTD:OutOfMemory"
    tiny.close()
    rid = srv.submit(sim, steps, sc.mode, source=sc.source)
    with pytest.raises(SimFDTDError):
        srv.submit(sim, steps, sc.mode, source=sc.source)  # already submitted
    with pytest.raises(SimFDTDError):
        srv.release(rid)  # not completed
    with pytest.raises(SimFDTDError):
        srv.release(rid + 100)  # unknown
    assert srv.collect() == []
    assert srv.stats() == (0, 0, 0)
    assert srv.request_bytes(rid) > 8 * 4 * sc.cells  # two field sets of six components at least
    srv.close()  # destroys the queued simulation it owns


def test_failed_admission_is_reported_and_releasable():
    """A request whose set-up fails at admission (here a TF/SF inset too large for its grid, which
    ReadySimulation's host part rejects before any device call) completes at once as failed: it is
    charged nothing, collect returns it with admission time -1, its error is reported, it can be
    released, and the requests queued behind it are still admitted."""
    import dataclasses
    from maxwell_solver_amd import Server, Simulation
    from maxwell_solver_amd import scenarios as S
    sc, steps, _ = _requests(1)[0]
    bad = dataclasses.replace(sc, inset=sc.I // 2)
    srv = Server(hbm_budget=1 << 40, max_active=1, steps_per_round=8)
    ids = []
    for q in range(2):
        sim = Simulation()
        S.setup_host(sim, bad)
        ids.append(srv.submit(sim, steps, bad.mode, source=bad.source))
    active, queued = srv.round()  # both admissions fail before touching the device
    assert (active, queued) == (0, 0)
    done = srv.collect()
    assert [d[0] for d in done] == ids
    for rid, ta, tadm, tdone in done:
        assert tadm == -1 and tdone >= ta
        code, msg = srv.request_error(rid)
        assert code == 1 and "inset" in msg
        srv.release(rid)
    assert srv.stats()[0] == 0 and srv.stats()[1] == 0
    srv.close()


def test_stale_and_unknown_handles_are_rejected():
    """Every entry point rejects a handle that was destroyed or never created."""
    import ctypes as C
    from maxwell_solver_amd import SimFDTDError, Simulation
    from maxwell_solver_amd.simfdtd import lib
    s = Simulation()
    s.set_const([3e9, 0.01, 1.9e-11], [10, 40, 40, 40, 0])
    h = s._h
    s.close()
    for fn in (lambda: lib().fdtd_update_h_all(h), lambda: lib().fdtd_destroy(h),
               lambda: lib().fdtd_reset_fields(C.c_void_p(0x1234560))):
        assert fn() != 0
        assert lib().fdtd_last_error_id().decode() == "simFDTD:InvalidInput"
    t = Simulation()
    t.set_const([3e9, 0.01, 1.9e-11], [10, 40, 40, 40, 0])
    with pytest.raises(SimFDTDError):
        t.set_const([3e9, 0.01, 1.9e-11], [10, 40, 40, 40, 0])
    out = t.output_ant(8)  # no antennas (and not ready): the reference returns the zero array
    assert out.shape == (8,) and not out.any()
    t.close()


def test_served_workload_generator_is_seeded():
    from maxwell_solver_amd import scenarios as S
    a, b = S.config5(8), S.config5(8)
    for (sa, na, ga), (sb, nb, gb) in zip(a, b):
        assert (sa.I, sa.J, sa.K, sa.mode, na, ga) == (sb.I, sb.J, sb.K, sb.mode, nb, gb)
        assert 48 <= min(sa.I, sa.J, sa.K) and max(sa.I, sa.J, sa.K) <= 400
        assert 500 <= na <= 3000 and ga > 0
        if sa.mat_index is not None:
            np.testing.assert_array_equal(sa.mat_index, sb.mat_index)
    kinds = {sc.mode for sc, _, _ in S.config5(40)}
    assert kinds == {2, 5}
