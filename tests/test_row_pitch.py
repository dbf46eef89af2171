"""Padded row pitch of the per-cell device arrays (fields, materials, per-cell coefficients).

By default the rows of a grid whose state streams to HBM (larger than the 256 MB Infinity Cache)
are padded to a multiple of 32 floats; the test grids here are small, so FDTD_ROW_ALIGN forces
the padding (16, 32 or 64 floats) and every path that addresses a per-cell array is compared with
the CPU oracle bit for bit: the per-op reference sequence, fused runs over the row layouts, index
and per-cell coefficient modes, per-cell mu, probes, snapshots, OutputVis, step currents, the
graded mesh, the batched engine with padded and unpadded grids together, the allocation-free
entry point (against an ordinary run) and the stream ceiling.  CPU tests check
the device-bytes bound the request server admits against covers the padded arrays.
"""
import numpy as np
import pytest

from tests.fdtd_compare import assert_close, compare_state
from tests.test_gpu_parity import CFG_CPML, _backends, _per_op_steps, _scenario


@pytest.fixture
def pitch(monkeypatch):
    def set_align(a):
        monkeypatch.setenv("FDTD_ROW_ALIGN", str(a))
    return set_align


def _bytes(sc, align, monkeypatch):
    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    monkeypatch.setenv("FDTD_ROW_ALIGN", str(align))
    g = Simulation()
    S.setup(g, sc)
    b = g.device_bytes()
    g.close()
    return b


@pytest.mark.gpu
def test_padding_is_applied(monkeypatch):
    """Forced padding adds the padded columns to the device arrays (I = 40 -> 64 floats per row);
    a multiple of 64 columns is never padded."""
    sc = _scenario(I=40)
    b1, b64 = _bytes(sc, 1, monkeypatch), _bytes(sc, 64, monkeypatch)
    assert b64 - b1 >= 2 * 6 * 4 * 24 * sc.J * sc.K  # two field sets of six components, 24 more columns
    sc64 = _scenario(I=64)
    assert _bytes(sc64, 1, monkeypatch) == _bytes(sc64, 32, monkeypatch)


@pytest.mark.gpu
@pytest.mark.parametrize("align", [16, 32, 64])
@pytest.mark.parametrize("mode", [3, 6])
def test_per_op_sequence_padded(pitch, align, mode):
    from maxwell_solver_amd import scenarios as S
    pitch(align)
    sc = _scenario(I=50, mode=mode)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    _per_op_steps(g, sc.T, mode, currents=(1, (12, 16, 10, 14, 11, 13)))
    _per_op_steps(o, sc.T, mode, currents=(1, (12, 16, 10, 14, 11, 13)))
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


@pytest.mark.gpu
@pytest.mark.parametrize("I,rows,cpl,mode", [(50, 0, 0, 5), (130, 2, 1, 3), (100, 1, 2, 1), (196, 0, 4, 5)])
def test_fused_layouts_padded(pitch, I, rows, cpl, mode):
    from maxwell_solver_amd import scenarios as S
    pitch(32)
    sc = _scenario(I=I, mode=mode, cpml=CFG_CPML, inset=12, oblique=(0.9, 2.3, 0.7) if mode == 5 else None)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.set_tuning(rows, 0, cpl)
    g.run(0, sc.T, mode)
    o.run(0, sc.T, mode)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


@pytest.mark.gpu
def test_full_coefficients_mu_snapshots_padded(pitch):
    """Per-cell coefficient arrays, per-cell mu, snapshots and OutputVis on a padded grid."""
    from maxwell_solver_amd import scenarios as S
    pitch(32)
    sc = _scenario(I=44, mode=3, cpml=CFG_CPML, spheres=False)
    I, J, K = sc.I, sc.J, sc.K
    rng = np.random.default_rng(3)
    eps = (1 + rng.integers(0, 300, size=I * J * K) * 0.01).astype(np.float32)
    g, o = _backends()
    for b in (g, o):
        S.setup(b, sc)
        b.update_material(4, eps)
        b.update_coeffs()
    g.run(0, sc.T, 3, snap_every=20, snap_comp="Ez")
    snap_o = o.run(0, sc.T, 3, snap_every=20, snap_comp="Ez")
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.snapshots(), snap_o[:3].astype(np.float32))
    for flag, name in [(1, "Ex"), (3, "Ez"), (5, "Hy")]:
        np.testing.assert_array_equal(g.output_vis(flag), o.field(name).reshape(-1))
    # per-cell mu on another padded grid, fused run
    sc2 = _scenario(I=52, mode=3, cpml=CFG_CPML)
    mu = [np.random.default_rng(11 + c).uniform(1.0, 4.0, sc2.I * sc2.J * sc2.K).astype(np.float32) for c in range(3)]
    g2, o2 = _backends()
    S.setup(g2, sc2, mu=mu)
    S.setup(o2, sc2)
    o2.create_mu(*mu)
    g2.run(0, sc2.T, 3)
    o2.run(0, sc2.T, 3)
    assert_close(compare_state(g2, o2), exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("strategy", [0, 3])
def test_batch_padded_and_unpadded_grids(monkeypatch, strategy):
    """Grids created with and without padding advance together in one batch."""
    from maxwell_solver_amd import Batch
    from maxwell_solver_amd import scenarios as S
    b = Batch()
    b.set_strategy(strategy)
    pairs = []
    for align, c in ((32, dict(I=50, mode=3)), (1, dict(I=52, mode=5, oblique=(0.9, 2.3, 0.7))),
                     (16, dict(I=130, mode=1, inset=12)), (64, dict(I=40, mode=0))):
        monkeypatch.setenv("FDTD_ROW_ALIGN", str(align))
        sc = _scenario(T=60, cpml=CFG_CPML, **c)
        g, o = _backends()
        S.setup(g, sc)
        S.setup(o, sc)
        b.add(g, 0, 45, sc.mode)
        pairs.append((g, o, sc))
    b.run()
    for g, o, sc in pairs:
        o.run(0, 45, sc.mode)
        assert_close(compare_state(g, o), exact=True)
        np.testing.assert_array_equal(g.probes(), o.probes())
    b.close()


@pytest.mark.gpu
def test_graded_mesh_and_noalloc_on_padded_grids(pitch):
    """The graded-mesh row kernel (fused run with step currents and probes) against the oracle, and
    the allocation-free entry point against an ordinary run, on padded grids."""
    from tests.test_mesh import _gpu_pair, _mesh3
    from tests.test_zz_no_alloc import _pair as noalloc_pair
    from tests.test_zz_no_alloc import _same
    from tests.test_zz_no_alloc import _scenario as noalloc_scenario
    pitch(32)
    I, J, K, P = 50, 42, 46, 6
    g, o = _gpu_pair(I, J, K, P, _mesh3(I, J, K, P))
    box = [I // 2 - 2, I // 2 + 1, 20, 21, 22, 24]
    g.set_step_currents(2, box)
    g.run(0, 30, 0)
    for n in range(30):
        o.update_h_all(); o.update_h_pml(); o.update_e_all(); o.update_e_pml()
        o.add_currents(n, 2, box)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())
    a, b = noalloc_pair(noalloc_scenario(52, 40, 44, 50, 3))
    a.run_noalloc(0, 40, 3)
    b.run(0, 40, 3)
    _same(a, b)
    for s in (g, o, a, b):
        s.close()


@pytest.mark.gpu
def test_stream_ceiling_on_a_padded_grid(pitch):
    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    pitch(32)
    g = Simulation()
    S.setup(g, _scenario(I=50))
    assert g.stream_ceiling(2) > 0.0
    g.close()


def test_device_bytes_bound_covers_the_widest_pitch(monkeypatch):
    """The server's admission bound (host set-up only, no device call) holds the arrays at a row
    pitch of 64 floats whatever FDTD_ROW_ALIGN says."""
    from maxwell_solver_amd import Server, Simulation
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=66, J=44, K=48)
    sizes = []
    for align in (1, 64):
        monkeypatch.setenv("FDTD_ROW_ALIGN", str(align))
        srv = Server(hbm_budget=1 << 40, max_active=1, steps_per_round=1)
        sim = Simulation()
        S.setup_host(sim, sc)
        sizes.append(srv.request_bytes(srv.submit(sim, sc.T, sc.mode, source=sc.source)))
        srv.close()
    assert sizes[0] == sizes[1]
    assert sizes[0] > 2 * 6 * 4 * 128 * sc.J * sc.K  # two sets of six components at 128 columns
