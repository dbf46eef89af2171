"""GPU parity: the HIP library against the CPU oracle on identical seeded inputs.

All device arithmetic is IEEE float with no contraction, as the oracle's, so the expected result
is equality; the acceptance bar is relative L2 <= 1e-5 (max-abs and the count of unequal
elements are reported in the failure message).
"""
import numpy as np
import pytest

from tests.fdtd_compare import FIELDS, assert_close, compare_state, rel_l2

pytestmark = pytest.mark.gpu

REF_CPML = (10, 3.0, 1.0, 0.08, 1.0, 0.8)
CFG_CPML = (10, 3.0, 5.0, 0.05, 1.0, 0.8)


def _backends():
    from maxwell_solver_amd import Simulation
    from oracle.oracle import OracleSim
    return Simulation(), OracleSim(np.float32)


def _scenario(I=40, J=44, K=48, T=60, mode=3, cpml=REF_CPML, inset=15, spheres=True, layered=False,
              oblique=None, seed=7):
    from maxwell_solver_amd import scenarios as S
    dx = 0.01
    dt = 0.99 * dx / (np.sqrt(3) * S.C0)
    rng = S.SplitMix64(seed)
    idx = tab = None
    if spheres:
        box = (inset, I - inset - 1, inset, J - inset - 1, inset, K - inset - 1)
        idx, tab = S.spheres(rng, I, J, K, box, 5, 1.5, 3.5, (2.0, 12.0), (0.0, 0.05))
    if layered:
        idx = np.zeros((K, J, I), np.uint8)
        idx[: K // 2] = 1
        idx[: K // 3] = 2
        tab = np.array([[0, 0, 0, 1, 1, 1], [0.01, 0.01, 0.01, 6, 6, 6], [0.2, 0.2, 0.2, 15, 15, 15]], np.float32)
    params = (1.0,) if layered else (oblique if oblique is not None else (0.0,))
    probes = np.array([[3, 4, 5], [I - 2 * cpml[0] - 2, 7, 9], [10, J - 2 * cpml[0] - 3, K - 2 * cpml[0] - 1]], np.int32)
    return S.Scenario("t", I, J, K, T, dx, dt, 3e9, cpml, inset, mode, params, S.gauss_sine(T, dt, 3e9),
                      mat_index=idx, mat_table=tab, probes=probes, probe_comps=(3, 5, 1))


def _per_op_steps(b, T, mode, currents=None):
    for n in range(T):
        b.update_h_all()
        if mode:
            b.update_h_planewave(n, mode)
        b.update_h_pml()
        b.update_e_all()
        if mode:
            b.update_e_planewave(n, mode)
        b.update_e_pml()
        if currents is not None:
            b.add_currents(n, currents[0], currents[1])


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 6, 7])
def test_per_op_reference_sequence(mode):
    """highfrequency.m:85-91 op sequence, reference CPML preset, modes 1..4 and the y modes 6, 7."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(mode=mode)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    _per_op_steps(g, sc.T, mode)
    _per_op_steps(o, sc.T, mode)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


def _mu(sc, seed):
    rng = np.random.default_rng(seed)
    n = sc.I * sc.J * sc.K
    return [rng.uniform(1.0, 4.0, n).astype(np.float32) for _ in range(3)]


@pytest.mark.parametrize("mode", [3, 6])
def test_per_op_with_mu(mode):
    """CreateMu (per-cell mu, K2 GPU_updateH_Full) in the reference op sequence, with spheres and a
    plane wave; CreateMu given before ReadySimulation on the GPU and after it on the oracle (the
    fixed initialisation order makes both take effect)."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(mode=mode, cpml=CFG_CPML)
    mu = _mu(sc, 11)
    g, o = _backends()
    S.setup(g, sc, mu=mu)  # before ReadySimulation
    S.setup(o, sc)
    o.create_mu(*mu)  # after ReadySimulation
    _per_op_steps(g, sc.T, mode)
    _per_op_steps(o, sc.T, mode)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


def test_fused_run_with_mu():
    """With per-cell mu the fused entry points (Run, Step) take the row kernel's per-cell H
    coefficients and a batch runs the grid on its own: same bits as the oracle's K2; CreateMu after
    ReadySimulation on the GPU."""
    from maxwell_solver_amd import Batch
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=48, J=40, K=44, mode=5, cpml=CFG_CPML, oblique=(0.8, 2.0, 0.5))
    mu = _mu(sc, 12)
    g, o = _backends()
    S.setup(g, sc)
    g.create_mu(*mu)
    S.setup(o, sc, mu=mu)
    g.run(0, 20, 5)
    g.step(20, 5)
    o.run(0, 21, 5)
    assert_close(compare_state(g, o), exact=True)
    b = Batch()
    b.add(g, 21, sc.T - 21, 5)
    b.run()
    g.synchronize()
    o.run(21, sc.T - 21, 5)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


# TF/SF box faces inside the 10-cell CPML (inset 6: the faces lie in slab rows, slab columns and
# slab planes; inset 10: the faces sit on the slabs' last cells), every plane-wave mode, per-op
# sequence and fused run + step
@pytest.mark.parametrize("inset", [6, 10])
@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5, 6, 7])
def test_tfsf_box_inside_cpml(mode, inset):
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=44, J=40, K=46, T=40, mode=mode, cpml=CFG_CPML, inset=inset,
                   oblique=(0.9, 2.3, 0.7) if mode == 5 else None)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    if mode != 5:  # oblique mode 5 has no per-op form (the reference dispatches modes 1-4)
        _per_op_steps(g, 20, mode)
        _per_op_steps(o, 20, mode)
        assert_close(compare_state(g, o), exact=True)
        g.reset()
        o.reset()
    g.run(0, sc.T - 1, mode)
    g.step(sc.T - 1, mode)
    o.run(0, sc.T, mode)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


def _gic_oracle_steps(o, n0, count, box):
    for n in range(n0, n0 + count):
        o.update_h_all()
        o.update_h_pml()
        o.update_e_all()
        o.update_e_pml()
        o.add_currents(n, 1, box)


# current boxes: in the interior part's cells (applied inside the slab part's launch), in a y-slab
# row and across z-slab planes (separate launches)
@pytest.mark.parametrize("box", [[1, 46, 26, 26, 26, 26], [1, 46, 5, 5, 26, 26], [20, 21, 20, 30, 2, 35],
                                 [-3, 46, 26, 26, 26, 26]])
def test_step_currents_graph_replays(box):
    """GICsimulation.m:119-124 loop (no plane wave, AddCurrents after every step) as step currents:
    runs replay one recorded graph of 8 steps (the current sample read on the device), with eager
    steps and a re-recording between the replays; every stage equals the oracle's op sequence."""
    import dataclasses
    from maxwell_solver_amd import scenarios as S
    sc = dataclasses.replace(_scenario(I=48, J=52, K=40, T=60, mode=0, cpml=CFG_CPML, spheres=False, layered=True),
                             probes=None)
    g, o = _backends()
    S.setup(g, sc, init_line=False)
    S.setup(o, sc, init_line=False)
    g.set_step_currents(1, box)
    g.set_graph_steps(8)
    g.run(0, 21, 0)           # two replays, five steps one by one
    _gic_oracle_steps(o, 0, 21, box)
    assert_close(compare_state(g, o), exact=True)
    g.step(21, 0)             # the current set flips: the next run records again
    _gic_oracle_steps(o, 21, 1, box)
    g.run(22, 17, 0)          # two replays of the new recording, one step
    _gic_oracle_steps(o, 22, 17, box)
    assert_close(compare_state(g, o), exact=True)
    g.run(39, 16, 0)          # replays of the same recording again, after eager work
    _gic_oracle_steps(o, 39, 16, box)
    assert_close(compare_state(g, o), exact=True)
    g.set_graph_steps(0)
    g.run(55, 5, 0)
    _gic_oracle_steps(o, 55, 5, box)
    assert_close(compare_state(g, o), exact=True)


def test_step_currents_with_probes():
    """Step currents in runs with probes (steps one by one): each record samples the step's E
    before its currents, as UpdateE_PML closes a step before AddCurrents."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=64, J=44, K=40, T=40, mode=0, cpml=CFG_CPML)
    box = [3, 60, 20, 20, 22, 22]
    g, o = _backends()
    S.setup(g, sc, init_line=False)
    S.setup(o, sc, init_line=False)
    g.set_step_currents(1, box)
    g.set_graph_steps(8)      # not used: the run records probes
    g.run(0, sc.T, 0)
    _gic_oracle_steps(o, 0, sc.T, box)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


# row layouts with per-cell mu: one column per lane (split and unsplit), two and four per lane,
# several rows per tile, a plain-wave grid and an oblique one
@pytest.mark.parametrize("I,rows,cpl,mode", [(64, 0, 1, 5), (52, 0, 0, 3), (200, 0, 2, 5), (264, 1, 4, 5),
                                             (128, 4, 1, 1), (128, 0, -1, 5), (384, 0, 0, 5)])
def test_fused_layouts_with_mu(I, rows, cpl, mode):
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=I, J=37, K=45, T=40, mode=mode, cpml=CFG_CPML, oblique=(0.9, 2.3, 0.7) if mode == 5 else None)
    mu = _mu(sc, 13)
    g, o = _backends()
    S.setup(g, sc, mu=mu)
    S.setup(o, sc, mu=mu)
    g.set_tuning(rows, 0, cpl)
    g.run(0, sc.T, mode)
    o.run(0, sc.T, mode)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


# a line, a block past the upper ends (clipped), and boxes with a negative start (the reference's
# size_t box writes nothing) -- the last two with a line that does write, so the state is not zero
@pytest.mark.parametrize("box", [[1, 30, 20, 20, 12, 12], [30, 60, 40, 70, 40, 60], [-1, 30, 20, 20, 12, 12],
                                 [4, 9, 20, 22, -2, 12]])
def test_per_op_gic_currents(box):
    """GICsimulation.m:119-124 order (no plane wave, AddCurrents after UpdateE_PML)."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(mode=0, cpml=CFG_CPML)
    g, o = _backends()
    S.setup(g, sc, init_line=False)
    S.setup(o, sc, init_line=False)
    if min(box[0], box[2], box[4]) < 0:
        for b in (g, o):
            b.add_currents(0, 3, [10, 12, 11, 11, 9, 9])
    _per_op_steps(g, sc.T, 0, currents=(1, box))
    _per_op_steps(o, sc.T, 0, currents=(1, box))
    assert_close(compare_state(g, o), exact=True)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 6, 7])
@pytest.mark.parametrize("cpml", [REF_CPML, CFG_CPML], ids=["refcpml", "cfgcpml"])
def test_fused_step_axis_modes(mode, cpml):
    """One fused kernel per step equals the per-op sequence of the oracle."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(mode=mode, cpml=cpml)
    g, o = _backends()
    S.setup(g, sc, init_line=bool(mode))
    S.setup(o, sc, init_line=bool(mode))
    for n in range(sc.T):
        g.step(n, mode)
    o.run(0, sc.T, mode)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


# (I, rows per group, z chunk, columns per lane, TF/SF inset): every row layout of the row kernel --
# 1, 2 and 4 columns per lane, one or several row segments (I = 52 with 1 column per lane and 200
# with 2 take 1 and 2 segments, 264 with 4 takes 2, 136 with 1 takes 3), several rows per
# workgroup; with 1 column per lane and I a multiple of 64 the split edge segment (both row ends in
# one wave: I = 64, 128, 192, 256 take 1 to 4 segments), with the x faces of the TF/SF box in the
# edge wave or (inset 40) in plain segments; -1 columns per lane: the same grids unsplit; rows of a
# workgroup exchange their neighbour rows through LDS (up to 16 rows, 1024 threads); rows of 300 to
# 512 columns take 5 to 8 one-column segments (split when I is a multiple of 64), or 2 columns.
@pytest.mark.parametrize("I,rows,z_chunk,cpl,inset", [
    (52, 0, 0, 0, 15), (52, 1, 5, 1, 15), (52, 2, 13, 4, 15), (52, 4, 3, 2, 15),
    (136, 0, 0, 1, 15), (200, 0, 7, 2, 15), (264, 0, 0, 4, 15), (264, 1, 4, 2, 15),
    (64, 0, 0, 1, 15), (64, 4, 6, 1, 13), (128, 0, 0, 1, 15), (128, 2, 5, 1, 40), (192, 1, 0, 1, 40),
    (192, 0, 3, 1, 15), (256, 0, 0, 1, 15), (256, 0, 9, 1, 40), (128, 0, 0, -1, 15), (256, 0, 0, -1, 15),
    (64, 16, 0, 1, 15), (128, 8, 5, 1, 15), (52, 8, 4, 1, 15), (256, 4, 0, 1, 15), (200, 2, 0, 2, 15),
    (384, 0, 0, 0, 15), (320, 1, 0, 1, 40), (300, 0, 6, 0, 15), (512, 0, 0, 1, 15), (384, 0, 0, 2, 15)])
def test_fused_tilings_oblique(I, rows, z_chunk, cpl, inset):
    """Oblique plane wave through every row layout: halo rows/planes recomputed from old data."""
    from maxwell_solver_amd import scenarios as S
    J, K = (37, 45) if inset < 20 else (90, 88)
    sc = _scenario(I=I, J=J, K=K, mode=5, cpml=CFG_CPML, oblique=(0.9, 2.3, 0.7), inset=inset)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.set_tuning(rows, z_chunk, cpl)
    g.run(0, sc.T, 5)
    o.run(0, sc.T, 5)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


# The row kernel's 64-bit global path (grids whose arrays exceed 4 GiB take it) forced on small
# grids: the row-layout families (1, 2 and 4 columns per lane, split and unsplit edge segments,
# several rows per tile), per-cell mu, a TF/SF box inside the CPML (the slab march with its TF/SF
# code) and a merged batch.  The buffer path reads 0 and drops stores out of range, the global
# path would not, so this also shows that no access of the row kernel leaves its array.
GLOBAL_CASES = [(52, 0, 0, 0, 15), (52, 2, 13, 4, 15), (136, 0, 0, 1, 15), (200, 0, 7, 2, 15), (264, 1, 4, 2, 15),
                (64, 4, 6, 1, 13), (128, 2, 5, 1, 40), (256, 0, 0, 1, 15), (256, 0, 0, -1, 15), (300, 0, 6, 0, 15)]


def _global_case(I, rows, z_chunk, cpl, inset):
    from maxwell_solver_amd import scenarios as S
    J, K = (37, 45) if inset < 20 else (90, 88)
    sc = _scenario(I=I, J=J, K=K, mode=5, cpml=CFG_CPML, oblique=(0.9, 2.3, 0.7), inset=inset)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.set_addressing(1)
    g.set_tuning(rows, z_chunk, cpl)
    g.run(0, sc.T, 5)
    o.run(0, sc.T, 5)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


def _global_mu_and_inset():
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=64, J=37, K=45, T=40, mode=5, cpml=CFG_CPML, oblique=(0.9, 2.3, 0.7))
    mu = _mu(sc, 13)
    g, o = _backends()
    S.setup(g, sc, mu=mu)
    S.setup(o, sc, mu=mu)
    g.set_addressing(1)
    g.run(0, sc.T, 5)
    o.run(0, sc.T, 5)
    assert_close(compare_state(g, o), exact=True)
    sc = _scenario(I=64, J=44, K=48, mode=5, cpml=CFG_CPML, inset=6, spheres=False, oblique=(0.8, 2.0, 0.5))
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.set_addressing(1)
    g.run(0, sc.T, 5)
    o.run(0, sc.T, 5)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


def _global_batch():
    from maxwell_solver_amd import Batch
    from maxwell_solver_amd import scenarios as S
    b = Batch()
    pairs = []
    for c in (dict(I=40, J=44, K=48, mode=3, cpml=REF_CPML, count=37),
              dict(I=52, J=37, K=45, mode=5, cpml=CFG_CPML, oblique=(0.9, 2.3, 0.7), count=60)):
        count = c.pop("count")
        sc = _scenario(T=60, **c)
        g, o = _backends()
        S.setup(g, sc)
        S.setup(o, sc)
        g.set_addressing(1)
        b.add(g, 0, count, sc.mode)
        pairs.append((g, o, sc, count))
    b.run()
    for g, o, sc, count in pairs:
        o.run(0, count, sc.mode)
        assert_close(compare_state(g, o), exact=True)
        np.testing.assert_array_equal(g.probes(), o.probes())
    b.close()


@pytest.mark.parametrize("I,rows,z_chunk,cpl,inset", GLOBAL_CASES)
def test_forced_global_addressing(I, rows, z_chunk, cpl, inset):
    _global_case(I, rows, z_chunk, cpl, inset)


def test_forced_global_addressing_mu_inset_batch():
    _global_mu_and_inset()
    _global_batch()


@pytest.mark.parametrize("I,slab,z_chunk,inset", [
    (64, 1, 0, 15), (64, 3, 0, 13), (128, 4, 6, 15), (256, 5, 0, 40), (52, 7, 3, 15), (136, 16, 0, 15),
    (256, 2, 9, 15)])
def test_slab_chunk_depths(I, slab, z_chunk, inset):
    """The slab part's own z-chunk depth (y-band rows over the interior planes, z slabs cut into
    chunks of at most that many planes): every depth equals the oracle, with the wave crossing the
    TF/SF faces in the slabs' chunks."""
    from maxwell_solver_amd import scenarios as S
    J, K = (37, 45) if inset < 20 else (90, 88)
    sc = _scenario(I=I, J=J, K=K, mode=5, cpml=CFG_CPML, oblique=(0.9, 2.3, 0.7), inset=inset)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.set_tuning(0, z_chunk, 0)
    g.set_slab_chunk(slab)
    assert g.slab_chunk() == slab
    g.run(0, sc.T, 5)
    o.run(0, sc.T, 5)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


def test_runs_and_steps_interleaved():
    """A run fills the next step's face sheets and gathers the previous step's probes inside the
    row kernel's slab-part launch: runs of one and two steps, single steps between runs, and a run
    that ends at the last time index keep every record and field equal to the oracle."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=64, J=40, K=44, mode=5, cpml=CFG_CPML, oblique=(0.8, 2.0, 0.5))
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    n = 0
    for kind, count in (("run", 1), ("step", 1), ("run", 2), ("step", 1), ("step", 1), ("run", 7),
                        ("run", 1), ("step", 1)):
        if kind == "run":
            g.run(n, count, 5)
        else:
            g.step(n, 5)
        n += count
    g.run(n, sc.T - n, 5)
    o.run(0, sc.T, 5)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


def test_sampled_timing_keeps_parity():
    """Timing events on every third step of a run: the run still equals the oracle, the timed step
    count is ceil(count / 3), and the interior part's time lies inside the step's."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=64, J=40, K=44, mode=5, cpml=CFG_CPML, oblique=(0.8, 2.0, 0.5))
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.set_timing(True, every=3)
    g.run(0, 10, 5)
    step_ms, launches = g.kernel_time()
    rows_ms = g.row_kernel_time()
    assert launches == 4
    assert 0 < rows_ms < step_ms
    g.set_timing(False)
    g.run(10, sc.T - 10, 5)
    assert g.kernel_time()[1] == 4  # an untimed run leaves the last figures
    o.run(0, sc.T, 5)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


def test_step_outside_fused_layouts():
    """Rows longer than 1024 columns are outside the row kernel's layouts: Step and Run take the
    per-op kernels in the reference op order instead, with the same result."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=1032, J=34, K=36, T=40, mode=3, cpml=CFG_CPML, inset=12)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.run(0, sc.T, 3)
    o.run(0, sc.T, 3)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


def test_autotune_leaves_state_and_keeps_parity():
    """The autotuner runs dry steps over every row layout and z-chunk: fields, psi and probes are
    unchanged, and the chosen layout then steps bit-exactly with the oracle."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=64, J=40, K=44, mode=5, cpml=CFG_CPML, oblique=(0.7, 1.9, 0.4))
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.run(0, 10, 5)
    o.run(0, 10, 5)
    before = compare_state(g, o)
    ms = g.autotune(10, 5, reps=3)
    assert ms > 0
    assert_close(before, exact=True)
    assert_close(compare_state(g, o), exact=True)
    g.run(10, sc.T - 10, 5)
    o.run(10, sc.T - 10, 5)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


@pytest.mark.parametrize("strategy", [0, 1, 2, 3])
def test_batch_of_grids_matches_each_alone(strategy):
    """Grids of different sizes, row layouts, materials, plane-wave modes and step counts advanced
    together by the batched engine equal the oracle run of each grid alone, bit for bit (fields,
    psi, probe records), and each grid can be stepped on afterwards."""
    from maxwell_solver_amd import Batch
    from maxwell_solver_amd import scenarios as S
    cases = [
        dict(I=40, J=44, K=48, mode=3, cpml=REF_CPML, count=37),
        dict(I=52, J=37, K=45, mode=5, cpml=CFG_CPML, oblique=(0.9, 2.3, 0.7), count=60),
        dict(I=200, J=34, K=36, mode=1, cpml=CFG_CPML, inset=12, count=23),
        dict(I=64, J=40, K=44, mode=0, cpml=CFG_CPML, count=60),
        dict(I=1032, J=34, K=36, mode=3, cpml=CFG_CPML, inset=12, count=11),  # outside the fused layouts
    ]
    b = Batch()
    b.set_strategy(strategy)
    pairs = []
    for c in cases:
        c = dict(c)
        count = c.pop("count")
        sc = _scenario(T=60, **c)
        g, o = _backends()
        S.setup(g, sc)
        S.setup(o, sc)
        b.add(g, 0, count, sc.mode)
        pairs.append((g, o, sc, count))
    b.run()
    for g, o, sc, count in pairs:
        o.run(0, count, sc.mode)
        assert_close(compare_state(g, o), exact=True)
        np.testing.assert_array_equal(g.probes(), o.probes())
    g, o, sc, count = pairs[0]
    g.run(count, 5, sc.mode)
    o.run(count, 5, sc.mode)
    assert_close(compare_state(g, o), exact=True)
    b.close()


def test_batch_full_coefficient_grid():
    """A grid with more than 256 materials (per-cell coefficient arrays) in a batch with an
    index-mode grid."""
    from maxwell_solver_amd import Batch
    from maxwell_solver_amd import scenarios as S
    b = Batch()
    pairs = []
    for full in (True, False):
        sc = _scenario(mode=3, cpml=CFG_CPML, spheres=not full)
        I, J, K = sc.I, sc.J, sc.K
        g, o = _backends()
        for be in (g, o):
            S.setup(be, sc)
            if full:
                rng = np.random.default_rng(5)
                be.update_material(4, (1 + rng.integers(0, 300, size=I * J * K) * 0.01).astype(np.float32))
                be.update_coeffs()
        b.add(g, 0, 30, 3)
        pairs.append((g, o))
    b.run()
    for g, o in pairs:
        o.run(0, 30, 3)
        assert_close(compare_state(g, o), exact=True)


def test_oblique_per_op_and_lines():
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(mode=5, cpml=CFG_CPML, oblique=(2.2, 4.0, 1.1))
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    np.testing.assert_array_equal(g.line(0), o.line(0))
    np.testing.assert_array_equal(g.line(1), o.line(1))
    _per_op_steps(g, sc.T, 5)
    _per_op_steps(o, sc.T, 5)
    assert_close(compare_state(g, o), exact=True)


def test_reference_lines_equal():
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(mode=3, cpml=CFG_CPML)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    for w in range(4):
        np.testing.assert_array_equal(g.line(w), o.line(w))


def test_layered_line_mode2():
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=40, J=40, K=56, mode=2, cpml=(12, 3.0, 5.0, 0.05, 1.0, 0.8), inset=17, spheres=False, layered=True)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    np.testing.assert_array_equal(g.line(0), o.line(0))
    g.run(0, sc.T, 2)
    o.run(0, sc.T, 2)
    assert_close(compare_state(g, o), exact=True)


def test_full_coefficient_mode_matches_index_mode():
    """More than 256 distinct materials switch to per-cell coefficient arrays; same result."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(mode=3, cpml=CFG_CPML, spheres=False)
    I, J, K = sc.I, sc.J, sc.K
    rng = np.random.default_rng(3)
    eps = (1 + rng.integers(0, 300, size=I * J * K) * 0.01).astype(np.float32)
    g, o = _backends()
    for b in (g, o):
        S.setup(b, sc)
        b.update_material(4, eps)
        b.update_material(5, eps)
        b.update_coeffs()
    for n in range(sc.T):
        g.step(n, 3)
    o.run(0, sc.T, 3)
    assert_close(compare_state(g, o), exact=True)


def test_snapshots_and_output_vis():
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(mode=3, cpml=CFG_CPML)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.run(0, sc.T, 3, snap_every=20, snap_comp="Ez")
    snap_o = o.run(0, sc.T, 3, snap_every=20, snap_comp="Ez")
    snaps = g.snapshots()
    assert snaps.shape[0] == 3
    np.testing.assert_array_equal(snaps, snap_o[:3].astype(np.float32))
    for flag, name in [(1, "Ex"), (2, "Ey"), (3, "Ez"), (4, "Hx"), (5, "Hy"), (6, "Hz")]:
        np.testing.assert_array_equal(g.output_vis(flag), o.field(name).reshape(-1))


def test_workload1_full():
    """64^3 vacuum workload, 600 steps: probe series and final fields vs the oracle."""
    from maxwell_solver_amd import scenarios as S
    sc = S.config1()
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.run(0, sc.T, sc.mode)
    o.run(0, sc.T, sc.mode)
    rep = compare_state(g, o)
    assert_close(rep)
    assert rel_l2(g.probes(), o.probes()) <= 1e-5


def test_workload2_reduced():
    """Workload-2 physics (spheres, oblique wave) on 64^3, 200 steps."""
    from maxwell_solver_amd import scenarios as S
    sc = S.small_variant(S.config2(T=200), 64, 200)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.run(0, sc.T, sc.mode)
    o.run(0, sc.T, sc.mode)
    assert_close(compare_state(g, o))
    assert rel_l2(g.probes(), o.probes()) <= 1e-5


def test_workload2_full_size_first_steps():
    """256^3 workload 2 at its real size: the first 12 steps vs the oracle."""
    from maxwell_solver_amd import scenarios as S
    sc = S.config2(T=12)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.run(0, sc.T, sc.mode)
    o.run(0, sc.T, sc.mode)
    assert_close(compare_state(g, o, psi=False))


def test_workload3_reduced():
    from maxwell_solver_amd import scenarios as S
    sc = S.config3(T=200, I=64, J=64, K=96, lattice=4)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.run(0, sc.T, sc.mode)
    o.run(0, sc.T, sc.mode)
    assert_close(compare_state(g, o))
    assert rel_l2(g.probes(), o.probes()) <= 1e-5


# FP32 error of each workload against an FP64 run of the oracle on the same seeded inputs (the 64^3
# workload in full, the 256^3 and layered-ground workloads reduced): the probe series relative L2,
# and the final fields' max-abs error against the largest recorded probe value (the 64^3 pulse has
# left the box by step 600, so its final fields are residue and a relative L2 of them says nothing).
@pytest.mark.parametrize("which", ["c1", "c2", "c3"])
def test_fp32_error_against_fp64(which):
    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    from oracle.oracle import OracleSim
    sc = {"c1": lambda: S.config1(T=600), "c2": lambda: S.small_variant(S.config2(T=200), 64, 200),
          "c3": lambda: S.config3(T=200, I=64, J=64, K=96, lattice=4)}[which]()
    g, o = Simulation(), OracleSim(np.float64)
    S.setup(g, sc)
    S.setup(o, sc)
    g.run(0, sc.T, sc.mode)
    o.run(0, sc.T, sc.mode)
    pg, po = g.probes().astype(np.float64), o.probes()
    err_p = rel_l2(pg, po)
    err_f = max(np.abs(g.field(n).astype(np.float64) - o.field(n)).max() for n in FIELDS)
    peak = np.abs(po).max()
    print(f"{which}: probes rel-L2 vs FP64 {err_p:.2e}, final-field max-abs / probe peak {err_f / peak:.2e}")
    assert peak > 0
    assert err_p <= 1e-5
    assert err_f <= 1e-5 * peak


def test_vacuum_energy_balance():
    """Lossless vacuum (the 64^3 workload, pulse inside the TF box at step 150): the field energy in
    the box's inner cells equals the energy of the 1D incident wave over the same cells (E^n and
    H^{n-1/2} from the line histories), and the scattered-field region holds none of it."""
    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    eps0, mu0 = 8.854187817e-12, 4e-7 * np.pi
    sc = S.config1(T=600)
    g = Simulation()
    S.setup(g, sc)
    n = 150
    g.run(0, n, sc.mode)
    F = {c: g.field(c).astype(np.float64) for c in FIELDS}
    u = sum(eps0 * F[c] ** 2 for c in ("Ex", "Ey", "Ez")) + sum(mu0 * F[c] ** 2 for c in ("Hx", "Hy", "Hz"))
    N, P, w = sc.I, sc.P, sc.inset
    tf = np.zeros(u.shape, bool)
    tf[w:N - w, w:N - w, w:N - w] = True
    near = np.zeros(u.shape, bool)  # the box and the cells next to its faces
    near[w - 2:N - w + 2, w - 2:N - w + 2, w - 2:N - w + 2] = True
    sf = np.zeros(u.shape, bool)
    sf[P:N - P, P:N - P, P:N - P] = True
    sf &= ~near
    assert u[sf].sum() <= 1e-9 * u[tf].sum()
    a, b = w + 1, N - w - 1
    u3 = u[a:b, a:b, a:b].sum()
    E1, H1 = g.line(0)[n].astype(np.float64), g.line(1)[n - 1].astype(np.float64)
    idx = np.arange(a, b) + E1.size - N
    u1 = (b - a) ** 2 * (eps0 * (E1[idx] ** 2).sum() + mu0 * (H1[idx] ** 2).sum())
    assert u1 > 0
    assert abs(u3 - u1) <= 1e-5 * u1


def test_error_paths():
    from maxwell_solver_amd import SimFDTDError, Simulation
    s = Simulation()
    with pytest.raises(SimFDTDError) as e:
        s.update_h_all()
    assert e.value.identifier == "simFDTD:InvalidInput"
    s.set_const([3e9, 0.01, 1.9e-11], [10, 40, 40, 40, 0])
    with pytest.raises(SimFDTDError) as e:
        s.set_const([3e9, 0.01, 1.9e-11], [10, 40, 40, 40, 0])
    assert e.value.identifier == "simFDTD:AlreadyInitialized"
    s.ready()
    with pytest.raises(SimFDTDError):
        s.update_h_planewave(0, 3)  # before Init1D
    s.update_h_planewave(0, 9)  # unknown modes are ignored, as the reference


@pytest.mark.parametrize("how", ["step", "run1", "run"])
def test_probe_record_is_the_field_after_each_step(how):
    """The GPU recorder against the GPU's own fields: record row n of component c and probe a is
    OutputVis(c) at the cell (P + x, P + y, P + z) after step n -- for single fused steps, runs of
    one step and one long run (whose slab-part launches gather the previous step's probes)."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=64, J=40, K=44, mode=5, cpml=CFG_CPML, oblique=(0.8, 2.0, 0.5))
    g, _ = _backends()
    S.setup(g, sc)
    P, comps = sc.P, sc.probe_comps
    loc = np.asarray(sc.probes)
    cells = (P + loc[:, 0]) + (P + loc[:, 1]) * sc.I + (P + loc[:, 2]) * sc.I * sc.J
    seen = []
    if how == "run":
        g.run(0, 25, 5)
        seen = [{c: g.output_vis(c)[cells] for c in comps}]
    else:
        for n in range(25):
            if how == "step":
                g.step(n, 5)
            else:
                g.run(n, 1, 5)
            seen.append({c: g.output_vis(c)[cells] for c in comps})
    rec = g.probes()
    rows = range(24, 25) if how == "run" else range(25)
    for q, c in enumerate(comps):
        for m, n in enumerate(rows):
            np.testing.assert_array_equal(rec[q, n], seen[m][c])
        assert not rec[q, 25:].any()
    assert np.abs(rec).max() > 0


def test_repeated_update_cuda_and_init_1d_reuse_device_buffers():
    """UpdateCuda copies into the material buffers it has (the reference copies into its arrays,
    src/simulation.h:1116-1127) and a repeated Init1D or SetAntennaLocations replaces its buffers:
    the device bytes stay constant and the results stay equal to the oracle's."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(mode=3, cpml=CFG_CPML)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.run(0, 10, 3)
    o.run(0, 10, 3)
    base = g.device_bytes()
    for _ in range(4):
        for b in (g, o):
            b.update_coeffs()
        assert g.device_bytes() == base
    for b in (g, o):
        b.init_1d(1.0, sc.source)
    assert g.device_bytes() == base
    loc = np.asarray(sc.probes, dtype=np.int32).T.reshape(-1)
    g.set_antenna_locations(loc)
    assert g.device_bytes() == base
    g.run(10, 20, 3)
    o.run(10, 20, 3)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes()[:, 10:], o.probes()[:, 10:])


# TF/SF box faces at or inside the CPML (inset <= P): the corrections then fall in slab rows, slab
# columns and slab planes, next to the inline CPML, in the same order as the per-op sequence
@pytest.mark.parametrize("inset,mode", [(10, 3), (6, 3), (10, 5), (4, 5), (8, 1)])
def test_tfsf_box_inside_the_cpml(inset, mode):
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=64, J=44, K=48, mode=mode, cpml=CFG_CPML, inset=inset, spheres=False,
                   oblique=(0.8, 2.0, 0.5) if mode == 5 else None)
    for fused in (False, True):
        g, o = _backends()
        S.setup(g, sc)
        S.setup(o, sc)
        if fused:
            g.run(0, sc.T, mode)
        else:
            _per_op_steps(g, sc.T, mode)
        o.run(0, sc.T, mode)
        assert_close(compare_state(g, o), exact=True)
        np.testing.assert_array_equal(g.probes(), o.probes())


@pytest.mark.parametrize("mode,graph", [(3, 8), (5, 8), (5, 2), (1, 16)])
def test_plane_wave_runs_replay_graphs(mode, graph):
    """Runs with a plane wave and probes replay a recorded graph of several fused steps (the time
    index of the face sheets and the probe record row read on the device): runs of several replays
    plus eager steps, single steps and runs between them, a recording per state set, and a run that
    ends on the last time index; fields, psi and every probe record equal the oracle's."""
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=64, J=40, K=44, T=60, mode=mode, cpml=CFG_CPML, oblique=(0.8, 2.0, 0.5) if mode == 5 else None)
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    g.set_graph_steps(graph)
    n = 0
    for kind, count in (("run", 21), ("step", 1), ("run", 17), ("run", 3), ("step", 1)):
        if kind == "run":
            g.run(n, count, mode)
        else:
            g.step(n, mode)
        n += count
    o.run(0, n, mode)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())
    g.run(n, sc.T - n, mode)  # to the last time index (no sheets past it)
    o.run(n, sc.T - n, mode)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


def test_size_class_tuning_table(tmp_path, monkeypatch):
    """fdtd_run consults the size-class table by default: an entry for the grid's class (from the
    JSON file FDTD_TUNE_CACHE names) replaces the automatic tuning; with policy 2 a class missing
    from the table is autotuned and written to the file; set_tuning overrides both; every run
    equals the oracle."""
    import json
    from maxwell_solver_amd import scenarios as S
    sc = _scenario(I=64, J=40, K=44, mode=5, cpml=CFG_CPML, oblique=(0.8, 2.0, 0.5))
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps({"classes": [{"I": 64, "log2J": 5, "log2K": 5, "P": 10, "index": 1, "plane_wave": 1,
                                             "mu": 0, "rows": 4, "z_chunk": 6, "cols_per_lane": -1}]}))
    monkeypatch.setenv("FDTD_TUNE_CACHE", str(path))
    g, o = _backends()
    S.setup(g, sc)
    S.setup(o, sc)
    assert g.get_tuning() == (0, 0, 0)
    g.run(0, 30, 5)
    assert g.get_tuning() == (4, 6, -1)
    g.set_tuning(2, 0, 1)  # the caller's choice wins over the table
    g.run(30, sc.T - 30, 5)
    assert g.get_tuning() == (2, 0, 1)
    o.run(0, sc.T, 5)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())
    # policy 2: a class not in the table (another I) is autotuned at its first run and saved
    sc2 = _scenario(I=68, J=40, K=44, mode=3, cpml=CFG_CPML)
    g2, o2 = _backends()
    S.setup(g2, sc2)
    S.setup(o2, sc2)
    g2.set_autotune(2)
    g2.run(0, sc2.T, 3)
    o2.run(0, sc2.T, 3)
    assert_close(compare_state(g2, o2), exact=True)
    saved = json.loads(path.read_text())["classes"]
    assert any(c["I"] == 68 and c["plane_wave"] == 1 for c in saved)
    assert any(c["I"] == 64 and c["rows"] == 4 for c in saved)
