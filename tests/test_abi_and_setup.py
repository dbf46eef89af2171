"""CPU tests of the C ABI library: exports, error behaviour without a device, and bit-exact
agreement of its host set-up (CPML grading, 1/kappa, material coefficients, oblique tables)
with the CPU oracle."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "fdtd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fdtd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    from maxwell_solver_amd import simfdtd
    L = simfdtd.lib()
    syms = _header_symbols()
    assert len(syms) >= 38
    for s in syms:
        assert hasattr(L, s), f"missing export {s}"
    assert sorted(simfdtd.EXPORTS) == syms


def test_no_device_calls_fail_cleanly():
    from maxwell_solver_amd import SimFDTDError, Simulation, simFDTD
    s = Simulation()
    with pytest.raises(SimFDTDError) as e:
        s.update_e_all()
    assert e.value.identifier == "simFDTD:InvalidInput"
    h = simFDTD("CreateSimulation")
    with pytest.raises(SimFDTDError) as e:
        simFDTD("NoSuchOp", h)
    assert e.value.identifier == "simFDTD:UnknownOperation"
    with pytest.raises(SimFDTDError) as e:
        simFDTD("SetConst", h)
    assert e.value.identifier == "simFDTD:InvalidInput"
    simFDTD("SetConst", h, np.float32([3e9, 0.01, 1.9e-11]), np.int32([10, 30, 30, 30, 0]))
    with pytest.raises(SimFDTDError) as e:
        simFDTD("SetConst", h, np.float32([3e9, 0.01, 1.9e-11]), np.int32([10, 30, 30, 30, 0]))
    assert e.value.identifier == "simFDTD:AlreadyInitialized"
    simFDTD("DeleteSimulation", h)


def _pair(I, J, K, cpml, inset=15, mat=None, mode=None, prm=None, arrays=False):
    from maxwell_solver_amd import Simulation
    from oracle.oracle import OracleSim
    dx, dt = 0.01, 0.99 * 0.01 / (np.sqrt(3) * 299792458.0)
    g, o = Simulation(), OracleSim(np.float32)
    for b in (g, o):
        b.set_const([3e9, dx, dt], [20, I, J, K, 0])
        b.set_cpml(cpml)
        b.set_tfsf(inset)
        if mat is not None:
            if arrays or b is o:
                for f in range(6):
                    b.update_material(f + 1, mat[1][mat[0].reshape(-1), f])
            else:
                b.set_material_index(mat[0], mat[1])
        if mode is not None:
            b.set_plane_wave(mode, prm)
    g.prepare()
    o.ready()
    return g, o


CPMLS = [(10, 3, 1, 0.08, 1, 0.8), (10, 3, 5, 0.05, 1, 0.8), (12, 3, 5, 0.05, 1, 0.8), (8, 4, 2.5, 0.2, 1.5, 1.1)]


@pytest.mark.parametrize("cpml", CPMLS)
def test_cpml_profiles_bit_exact(cpml):
    g, o = _pair(36, 33, 40, cpml)
    for w in range(12):
        a = g.read_setup(w)
        b = o.setup(w).astype(np.float32)
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f"setup array {w}")
    sg, so = g.read_setup(12), o.scalars()
    assert sg[0] == np.float32(so["Chh"]) and sg[1] == np.float32(so["Che"]) and sg[2] == np.float32(so["sig_max"])


@pytest.mark.parametrize("arrays", [False, True])
def test_material_coefficients_bit_exact(arrays):
    from maxwell_solver_amd import scenarios as S
    rng = S.SplitMix64(11)
    I, J, K = 41, 37, 39
    idx, tab = S.spheres(rng, I, J, K, (15, 25, 15, 21, 15, 23), 7, 1.0, 4.0, (2.0, 12.0), (0.0, 0.05))
    g, o = _pair(I, J, K, CPMLS[1], mat=(idx, tab), arrays=arrays)
    ce, ch = o.coeffs()
    for c in range(3):
        np.testing.assert_array_equal(g.read_setup(13 + c), ce[c])
        np.testing.assert_array_equal(g.read_setup(16 + c), ch[c])
    assert np.array_equal(g.read_setup(19), idx.reshape(-1).astype(np.float32)) or arrays


def test_many_materials_use_per_cell_arrays():
    I, J, K = 34, 34, 34
    rng = np.random.default_rng(0)
    eps = (1 + rng.integers(0, 400, size=I * J * K) * 0.01).astype(np.float32)
    from maxwell_solver_amd import Simulation
    from oracle.oracle import OracleSim
    g, o = Simulation(), OracleSim(np.float32)
    for b in (g, o):
        b.set_const([3e9, 0.01, 1.9e-11], [5, I, J, K, 0])
        b.update_material(5, eps)
    g.prepare()
    o.ready()
    assert np.all(g.read_setup(19) == -1.0)  # per-cell mode
    ce, ch = o.coeffs()
    for c in range(3):
        np.testing.assert_array_equal(g.read_setup(13 + c), ce[c])
        np.testing.assert_array_equal(g.read_setup(16 + c), ch[c])


@pytest.mark.parametrize("angles", [(0.7, 0.4, 0.3), (2.2, 4.0, 1.1), (0.0, 0.0, 0.0), (np.pi / 2, np.pi, 2.0)])
def test_oblique_tables_bit_exact(angles):
    g, o = _pair(40, 44, 36, CPMLS[1], mode=5, prm=angles)
    o.init_1d(1.0, np.zeros(20, np.float32))
    sg, so = g.read_setup(12), o.scalars()
    assert int(sg[3]) == int(so["N1"]) and int(sg[12]) == int(so["src"])
    assert sg[11] == np.float32(so["ratio"])
    for w in list(range(20, 33)):
        np.testing.assert_array_equal(g.read_setup(w), o.setup(w).astype(np.float32), err_msg=f"table {w}")


def test_simfdtd_gateway_op_names():
    """Every op of src/simFDTD.cu:30-182 is accepted (case-insensitively) by the mirror."""
    from maxwell_solver_amd import SimFDTDError, simFDTD
    ops = ["CreateSimulation", "DeleteSimulation", "UpdateE_All", "UpdateH_All", "UpdateE_PML", "UpdateH_PML",
           "UpdateCuda", "AddCurrents", "ReadySimulation", "ResetFieldAll", "UpdateE_PlaneWave",
           "UpdateH_PlaneWave", "SetConst", "SetSource", "SetAntennaLocations", "Init1D", "UpdateMaterial",
           "OutputAnt", "OutputVis"]
    h = simFDTD("createsimulation")
    for op in ops[2:]:
        try:
            simFDTD(op.upper(), h)
        except SimFDTDError as e:
            assert e.identifier != "simFDTD:UnknownOperation", op
    simFDTD("deletesimulation", h)


def test_gic_workload_transcribes_the_script():
    """The GIC workload's ground model is GICsimulation.m:70-94 (1-based MATLAB ranges): sigma
    0.01586 S/m for x <= 108 and 0.001 beyond x = 135 in the 16 ground layers, halves in layer 17,
    the 27 columns between graded by 1.1^-n, air above; sigma_z not halved in layer 17."""
    from maxwell_solver_amd import scenarios as S
    sc = S.config_gic(T=8)
    assert (sc.I, sc.J, sc.K) == (250, 280, 50)
    sx, sy, sz = (sc.mat_table[sc.mat_index, c] for c in range(3))
    f32 = np.float32
    assert sx[0, 5, 0] == f32(0.01586) and sx[15, 5, 107] == f32(0.01586)
    assert sx[3, 7, 135] == f32(0.001) and sz[3, 7, 249] == f32(0.001)
    assert sx[16, 0, 0] == f32(0.01586 / 2) and sy[16, 0, 140] == f32(0.001 / 2)
    assert sz[16, 0, 0] == 0.0 and sx[17, 0, 0] == 0.0
    fact = 1.1
    for i in range(108, 135):
        assert sx[0, 9, i] == f32(0.01586 / fact) and sz[0, 9, i] == f32(0.01586 / fact)
        assert sy[16, 9, i] == f32(0.01586 / fact / 2)
        fact *= 1.1
    assert np.all(sc.mat_table[:, 3:] == 1.0)
    assert sc.currents == (1, (1, 249, 139, 139, 26, 26))
    assert abs(sc.dt - 0.99 * 10e3 / (np.sqrt(3) * S.C0)) < 1e-18


def test_new_entry_points_validate_before_the_device():
    """CreateMu, step currents, graph steps and plane-wave modes reject bad input on the host,
    before any device call (these run without a GPU)."""
    from maxwell_solver_amd import SimFDTDError, Simulation
    s = Simulation()
    n = 30 * 30 * 30
    ones = np.ones(n, np.float32)
    for call in (lambda: s.create_mu(ones, ones, ones), lambda: s.set_step_currents(1, [0, 1, 0, 1, 0, 1])):
        with pytest.raises(SimFDTDError) as e:
            call()  # before SetConst
        assert e.value.identifier == "simFDTD:InvalidInput"
    s.set_const([3e9, 0.01, 1.9e-11], [10, 30, 30, 30, 0])
    bad = ones.copy()
    bad[123] = 0.0
    with pytest.raises(SimFDTDError):
        s.create_mu(ones, bad, ones)  # mu must be positive
    bad[123] = np.nan
    with pytest.raises(SimFDTDError):
        s.create_mu(bad, ones, ones)
    s.create_mu(ones, ones, 2 * ones)  # taken before ReadySimulation (the fixed order)
    with pytest.raises(SimFDTDError):
        s.set_step_currents(1, [0, 1, 0, 1, 0, 1])  # before SetSource
    s.set_source(np.zeros(10, np.float32))
    with pytest.raises(SimFDTDError):
        s.set_step_currents(4, [0, 1, 0, 1, 0, 1])  # pol must be 0..3
    s.set_step_currents(1, [0, 1, 0, 1, 0, 1])
    s.set_step_currents(0)
    with pytest.raises(SimFDTDError):
        s.set_graph_steps(7)  # an odd count would end a replay on the other state set
    with pytest.raises(SimFDTDError):
        s.set_graph_steps(-2)
    s.set_graph_steps(16)
    s.set_graph_steps(0)
    for mode in (6, 7):
        s.set_plane_wave(mode, [0.0])
    with pytest.raises(SimFDTDError):
        s.set_plane_wave(8, [0.0])
    s.set_addressing(1)  # the row kernel's 64-bit global path on any grid
    s.set_addressing(0)
    with pytest.raises(SimFDTDError):
        s.set_addressing(2)
    s.close()
