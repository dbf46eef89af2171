"""Graded (tensor-product) mesh: fdtd_set_mesh and the oracle's set_mesh.

The reference allocates the den_ey/den_ez/den_hy/den_hz tables of a non-uniform mesh and never
reads them (src/simulation.h:2, 679-685, 718-724), so no reference arithmetic fixes the graded
update; the oracle's graded tables are pinned here by what the mathematics fixes:

* spacings all equal to dx reproduce the uniform grid bit for bit (the factors are exactly 1);
* the H update keeps the discrete divergence of H weighted by the primal spacings at zero, and the
  E update that of E weighted by the dual spacings (div curl = 0 holds for the graded operators
  only with the spacing each difference is taken over: a primal / dual swap breaks it);
* a lossless blob keeps the discrete Yee energy weighted by the primal / dual cell volumes of each
  component (the graded curls are adjoint only under those weights).

The GPU tests compare the HIP path with the oracle, bit for bit: the per-op operations with
AddCurrents, fused runs with step currents and probes, a batch, per-cell mu (per-op fallback) and
uniform spacings given explicitly against no mesh.
"""
import numpy as np
import pytest

from oracle.oracle import OracleError, OracleSim

C0 = 299792458.0
PI_REF = 3.1415926535897
MU0 = 4 * PI_REF * 1e-7
EPS0 = 1 / C0 / C0 / MU0
DX = 0.01
DT = 0.99 * DX / (np.sqrt(3) * C0)
DXF = np.float32(DX)


def graded(Q, P, amp=0.4, phase=0.0):
    """Spacings of an axis of Q planes: dx in the CPML cells (n <= P, n >= Q-P-2), a smooth
    grading from dx up to (1 + amp) dx inside (no cell finer than dx: the Courant limit holds)."""
    from maxwell_solver_amd.scenarios import graded_spacings
    return graded_spacings(Q, P, DX, amp, phase)


def _oracle(I, J, K, P, mesh=None, dtype=np.float64, T=40, inset=None):
    s = OracleSim(dtype)
    s.set_const([3e9, DX, DT], [T, I, J, K, 0])
    s.set_cpml([P, 3, 5, 0.05, 1, 0.8])
    s.set_tfsf(inset if inset is not None else P + 1)
    if mesh is not None:
        s.set_mesh(*mesh)
    s.set_source(np.sin(0.3 * np.arange(T)).astype(np.float32))
    s.ready()
    return s


def _dual(d):
    """Dual spacings (d[n-1] + d[n]) / 2 (entry 0: d[0])."""
    d = d.astype(np.float64)
    out = d.copy()
    out[1:] = 0.5 * (d[:-1] + d[1:])
    return out


def _random_interior(s, names, lo, seed):
    I, J, K = s.dims
    rng = np.random.default_rng(seed)
    for n in names:
        f = np.zeros((K, J, I))
        f[lo:K - lo, lo:J - lo, lo:I - lo] = rng.standard_normal((K - 2 * lo, J - 2 * lo, I - 2 * lo))
        s.set_field(n, f)


def test_uniform_spacings_equal_the_uniform_grid():
    I, J, K, P = 26, 24, 28, 4
    mesh = [np.full(q, DXF, np.float32) for q in (I, J, K)]
    runs = []
    for m in (None, mesh):
        s = _oracle(I, J, K, P, m, np.float32)
        runs.append(s)
    for w in range(6, 12):
        np.testing.assert_array_equal(runs[0].setup(w), runs[1].setup(w))
    for s in runs:
        _random_interior(s, ["Ex", "Ey", "Ez", "Hy"], 6, 3)
        for n in range(6):
            s.update_h_all(); s.update_h_pml(); s.update_e_all(); s.update_e_pml()
            s.add_currents(n, 3, [12, 12, 11, 12, 13, 14])
    for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(runs[0].field(c), runs[1].field(c))


def test_graded_tables_are_spacing_ratios():
    """The tables away from the CPML cells hold dx / spacing (H: primal, E: dual), within the
    CPML cells the uniform grid's 1/kappa."""
    I, J, K, P = 30, 28, 26, 4
    mesh = [graded(I, P), graded(J, P, 0.3, 0.5), graded(K, P, 0.25, 1.0)]
    s = _oracle(I, J, K, P, mesh)
    u = _oracle(I, J, K, P)
    for a, q in enumerate((I, J, K)):
        d = mesh[a].astype(np.float64)
        kH, kE = s.setup(9 + a), s.setup(6 + a)
        uH, uE = u.setup(9 + a), u.setup(6 + a)
        inner = slice(P + 1, q - P - 2)
        np.testing.assert_allclose(kH[inner], DXF / d[inner], rtol=1e-15)
        np.testing.assert_allclose(kE[P + 2:q - P - 2], DXF / _dual(d)[P + 2:q - P - 2], rtol=1e-15)
        np.testing.assert_array_equal(kH[:P + 1], uH[:P + 1])
        np.testing.assert_array_equal(kE[:P + 2], uE[:P + 2])
        np.testing.assert_array_equal(kH[q - P - 2:], uH[q - P - 2:])
        np.testing.assert_array_equal(kE[q - P - 1:], uE[q - P - 1:])


def test_graded_h_update_keeps_weighted_div_h_zero():
    I, J, K, P = 27, 25, 29, 4
    mesh = [graded(I, P), graded(J, P, 0.3, 0.5), graded(K, P, 0.35, 1.0)]
    s = _oracle(I, J, K, P, mesh)
    _random_interior(s, ["Ex", "Ey", "Ez"], 6, 1)
    s.update_h_all()
    s.update_h_pml()
    Hx, Hy, Hz = s.field("Hx"), s.field("Hy"), s.field("Hz")
    gx, gy, gz = (DX / mesh[a].astype(np.float64) for a in range(3))
    # div H at the cell centre (i+1/2, j+1/2, k+1/2), each difference over its primal cell
    dHx = (Hx[:-1, :-1, 1:] - Hx[:-1, :-1, :-1]) * gx[None, None, :-1]
    dHy = (Hy[:-1, 1:, :-1] - Hy[:-1, :-1, :-1]) * gy[None, :-1, None]
    dHz = (Hz[1:, :-1, :-1] - Hz[:-1, :-1, :-1]) * gz[:-1, None, None]
    div = dHx + dHy + dHz
    scale = max(np.abs(Hx).max(), np.abs(Hy).max(), np.abs(Hz).max())
    assert np.abs(div[5:-5, 5:-5, 5:-5]).max() <= 1e-13 * scale
    # the same divergence without the spacings (a uniform-grid operator) is far from zero
    plain = (Hx[:-1, :-1, 1:] - Hx[:-1, :-1, :-1]) + (Hy[:-1, 1:, :-1] - Hy[:-1, :-1, :-1]) + (Hz[1:, :-1, :-1] - Hz[:-1, :-1, :-1])
    assert np.abs(plain[5:-5, 5:-5, 5:-5]).max() > 1e-3 * scale


def test_graded_e_update_keeps_weighted_div_e_zero():
    I, J, K, P = 27, 25, 29, 4
    mesh = [graded(I, P), graded(J, P, 0.3, 0.5), graded(K, P, 0.35, 1.0)]
    s = _oracle(I, J, K, P, mesh)
    _random_interior(s, ["Hx", "Hy", "Hz"], 6, 2)
    s.update_e_all()
    s.update_e_pml()
    Ex, Ey, Ez = s.field("Ex"), s.field("Ey"), s.field("Ez")
    hx, hy, hz = (DX / _dual(mesh[a]) for a in range(3))
    # div E at the node (i, j, k), each backward difference over its dual cell
    div = ((Ex[1:, 1:, 1:] - Ex[1:, 1:, :-1]) * hx[None, None, 1:] + (Ey[1:, 1:, 1:] - Ey[1:, :-1, 1:]) * hy[None, 1:, None]
           + (Ez[1:, 1:, 1:] - Ez[:-1, 1:, 1:]) * hz[1:, None, None])
    scale = max(np.abs(Ex).max(), np.abs(Ey).max(), np.abs(Ez).max())
    assert np.abs(div[5:-5, 5:-5, 5:-5]).max() <= 1e-13 * scale
    # the primal spacings in place of the dual ones do not cancel
    px, py, pz = (DX / mesh[a].astype(np.float64) for a in range(3))
    wrong = ((Ex[1:, 1:, 1:] - Ex[1:, 1:, :-1]) * px[None, None, 1:] + (Ey[1:, 1:, 1:] - Ey[1:, :-1, 1:]) * py[None, 1:, None]
             + (Ez[1:, 1:, 1:] - Ez[:-1, 1:, 1:]) * pz[1:, None, None])
    assert np.abs(wrong[5:-5, 5:-5, 5:-5]).max() > 1e-3 * scale


def _volumes(mesh):
    """Cell volumes [k, j, i] of each component: the component's own axis primal, the others dual
    for E; the opposite for H."""
    p = [m.astype(np.float64) for m in mesh]
    d = [_dual(m) for m in mesh]
    def vol(ax, ay, az):
        return az[:, None, None] * ay[None, :, None] * ax[None, None, :]
    VE = [vol(p[0], d[1], d[2]), vol(d[0], p[1], d[2]), vol(d[0], d[1], p[2])]
    VH = [vol(d[0], p[1], p[2]), vol(p[0], d[1], p[2]), vol(p[0], p[1], d[2])]
    return VE, VH


def _energies(s, VE, VH, steps):
    out = []
    for _ in range(steps):
        E2 = sum(np.sum(s.field(c) ** 2 * v) for c, v in zip(("Ex", "Ey", "Ez"), VE))
        Hm = [s.field(c).copy() for c in ("Hx", "Hy", "Hz")]
        s.update_h_all()
        s.update_h_pml()
        Hp = [s.field(c) for c in ("Hx", "Hy", "Hz")]
        out.append(EPS0 * E2 + MU0 * sum(np.sum(a * b * v) for a, b, v in zip(Hm, Hp, VH)))
        s.update_e_all()
        s.update_e_pml()
    return np.array(out)


def test_graded_yee_energy_conserved_without_loss():
    """A Gaussian E blob in vacuum on a graded mesh keeps the volume-weighted discrete energy
    eps0 sum V_E |E^n|^2 + mu0 sum V_H H^{n-1/2} H^{n+1/2} constant while it stays clear of the
    CPML; weighting with uniform volumes instead does not."""
    N, P = 64, 4
    mesh = [graded(N, P, 0.3), graded(N, P, 0.25, 0.5), graded(N, P, 0.2, 1.0)]
    z, y, x = np.mgrid[0:N, 0:N, 0:N]
    blob = np.exp(-((x - 32) ** 2 + (y - 31) ** 2 + (z - 33) ** 2) / 4.0)
    VE, VH = _volumes(mesh)
    ones = [np.ones((N, N, N))] * 3
    res = {}
    for name, (ve, vh) in (("graded", (VE, VH)), ("uniform weights", (ones, ones))):
        s = _oracle(N, N, N, P, mesh)
        s.set_field("Ex", blob)
        s.set_field("Ey", 0.5 * np.roll(blob, 1, axis=0))
        res[name] = _energies(s, ve, vh, 16)
    g, u = res["graded"], res["uniform weights"]
    assert np.abs(g / g[0] - 1).max() < 1e-12
    assert np.abs(u / u[0] - 1).max() > 1e-4


def test_mesh_validation_oracle():
    I, J, K, P = 26, 24, 28, 4
    s = OracleSim(np.float64)
    with pytest.raises(ValueError):
        s.set_mesh(np.full(I, DXF), None, None)  # before SetConst (no axis lengths yet)
    s.set_const([3e9, DX, DT], [20, I, J, K, 0])
    s.set_cpml([P, 3, 5, 0.05, 1, 0.8])
    s.set_tfsf(P + 1)
    bad = np.full(I, DXF)
    bad[7] = 0.0
    with pytest.raises(OracleError, match="positive"):
        s.set_mesh(bad, None, None)
    cp = np.full(J, DXF)
    cp[P] = np.float32(1.1 * DX)  # a CPML cell
    s.set_mesh(None, cp, None)
    with pytest.raises(OracleError, match="CPML"):
        s.ready()
    s2 = _oracle(I, J, K, P, [graded(I, P), None, None])
    with pytest.raises(OracleError, match="plane wave"):
        s2.update_h_planewave(0, 3)
    fine = np.full(K, DXF)
    fine[P + 3] = np.float32(0.8 * DX)  # finer than the cells dt was chosen for
    s3 = OracleSim(np.float64)
    s3.set_const([3e9, DX, DT], [20, I, J, K, 0])
    s3.set_cpml([P, 3, 5, 0.05, 1, 0.8])
    s3.set_tfsf(P + 1)
    s3.set_mesh(None, None, fine)
    with pytest.raises(OracleError, match="Courant"):
        s3.ready()


def test_host_tables_equal_oracle_and_validation():
    """The library's host set-up (fdtd_prepare, no GPU) builds the same graded tables as the
    oracle, bit for bit, and refuses the same inputs."""
    from maxwell_solver_amd import Simulation, SimFDTDError
    I, J, K, P = 30, 28, 26, 4
    mesh = [graded(I, P), graded(J, P, 0.3, 0.5), graded(K, P, 0.25, 1.0)]
    o = _oracle(I, J, K, P, mesh, np.float32)
    g = Simulation()
    g.set_const([3e9, DX, DT], [40, I, J, K, 0])
    g.set_cpml([P, 3, 5, 0.05, 1, 0.8])
    g.set_tfsf(P + 1)
    g.set_mesh(*mesh)
    g.prepare()
    for w in range(6, 12):
        np.testing.assert_array_equal(g.read_setup(w), o.setup(w))
    g.close()
    g = Simulation()
    g.set_const([3e9, DX, DT], [40, I, J, K, 0])
    g.set_cpml([P, 3, 5, 0.05, 1, 0.8])
    g.set_tfsf(P + 1)
    bad = mesh[0].copy()
    bad[3] = np.float32(-DX)
    with pytest.raises(SimFDTDError, match="positive"):
        g.set_mesh(bad, None, None)
    cp = mesh[0].copy()
    cp[I - 2] = np.float32(2 * DX)  # the last-but-one cell belongs to the CPML
    g.set_mesh(cp, None, None)
    with pytest.raises(SimFDTDError, match="CPML"):
        g.prepare()
    fine = mesh[1].copy()
    fine[P + 3] = np.float32(0.8 * DX)
    g.set_mesh(None, fine, None)
    with pytest.raises(SimFDTDError, match="Courant"):
        g.prepare()
    g.close()


# ---------------------------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------------------------

def _gpu_pair(I, J, K, P, mesh, T=60, mu=None, probes=True):
    from maxwell_solver_amd import Simulation
    idx = np.zeros((K, J, I), np.uint8)
    idx[K // 3:2 * K // 3, :, :] = 1
    idx[:, J // 2:, I // 3:] = 2
    tab = np.array([[0, 0, 0, 1, 1, 1], [0.02, 0.02, 0.02, 4, 4, 4], [0.3, 0.3, 0.3, 9, 9, 9]], np.float32)
    src = (np.sin(np.arange(T) * 0.37) * np.exp(-((np.arange(T) - 20) / 9.0) ** 2)).astype(np.float32)
    loc = np.array([[P + 3, P + 4, P + 5], [I - P - 6, J // 2, K // 2], [I // 2, J - P - 5, K - P - 4]], np.int32)
    pair = []
    for b in (Simulation(), OracleSim(np.float32)):
        b.set_const([3e9, DX, DT], [T, I, J, K, len(loc) if probes else 0])
        b.set_cpml([P, 3, 5, 0.05, 1, 0.8])
        b.set_tfsf(P + 1)
        b.set_mesh(*mesh)
        b.set_source(src)
        for f in range(6):
            b.update_material(f + 1, tab[idx.reshape(-1), f])
        if mu is not None:
            b.create_mu(*mu)
        if probes:
            b.set_antenna_locations(loc.T.reshape(-1))  # column-major: locations[a + c nAnt]
            b.set_probe_components([3, 5, 1])
        b.ready()
        pair.append(b)
    return pair


def _mesh3(I, J, K, P):
    return [graded(I, P), graded(J, P, 0.3, 0.5), graded(K, P, 0.35, 1.0)]


@pytest.mark.gpu
def test_gpu_graded_per_op_with_currents():
    from tests.fdtd_compare import assert_close, compare_state
    I, J, K, P = 48, 44, 40, 6
    g, o = _gpu_pair(I, J, K, P, _mesh3(I, J, K, P), probes=False)
    box = [20, 22, 18, 19, 17, 21]
    for b in (g, o):
        for n in range(40):
            b.update_h_all(); b.update_h_pml(); b.update_e_all(); b.update_e_pml()
            b.add_currents(n, 3, box)
    assert_close(compare_state(g, o), exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("I", [64, 50, 130])
def test_gpu_graded_fused_run_with_step_currents(I):
    """Fused runs (the row kernel's full-slab march on every row; I = 64 split segment, 50 a
    partial segment, 130 three segments) with step currents and probes, against the oracle's op
    sequence with AddCurrents after every step."""
    from tests.fdtd_compare import assert_close, compare_state
    J, K, P = 42, 46, 6
    g, o = _gpu_pair(I, J, K, P, _mesh3(I, J, K, P))
    box = [I // 2 - 2, I // 2 + 1, 20, 21, 22, 24]
    g.set_step_currents(2, box)
    g.run(0, 33, 0)
    g.step(33, 0)
    g.run(34, 26, 0)
    for n in range(60):
        o.update_h_all(); o.update_h_pml(); o.update_e_all(); o.update_e_pml()
        o.add_currents(n, 2, box)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())


@pytest.mark.gpu
def test_gpu_graded_fused_probes_and_batch():
    """A graded grid in a batch (it runs on its own), fields and probes against the oracle's op
    sequence (the probes sample each step's E before its currents)."""
    from maxwell_solver_amd import Batch
    from tests.fdtd_compare import assert_close, compare_state
    I, J, K, P = 56, 40, 44, 6
    g, o = _gpu_pair(I, J, K, P, _mesh3(I, J, K, P))
    box = [26, 28, 18, 19, 20, 22]
    g.set_step_currents(1, box)
    bt = Batch()
    bt.add(g, 0, 60, 0)
    bt.run()
    g.synchronize()
    for n in range(60):
        o.update_h_all(); o.update_h_pml(); o.update_e_all(); o.update_e_pml()
        o.add_currents(n, 1, box)
    assert_close(compare_state(g, o), exact=True)
    np.testing.assert_array_equal(g.probes(), o.probes())
    bt.close()


@pytest.mark.gpu
def test_gpu_graded_with_mu_takes_per_op_path():
    from tests.fdtd_compare import assert_close, compare_state
    I, J, K, P = 40, 36, 38, 6
    rng = np.random.default_rng(23)
    mu = [rng.uniform(1.0, 3.0, I * J * K).astype(np.float32) for _ in range(3)]
    g, o = _gpu_pair(I, J, K, P, _mesh3(I, J, K, P), mu=mu, probes=False)
    box = [18, 20, 17, 18, 18, 20]
    g.set_step_currents(3, box)
    g.run(0, 30, 0)
    for n in range(30):
        o.update_h_all(); o.update_h_pml(); o.update_e_all(); o.update_e_pml()
        o.add_currents(n, 3, box)
    assert_close(compare_state(g, o), exact=True)


@pytest.mark.gpu
def test_gpu_uniform_spacings_equal_no_mesh():
    """Spacings all dx through fdtd_set_mesh (the full-slab march everywhere) equal the uniform
    grid's fused run (its plain, edge and slab marches), bit for bit."""
    from tests.fdtd_compare import compare_state, assert_close
    I, J, K, P = 64, 40, 44, 6
    mesh = [np.full(q, DXF, np.float32) for q in (I, J, K)]
    g1, _ = _gpu_pair(I, J, K, P, mesh)
    g0, _ = _gpu_pair(I, J, K, P, [None, None, None])
    box = [30, 33, 18, 19, 20, 22]
    for g in (g0, g1):
        g.set_step_currents(2, box)
        g.run(0, 50, 0)
    assert_close(compare_state(g1, g0), exact=True)
    np.testing.assert_array_equal(g1.probes(), g0.probes())


@pytest.mark.gpu
def test_gpu_graded_refuses_plane_wave():
    from maxwell_solver_amd import SimFDTDError
    I, J, K, P = 40, 36, 38, 6
    g, _ = _gpu_pair(I, J, K, P, _mesh3(I, J, K, P), probes=False)
    with pytest.raises(SimFDTDError, match="plane wave"):
        g.run(0, 3, 3)


@pytest.mark.gpu
def test_gpu_graded_request_served_matches_oracle():
    """A graded-mesh request (the graded sphere workload at 48^3, a line current after every
    step) served by the request server beside a uniform one: fields and probes equal the oracle's
    op sequence, bit for bit."""
    from maxwell_solver_amd import Server, Simulation
    from maxwell_solver_amd import scenarios as S
    from tests.fdtd_compare import assert_close, compare_state
    srv = Server(hbm_budget=64 << 30, max_active=2, steps_per_round=16)
    cases = {}
    for graded in (True, False):
        sc = S.config_graded(T=50, N=48, graded=graded)
        sim = Simulation()
        S.setup_host(sim, sc)
        sim.set_step_currents(sc.currents[0], list(sc.currents[1]))
        cases[srv.submit(sim, 50, 0, source=sc.source)] = sc
    done = []
    while len(done) < len(cases):
        srv.round()
        done += srv.collect()
    for rid, *_ in done:
        sc = cases[rid]
        o = OracleSim(np.float32)
        S.setup(o, sc, init_line=False)
        pol, box = sc.currents[0], list(sc.currents[1])
        for n in range(50):
            o.update_h_all(); o.update_h_pml(); o.update_e_all(); o.update_e_pml()
            o.add_currents(n, pol, box)
        g = srv.simulation(rid)
        assert_close(compare_state(g, o), exact=True)
        np.testing.assert_array_equal(g.probes(), o.probes())
        assert np.abs(g.field("Ex")).max() > 0
        srv.release(rid)
    srv.close()


@pytest.mark.gpu
def test_gpu_graded_graph_replay_with_step_currents():
    """Runs of a graded grid replayed from recorded HIP graphs (8 steps each), with eager steps
    between the replays, equal to the oracle's op sequence."""
    from tests.fdtd_compare import assert_close, compare_state
    I, J, K, P = 48, 40, 44, 6
    g, o = _gpu_pair(I, J, K, P, _mesh3(I, J, K, P), probes=False)
    box = [20, 26, 18, 18, 21, 21]
    g.set_step_currents(1, box)
    g.set_graph_steps(8)
    g.run(0, 21, 0)
    g.step(21, 0)
    g.run(22, 18, 0)
    for n in range(40):
        o.update_h_all(); o.update_h_pml(); o.update_e_all(); o.update_e_pml()
        o.add_currents(n, 1, box)
    assert_close(compare_state(g, o), exact=True)
