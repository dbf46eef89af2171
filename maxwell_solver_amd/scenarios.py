"""Seeded synthetic workloads (grids, materials, plane waves, sources, probes).

The reference fixes no seeds; these generators use SplitMix64 with seed 0x5EED0000 + workload
number so that every run and every backend (HIP library, CPU oracle) sees the same bytes.
Workloads: 1 = 64^3 vacuum, axis plane wave; 2 = 256^3 with 40 dielectric spheres and an oblique
plane wave; 3 = 384x384x256 lossy layered ground under a vertical plane wave; 4 = 96 independent
grids of 32..384 cells per edge (the batched engine).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

C0 = 299792458.0
MU0 = 4 * 3.1415926535897 * 1e-7
EPS0 = 1 / C0 / C0 / MU0
MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo: int, hi: int) -> int:
        """Integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo

    def log_uniform(self, lo, hi) -> float:
        return math.exp(self.uniform(math.log(lo), math.log(hi)))


# ---------------------------------------------------------------------------------------------
# source waveforms (inputs to Init1D / SetSource)
# ---------------------------------------------------------------------------------------------
def gauss_sine(T, dt, f0, tau=None, t0=None):
    """Gaussian-modulated sinusoid exp(-((t-t0)/tau)^2) sin(2 pi f0 (t-t0)), t = n dt."""
    tau = 1.0 / f0 if tau is None else tau
    t0 = 4.0 * tau if t0 is None else t0
    t = np.arange(T, dtype=np.float64) * dt
    return (np.exp(-(((t - t0) / tau) ** 2)) * np.sin(2 * np.pi * f0 * (t - t0))).astype(np.float32)


def gauss_pulse(T, dt, fc):
    """Gaussian pulse exp(-((t-t0)/tau)^2) with tau = 1/(pi fc) (spectrum centred at 0, -3 dB near fc)."""
    tau = 1.0 / (math.pi * fc)
    t0 = 4.0 * tau
    t = np.arange(T, dtype=np.float64) * dt
    return np.exp(-(((t - t0) / tau) ** 2)).astype(np.float32)


def sinc_source(fctr, dt, bw, lobes, T):
    """The reference's windowed-sinc waveform (src/cpucode/FunLib.cpp:19-44)."""
    fmaxx = fctr * bw / 2.0
    spread = 1.0 / (2.0 * dt * fmaxx)
    tto = lobes * spread
    out = np.zeros(T)
    for n in range(T):
        nn = n + 1.0
        v = 0.0
        if nn < 2.0 * tto:
            arg = 2.0 * math.pi * fmaxx * (nn - tto) * dt
            den = 1e-14 if abs(arg) < 1e-14 else arg
            v = math.sin(arg) / den * math.sin(2.0 * math.pi * fctr * (nn - tto) * dt)
        v *= math.cos(((nn - 1.0) / T) * (math.pi / 2.0))
        out[n] = v
    return out.astype(np.float32)


# ---------------------------------------------------------------------------------------------
@dataclasses.dataclass
class Scenario:
    name: str
    I: int
    J: int
    K: int
    T: int
    dx: float
    dt: float
    fctr: float
    cpml: tuple            # [P, m, kappa_max, alpha_max, alpha_m, sigma_factor]
    inset: int
    mode: int              # plane-wave mode 1..7 (0: none)
    pw_params: tuple       # SetPlaneWave parameters
    source: np.ndarray     # PW source samples [T]
    mat_index: np.ndarray | None = None  # uint8 [K, J, I] (x fastest in memory)
    mat_table: np.ndarray | None = None  # float32 [n, 6] sigma x,y,z eps x,y,z
    probes: np.ndarray | None = None     # int32 [nAnt, 3] offsets from the CPML
    probe_comps: tuple = (3,)
    snap_every: int = 0
    snap_comp: int = 3
    seed: int = 0
    currents: tuple | None = None        # step currents (pol, box[6]): AddCurrents after every step
    mesh: tuple | None = None            # graded mesh: spacings along x, y, z (None: uniform dx)

    @property
    def cells(self) -> int:
        return self.I * self.J * self.K

    @property
    def P(self) -> int:
        return int(self.cpml[0])

    def material_arrays(self):
        """The six UpdateMaterial arrays (flags 1..6), flat x-fastest."""
        if self.mat_index is None:
            n = self.cells
            return [np.zeros(n, np.float32)] * 3 + [np.ones(n, np.float32)] * 3
        flat = self.mat_index.reshape(-1)
        return [self.mat_table[flat, c].astype(np.float32) for c in range(6)]


def _dt(dx, courant=0.99):
    return courant * dx / (math.sqrt(3.0) * C0)


CONFIG_CPML = (10, 3.0, 5.0, 0.05, 1.0, 0.8)


def _probes(rng, n, I, J, K, P):
    """n probe cells uniform in the interior (offsets from the CPML, as SetAntennaLocations)."""
    return np.array([[rng.randint(0, I - 2 * P - 1), rng.randint(0, J - 2 * P - 1), rng.randint(0, K - 2 * P - 1)]
                     for _ in range(n)], dtype=np.int32)


def config1(T: int = 600, N: int = 64) -> Scenario:
    """64^3 vacuum, 10-cell CPML (m=3, kappa_max=5, alpha_max=0.05), inset 15, x-propagating
    z-polarised plane wave (reference mode 3), Gaussian-modulated sine f0 = 3 GHz, 8 probes Ez/Hy."""
    seed = 0x5EED0000 + 1
    rng = SplitMix64(seed)
    dx = 0.01
    dt = _dt(dx)
    f0 = 3e9
    P = 10
    return Scenario("config1", N, N, N, T, dx, dt, f0, CONFIG_CPML, P + 5, 3, (0.0,),
                    gauss_sine(T, dt, f0), probes=_probes(rng, 8, N, N, N, P), probe_comps=(3, 5), seed=seed)


def spheres(rng, I, J, K, box, n, rmin, rmax, eps_rng, sig_rng):
    """n dielectric spheres, centres uniform inside the TF box (no material crosses the TF/SF
    surface); later spheres overwrite earlier ones.  Returns uint8 index [K,J,I] and table."""
    i0, i1, j0, j1, k0, k1 = box
    idx = np.zeros((K, J, I), dtype=np.uint8)
    table = [[0, 0, 0, 1, 1, 1]]
    z, y, x = np.ogrid[0:K, 0:J, 0:I]
    for m in range(n):
        r = rng.uniform(rmin, rmax)
        ri = int(math.ceil(r))
        cx = rng.uniform(i0 + ri + 1, i1 - ri - 1)
        cy = rng.uniform(j0 + ri + 1, j1 - ri - 1)
        cz = rng.uniform(k0 + ri + 1, k1 - ri - 1)
        er = rng.uniform(*eps_rng)
        sg = rng.uniform(*sig_rng)
        table.append([sg, sg, sg, er, er, er])
        zl, zh = max(0, int(cz - r) - 1), min(K, int(cz + r) + 2)
        yl, yh = max(0, int(cy - r) - 1), min(J, int(cy + r) + 2)
        xl, xh = max(0, int(cx - r) - 1), min(I, int(cx + r) + 2)
        sub = ((x[:, :, xl:xh] - cx) ** 2 + (y[:, yl:yh, :] - cy) ** 2 + (z[zl:zh] - cz) ** 2) <= r * r
        idx[zl:zh, yl:yh, xl:xh][sub] = m + 1
    return idx, np.asarray(table, dtype=np.float32)


def config2(T: int = 2000, N: int = 256, n_spheres: int = 40, n_probes: int = 64) -> Scenario:
    """256^3, 40 spheres (r in [4,20], eps_r in [2,12], sigma in [0,0.05] S/m) as uint8 index,
    oblique plane wave at seeded angles, 2000 steps, 64 probes (Ez, Hy), Ez snapshot every 250."""
    seed = 0x5EED0000 + 2
    rng = SplitMix64(seed)
    dx = 0.01
    dt = _dt(dx)
    f0 = 3e9
    P = 10
    inset = P + 5
    box = (inset, N - inset - 1) * 3
    rmax = min(20.0, (N - 2 * inset) / 2 - 3)
    rmin = min(4.0, rmax)
    idx, tab = spheres(rng, N, N, N, box, n_spheres, rmin, rmax, (2.0, 12.0), (0.0, 0.05))
    theta, phi, psi = rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi)
    return Scenario("config2", N, N, N, T, dx, dt, f0, CONFIG_CPML, inset, 5, (theta, phi, psi),
                    gauss_sine(T, dt, f0), mat_index=idx, mat_table=tab,
                    probes=_probes(rng, n_probes, N, N, N, P), probe_comps=(3, 5), snap_every=250,
                    snap_comp=3, seed=seed)


def graded_spacings(Q: int, P: int, dx: float, amp: float, phase: float = 0.0) -> np.ndarray:
    """Spacings of an axis of Q grid planes for fdtd_set_mesh: dx in the CPML cells (n <= P and
    n >= Q-P-2), between them a smooth grading from dx up to (1 + amp) dx (no cell finer than dx,
    so that the time step chosen for dx stays within the Courant limit)."""
    d = np.full(Q, np.float32(dx), np.float32)
    n = np.arange(P + 1, Q - P - 2)
    t = (n - (P + 1)) / max(1, Q - 2 * P - 4)
    w = np.abs(np.sin(math.pi * t + phase) * np.sin(3.1 * math.pi * t))
    d[P + 1:Q - P - 2] = (dx * (1 + amp * w)).astype(np.float32)
    return d


def config_graded(T: int = 2000, N: int = 256, graded: bool = True) -> Scenario:
    """The 256^3 sphere grid of config2 on a graded mesh (spacings from dx to 1.3 dx along every
    axis, dx in the CPML), driven by an x-directed line current through the centre (AddCurrents
    after every step, a Gaussian-modulated 3 GHz pulse) instead of a plane wave, which a graded
    mesh does not take; 64 probes (Ez, Hy).  graded=False: the same grid and source, uniform."""
    base = config2(T, N)
    P = base.cpml[0]
    mesh = ((graded_spacings(N, P, base.dx, 0.3), graded_spacings(N, P, base.dx, 0.25, 0.5),
             graded_spacings(N, P, base.dx, 0.2, 1.0)) if graded else None)
    c = N // 2
    return Scenario("graded" if graded else "graded-uniform", N, N, N, T, base.dx, base.dt, base.fctr, base.cpml,
                    base.inset, 0, (0.0,), gauss_sine(T, base.dt, base.fctr), mat_index=base.mat_index,
                    mat_table=base.mat_table, probes=base.probes, probe_comps=(3, 5), seed=base.seed,
                    currents=(1, (P + 2, N - P - 3, c, c, c, c)), mesh=mesh)


def config3(T: int = 4000, I: int = 384, J: int = 384, K: int = 256, lattice: int = 16) -> Scenario:
    """Lossy layered ground: five horizontal layers (20-80 cells, sigma log-uniform 1e-4..1 S/m,
    eps_r 4..30) under air, vertical x-polarised Gaussian pulse (reference mode 2, layered 1D line),
    dx for >= 10 cells per wavelength in the most conductive layer, 12-cell CPML, 16x16 surface
    probe lattice recording Ex and Hy."""
    seed = 0x5EED0000 + 3
    rng = SplitMix64(seed)
    P = 12
    fc = rng.uniform(1e6, 100e6)
    layers = []
    for _ in range(5):
        layers.append((rng.randint(20, 80), rng.log_uniform(1e-4, 1.0), rng.uniform(4.0, 30.0)))
    # wavelength in the most conductive layer at 2 fc (good-conductor skin-depth estimate vs
    # dielectric wavelength, whichever is shorter)
    f_hi = 2 * fc
    lam = []
    for _, sg, er in layers:
        w = 2 * math.pi * f_hi
        eps_c = complex(er * EPS0, -sg / w)
        kk = w * (MU0 * eps_c) ** 0.5
        lam.append(2 * math.pi / abs(kk.real))
    dx = min(lam) / 10.0
    dt = _dt(dx)
    inset = P + 5
    # layers from the bottom CPML upward, clipped to leave an air half-space above
    idx = np.zeros((K, J, I), dtype=np.uint8)
    table = [[0, 0, 0, 1, 1, 1]]
    ztop = K // 2
    z = ztop
    for m, (th, sg, er) in enumerate(layers):
        zl = max(0, z - th)
        idx[zl:z] = m + 1
        table.append([sg, sg, sg, er, er, er])
        z = zl
        if z == 0:
            break
    idx[:z] = idx[z] if z < K else 0
    tab = np.asarray(table, dtype=np.float32)
    # 16x16 probe lattice one cell above the ground surface
    xs = np.linspace(0, I - 2 * P - 1, lattice).astype(int)
    ys = np.linspace(0, J - 2 * P - 1, lattice).astype(int)
    zp = ztop - P + 1
    probes = np.array([[x, y, zp] for y in ys for x in xs], dtype=np.int32)
    return Scenario("config3", I, J, K, T, dx, dt, fc, (P, 3.0, 5.0, 0.05, 1.0, 0.8), inset, 2, (1.0,),
                    gauss_pulse(T, dt, fc), mat_index=idx, mat_table=tab, probes=probes, probe_comps=(1, 5),
                    seed=seed)


def config_gic(T: int = 300000) -> Scenario:
    """GICsimulation.m: geomagnetically induced currents.  dx = 10 km, dt = 0.99 dx / (sqrt(3) c),
    a 230 x 260 x 30 domain extended by a 10-cell CPML on every side (250 x 280 x 50 cells; the
    script's SetConst passes the unextended size, :56, while its material arrays are extended,
    :57-61 -- the extended grid is the one it describes), eps_r = 1, a layered and laterally
    graded conductivity (:70-94), no plane wave, and an x-directed line current along
    i = 1..249 at (j, k) = (139, 26) driven by a 100 kA Gaussian pulse (t0 = 5 s, FWHM 10 s)
    through AddCurrents after every step (:119-124)."""
    P = 10
    I, J, K = 230 + 2 * P, 260 + 2 * P, 30 + 2 * P
    dx = 10e3
    dt = _dt(dx)
    x1, x2 = 108, 135  # 1-based, :73-74
    sx = np.zeros((K, J, I), np.float64)
    sy = np.zeros_like(sx)
    sz = np.zeros_like(sx)
    for s in (sx, sy, sz):
        s[0:16, :, 0:x1] = 0.01586
        s[0:16, :, x2:] = 0.001
    for s in (sx, sy):
        s[16, :, 0:x1] = 0.01586 / 2
        s[16, :, x2:] = 0.001 / 2
    fact = 1.1
    for i in range(x1, x2):  # 1-based x1+1 .. x2
        sx[0:16, :, i] = 0.01586 / fact
        sy[0:16, :, i] = 0.01586 / fact
        sx[16, :, i] = 0.01586 / fact / 2
        sy[16, :, i] = 0.01586 / fact / 2
        sz[0:16, :, i] = 0.01586 / fact
        fact *= 1.1
    rows = np.stack([sx.reshape(-1), sy.reshape(-1), sz.reshape(-1)], axis=1).astype(np.float32)
    table, index = np.unique(rows, axis=0, return_inverse=True)
    mat_table = np.concatenate([table, np.ones((len(table), 3), np.float32)], axis=1)
    t = np.arange(T) * dt
    sig = 10.0 / 4 / 2.355
    source = (100e3 * np.exp(-(t - 5.0) ** 2 / (2 * sig ** 2))).astype(np.float32)
    return Scenario("gic", I, J, K, T, dx, dt, 5.5e9, (P, 3.0, 1.0, 0.08, 1.0, 0.8), 15, 0, (0.0,), source,
                    mat_index=index.reshape(K, J, I).astype(np.uint8), mat_table=mat_table.astype(np.float32),
                    currents=(1, (1, 249, 139, 139, 26, 26)))


def config4(n_grids: int = 96, edge_lo: float = 32, edge_hi: float = 192) -> list:
    """Many independent grids: base edge log-uniform in [32, 192], axes base * a_x, a_y, a_z with
    a in U[1, 2] normalised to a unit product (volume base^3, max/min <= 2), edges rounded to
    multiples of 4 (at least 32); 300..1500 steps; 0..20 spheres as in
    workload 2; 8-cell CPML; each grid its own oblique plane wave (seeded angles) and
    f0 in [1, 3] GHz; 4 probes recording Ez.  Returns (Scenario, steps) pairs."""
    seed = 0x5EED0000 + 4
    rng = SplitMix64(seed)
    P = 8
    cpml = (P, 3.0, 5.0, 0.05, 1.0, 0.8)
    dx = 0.01
    dt = _dt(dx)
    out = []
    for g in range(n_grids):
        base = rng.log_uniform(edge_lo, edge_hi)
        a = [rng.uniform(1.0, 2.0) for _ in range(3)]
        gm = (a[0] * a[1] * a[2]) ** (1.0 / 3.0)
        I, J, K = (max(32, 4 * int(round(base * x / gm / 4))) for x in a)
        steps = rng.randint(300, 1500)
        f0 = rng.uniform(1e9, 3e9)
        inset = P + 5
        box = (inset, I - inset - 1, inset, J - inset - 1, inset, K - inset - 1)
        half = min(box[1] - box[0], box[3] - box[2], box[5] - box[4]) / 2
        rmax = min(20.0, half - 3)
        n_sph = rng.randint(0, 20) if rmax >= 2.0 else 0
        idx = tab = None
        if n_sph:
            idx, tab = spheres(rng, I, J, K, box, n_sph, min(4.0, rmax), rmax, (2.0, 12.0), (0.0, 0.05))
        theta, phi, psi = rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi)
        sc = Scenario(f"config4_{g}", I, J, K, steps, dx, dt, f0, cpml, inset, 5, (theta, phi, psi),
                      gauss_sine(steps, dt, f0), mat_index=idx, mat_table=tab,
                      probes=_probes(rng, 4, I, J, K, P), probe_comps=(3,), seed=seed + g)
        out.append((sc, steps))
    return out


def _layered(rng, I, J, K, P, T, n_layers=5):
    """Horizontal lossy layers under air as in workload 3, thicknesses scaled to the grid height
    (U[20, 80] cells at K = 256), dx for >= 10 cells per wavelength in the most conductive layer at
    twice the pulse centre frequency.  Returns (dx, dt, fc, index, table, ground surface plane)."""
    fc = rng.uniform(1e6, 100e6)
    layers = []
    for _ in range(n_layers):
        th = max(2, int(round(rng.randint(20, 80) * K / 256.0)))
        layers.append((th, rng.log_uniform(1e-4, 1.0), rng.uniform(4.0, 30.0)))
    w = 2 * math.pi * 2 * fc
    lam = []
    for _, sg, er in layers:
        kk = w * (MU0 * complex(er * EPS0, -sg / w)) ** 0.5
        lam.append(2 * math.pi / abs(kk.real))
    dx = min(lam) / 10.0
    idx = np.zeros((K, J, I), dtype=np.uint8)
    table = [[0, 0, 0, 1, 1, 1]]
    ztop = K // 2
    z = ztop
    for m, (th, sg, er) in enumerate(layers):
        zl = max(0, z - th)
        idx[zl:z] = m + 1
        table.append([sg, sg, sg, er, er, er])
        z = zl
        if z == 0:
            break
    idx[:z] = idx[z] if z < K else 0
    return dx, _dt(dx), fc, idx, np.asarray(table, dtype=np.float32), ztop


def config5(n_requests: int = 2000, edge_lo: float = 48, edge_hi: float = 320, steps_lo: int = 500,
            steps_hi: int = 3000) -> list:
    """Served workload: a stream of solve requests.  Per request: base edge log-uniform in
    [edge_lo, edge_hi] with per-axis aspect factors in U[1, 1.25] normalised to a unit product
    (edges multiples of 4); steps uniform in [steps_lo, steps_hi]; materials with probability 0.5
    from the sphere generator of workload 2 (0..40 spheres, 10-cell CPML, oblique plane wave at
    seeded angles, Gaussian-modulated sine f0 in [1, 3] GHz, 8 probes Ez / Hy) and otherwise from
    the layered generator of workload 3 (12-cell CPML, vertical x-polarised Gaussian pulse through
    a layered 1D line, 8 probes Ex / Hy on the ground surface).  Returns (Scenario, steps, gap)
    triples: gap is an exponential inter-arrival time of unit mean (a Poisson stream once scaled by
    the mean request time the arrival rate is sized for)."""
    seed = 0x5EED0000 + 5
    rng = SplitMix64(seed)
    out = []
    for q in range(n_requests):
        base = rng.log_uniform(edge_lo, edge_hi)
        a = [rng.uniform(1.0, 1.25) for _ in range(3)]
        gm = (a[0] * a[1] * a[2]) ** (1.0 / 3.0)
        I, J, K = (max(48, 4 * int(round(base * x / gm / 4))) for x in a)
        steps = rng.randint(steps_lo, steps_hi)
        gap = -math.log(max(1e-12, 1.0 - rng.uniform()))
        if rng.uniform() < 0.5:
            P = 10
            cpml = (P, 3.0, 5.0, 0.05, 1.0, 0.8)
            dx = 0.01
            dt = _dt(dx)
            f0 = rng.uniform(1e9, 3e9)
            inset = P + 5
            box = (inset, I - inset - 1, inset, J - inset - 1, inset, K - inset - 1)
            half = min(box[1] - box[0], box[3] - box[2], box[5] - box[4]) / 2
            rmax = min(20.0, half - 3)
            n_sph = rng.randint(0, 40) if rmax >= 2.0 else 0
            idx = tab = None
            if n_sph:
                idx, tab = spheres(rng, I, J, K, box, n_sph, min(4.0, rmax), rmax, (2.0, 12.0), (0.0, 0.05))
            theta, phi, psi = rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi)
            sc = Scenario(f"serve_{q}_spheres", I, J, K, steps, dx, dt, f0, cpml, inset, 5, (theta, phi, psi),
                          gauss_sine(steps, dt, f0), mat_index=idx, mat_table=tab,
                          probes=_probes(rng, 8, I, J, K, P), probe_comps=(3, 5), seed=seed + q)
        else:
            P = 12
            cpml = (P, 3.0, 5.0, 0.05, 1.0, 0.8)
            dx, dt, fc, idx, tab, ztop = _layered(rng, I, J, K, P, steps)
            zp = max(0, min(K - 2 * P - 1, ztop - P + 1))
            probes = np.array([[rng.randint(0, I - 2 * P - 1), rng.randint(0, J - 2 * P - 1), zp] for _ in range(8)],
                              dtype=np.int32)
            sc = Scenario(f"serve_{q}_layers", I, J, K, steps, dx, dt, fc, cpml, P + 5, 2, (1.0,),
                          gauss_pulse(steps, dt, fc), mat_index=idx, mat_table=tab, probes=probes,
                          probe_comps=(1, 5), seed=seed + q)
        out.append((sc, steps, gap))
    return out


def small_variant(sc: Scenario, N: int, T: int) -> Scenario:
    """The same physics on an N^3 grid (materials regenerated, probes clipped) for parity tests."""
    d = dataclasses.replace(sc, I=N, J=N, K=N, T=T, source=sc.source[:T].copy())
    P = sc.P
    if sc.mat_index is not None:
        rng = SplitMix64(sc.seed ^ 0xABCDEF)
        box = (sc.inset, N - sc.inset - 1) * 3
        rmax = max(1.5, min(6.0, (N - 2 * sc.inset) / 2 - 3))
        idx, tab = spheres(rng, N, N, N, box, 6, 1.0, rmax, (2.0, 12.0), (0.0, 0.05))
        d.mat_index, d.mat_table = idx, tab
    if sc.probes is not None:
        rng = SplitMix64(sc.seed ^ 0x1234)
        d.probes = _probes(rng, min(8, len(sc.probes)), N, N, N, P)
    d.snap_every = 0
    return d


def setup_host(backend, sc: Scenario, use_arrays: bool = False):
    """The host part of setup(): every op before ReadySimulation (what a request submitted to the
    server carries; the server runs ReadySimulation and Init1D at admission)."""
    nAnt = 0 if sc.probes is None else len(sc.probes)
    backend.set_const([sc.fctr, sc.dx, sc.dt], [sc.T, sc.I, sc.J, sc.K, nAnt])
    backend.set_cpml(list(sc.cpml))
    backend.set_tfsf(sc.inset)
    backend.set_source(sc.source)
    if nAnt:
        backend.set_probe_components(list(sc.probe_comps))
        backend.set_antenna_locations(np.asarray(sc.probes, dtype=np.int32).T.reshape(-1))
    if sc.mat_index is not None:
        if use_arrays or not hasattr(backend, "set_material_index"):
            arrs = sc.material_arrays()
            for f in range(6):
                backend.update_material(f + 1, arrs[f])
        else:
            backend.set_material_index(sc.mat_index, sc.mat_table)
    if sc.mesh is not None:
        backend.set_mesh(*sc.mesh)
    if sc.mode == 5 or (sc.mode in (1, 2, 3, 4, 6, 7) and sc.pw_params and sc.pw_params[0]):
        backend.set_plane_wave(sc.mode, sc.pw_params)


def setup(backend, sc: Scenario, use_arrays: bool = False, init_line: bool = True, mu=None):
    """Apply the op sequence of highfrequency.m:63-84 (plus the extensions a workload needs) to a
    backend exposing the op methods (HIP Simulation or CPU oracle).  mu: three per-cell relative
    permeability arrays given by CreateMu before ReadySimulation."""
    nAnt = 0 if sc.probes is None else len(sc.probes)
    backend.set_const([sc.fctr, sc.dx, sc.dt], [sc.T, sc.I, sc.J, sc.K, nAnt])
    backend.set_cpml(list(sc.cpml))
    backend.set_tfsf(sc.inset)
    backend.set_source(sc.source)
    if nAnt:
        backend.set_probe_components(list(sc.probe_comps))
        backend.set_antenna_locations(np.asarray(sc.probes, dtype=np.int32).T.reshape(-1))
    if sc.mat_index is not None:
        if use_arrays or not hasattr(backend, "set_material_index"):
            arrs = sc.material_arrays()
            for f in range(6):
                backend.update_material(f + 1, arrs[f])
        else:
            backend.set_material_index(sc.mat_index, sc.mat_table)
    if sc.mesh is not None:
        backend.set_mesh(*sc.mesh)
    if sc.mode == 5 or (sc.mode in (1, 2, 3, 4, 6, 7) and sc.pw_params and sc.pw_params[0]):
        backend.set_plane_wave(sc.mode, sc.pw_params)
    if mu is not None:
        backend.create_mu(*mu)
    backend.ready()
    if init_line and sc.mode:
        backend.init_1d(1.0, sc.source)
    backend.update_coeffs()
    backend.reset()
    if sc.currents is not None and hasattr(backend, "set_step_currents"):
        backend.set_step_currents(sc.currents[0], list(sc.currents[1]))
