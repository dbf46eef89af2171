"""Python mirror of the reference's MATLAB MEX operator API over the C ABI of libfdtd_hip.so.

`simFDTD(op, handle, ...)` keeps the gateway's op names (case-insensitive), argument order and
error identifiers (src/simFDTD.cu:19-187); `Simulation` is the object form used by the tests and
the bench.  The library is loaded from this package directory; there is no fallback: a missing
library raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfdtd_hip.so")
_lib = None

# Every symbol the header declares (include/fdtd.h).
EXPORTS = [
    "fdtd_last_error", "fdtd_last_error_id", "fdtd_version", "fdtd_create", "fdtd_destroy",
    "fdtd_set_const", "fdtd_set_source", "fdtd_set_antenna_locations", "fdtd_update_material", "fdtd_create_mu",
    "fdtd_set_step_currents", "fdtd_set_graph_steps",
    "fdtd_ready", "fdtd_init_1d", "fdtd_update_coeffs", "fdtd_reset_fields", "fdtd_update_h_all",
    "fdtd_update_e_all", "fdtd_update_h_pml", "fdtd_update_e_pml", "fdtd_update_h_planewave",
    "fdtd_update_e_planewave", "fdtd_add_currents", "fdtd_output_ant", "fdtd_output_vis",
    "fdtd_set_cpml", "fdtd_set_tfsf", "fdtd_set_plane_wave", "fdtd_set_material_index",
    "fdtd_set_probe_components", "fdtd_step", "fdtd_run", "fdtd_snapshot_count",
    "fdtd_output_snapshots", "fdtd_synchronize", "fdtd_read_array", "fdtd_prepare",
    "fdtd_read_setup", "fdtd_set_tuning", "fdtd_autotune", "fdtd_batch_create", "fdtd_batch_destroy",
    "fdtd_batch_add", "fdtd_batch_clear", "fdtd_batch_run", "fdtd_batch_set_strategy", "fdtd_set_timing", "fdtd_kernel_time", "fdtd_row_kernel_time",
    "fdtd_server_create", "fdtd_server_destroy", "fdtd_server_now", "fdtd_server_submit", "fdtd_server_round",
    "fdtd_server_collect", "fdtd_server_release", "fdtd_server_request_bytes", "fdtd_server_stats",
    "fdtd_server_request_error", "fdtd_device_bytes", "fdtd_run_noalloc",
    "fdtd_stream_ceiling", "fdtd_get_tuning", "fdtd_set_autotune", "fdtd_set_slab_chunk", "fdtd_set_addressing",
    "fdtd_set_mesh",
]


class SimFDTDError(RuntimeError):
    """Raised with the MEX error identifier of the reference gateway."""

    def __init__(self, identifier: str, message: str):
        super().__init__(f"{identifier}: {message}")
        self.identifier = identifier


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run maxwell_solver_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, i32, u32, u64 = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64
        fp, ip, dp, u8p, u64p = (C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                 C.POINTER(C.c_uint8), C.POINTER(C.c_uint64))
        sig = {
            "fdtd_create": [C.POINTER(vp)], "fdtd_destroy": [vp],
            "fdtd_set_const": [vp, fp, ip], "fdtd_set_source": [vp, fp, i32],
            "fdtd_set_antenna_locations": [vp, ip], "fdtd_update_material": [vp, i32, fp], "fdtd_create_mu": [vp, fp, fp, fp],
            "fdtd_set_step_currents": [vp, i32, ip], "fdtd_set_graph_steps": [vp, i32],
            "fdtd_ready": [vp], "fdtd_init_1d": [vp, C.c_float, fp], "fdtd_update_coeffs": [vp],
            "fdtd_reset_fields": [vp], "fdtd_update_h_all": [vp], "fdtd_update_e_all": [vp],
            "fdtd_update_h_pml": [vp], "fdtd_update_e_pml": [vp],
            "fdtd_update_h_planewave": [vp, ip], "fdtd_update_e_planewave": [vp, ip],
            "fdtd_add_currents": [vp, ip, ip], "fdtd_output_ant": [vp, fp, u32],
            "fdtd_output_vis": [vp, fp, u32, i32], "fdtd_set_cpml": [vp, dp], "fdtd_set_tfsf": [vp, i32],
            "fdtd_set_plane_wave": [vp, i32, dp], "fdtd_set_material_index": [vp, u8p, fp, i32],
            "fdtd_set_probe_components": [vp, ip, i32], "fdtd_step": [vp, i32, i32],
            "fdtd_run": [vp, i32, i32, i32, i32, i32], "fdtd_snapshot_count": [vp, ip],
            "fdtd_output_snapshots": [vp, fp, u64], "fdtd_synchronize": [vp],
            "fdtd_read_array": [vp, i32, fp, u64p], "fdtd_prepare": [vp],
            "fdtd_read_setup": [vp, i32, fp, u64p], "fdtd_set_tuning": [vp, i32, i32, i32],
            "fdtd_set_timing": [vp, i32], "fdtd_kernel_time": [vp, dp, C.POINTER(C.c_int64)],
            "fdtd_row_kernel_time": [vp, dp], "fdtd_autotune": [vp, i32, i32, i32, dp],
            "fdtd_batch_create": [C.POINTER(vp)], "fdtd_batch_destroy": [vp],
            "fdtd_batch_add": [vp, vp, i32, i32, i32], "fdtd_batch_clear": [vp], "fdtd_batch_run": [vp],
            "fdtd_batch_set_strategy": [vp, i32],
            "fdtd_server_create": [C.POINTER(vp), u64, i32, i32], "fdtd_server_destroy": [vp],
            "fdtd_server_now": [vp, dp],
            "fdtd_server_submit": [vp, vp, C.c_float, fp, i32, i32, C.c_double, C.POINTER(C.c_int64)],
            "fdtd_server_round": [vp, ip, ip],
            "fdtd_server_collect": [vp, C.POINTER(C.c_int64), dp, i32, ip],
            "fdtd_server_release": [vp, C.c_int64],
            "fdtd_server_stats": [vp, u64p, u64p, C.POINTER(C.c_int64)],
            "fdtd_server_request_bytes": [vp, C.c_int64, u64p],
            "fdtd_server_request_error": [vp, C.c_int64, ip, C.c_char_p, i32],
            "fdtd_device_bytes": [vp, u64p],
            "fdtd_run_noalloc": [vp, i32, i32, i32],
            "fdtd_stream_ceiling": [vp, i32, dp],
            "fdtd_get_tuning": [vp, ip],
            "fdtd_set_autotune": [vp, i32],
            "fdtd_set_slab_chunk": [vp, i32], "fdtd_set_addressing": [vp, i32],
            "fdtd_set_mesh": [vp, fp, fp, fp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        for name in ("fdtd_last_error", "fdtd_last_error_id", "fdtd_version"):
            getattr(L, name).restype = C.c_char_p
            getattr(L, name).argtypes = []
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


FIELD = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}


class Simulation:
    """One GPU simulation (reference: class simulation, src/simulation.h:24)."""

    def __init__(self):
        h = C.c_void_p()
        self._chk(lib().fdtd_create(C.byref(h)))
        self._h = h
        self.dims = None
        self.T = 0
        self.nAnt = 0
        self.ncomp = 1

    @staticmethod
    def _chk(rc):
        if rc:
            L = lib()
            raise SimFDTDError(L.fdtd_last_error_id().decode(), L.fdtd_last_error().decode())

    def close(self):
        if getattr(self, "_h", None) and getattr(self, "_owned", True):
            lib().fdtd_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- ops of src/simFDTD.cu ----
    def set_const(self, freqRes, timeDim):
        fr, td = _f32(freqRes), _i32(timeDim)
        self._chk(lib().fdtd_set_const(self._h, _p(fr, C.c_float), _p(td, C.c_int32)))
        self.T = int(td[0])
        self.dims = (int(td[1]), int(td[2]), int(td[3]))
        self.nAnt = int(td[4])

    def set_cpml(self, params):
        p = _f64(params)
        self._chk(lib().fdtd_set_cpml(self._h, _p(p, C.c_double)))

    def set_mesh(self, dx=None, dy=None, dz=None):
        """Graded mesh: distances between consecutive grid planes along x, y, z (None: uniform)."""
        ptrs = []
        self._mesh = []
        for q, d in enumerate((dx, dy, dz)):
            if d is None:
                ptrs.append(None)
                continue
            a = np.ascontiguousarray(d, dtype=np.float32)
            if self.dims is None or a.size != self.dims[q]:
                raise ValueError("set_mesh: one spacing per grid plane of the axis (after set_const)")
            self._mesh.append(a)
            ptrs.append(_p(a, C.c_float))
        self._chk(lib().fdtd_set_mesh(self._h, *ptrs))

    def set_tfsf(self, inset):
        self._chk(lib().fdtd_set_tfsf(self._h, int(inset)))

    def set_source(self, src, update=False):
        s = _f32(src)
        self._chk(lib().fdtd_set_source(self._h, _p(s, C.c_float), 1 if update else 0))

    def set_antenna_locations(self, loc):
        a = _i32(loc)
        self._chk(lib().fdtd_set_antenna_locations(self._h, _p(a, C.c_int32)))

    def set_probe_components(self, comps):
        c = _i32(comps)
        self._chk(lib().fdtd_set_probe_components(self._h, _p(c, C.c_int32), len(c)))
        self.ncomp = len(c)

    def update_material(self, flag, values):
        v = _f32(values)
        self._chk(lib().fdtd_update_material(self._h, int(flag), _p(v, C.c_float)))

    def create_mu(self, mu_x, mu_y, mu_z):
        """CreateMu: per-cell relative permeabilities (three arrays of I*J*K floats)."""
        m = [_f32(v) for v in (mu_x, mu_y, mu_z)]
        self._chk(lib().fdtd_create_mu(self._h, *(_p(v, C.c_float) for v in m)))

    def set_step_currents(self, pol, box=(0, 0, 0, 0, 0, 0)):
        """AddCurrents(n, pol, box) after every step of step() / run() (pol 0: off)."""
        b = _i32(box)
        self._chk(lib().fdtd_set_step_currents(self._h, int(pol), _p(b, C.c_int32)))

    def set_graph_steps(self, steps):
        """Record `steps` fused steps into one HIP graph and replay it in qualifying runs (0: off)."""
        self._chk(lib().fdtd_set_graph_steps(self._h, int(steps)))

    def set_material_index(self, index, table):
        idx = np.ascontiguousarray(np.asarray(index, dtype=np.uint8).reshape(-1))
        tab = np.ascontiguousarray(np.asarray(table, dtype=np.float32).reshape(-1, 6))
        self._chk(lib().fdtd_set_material_index(self._h, _p(idx, C.c_uint8), _p(tab, C.c_float), tab.shape[0]))

    def set_plane_wave(self, mode, params=(0.0, 0.0, 0.0)):
        p = _f64(list(params) + [0.0] * (4 - len(params)))
        self._chk(lib().fdtd_set_plane_wave(self._h, int(mode), _p(p, C.c_double)))

    def prepare(self):
        self._chk(lib().fdtd_prepare(self._h))

    def ready(self):
        self._chk(lib().fdtd_ready(self._h))

    def init_1d(self, rate, pw_source):
        s = _f32(pw_source)
        self._chk(lib().fdtd_init_1d(self._h, float(rate), _p(s, C.c_float)))

    def update_coeffs(self):
        self._chk(lib().fdtd_update_coeffs(self._h))

    def reset(self):
        self._chk(lib().fdtd_reset_fields(self._h))

    def update_h_all(self):
        self._chk(lib().fdtd_update_h_all(self._h))

    def update_e_all(self):
        self._chk(lib().fdtd_update_e_all(self._h))

    def update_h_pml(self):
        self._chk(lib().fdtd_update_h_pml(self._h))

    def update_e_pml(self):
        self._chk(lib().fdtd_update_e_pml(self._h))

    def update_h_planewave(self, n, mode):
        t = _i32([n, mode])
        self._chk(lib().fdtd_update_h_planewave(self._h, _p(t, C.c_int32)))

    def update_e_planewave(self, n, mode):
        t = _i32([n, mode])
        self._chk(lib().fdtd_update_e_planewave(self._h, _p(t, C.c_int32)))

    def add_currents(self, n, pol, box):
        t, b = _i32([n, pol]), _i32(box)
        self._chk(lib().fdtd_add_currents(self._h, _p(t, C.c_int32), _p(b, C.c_int32)))

    def step(self, n, mode):
        self._chk(lib().fdtd_step(self._h, int(n), int(mode)))

    def run(self, n0, count, mode, snap_every=0, snap_comp="Ez"):
        comp = FIELD[snap_comp] + 1 if isinstance(snap_comp, str) else int(snap_comp)
        self._chk(lib().fdtd_run(self._h, int(n0), int(count), int(mode), int(snap_every), comp))

    def run_noalloc(self, n0, count, mode):
        """fdtd_run_noalloc: the steps of run() with no allocation (device scratch of the library)."""
        self._chk(lib().fdtd_run_noalloc(self._h, int(n0), int(count), int(mode)))

    def stream_ceiling(self, reps=20):
        """Average ms of the 49-byte-per-cell stream kernel over this grid (overwrites the fields)."""
        ms = C.c_double()
        self._chk(lib().fdtd_stream_ceiling(self._h, int(reps), C.byref(ms)))
        return ms.value

    def synchronize(self):
        self._chk(lib().fdtd_synchronize(self._h))

    def device_bytes(self) -> int:
        b = C.c_uint64()
        self._chk(lib().fdtd_device_bytes(self._h, C.byref(b)))
        return b.value

    def set_autotune(self, policy):
        """Size-class tuning table: 0 off, 1 apply a class's entry (default), 2 also autotune a new class."""
        self._chk(lib().fdtd_set_autotune(self._h, int(policy)))

    def get_tuning(self):
        """(rows per group, z chunk, columns per lane) in force; 0 = automatic."""
        out = np.zeros(4, np.int32)
        self._chk(lib().fdtd_get_tuning(self._h, _p(out, C.c_int32)))
        return tuple(int(v) for v in out[:3])

    def set_tuning(self, rows_per_group=0, z_chunk=0, cols_per_lane=0):
        self._chk(lib().fdtd_set_tuning(self._h, int(rows_per_group), int(z_chunk), int(cols_per_lane)))

    def slab_chunk(self):
        """Planes per z chunk of the row kernel's slab part in force (0 = automatic)."""
        out = np.zeros(4, np.int32)
        self._chk(lib().fdtd_get_tuning(self._h, _p(out, C.c_int32)))
        return int(out[3])

    def set_slab_chunk(self, planes=0):
        self._chk(lib().fdtd_set_slab_chunk(self._h, int(planes)))

    def set_addressing(self, mode=0):
        """0: buffer addressing where arrays fit (default); 1: the 64-bit global path."""
        self._chk(lib().fdtd_set_addressing(self._h, int(mode)))

    def autotune(self, n=0, mode=0, reps=10):
        """Pick the fastest row layout and z-chunk for this grid by timing dry steps (state
        unchanged); returns the best average step time in ms."""
        ms = C.c_double()
        self._chk(lib().fdtd_autotune(self._h, int(n), int(mode), int(reps), C.byref(ms)))
        return ms.value

    def set_timing(self, on=True, every=1):
        """HIP events around every `every`-th step of the following runs (off: none)."""
        self._chk(lib().fdtd_set_timing(self._h, int(every) if on else 0))

    def kernel_time(self):
        ms, n = C.c_double(), C.c_int64()
        self._chk(lib().fdtd_kernel_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def row_kernel_time(self):
        """Average device time (ms) of the row kernel over the last timed run."""
        ms = C.c_double()
        self._chk(lib().fdtd_row_kernel_time(self._h, C.byref(ms)))
        return ms.value

    # ---- readback ----
    def output_vis(self, flag):
        I, J, K = self.dims
        out = np.zeros(I * J * K, dtype=np.float32)
        self._chk(lib().fdtd_output_vis(self._h, _p(out, C.c_float), out.size, int(flag)))
        return out

    def output_ant(self, size=None):
        n = self.ncomp * self.T * self.nAnt if size is None else int(size)
        out = np.zeros(n, dtype=np.float32)
        self._chk(lib().fdtd_output_ant(self._h, _p(out, C.c_float), n))
        return out

    def probes(self):
        return self.output_ant().reshape(self.ncomp, self.T, self.nAnt)

    def field(self, name):
        I, J, K = self.dims
        return self.read_array(FIELD[name]).reshape(K, J, I)

    def read_array(self, what):
        n = C.c_uint64()
        self._chk(lib().fdtd_read_array(self._h, int(what), None, C.byref(n)))
        out = np.empty(n.value, dtype=np.float32)
        self._chk(lib().fdtd_read_array(self._h, int(what), _p(out, C.c_float), C.byref(n)))
        return out

    def psi(self, q):
        return self.read_array(20 + q)

    def line(self, which):
        arr = self.read_array(40 + which)
        return arr.reshape(self.T, -1)

    def read_setup(self, what):
        n = C.c_uint64()
        self._chk(lib().fdtd_read_setup(self._h, int(what), None, C.byref(n)))
        out = np.empty(n.value, dtype=np.float32)
        self._chk(lib().fdtd_read_setup(self._h, int(what), _p(out, C.c_float), C.byref(n)))
        return out

    def snapshots(self):
        n = C.c_int32()
        self._chk(lib().fdtd_snapshot_count(self._h, C.byref(n)))
        I, J, K = self.dims
        out = np.empty((n.value, I * J * K), dtype=np.float32)
        self._chk(lib().fdtd_output_snapshots(self._h, _p(out, C.c_float), out.size))
        return out


class Batch:
    """Many independent simulations advanced together (fdtd_batch_*): each batch step launches every
    kernel of the fused step once for all grids still running."""

    def __init__(self):
        h = C.c_void_p()
        Simulation._chk(lib().fdtd_batch_create(C.byref(h)))
        self._h = h
        self._sims = []

    def add(self, sim, n0, count, mode=0):
        Simulation._chk(lib().fdtd_batch_add(self._h, sim._h, int(n0), int(count), int(mode)))
        self._sims.append(sim)

    def clear(self):
        Simulation._chk(lib().fdtd_batch_clear(self._h))
        self._sims = []

    def run(self):
        Simulation._chk(lib().fdtd_batch_run(self._h))

    def set_strategy(self, strategy):
        """0 automatic (merged launches), 1 merged launches, 2 one stream per grid, 3 device work queue."""
        Simulation._chk(lib().fdtd_batch_set_strategy(self._h, int(strategy)))

    def close(self):
        if getattr(self, "_h", None):
            lib().fdtd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server (fdtd_server_*): solves submitted one at a time, advanced together with
    continuous batching.  A submitted Simulation is owned by the server until release()."""

    def __init__(self, hbm_budget: int, max_active: int = 8, steps_per_round: int = 64):
        h = C.c_void_p()
        Simulation._chk(lib().fdtd_server_create(C.byref(h), int(hbm_budget), int(max_active), int(steps_per_round)))
        self._h = h
        self._reqs = {}

    def now(self) -> float:
        t = C.c_double()
        Simulation._chk(lib().fdtd_server_now(self._h, C.byref(t)))
        return t.value

    def submit(self, sim, count, mode, source=None, rate=1.0, t_arrive=-1.0) -> int:
        """Queue `count` steps of `sim` (host set-up done, not ready) in plane-wave mode `mode`;
        `source`: the Init1D samples (mode != 0).  Returns the request id."""
        rid = C.c_int64()
        src = _f32(source) if source is not None else None
        Simulation._chk(lib().fdtd_server_submit(self._h, sim._h, float(rate),
                                                  _p(src, C.c_float) if src is not None else None,
                                                  int(count), int(mode), float(t_arrive), C.byref(rid)))
        sim._owned = False
        self._reqs[rid.value] = sim
        return rid.value

    def round(self):
        """One round; returns (active, queued) after it."""
        a, q = C.c_int32(), C.c_int32()
        Simulation._chk(lib().fdtd_server_round(self._h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def collect(self, max_n: int = 1024):
        """Completed requests: list of (id, t_arrive, t_admit, t_done)."""
        ids = np.zeros(max_n, np.int64)
        ts = np.zeros(3 * max_n, np.float64)
        n = C.c_int32()
        Simulation._chk(lib().fdtd_server_collect(self._h, _p(ids, C.c_int64), _p(ts, C.c_double), int(max_n),
                                                   C.byref(n)))
        return [(int(ids[k]), float(ts[3 * k]), float(ts[3 * k + 1]), float(ts[3 * k + 2])) for k in range(n.value)]

    def simulation(self, rid: int):
        return self._reqs[rid]

    def release(self, rid: int):
        Simulation._chk(lib().fdtd_server_release(self._h, int(rid)))
        sim = self._reqs.pop(rid)
        sim._h = None

    def request_error(self, rid: int):
        """(FDTD_ERR_* code, message) of a request whose set-up failed at admission; (0, "")
        otherwise."""
        code = C.c_int32()
        msg = C.create_string_buffer(512)
        Simulation._chk(lib().fdtd_server_request_error(self._h, int(rid), C.byref(code), msg, 512))
        return code.value, msg.value.decode()

    def request_bytes(self, rid: int) -> int:
        b = C.c_uint64()
        Simulation._chk(lib().fdtd_server_request_bytes(self._h, int(rid), C.byref(b)))
        return b.value

    def stats(self):
        """(bytes in use, peak bytes, cell updates)."""
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_int64()
        Simulation._chk(lib().fdtd_server_stats(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def close(self):
        if getattr(self, "_h", None):
            lib().fdtd_server_destroy(self._h)
            self._h = None
            for sim in self._reqs.values():
                sim._h = None
            self._reqs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------------
# MEX-style gateway: simFDTD(op, handle, ...)   (src/simFDTD.cu:19-187)
# ---------------------------------------------------------------------------------------------
_handles: dict[int, Simulation] = {}


def simFDTD(op: str, *args):
    if not isinstance(op, str):
        raise SimFDTDError("simFDTD:InvalidInput", "At least one input required: operation string.")
    o = op.lower()
    if o == "createsimulation":
        s = Simulation()
        key = id(s)
        _handles[key] = s
        return np.uint64(key)
    if len(args) < 1:
        raise SimFDTDError("simFDTD:InvalidInput", "Simulation pointer required for this operation.")
    s = _handles.get(int(args[0]))
    if s is None:
        raise SimFDTDError("simFDTD:InvalidInput", "unknown simulation handle")
    a = args[1:]

    def need(n, msg):
        if len(a) < n:
            raise SimFDTDError("simFDTD:InvalidInput", msg)

    if o == "deletesimulation":
        _handles.pop(int(args[0])).close()
    elif o == "updatee_all":
        s.update_e_all()
    elif o == "updateh_all":
        s.update_h_all()
    elif o == "updatee_pml":
        s.update_e_pml()
    elif o == "updateh_pml":
        s.update_h_pml()
    elif o == "updatecuda":
        s.update_coeffs()
    elif o == "addcurrents":
        need(2, "time and indices required.")
        t = np.asarray(a[0]).reshape(-1)
        s.add_currents(int(t[0]), int(t[1]), a[1])
    elif o == "readysimulation":
        s.ready()
    elif o == "resetfieldall":
        s.reset()
    elif o in ("updatee_planewave", "updateh_planewave"):
        need(1, "time required.")
        t = np.asarray(a[0]).reshape(-1)
        (s.update_e_planewave if o == "updatee_planewave" else s.update_h_planewave)(int(t[0]), int(t[1]))
    elif o == "setconst":
        need(2, "freqRes and TimeDim required.")
        s.set_const(a[0], a[1])
    elif o == "setsource":
        need(1, "SourceIn required.")
        s.set_source(a[0], update=len(a) > 1)
    elif o == "setantennalocations":
        need(1, "Locations required.")
        s.set_antenna_locations(a[0])
    elif o == "createmu":
        need(3, "mu_x, mu_y, mu_z required.")
        s.create_mu(a[0], a[1], a[2])
    elif o == "init1d":
        need(2, "rate and PW_source required.")
        s.init_1d(float(np.asarray(a[0]).reshape(-1)[0]), a[1])
    elif o == "updatematerial":
        need(2, "flag, values required.")
        s.update_material(int(np.asarray(a[0]).reshape(-1)[0]), a[1])
    elif o == "outputant":
        need(1, "OutAntSize required.")
        return s.output_ant(int(np.asarray(a[0]).reshape(-1)[0])).reshape(1, -1)
    elif o == "outputvis":
        need(2, "OutAntSize and flag required.")
        size = int(np.asarray(a[0]).reshape(-1)[0])
        out = s.output_vis(int(np.asarray(a[1]).reshape(-1)[0]))
        res = np.zeros(size, dtype=np.float32)
        res[: min(size, out.size)] = out[: min(size, out.size)]
        return res.reshape(1, -1)
    # extensions
    elif o == "setcpml":
        s.set_cpml(a[0])
    elif o == "settfsf":
        s.set_tfsf(int(np.asarray(a[0]).reshape(-1)[0]))
    elif o == "setplanewave":
        s.set_plane_wave(int(np.asarray(a[0]).reshape(-1)[0]), a[1] if len(a) > 1 else (0.0,))
    elif o == "setmaterialindex":
        s.set_material_index(a[0], a[1])
    elif o == "setprobecomponents":
        s.set_probe_components(a[0])
    elif o == "step":
        t = np.asarray(a[0]).reshape(-1)
        s.step(int(t[0]), int(t[1]))
    elif o == "run":
        t = np.asarray(a[0]).reshape(-1)
        s.run(int(t[0]), int(t[1]), int(t[2]), int(t[3]) if t.size > 3 else 0, int(t[4]) if t.size > 4 else 3)
    else:
        raise SimFDTDError("simFDTD:UnknownOperation", f"Unknown operation: {o}")
    return None
