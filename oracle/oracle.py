"""ctypes front-end of the CPU oracle (fdtd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, to check results against and as the timed CPU baseline.  It mirrors the op vocabulary of the reference MEX gateway
(src/simFDTD.cu:19-187) so that one driver can replay the same op sequence on the oracle and on
the HIP library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class OracleError(RuntimeError):
    def __init__(self, ident: str, msg: str):
        super().__init__(f"{ident}: {msg}")
        self.identifier = ident


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        args = ["make", "-s", "-C", _HERE] + (["-B"] if force else [])
        subprocess.run(args, check=True)
    return _LIB_PATH


def build_native(out_dir: str) -> str:
    """Compile the oracle for this host's CPU (-march=native; otherwise the flags of the Makefile)
    into out_dir; returns the library path.  Used by the CPU baseline of bench.py."""
    out = os.path.join(out_dir, "liboracle_native.so")
    cmd = [os.environ.get("CXX", "g++"), "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           "-fPIC", "-std=c++17", "-shared", "-o", out, os.path.join(_HERE, "fdtd_oracle.cpp")]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


def use_library(path: str) -> None:
    """Load the oracle from `path` instead of the in-tree build (before the first use)."""
    global _LIB_PATH, _lib
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    _LIB_PATH = path


def set_threads(n: int) -> None:
    if lib().orc_set_threads(int(n)):
        raise ValueError("thread count must be >= 1")


def max_threads() -> int:
    return int(lib().orc_max_threads())


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        L.orc_max_threads.argtypes = []
        L.orc_max_threads.restype = C.c_int
        vp, i32, f32p, i32p, f64p = C.c_void_p, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_double)
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int]
        L.orc_destroy.argtypes = [vp]
        sig = {
            "orc_set_const": [vp, f32p, i32p],
            "orc_set_cpml": [vp, f64p],
            "orc_set_tfsf": [vp, i32],
            "orc_set_source": [vp, f32p, C.c_int],
            "orc_set_mesh": [vp, f32p, f32p, f32p],
            "orc_set_antenna": [vp, i32p],
            "orc_set_probe_comps": [vp, i32p, C.c_int],
            "orc_update_material": [vp, i32, f32p],
            "orc_create_mu": [vp, f32p, f32p, f32p],
            "orc_ready": [vp],
            "orc_set_plane_wave": [vp, i32, f64p],
            "orc_init_1d": [vp, C.c_float, f32p],
            "orc_update_coeffs": [vp],
            "orc_reset": [vp],
            "orc_update_h_all": [vp],
            "orc_update_e_all": [vp],
            "orc_update_h_pml": [vp],
            "orc_update_e_pml": [vp],
            "orc_update_h_pw": [vp, i32p],
            "orc_update_e_pw": [vp, i32p],
            "orc_add_currents": [vp, i32p, i32p],
            "orc_run": [vp, i32, i32, i32, i32, i32, f64p],
            "orc_get_array": [vp, i32, f64p],
            "orc_get_setup": [vp, i32, f64p],
            "orc_get_line": [vp, i32, f64p],
            "orc_get_probes": [vp, f64p],
            "orc_set_field": [vp, i32, f64p],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.orc_array_len.argtypes = [vp, i32]
        L.orc_array_len.restype = C.c_int64
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


FIELD = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}


class OracleSim:
    """One oracle simulation (float32 or float64 arithmetic)."""

    def __init__(self, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self._h = lib().orc_create(1 if self.dtype == np.float64 else 0)
        self.dims = None
        self.T = 0
        self.nAnt = 0
        self.ncomp = 1

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            msg = lib().orc_last_error().decode()
            ident, _, text = msg.partition("|")
            raise OracleError(ident or "simFDTD:Error", text or msg)

    # ---- op vocabulary of src/simFDTD.cu ----
    def set_const(self, freqRes, timeDim):
        fr, td = _f32(freqRes), _i32(timeDim)
        self._chk(lib().orc_set_const(self._h, _p(fr, C.c_float), _p(td, C.c_int32)))
        self.T = int(td[0])
        self.dims = (int(td[1]), int(td[2]), int(td[3]))
        self.nAnt = int(td[4])

    def set_cpml(self, params):
        p = _f64(params)
        self._chk(lib().orc_set_cpml(self._h, _p(p, C.c_double)))

    def set_tfsf(self, inset):
        self._chk(lib().orc_set_tfsf(self._h, int(inset)))

    def set_mesh(self, dx=None, dy=None, dz=None):
        """Graded mesh: cell sizes along x, y, z (None: uniform dx along that axis)."""
        ptrs = []
        self._mesh = []
        for q, d in enumerate((dx, dy, dz)):
            if d is None:
                ptrs.append(None)
            else:
                a = np.ascontiguousarray(d, dtype=np.float32)
                if self.dims is None or a.size != self.dims[q]:
                    raise ValueError("set_mesh: one spacing per grid plane of the axis (after set_const)")
                self._mesh.append(a)
                ptrs.append(_p(a, C.c_float))
        self._chk(lib().orc_set_mesh(self._h, *ptrs))

    def set_source(self, src, update=False):
        s = _f32(src)
        self._chk(lib().orc_set_source(self._h, _p(s, C.c_float), 1 if update else 0))

    def set_antenna_locations(self, loc):
        a = _i32(np.asarray(loc).reshape(-1, order="F") if np.asarray(loc).ndim == 2 else loc)
        self._chk(lib().orc_set_antenna(self._h, _p(a, C.c_int32)))

    def set_probe_components(self, comps):
        c = _i32(comps)
        self._chk(lib().orc_set_probe_comps(self._h, _p(c, C.c_int32), len(c)))
        self.ncomp = len(c)

    def update_material(self, flag, values):
        v = _f32(values)
        self._chk(lib().orc_update_material(self._h, int(flag), _p(v, C.c_float)))

    def create_mu(self, mu_x, mu_y, mu_z):
        m = [_f32(v) for v in (mu_x, mu_y, mu_z)]
        self._chk(lib().orc_create_mu(self._h, *(_p(v, C.c_float) for v in m)))

    def ready(self):
        self._chk(lib().orc_ready(self._h))

    def set_plane_wave(self, mode, params=(0.0, 0.0, 0.0)):
        p = _f64(list(params) + [0.0] * (4 - len(params)))
        self._chk(lib().orc_set_plane_wave(self._h, int(mode), _p(p, C.c_double)))

    def init_1d(self, rate, pw_source):
        s = _f32(pw_source)
        self._chk(lib().orc_init_1d(self._h, float(rate), _p(s, C.c_float)))

    def update_coeffs(self):
        self._chk(lib().orc_update_coeffs(self._h))

    def reset(self):
        self._chk(lib().orc_reset(self._h))

    def update_h_all(self):
        self._chk(lib().orc_update_h_all(self._h))

    def update_e_all(self):
        self._chk(lib().orc_update_e_all(self._h))

    def update_h_pml(self):
        self._chk(lib().orc_update_h_pml(self._h))

    def update_e_pml(self):
        self._chk(lib().orc_update_e_pml(self._h))

    def update_h_planewave(self, n, mode):
        t = _i32([n, mode])
        self._chk(lib().orc_update_h_pw(self._h, _p(t, C.c_int32)))

    def update_e_planewave(self, n, mode):
        t = _i32([n, mode])
        self._chk(lib().orc_update_e_pw(self._h, _p(t, C.c_int32)))

    def add_currents(self, n, pol, box):
        t, b = _i32([n, pol]), _i32(box)
        self._chk(lib().orc_add_currents(self._h, _p(t, C.c_int32), _p(b, C.c_int32)))

    def run(self, n0, count, mode, snap_every=0, snap_comp="Ez"):
        I, J, K = self.dims
        out = None
        nsnap = 0
        if snap_every > 0:
            nsnap = count // snap_every + 1
            out = np.zeros((nsnap, I * J * K), dtype=np.float64)
        ptr = _p(out, C.c_double) if out is not None else None
        self._chk(lib().orc_run(self._h, int(n0), int(count), int(mode), int(snap_every),
                                FIELD[snap_comp] if isinstance(snap_comp, str) else int(snap_comp), ptr))
        return out

    # ---- readback ----
    def field(self, name):
        I, J, K = self.dims
        out = np.empty(I * J * K, dtype=np.float64)
        self._chk(lib().orc_get_array(self._h, FIELD[name], _p(out, C.c_double)))
        return out.reshape((K, J, I)).astype(self.dtype)

    def array(self, what):
        n = lib().orc_array_len(self._h, int(what))
        out = np.empty(int(n), dtype=np.float64)
        self._chk(lib().orc_get_array(self._h, int(what), _p(out, C.c_double)))
        return out.astype(self.dtype)

    def coeffs(self):
        """(Ce[3], Ch[3]) as flat arrays in the reference x-fastest layout."""
        return [self.array(6 + c) for c in range(3)], [self.array(9 + c) for c in range(3)]

    def psi(self, q):
        return self.array(20 + q)

    def setup(self, what):
        I, J, K = self.dims
        P = int(self.scalars()["P"]) if what != 12 else 0
        n = {0: 2 * P, 1: 2 * P, 2: 2 * P, 3: 2 * P, 4: P, 5: P, 6: I, 7: J, 8: K, 9: I, 10: J, 11: K, 12: 13,
             32: 6}.get(what) or (I, J, K)[(what - 20) // 4]
        out = np.empty(n, dtype=np.float64)
        self._chk(lib().orc_get_setup(self._h, int(what), _p(out, C.c_double)))
        return out

    def scalars(self):
        v = np.empty(13, dtype=np.float64)
        self._chk(lib().orc_get_setup(self._h, 12, _p(v, C.c_double)))
        keys = ["Chh", "Che", "sig_max", "N1", "i0", "i1", "j0", "j1", "k0", "k1", "P", "ratio", "src"]
        return dict(zip(keys, v.tolist()))

    def line(self, which):
        N1 = int(self.scalars()["N1"])
        out = np.empty(self.T * N1, dtype=np.float64)
        self._chk(lib().orc_get_line(self._h, int(which), _p(out, C.c_double)))
        return out.reshape(self.T, N1).astype(self.dtype)

    def probes(self):
        out = np.empty(self.ncomp * self.T * self.nAnt, dtype=np.float64)
        self._chk(lib().orc_get_probes(self._h, _p(out, C.c_double)))
        return out.reshape(self.ncomp, self.T, self.nAnt).astype(self.dtype)

    def set_field(self, name, values):
        v = _f64(np.asarray(values).reshape(-1))
        self._chk(lib().orc_set_field(self._h, FIELD[name], _p(v, C.c_double)))
