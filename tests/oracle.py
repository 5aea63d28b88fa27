"""ctypes binding of the refcpu oracle (oracle/build/librefcpu.so).

Test infrastructure: the oracle is the CPU restatement of the reference path
the GPU results are checked against.  Only tests, the smoke check and the
benchmark's CPU-baseline leg load it.
"""

from __future__ import annotations

import ctypes
from pathlib import Path
from typing import List, Tuple

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "build" / "librefcpu.so"
FIELD_IDS = {"EX": 0, "EY": 1, "EZ": 2, "HX": 3, "HY": 4, "HZ": 5}
_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    L = ctypes.CDLL(str(LIB))
    vp, i64, dbl, cp = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_char_p
    p64 = ctypes.POINTER(ctypes.c_int64)
    sig = {
        "ref_last_error": (cp, []),
        "ref_create": (vp, [cp]),
        "ref_destroy": (None, [vp]),
        "ref_set_write_files": (ctypes.c_int, [vp, ctypes.c_int]),
        "ref_set_threads": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "ref_run": (ctypes.c_int, [vp, i64]),
        "ref_init": (ctypes.c_int, [vp, i64]),
        "ref_loop": (ctypes.c_int, [vp]),
        "ref_write_field": (ctypes.c_int, [vp, ctypes.c_int, vp]),
        "ref_dt": (dbl, [vp]),
        "ref_loop_seconds": (dbl, [vp]),
        "ref_step_end_seconds": (i64, [vp, ctypes.POINTER(dbl), i64]),
        "ref_grid": (ctypes.c_int, [vp, p64]),
        "ref_grid_array": (i64, [vp, ctypes.c_int, ctypes.c_int, vp]),
        "ref_field": (ctypes.c_int, [vp, ctypes.c_int, p64, vp]),
        "ref_coeff": (ctypes.c_int, [vp, cp, p64, vp]),
        "ref_material": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, p64, vp]),
        "ref_pml_profile": (i64, [vp, ctypes.c_int, ctypes.c_int, vp]),
        "ref_tfsf_table": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, p64, vp]),
        "ref_nf2ff_count": (ctypes.c_int, [vp]),
        "ref_nf2ff_surface": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, p64, p64]),
        "ref_nf2ff_current": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]),
        "ref_nf2ff_potential": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, vp, ctypes.c_int, vp, vp, vp]),
        "ref_nf2ff_power": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(dbl)]),
        "ref_objective": (ctypes.c_int, [vp, ctypes.POINTER(dbl), vp]),
        "ref_monitor_count": (ctypes.c_int, [vp]),
        "ref_monitor_records": (ctypes.c_int, [vp, ctypes.c_int]),
        "ref_monitor_record": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, p64, vp]),
        "ref_network_output": (ctypes.c_int, [vp, ctypes.c_int]),
        "ref_nf2ff_td_num": (i64, [vp, ctypes.c_int]),
        "ref_nf2ff_td_potential": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]),
        "ref_nf2ff_td_far_field": (ctypes.c_int, [vp, ctypes.c_int, vp]),
        "ref_network_s": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, vp]),
        "ref_series_dft": (ctypes.c_int, [vp, i64, dbl, vp, i64, dbl, vp]),
        "ref_port_waves": (ctypes.c_int, [vp, vp, i64, dbl, vp, i64, dbl, dbl, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def series_dft(x, dt: float, freqs, shift: float = 0.0) -> np.ndarray:
    """The oracle's dft (util/dft.h:9-26) of a real series."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    out = np.zeros(f.size, dtype=np.complex128)
    L = lib()
    if L.ref_series_dft(x.ctypes.data, x.size, dt, f.ctypes.data, f.size, shift, out.ctypes.data) != 0:
        raise RuntimeError(L.ref_last_error().decode())
    return out


def port_waves(current, voltage, dt: float, freqs, z: complex):
    """The oracle's Port::calculateSParameters on given records: (a, b)."""
    c = np.ascontiguousarray(current, dtype=np.float64)
    v = np.ascontiguousarray(voltage, dtype=np.float64)
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    a = np.zeros(f.size, dtype=np.complex128)
    b = np.zeros(f.size, dtype=np.complex128)
    L = lib()
    if L.ref_port_waves(c.ctypes.data, v.ctypes.data, c.size, dt, f.ctypes.data, f.size, complex(z).real,
                        complex(z).imag, a.ctypes.data, b.ctypes.data) != 0:
        raise RuntimeError(L.ref_last_error().decode())
    return a, b


class RefSim:
    def __init__(self, spec: str, threads: Tuple[int, int, int] = None):
        self.L = lib()
        self.h = self.L.ref_create(spec.encode())
        if not self.h:
            raise ValueError(self.L.ref_last_error().decode())
        if threads is not None:
            self.L.ref_set_threads(self.h, *threads)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.ref_last_error().decode())

    def nf2ff_td_potential(self, n: int, d: int, which: int) -> np.ndarray:
        """Raw potential `which` (0 ua, 1 ub, 2 wa, 3 wb) of surface d."""
        num = int(self.L.ref_nf2ff_td_num(self.h, n))
        out = np.zeros(num)
        self._check(self.L.ref_nf2ff_td_potential(self.h, n, d, which, out.ctypes.data))
        return out

    def nf2ff_td_far_field(self, n: int) -> np.ndarray:
        """processFarField: rows w_theta, w_phi, u_theta, u_phi."""
        num = int(self.L.ref_nf2ff_td_num(self.h, n))
        out = np.zeros((4, num))
        self._check(self.L.ref_nf2ff_td_far_field(self.h, n, out.ctypes.data))
        return out

    def network_s(self, n: int, keys, nf: int) -> dict:
        """Network::output on the last run, then {key: S} for `keys`."""
        self._check(self.L.ref_network_output(self.h, n))
        out = {}
        for k in keys:
            a = np.zeros(nf, dtype=np.complex128)
            self._check(self.L.ref_network_s(self.h, n, int(k), a.ctypes.data))
            out[int(k)] = a
        return out

    def set_threads(self, tx: int, ty: int, tz: int):
        """ThreadConfig of the next loop (Simulation's thread layout)."""
        self._check(self.L.ref_set_threads(self.h, tx, ty, tz))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_destroy(self.h)
            self.h = None

    def run(self, steps: int):
        self._check(self.L.ref_run(self.h, steps))

    def init(self, steps: int):
        self._check(self.L.ref_init(self.h, steps))

    def loop(self):
        self._check(self.L.ref_loop(self.h))

    def write_field(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        shape = self.field(name).shape
        assert a.shape == shape, (a.shape, shape)
        self._check(self.L.ref_write_field(self.h, FIELD_IDS[name], a.ctypes.data))

    @property
    def dt(self):
        return self.L.ref_dt(self.h)

    @property
    def loop_seconds(self):
        return self.L.ref_loop_seconds(self.h)

    def step_end_seconds(self):
        """Seconds from the start of the last loop to the end of each step."""
        n = self.L.ref_step_end_seconds(self.h, None, 0)
        out = (ctypes.c_double * max(1, n))()
        self.L.ref_step_end_seconds(self.h, out, n)
        return list(out)[:n]

    def grid(self):
        n = (ctypes.c_int64 * 3)()
        self._check(self.L.ref_grid(self.h, n))
        return tuple(n)

    def grid_array(self, which: int, axis: int) -> np.ndarray:
        n = self.L.ref_grid_array(self.h, which, axis, None)
        a = np.empty(n)
        self.L.ref_grid_array(self.h, which, axis, a.ctypes.data)
        return a

    def _arr3(self, fn, *key):
        shape = (ctypes.c_int64 * 3)()
        self._check(fn(self.h, *key, shape, None))
        a = np.empty(tuple(shape))
        self._check(fn(self.h, *key, shape, a.ctypes.data))
        return a

    def field(self, name: str) -> np.ndarray:
        return self._arr3(self.L.ref_field, FIELD_IDS[name])

    def coeff(self, name: str) -> np.ndarray:
        return self._arr3(self.L.ref_coeff, name.encode())

    def material(self, kind: int, comp: int) -> np.ndarray:
        return self._arr3(self.L.ref_material, kind, comp)

    def pml_profile(self, b: int, which: int) -> np.ndarray:
        n = self.L.ref_pml_profile(self.h, b, which, None)
        a = np.empty(n)
        self.L.ref_pml_profile(self.h, b, which, a.ctypes.data)
        return a

    def tfsf_table(self, i: int, which: int) -> np.ndarray:
        shape = (ctypes.c_int64 * 3)()
        self._check(self.L.ref_tfsf_table(self.h, i, which, shape, None))
        a = np.empty((shape[0], shape[1]))
        self._check(self.L.ref_tfsf_table(self.h, i, which, shape, a.ctypes.data))
        return a

    # ------------------------------------------------------------ NF2FF --
    def nf2ff_surface(self, n: int, d: int):
        lo = (ctypes.c_int64 * 3)()
        hi = (ctypes.c_int64 * 3)()
        self._check(self.L.ref_nf2ff_surface(self.h, n, d, lo, hi))
        return tuple(lo), tuple(hi)

    def nf2ff_current(self, n: int, f: int, d: int, which: int) -> np.ndarray:
        """Surface current array (complex) on surface d: which 0 ja, 1 jb, 2 ma, 3 mb."""
        lo, hi = self.nf2ff_surface(n, d)
        shape = tuple(h - l for l, h in zip(lo, hi))
        out = np.zeros(int(np.prod(shape)) * 2)
        self._check(self.L.ref_nf2ff_current(self.h, n, f, d, which, out.ctypes.data))
        return (out[0::2] + 1j * out[1::2]).reshape(shape)

    def nf2ff_potential(self, n: int, f: int, a_pot: bool, theta_comp: bool, theta, phi, origin=(0.0, 0.0, 0.0)):
        th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        ph = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
        org = np.ascontiguousarray(origin, dtype=np.float64)
        out = np.zeros(2 * th.size * ph.size)
        self._check(self.L.ref_nf2ff_potential(self.h, n, f, int(a_pot), int(theta_comp), th.size, th.ctypes.data,
                                               ph.size, ph.ctypes.data, org.ctypes.data, out.ctypes.data))
        return out[0::2] + 1j * out[1::2]

    def nf2ff_power(self, n: int, f: int) -> float:
        v = ctypes.c_double()
        self._check(self.L.ref_nf2ff_power(self.h, n, f, ctypes.byref(v)))
        return v.value

    def objective(self):
        """(J, per-point DFT values) of the scenario's DFT objective."""
        n = self.L.ref_objective(self.h, None, None)
        if n < 0:
            raise RuntimeError("no objective in the scenario")
        v = ctypes.c_double()
        d = np.zeros(n, dtype=np.complex128)
        self.L.ref_objective(self.h, ctypes.byref(v), d.ctypes.data)
        return float(v.value), d

    def monitor_records(self, m: int) -> List[np.ndarray]:
        out = []
        for r in range(self.L.ref_monitor_records(self.h, m)):
            shape = (ctypes.c_int64 * 3)()
            self._check(self.L.ref_monitor_record(self.h, m, r, shape, None))
            a = np.empty(tuple(shape))
            self._check(self.L.ref_monitor_record(self.h, m, r, shape, a.ctypes.data))
            out.append(a[:, :, 0] if (shape[0] == 2 and shape[2] == 1) else a)
        return out


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return float(num / den) if den > 0 else float(num)
