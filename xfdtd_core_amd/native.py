"""ctypes binding of libxfdtd_core_amd.so.

The library is built in-tree (``xfdtd_core_amd/build.py``) and loaded from
``xfdtd_core_amd/lib``.  Nothing here computes fields: every step runs in the
HIP kernels behind the C-ABI, and a missing library raises instead of falling
back to anything else.
"""

from __future__ import annotations

import ctypes
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

_PKG = Path(__file__).resolve().parent
_LIB_PATH = _PKG / "lib" / "libxfdtd_core_amd.so"
_lib: Optional[ctypes.CDLL] = None

FIELD_IDS = {"EX": 0, "EY": 1, "EZ": 2, "HX": 3, "HY": 4, "HZ": 5}
COEFF_NAMES = ("cexe", "cexhy", "cexhz", "ceye", "ceyhz", "ceyhx", "ceze", "cezhx", "cezhy",
               "chxh", "chxey", "chxez", "chyh", "chyez", "chyex", "chzh", "chzex", "chzey")


def library_path() -> Path:
    return _LIB_PATH


def load_library() -> ctypes.CDLL:
    """Loads the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise RuntimeError(f"{_LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(str(_LIB_PATH))
    vp, i32, i64, dbl, cp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_char_p
    p64 = ctypes.POINTER(ctypes.c_int64)
    p32 = ctypes.POINTER(ctypes.c_int32)
    sig = {
        "xfd_sim_last_error": (cp, []),
        "xfd_sim_create": (vp, [cp]),
        "xfd_sim_destroy": (None, [vp]),
        "xfd_sim_set_precision": (ctypes.c_int, [vp, i32]),
        "xfd_sim_set_profiling": (ctypes.c_int, [vp, i32]),
        "xfd_sim_set_profiling_classes": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "xfd_sim_precompute": (ctypes.c_int, [vp, i64]),
        "xfd_sim_init": (ctypes.c_int, [vp, i64]),
        "xfd_sim_run_loop": (ctypes.c_int, [vp]),
        "xfd_sim_run": (ctypes.c_int, [vp, i64]),
        "xfd_sim_advance": (ctypes.c_int, [vp, i64, i64]),
        "xfd_sim_finish": (ctypes.c_int, [vp]),
        "xfd_sim_dt": (dbl, [vp]),
        "xfd_sim_loop_seconds": (dbl, [vp]),
        "xfd_sim_grid": (ctypes.c_int, [vp, p64]),
        "xfd_sim_device_bytes": (i64, [vp]),
        "xfd_sim_coef_classed": (ctypes.c_int32, [vp]),
        "xfd_sim_cpml_fused": (ctypes.c_int32, [vp]),
        "xfd_cpml_fused": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int32)]),
        "xfd_sim_nf2ff_count": (ctypes.c_int32, [vp]),
        "xfd_sim_nf2ff_td_count": (ctypes.c_int32, [vp]),
        "xfd_sim_nf2ff_td_num": (ctypes.c_int64, [vp, i32]),
        "xfd_sim_nf2ff_td_potential": (ctypes.c_int, [vp, i32, i32, i32, vp]),
        "xfd_sim_nf2ff_td_far_field": (ctypes.c_int, [vp, i32, vp]),
        "xfd_sim_network_count": (ctypes.c_int32, [vp]),
        "xfd_sim_network_nf": (ctypes.c_int32, [vp, i32]),
        "xfd_sim_network_output": (ctypes.c_int, [vp, i32]),
        "xfd_sim_network_keys": (ctypes.c_int32, [vp, i32, vp, i32]),
        "xfd_sim_network_s": (ctypes.c_int, [vp, i32, i32, vp]),
        "xfd_sim_nf2ff_surface": (ctypes.c_int, [vp, i32, i32, p64, p64]),
        "xfd_sim_nf2ff_current": (ctypes.c_int, [vp, i32, i32, i32, i32, vp]),
        "xfd_sim_nf2ff_potential": (ctypes.c_int, [vp, i32, i32, i32, i32, i32, vp, i32, vp, vp, vp]),
        "xfd_sim_nf2ff_power": (ctypes.c_int, [vp, i32, i32, ctypes.POINTER(ctypes.c_double)]),
        "xfd_sim_nf2ff_process_far_field": (ctypes.c_int, [vp, i32, i32, vp, i32, vp, ctypes.c_char_p, vp]),
        "xfd_sim_nf2ff_output_power": (ctypes.c_int, [vp, i32]),
        "xfd_sim_grid_array": (i64, [vp, i32, i32, vp]),
        "xfd_sim_field": (ctypes.c_int, [vp, i32, p64, vp]),
        "xfd_sim_coeff": (ctypes.c_int, [vp, cp, p64, vp]),
        "xfd_sim_monitor_count": (ctypes.c_int, [vp]),
        "xfd_sim_monitor_output": (ctypes.c_int, [vp, i32]),
        "xfd_sim_monitor_records": (ctypes.c_int, [vp, i32]),
        "xfd_sim_monitor_record": (ctypes.c_int, [vp, i32, i32, p64, vp]),
        "xfd_sim_kernel_count": (ctypes.c_int, [vp]),
        "xfd_sim_kernel_stat": (ctypes.c_int, [vp, i32, cp, i32, p64, ctypes.POINTER(ctypes.c_double)]),
        "xfd_last_error": (cp, []),
        "xfd_version": (cp, []),
        "xfd_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32)]),
        "xfd_device_synchronize": (ctypes.c_int, []),
        "xfd_device_info": (ctypes.c_int, [cp, i32, p64]),
        "xfd_create": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
        "xfd_destroy": (None, [vp]),
        "xfd_step": (ctypes.c_int, [vp, i64, i64]),
        "xfd_read_field": (ctypes.c_int, [vp, i32, vp]),
        "xfd_write_field": (ctypes.c_int, [vp, i32, vp]),
        "xfd_read_gather": (ctypes.c_int, [vp, i32, vp]),
        "xfd_set_profiling": (ctypes.c_int, [vp, i32]),
        "xfd_set_profiling_classes": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "xfd_kernel_stats": (ctypes.c_int, [vp, i32, ctypes.POINTER(cp), p64, ctypes.POINTER(ctypes.c_double)]),
        "xfd_device_bytes": (ctypes.c_int, [vp, p64]),
        "xfd_coef_layout": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int32), p64]),
        "xfd_reset": (ctypes.c_int, [vp]),
        "xfd_read_line": (ctypes.c_int, [vp, i32, vp]),
        "xfd_nf2ff_far_field": (ctypes.c_int, [vp, i32, i32, i32, vp, vp, vp, dbl, vp, vp, vp, vp]),
        "xfd_nf2ff_far_field_noalloc": (ctypes.c_int, [vp, i32, i32, i32, vp, vp, vp, dbl, vp, vp, vp, vp]),
        "xfd_sim_device_context": (vp, [vp]),
        "xfd_batch_create": (ctypes.c_int, [i32, vp, ctypes.POINTER(vp)]),
        "xfd_batch_destroy": (None, [vp]),
        "xfd_batch_case": (ctypes.c_int, [vp, i32, ctypes.POINTER(vp)]),
        "xfd_batch_info": (ctypes.c_int, [vp, p32, p32, p32]),
        "xfd_batch_tile": (ctypes.c_int, [vp, p32, p32]),
        "xfd_batch_set_graph": (ctypes.c_int, [vp, i32]),
        "xfd_batch_tune": (ctypes.c_int, [vp, i32]),
        "xfd_batch_step": (ctypes.c_int, [vp, i64, i64]),
        "xfd_batch_last_ms": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double)]),
        "xfd_sim_set_design": (ctypes.c_int, [vp, p64, p64, vp]),
        "xfd_sim_forward": (ctypes.c_int, [vp, i64, i64, ctypes.POINTER(ctypes.c_double)]),
        "xfd_sim_backward": (ctypes.c_int, [vp, vp]),
        "xfd_sim_objective_dft": (ctypes.c_int, [vp, vp]),
        "xfd_forward_checkpointed": (ctypes.c_int, [vp, i64, i64]),
        "xfd_design_edges": (ctypes.c_int, [vp, vp, vp]),
        "xfd_backward": (ctypes.c_int, [vp, vp, vp]),
        "xfd_sim_batch_create": (vp, [i32, ctypes.POINTER(vp)]),
        "xfd_sim_batch_destroy": (None, [vp]),
        "xfd_sim_batch_init": (ctypes.c_int, [vp, i64]),
        "xfd_sim_batch_set_graph": (ctypes.c_int, [vp, i32]),
        "xfd_sim_batch_tune": (ctypes.c_int, [vp, i32]),
        "xfd_sim_batch_advance": (ctypes.c_int, [vp, i64, i64]),
        "xfd_sim_batch_finish": (ctypes.c_int, [vp]),
        "xfd_sim_batch_info": (ctypes.c_int, [vp, p32, p32, p32, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double)]),
        "xfd_sim_batch_tile": (ctypes.c_int, [vp, p32, p32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def device_count() -> int:
    n = ctypes.c_int32(0)
    if load_library().xfd_device_count(ctypes.byref(n)) != 0:
        return 0
    return int(n.value)


def device_synchronize():
    L = load_library()
    if L.xfd_device_synchronize() != 0:
        raise RuntimeError(L.xfd_last_error().decode())


def device_info():
    """(name, HBM bytes) of the current device."""
    L = load_library()
    name = ctypes.create_string_buffer(128)
    total = ctypes.c_int64(0)
    if L.xfd_device_info(name, 128, ctypes.byref(total)) != 0:
        raise RuntimeError(L.xfd_last_error().decode())
    return name.value.decode(), int(total.value)


_dp = ctypes.POINTER(ctypes.c_double)


class CpmlDesc(ctypes.Structure):
    """xfd_cpml_desc (include/xfdtd_core_amd.h)."""

    _fields_ = [("axis", ctypes.c_int32), ("n", ctypes.c_int32), ("e_start", ctypes.c_int32),
                ("h_start", ctypes.c_int32), ("a_e", _dp), ("b_e", _dp), ("a_h", _dp), ("b_h", _dp),
                ("c_ea_psi_hb", _dp), ("c_eb_psi_ha", _dp), ("c_ha_psi_eb", _dp), ("c_hb_psi_ea", _dp)]


class TfsfDesc(ctypes.Structure):
    """xfd_tfsf_desc."""

    _fields_ = [("origin", ctypes.c_int32 * 3), ("size", ctypes.c_int32 * 3), ("px_int", _dp), ("py_int", _dp),
                ("pz_int", _dp), ("px_half", _dp), ("py_half", _dp), ("pz_half", _dp), ("nt", ctypes.c_int64),
                ("aux", ctypes.c_int64), ("e_inc", _dp), ("h_inc", _dp), ("transform_e", ctypes.c_double * 3),
                ("transform_h", ctypes.c_double * 3), ("ca", ctypes.c_double * 3), ("cb", ctypes.c_double * 3),
                ("kind", ctypes.c_int32), ("forward", ctypes.c_int32)]


class LumpedDesc(ctypes.Structure):
    """xfd_lumped_desc."""

    _fields_ = [("component", ctypes.c_int32), ("origin", ctypes.c_int32 * 3), ("size", ctypes.c_int32 * 3),
                ("coff", _dp), ("nt", ctypes.c_int64), ("waveform", _dp), ("kind", ctypes.c_int32),
                ("coff_j", _dp)]


class GatherDesc(ctypes.Structure):
    """xfd_gather_desc."""

    _fields_ = [("field", ctypes.c_int32), ("origin", ctypes.c_int32 * 3), ("end", ctypes.c_int32 * 3),
                ("first", ctypes.c_int64), ("interval", ctypes.c_int64), ("count", ctypes.c_int64)]


# kernel classes of xfd_kernel_stats, in bit order
KERNEL_CLASSES = ("update_h", "update_e", "cpml_h", "cpml_e", "records_h", "records_e", "gather", "nf2ff", "step")


class Desc(ctypes.Structure):
    """xfd_desc."""

    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("coeff", _dp * 18), ("n_cpml", ctypes.c_int32), ("cpml", ctypes.POINTER(CpmlDesc)),
                ("n_lumped", ctypes.c_int32), ("lumped", ctypes.POINTER(LumpedDesc)),
                ("n_tfsf", ctypes.c_int32), ("tfsf", ctypes.POINTER(TfsfDesc)),
                ("n_gather", ctypes.c_int32), ("gather", ctypes.POINTER(GatherDesc)),
                ("ade", ctypes.c_void_p), ("n_nf2ff", ctypes.c_int32), ("nf2ff", ctypes.c_void_p),
                ("n_line", ctypes.c_int32), ("line", ctypes.c_void_p),
                ("n_tdnf", ctypes.c_int32), ("tdnf", ctypes.c_void_p), ("dim", ctypes.c_int32)]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class Context:
    """The hot-path C-ABI (xfd_create / xfd_step / xfd_read_field) driven
    with coefficient arrays supplied by the caller (no CPML, TFSF or lumped
    elements): the Yee loop of BasicUpdator3D alone."""

    def __init__(self, coeffs, dtype: str = "f64", gathers=(), dim: int = 3):
        """`dim` 2 steps a 2D (TE, nz = 1) grid, 1 a 1D (TEM, nx = ny = 1)
        one: BasicUpdatorTE / BasicUpdatorTEM instead of BasicUpdator3D."""
        self._L = load_library()
        self._keep = [np.ascontiguousarray(coeffs[n], dtype=np.float64) for n in COEFF_NAMES]
        d = Desc()
        shape_ex = self._keep[0].shape
        d.nx, d.ny, d.nz = shape_ex[0], shape_ex[1] - 1, shape_ex[2] - 1
        d.dtype = 0 if dtype == "f64" else 1
        d.dim = dim
        for i, a in enumerate(self._keep):
            d.coeff[i] = _ptr(a)
        gl = (GatherDesc * max(1, len(gathers)))()
        for i, g in enumerate(gathers):
            gl[i] = g
        d.n_gather = len(gathers)
        d.gather = ctypes.cast(gl, ctypes.POINTER(GatherDesc))
        self._gl = gl
        self.shape = (d.nx, d.ny, d.nz)
        h = ctypes.c_void_p()
        if self._L.xfd_create(ctypes.byref(d), ctypes.byref(h)) != 0:
            raise RuntimeError(self._L.xfd_last_error().decode())
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self._L.xfd_last_error().decode())

    def extent(self, field: str):
        n = list(self.shape)
        c = FIELD_IDS[field] % 3
        h = FIELD_IDS[field] >= 3
        return tuple((n[a] + 1 if a == c else n[a]) if h else (n[a] if a == c else n[a] + 1) for a in range(3))

    def write(self, field: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.extent(field)
        self._check(self._L.xfd_write_field(self._h, FIELD_IDS[field], a.ctypes.data))

    def read(self, field: str) -> np.ndarray:
        a = np.empty(self.extent(field))
        self._check(self._L.xfd_read_field(self._h, FIELD_IDS[field], a.ctypes.data))
        return a

    def step(self, t0: int, n: int):
        self._check(self._L.xfd_step(self._h, t0, n))

    def coef_layout(self):
        """(classed, [table size of Ex..Hz]) as chosen by xfd_create."""
        cl = ctypes.c_int32()
        n = (ctypes.c_int64 * 6)()
        self._check(self._L.xfd_coef_layout(self._h, ctypes.byref(cl), n))
        return bool(cl.value), list(n)

    def close(self):
        if getattr(self, "_h", None):
            self._L.xfd_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Simulation:
    """A scenario (spec text) run by the native library."""

    def __init__(self, spec: str, precision: str = "f64"):
        self._L = load_library()
        self._h = self._L.xfd_sim_create(spec.encode())
        if not self._h:
            raise ValueError(self._L.xfd_sim_last_error().decode())
        self.set_precision(precision)

    def _check(self, rc: int):
        if rc != 0:
            raise RuntimeError(self._L.xfd_sim_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.xfd_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_precision(self, precision: str):
        if precision not in ("f64", "f32"):
            raise ValueError("precision must be 'f64' or 'f32'")
        self._check(self._L.xfd_sim_set_precision(self._h, 1 if precision == "f32" else 0))

    def set_profiling(self, on: bool = True, classes=None):
        """HIP-event timing of every kernel class, or only of `classes`
        (names as in kernel_stats(), e.g. ["update_e"])."""
        if classes is None:
            self._check(self._L.xfd_sim_set_profiling(self._h, int(on)))
            return
        mask = 0
        for c in classes:
            mask |= 1 << KERNEL_CLASSES.index(c)
        self._check(self._L.xfd_sim_set_profiling_classes(self._h, mask if on else 0))

    def precompute(self, steps: int):
        self._check(self._L.xfd_sim_precompute(self._h, steps))

    def init(self, steps: int):
        self._check(self._L.xfd_sim_init(self._h, steps))

    def run_loop(self):
        self._check(self._L.xfd_sim_run_loop(self._h))

    def run(self, steps: int):
        self._check(self._L.xfd_sim_run(self._h, steps))

    def advance(self, t0: int, n: int):
        """Runs steps t0 .. t0+n-1 on the GPU; returns when they are done."""
        self._check(self._L.xfd_sim_advance(self._h, t0, n))

    def finish(self):
        """Copies monitor recordings and final fields back."""
        self._check(self._L.xfd_sim_finish(self._h))

    def set_design(self, origin, eps_r: np.ndarray):
        """Replaces the design cells' relative permittivity (shape = design
        box size, cells from `origin`)."""
        e = np.ascontiguousarray(eps_r, dtype=np.float64)
        o = (ctypes.c_int64 * 3)(*[int(v) for v in origin])
        n = (ctypes.c_int64 * 3)(*[int(v) for v in e.shape])
        self._check(self._L.xfd_sim_set_design(self._h, o, n, e.ctypes.data))

    def forward(self, steps: int, checkpoint_interval: int) -> float:
        """Adjoint-ready run: init + time loop keeping a state checkpoint every
        `checkpoint_interval` steps; returns the DFT objective J (needs a
        design region and an objective in the scenario)."""
        v = ctypes.c_double()
        self._check(self._L.xfd_sim_forward(self._h, steps, checkpoint_interval, ctypes.byref(v)))
        return float(v.value)

    def backward(self, design_shape) -> np.ndarray:
        """dJ/d eps_r of every design cell, shaped `design_shape`."""
        g = np.zeros(tuple(int(n) for n in design_shape), dtype=np.float64)
        self._check(self._L.xfd_sim_backward(self._h, g.ctypes.data))
        return g

    def objective_dft(self, n_points: int) -> np.ndarray:
        d = np.zeros(n_points, dtype=np.complex128)
        self._check(self._L.xfd_sim_objective_dft(self._h, d.ctypes.data))
        return d

    @property
    def dt(self) -> float:
        return self._L.xfd_sim_dt(self._h)

    @property
    def loop_seconds(self) -> float:
        return self._L.xfd_sim_loop_seconds(self._h)

    def grid(self) -> Tuple[int, int, int]:
        n = (ctypes.c_int64 * 3)()
        self._check(self._L.xfd_sim_grid(self._h, n))
        return tuple(n)

    def device_bytes(self) -> int:
        return int(self._L.xfd_sim_device_bytes(self._h))

    def cpml_fused(self) -> int:
        """Which CPML slabs the Yee sweeps apply in the same pass: 0 none,
        1 the leading x/y slabs (XFD_CPML_FUSED=1), 2 the z slabs (FP32
        when nz > 192; either precision with XFD_CPML_ZFUSED=1; =0 turns it
        off)."""
        return int(self._L.xfd_sim_cpml_fused(self._h))

    def coef_classed(self) -> bool:
        """True when the device stores coefficient class ids, not 18 arrays."""
        return self.coef_layout() != 0

    def coef_layout(self) -> int:
        """0: the 18 dense arrays; 1: class ids; 2: 32-bit class ids, when a
        component has more than 65535 distinct triples (xfd_coef_layout)."""
        return int(self._L.xfd_sim_coef_classed(self._h))

    # ------------------------------------------------------- time-domain NF2FF --
    def nf2ff_td_potential(self, n: int, d: int, which: int) -> np.ndarray:
        """Potential `which` (0 ua, 1 ub, 2 wa, 3 wb) of surface d after a run."""
        num = int(self._L.xfd_sim_nf2ff_td_num(self._h, n))
        if num < 0:
            raise IndexError("bad time-domain NF2FF index")
        out = np.zeros(num)
        self._check(self._L.xfd_sim_nf2ff_td_potential(self._h, n, d, which, out.ctypes.data))
        return out

    def nf2ff_td_far_field(self, n: int = 0) -> np.ndarray:
        """processFarField: rows w_theta, w_phi, u_theta, u_phi (and the .npy
        files when the NF2FF has an output directory)."""
        num = int(self._L.xfd_sim_nf2ff_td_num(self._h, n))
        if num < 0:
            raise IndexError("bad time-domain NF2FF index")
        out = np.zeros((4, num))
        self._check(self._L.xfd_sim_nf2ff_td_far_field(self._h, n, out.ctypes.data))
        return out

    # -------------------------------------------------------------- networks --
    def network_count(self) -> int:
        return int(self._L.xfd_sim_network_count(self._h))

    def network_output(self, n: int = 0) -> dict:
        """Network::output on the last run's recordings (writes the .npy
        files when the network has an output directory); returns
        {10 j + i: S(j, i) over the frequencies}."""
        self._check(self._L.xfd_sim_network_output(self._h, n))
        cnt = int(self._L.xfd_sim_network_keys(self._h, n, None, 0))
        if cnt < 0:
            self._check(-1)
        keys = (ctypes.c_int32 * max(cnt, 1))()
        self._L.xfd_sim_network_keys(self._h, n, keys, cnt)
        out = {}
        for q in range(cnt):
            nf = self._network_nf(n)
            a = np.zeros(nf, dtype=np.complex128)
            self._check(self._L.xfd_sim_network_s(self._h, n, int(keys[q]), a.ctypes.data))
            out[int(keys[q])] = a
        return out

    def _network_nf(self, n: int) -> int:
        return int(self._L.xfd_sim_network_nf(self._h, n))

    # ---------------------------------------------------------------- NF2FF --
    def nf2ff_count(self) -> int:
        return int(self._L.xfd_sim_nf2ff_count(self._h))

    def nf2ff_surface(self, n: int, d: int):
        lo = (ctypes.c_int64 * 3)()
        hi = (ctypes.c_int64 * 3)()
        self._check(self._L.xfd_sim_nf2ff_surface(self._h, n, d, lo, hi))
        return tuple(lo), tuple(hi)

    def nf2ff_current(self, n: int, f: int, d: int, which: int) -> np.ndarray:
        """Accumulated surface current on surface d: which 0 ja, 1 jb, 2 ma, 3 mb."""
        lo, hi = self.nf2ff_surface(n, d)
        shape = tuple(h - l for l, h in zip(lo, hi))
        out = np.zeros(int(np.prod(shape)), dtype=np.complex128)
        self._check(self._L.xfd_sim_nf2ff_current(self._h, n, f, d, which, out.ctypes.data))
        return out.reshape(shape)

    def nf2ff_potential(self, n: int, f: int, a_pot: bool, theta_comp: bool, theta, phi, origin=(0.0, 0.0, 0.0)):
        th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        ph = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
        org = np.ascontiguousarray(origin, dtype=np.float64)
        out = np.zeros(th.size * ph.size, dtype=np.complex128)
        self._check(self._L.xfd_sim_nf2ff_potential(self._h, n, f, int(a_pot), int(theta_comp), th.size,
                                                    th.ctypes.data, ph.size, ph.ctypes.data, org.ctypes.data,
                                                    out.ctypes.data))
        return out

    def nf2ff_power(self, n: int, f: int) -> float:
        v = ctypes.c_double()
        self._check(self._L.xfd_sim_nf2ff_power(self._h, n, f, ctypes.byref(v)))
        return v.value

    def nf2ff_process_far_field(self, n: int, theta, phi, sub_dir: str, origin=(0.0, 0.0, 0.0)):
        """processFarField: one of theta / phi is a scalar."""
        th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        ph = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
        org = np.ascontiguousarray(origin, dtype=np.float64)
        self._check(self._L.xfd_sim_nf2ff_process_far_field(self._h, n, th.size, th.ctypes.data, ph.size,
                                                            ph.ctypes.data, sub_dir.encode(), org.ctypes.data))

    def nf2ff_output_power(self, n: int):
        self._check(self._L.xfd_sim_nf2ff_output_power(self._h, n))

    def device_context(self):
        """The xfd_ctx of the initialised scenario (for direct C-ABI calls)."""
        return self._L.xfd_sim_device_context(self._h)

    def grid_array(self, which: int, axis: int) -> np.ndarray:
        n = self._L.xfd_sim_grid_array(self._h, which, axis, None)
        a = np.empty(n)
        self._L.xfd_sim_grid_array(self._h, which, axis, a.ctypes.data)
        return a

    def _arr(self, fn, key) -> np.ndarray:
        shape = (ctypes.c_int64 * 3)()
        self._check(fn(self._h, key, shape, None))
        a = np.empty(tuple(shape))
        self._check(fn(self._h, key, shape, a.ctypes.data))
        return a

    def field(self, name: str) -> np.ndarray:
        return self._arr(self._L.xfd_sim_field, FIELD_IDS[name])

    def coeff(self, name: str) -> np.ndarray:
        return self._arr(self._L.xfd_sim_coeff, name.encode())

    def monitor_records(self, m: int) -> List[np.ndarray]:
        out = []
        for r in range(self._L.xfd_sim_monitor_records(self._h, m)):
            shape = (ctypes.c_int64 * 3)()
            self._check(self._L.xfd_sim_monitor_record(self._h, m, r, shape, None))
            a = np.empty(tuple(shape))
            self._check(self._L.xfd_sim_monitor_record(self._h, m, r, shape, a.ctypes.data))
            out.append(a if shape[2] != 1 or shape[0] != 2 else a[:, :, 0])
        return out

    def monitor_count(self) -> int:
        return self._L.xfd_sim_monitor_count(self._h)

    def monitor_output(self, m: int):
        """Monitor::output() of monitor m: writes its .npy file(s)."""
        self._check(self._L.xfd_sim_monitor_output(self._h, m))

    def kernel_stats(self):
        out = {}
        for k in range(self._L.xfd_sim_kernel_count(self._h)):
            name = ctypes.create_string_buffer(64)
            launches = ctypes.c_int64()
            ms = ctypes.c_double()
            self._check(self._L.xfd_sim_kernel_stat(self._h, k, name, 64, ctypes.byref(launches), ctypes.byref(ms)))
            out[name.value.decode()] = (launches.value, ms.value)
        return out


class SimulationBatch:
    """Several scenarios stepped together on one GPU (xfdtd::SimulationBatch).

    Every phase of a time step is one launch over the work of all cases; with
    ``graph=True`` one step is captured as a HIP graph and replayed.  Results
    are read from each Simulation as after ``run()``.
    """

    def __init__(self, sims: Sequence["Simulation"]):
        self._L = load_library()
        self.sims = list(sims)  # keeps the handles alive
        arr = (ctypes.c_void_p * len(self.sims))(*[s._h for s in self.sims])
        self._h = self._L.xfd_sim_batch_create(len(self.sims), arr)
        if not self._h:
            raise ValueError(self._L.xfd_sim_last_error().decode())

    def _check(self, rc: int):
        if rc != 0:
            raise RuntimeError(self._L.xfd_sim_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.xfd_sim_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def init(self, steps: int):
        self._check(self._L.xfd_sim_batch_init(self._h, steps))

    def set_graph(self, on: bool = True):
        self._check(self._L.xfd_sim_batch_set_graph(self._h, int(on)))

    def tune(self, reps: int = 5):
        self._check(self._L.xfd_sim_batch_tune(self._h, reps))

    def advance(self, t0: int, n: int):
        self._check(self._L.xfd_sim_batch_advance(self._h, t0, n))

    def finish(self):
        self._check(self._L.xfd_sim_batch_finish(self._h))

    def info(self) -> dict:
        sh, ic, gr = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        ls, ms = ctypes.c_double(), ctypes.c_double()
        self._check(self._L.xfd_sim_batch_info(self._h, ctypes.byref(sh), ctypes.byref(ic), ctypes.byref(gr),
                                               ctypes.byref(ls), ctypes.byref(ms)))
        lk, rj = ctypes.c_int32(), ctypes.c_int32()
        self._check(self._L.xfd_sim_batch_tile(self._h, ctypes.byref(lk), ctypes.byref(rj)))
        return {"shared": bool(sh.value), "ichunk": ic.value, "grid": gr.value, "loop_seconds": ls.value,
                "device_ms": ms.value, "tile": (lk.value, rj.value)}
