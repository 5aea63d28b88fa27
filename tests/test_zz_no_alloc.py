"""xfd_nf2ff_far_field_noalloc, the far-field entry point that allocates
nothing (its staging lives in the library's file-scope device array),
against xfd_nf2ff_far_field on the same inputs: bitwise equal results on two
calls in a row, the second with fewer directions, and a clear error above
the scratch array's maximum."""

import ctypes
import sys

import numpy as np
import pytest

from xfdtd_core_amd import native

pytestmark = pytest.mark.gpu


def _call(L, name, ctx, f, theta, phi, grid, k):
    na = len(theta)
    dst = np.full(na * 8, np.nan)
    origin = np.zeros(3)
    arrs = [[np.ascontiguousarray(a, dtype=np.float64) for a in grid[w]] for w in range(3)]
    ptrs = [(ctypes.c_void_p * 3)(*[a.ctypes.data for a in arrs[w]]) for w in range(3)]
    rc = getattr(L, name)(ctx, 0, f, na, theta.ctypes.data, phi.ctypes.data, origin.ctypes.data, k, ptrs[0],
                          ptrs[1], ptrs[2], dst.ctypes.data)
    return rc, dst


def test_far_field_noalloc_matches_allocating_entry_point():
    sys.path.insert(0, "tests")
    from test_oracle_nf2ff import _dipole

    print("[gpu] no-alloc far field: 300 FP64 steps of the dipole with NF2FF", flush=True)
    sim = native.Simulation(_dipole([6e9, 9e9]), "f64")
    sim.run(300)
    L = native.load_library()
    ctx = sim.device_context()
    assert ctx
    e_node = [sim.grid_array(0, a) for a in range(3)]
    h_node = [(e[:-1] + e[1:]) / 2 for e in e_node]
    e_size = [e[1:] - e[:-1] for e in e_node]
    grid = (e_node, h_node, e_size)
    rng = np.random.default_rng(4)
    for f, na in ((0, 181), (1, 37)):  # the second call smaller than the first
        theta = np.ascontiguousarray(rng.uniform(0, np.pi, na))
        phi = np.ascontiguousarray(rng.uniform(0, 2 * np.pi, na))
        k = 2 * np.pi * (6e9, 9e9)[f] / 299792458.0
        rc_a, a = _call(L, "xfd_nf2ff_far_field", ctx, f, theta, phi, grid, k)
        rc_b, b = _call(L, "xfd_nf2ff_far_field_noalloc", ctx, f, theta, phi, grid, k)
        assert rc_a == 0 and rc_b == 0, L.xfd_last_error().decode()
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        np.testing.assert_array_equal(a, b)
        print(f"  frequency {f}: {na} directions, bitwise equal", flush=True)
    # above the maximum: an error, no call
    na = 80000
    theta = np.zeros(na)
    phi = np.zeros(na)
    rc, _ = _call(L, "xfd_nf2ff_far_field_noalloc", ctx, 0, theta, phi, grid, 1.0)
    assert rc != 0
    assert "scratch" in L.xfd_last_error().decode()
