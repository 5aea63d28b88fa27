"""Pins the oracle's VoltageMonitor and CurrentMonitor in a PEC cavity
holding a discrete eigenmode (closed form in eigenmode.py):

* CurrentMonitor: the H loop integral around the cube's last cross-section
  equals the discrete Ampere law of the Yee update, eps0/dt * dA * sum of
  (E^{t+1} - E^t) over the enclosed edges (the edges [o, e] x [o, e] of the
  two transverse axes): a wrong side, sign, size or plane fails it.
* VoltageMonitor: -(1/(na nb)) sum of dl * E^{t+1} over the edges of the
  cube's range extended by one along the transverse axes; the negative
  direction flips both.
"""

import numpy as np
import pytest

import eigenmode as em
from oracle import RefSim
from xfdtd_core_amd import spec as S

N = (24, 20, 16)
DL = (1e-3, 1.2e-3, 0.9e-3)
STEPS = 60
# monitor cube in grid nodes: o .. e per axis
O = (8, 6, 4)
E = (12, 11, 10)
FIELD_OF_AXIS = {0: "EX", 1: "EY", 2: "EZ"}
DIRS = {0: ("XN", "XP"), 1: ("YN", "YP"), 2: ("ZN", "ZP")}


def _cube():
    return S.Cube(S.Vector(*[O[a] * DL[a] for a in range(3)]),
                  S.Vector(*[(E[a] - O[a]) * DL[a] for a in range(3)]))


def _run(axis, direction, kind):
    field = FIELD_OF_AXIS[axis]
    sp = em.cavity_spec(N, DL)
    mon = (S.VoltageMonitor if kind == "v" else S.CurrentMonitor)("m", _cube(), direction)
    sp += mon.line() + "\n"
    ref = RefSim(sp)
    ref.init(STEPS)
    e0, _, _ = em.mode(field, N)
    ref.write_field(field, e0)
    ref.loop()
    rec = ref.monitor_records(0)[0]
    g = np.array([em.growth(field, N, DL, ref.dt, t) for t in range(STEPS + 1)])
    return rec[1], e0, g, ref.dt


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("neg", [False, True])
def test_current_monitor_is_discrete_ampere(axis, neg):
    a, b = (axis + 1) % 3, (axis + 2) % 3
    data, e0, g, dt = _run(axis, DIRS[axis][0 if neg else 1], "i")
    idx = [slice(None)] * 3
    idx[a] = slice(O[a], E[a] + 1)
    idx[b] = slice(O[b], E[b] + 1)
    idx[axis] = E[axis] - 1
    enclosed = e0[tuple(idx)].sum()
    want = em.EPS0 / dt * DL[a] * DL[b] * (g[1:] - g[:-1]) * enclosed
    if neg:
        want = -want
    np.testing.assert_allclose(data, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())
    assert np.abs(want).max() > 0


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("neg", [False, True])
def test_voltage_monitor_is_line_integral(axis, neg):
    a, b = (axis + 1) % 3, (axis + 2) % 3
    data, e0, g, dt = _run(axis, DIRS[axis][0 if neg else 1], "v")
    idx = [slice(None)] * 3
    idx[a] = slice(O[a], E[a] + 1)
    idx[b] = slice(O[b], E[b] + 1)
    idx[axis] = slice(O[axis], E[axis])
    box = e0[tuple(idx)]
    na, nb = E[a] + 1 - O[a], E[b] + 1 - O[b]
    want = -DL[axis] * box.sum() / (na * nb) * g[1:]
    if neg:
        want = -want
    np.testing.assert_allclose(data, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())
    assert np.abs(want).max() > 0
