"""Pins the oracle's CPML corrector, oblique TFSF, lumped sources, resistor
and monitor index rules against what physics and geometry fix.

* Lumped circuits (simple_capacitor.cpp:63-70 and simple_circuit.cpp of the
  reference's examples): two PEC plates close a loop through a source edge
  and a load edge.  A 1 V step VoltageSource (50 ohm) charging a 10 pF
  Capacitor follows V = 1 - exp(-(t - t0)/RC); the same source into a
  Resistor R_L settles at the divider value R_L/(R_s + R_L); a CurrentSource
  I with internal R_s into R_L settles at I (R_s || R_L).  The loads span
  several edges in series and in parallel, which fixes the node-count scaling
  of every element's coefficient (resistor.cpp:31-127, voltage_source.cpp:
  109-175, current_source.cpp:111-196, capacitor.cpp:16-112).
* CPML (pml_corrector.h:14-52, pml.cpp:286-633): a dipole pulse in a box
  with 8-cell CPML equals the same pulse in a box 4x larger (whose own
  boundary is too far to be seen in the window) to the CPML's reflection
  level; a PEC box misses by O(1).  c_psi taken after the kappa division
  (0.03) or a flipped psi recursion sign (0.08) fail the bound.  The fields
  then decay monotonically (no late-time growth).  A field uniform along the
  slab axis drives no psi, so a cavity mode with two x slabs evolves by the
  closed form.
* Oblique TFSF (tfsf.cpp:118-302, tfsf_corrector.cpp:16-232): in vacuum the
  scattered-field region stays below 1e-3 of the plane wave, the total field
  is polarised along theta-hat (psi = 0), and the pulse reaches points of the
  box at the delays k-hat . r / c of the propagation direction.
* Monitor index rules (grid_space.cpp:225-230, 383-396; field_monitor.cpp:
  17-78; field_time_monitor.cpp:43-95): a coordinate maps to the nearest
  e-node, a box is [origin node, end node), a box starting at node 0 is
  shifted to 1, and a FieldTimeMonitor records the last cell of its box.
  Checked on a cavity eigenmode whose every cell value is known in closed
  form.
"""

import numpy as np
import pytest

import eigenmode as em
from oracle import RefSim
from xfdtd_core_amd import spec as S

MM = 1e-3


# --------------------------------------------------------------- circuits --
def _plates_circuit(axis, source, rs, load, value, steps):
    """simple_circuit.cpp's loop along `axis`: plates 8 x 2 cells at c = 0
    and c = 4, source edge box at a = 0 (4 series x 3 parallel edges), load
    at a = 8, a VoltageMonitor between them at a = 5."""
    s = S.Simulation(MM, MM, MM, 0.98)

    def V(c, a, b):  # (main c, tangential a, b) -> (x, y, z)
        return S.Vector(*[(c, a, b), (b, c, a), (a, b, c)][axis])

    d = "XYZ"[axis]
    s.addObject(S.Object("domain", S.Cube(V(-3 * MM, -3 * MM, -3 * MM), V(10 * MM, 14 * MM, 8 * MM)),
                         S.Material.createAir()))
    s.addObject(S.PecPlane("p0", S.Cube(V(0, 0, 0), V(0, 8 * MM, 2 * MM))))
    s.addObject(S.PecPlane("p1", S.Cube(V(4 * MM, 0, 0), V(0, 8 * MM, 2 * MM))))
    dt = 0.98 / (em.C0 * np.sqrt(3) / MM)
    src = S.VoltageSource if source == "v" else S.CurrentSource
    s.addObject(src("src", S.Cube(V(0, 0, 0), V(4 * MM, 0, 2 * MM)), d + "P", rs, S.Waveform.step(100 * dt)))
    if load == "r":
        s.addObject(S.Resistor("load", S.Cube(V(0, 8 * MM, 0), V(4 * MM, 0, 2 * MM)), d, value))
    else:
        s.addObject(S.Capacitor("load", S.Cube(V(0, 8 * MM, 0), V(4 * MM, 0, 2 * MM)), d, value))
    s.addMonitor(S.VoltageMonitor("vm", S.Cube(V(0, 5 * MM, 0), V(4 * MM, 0, 2 * MM)), d + "P"))
    ref = RefSim(s.spec())
    ref.run(steps)
    t, v = ref.monitor_records(0)[0]
    return t, v, ref.dt


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("rl", [50.0, 150.0])
def test_voltage_source_into_resistor_settles_at_divider(axis, rl):
    _, v, _ = _plates_circuit(axis, "v", 50.0, "r", rl, 3000)
    settled = v[-300:].mean()
    assert abs(settled - rl / (50.0 + rl)) < 2e-4
    assert np.ptp(v[-300:]) < 0.05  # ringing has died down


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_current_source_into_resistor_settles_at_norton_value(axis):
    rs, rl = 100.0, 50.0
    _, v, _ = _plates_circuit(axis, "i", rs, "r", rl, 3000)
    want = rs * rl / (rs + rl)
    assert abs(v[-300:].mean() - want) < 2e-4 * want


def _capacitor_circuit(steps, rs, cap):
    """simple_capacitor.cpp capacitorX: 1 x 1 cell plates at x = 0 and x = 1,
    the source edge at y = 0 and the capacitor and monitor at y = 1."""
    s = S.Simulation(MM, MM, MM, 0.98)
    s.addObject(S.Object("domain", S.Cube(S.Vector(-MM, -MM, -MM), S.Vector(4 * MM, 6 * MM, 4 * MM)),
                         S.Material.createAir()))
    s.addObject(S.PecPlane("p0", S.Cube(S.Vector(0, 0, 0), S.Vector(0, MM, MM))))
    s.addObject(S.PecPlane("p1", S.Cube(S.Vector(MM, 0, 0), S.Vector(0, MM, MM))))
    dt = 0.98 / (em.C0 * np.sqrt(3) / MM)
    s.addObject(S.VoltageSource("v", S.Cube(S.Vector(0, 0, 0), S.Vector(MM, 0, MM)), "XP", rs,
                                S.Waveform.step(100 * dt)))
    s.addObject(S.Capacitor("c", S.Cube(S.Vector(0, MM, 0), S.Vector(MM, 0, MM)), "X", cap))
    s.addMonitor(S.VoltageMonitor("vm", S.Cube(S.Vector(0, MM, 0), S.Vector(MM, 0, MM)), "XP"))
    ref = RefSim(s.spec())
    ref.run(steps)
    t, v = ref.monitor_records(0)[0]
    return t, v, ref.dt


def test_rc_step_closed_form():
    """simple_capacitor.cpp:63-70: V(t) = 1 - exp(-(t - t0)/RC), R = 50 ohm,
    C = 10 pF (RC = 265 steps); the plates' own capacitance and the loop's
    inductance leave a percent-level difference during the charging."""
    rs, cap = 50.0, 1e-11
    t, v, dt = _capacitor_circuit(2000, rs, cap)
    t0 = 100 * dt
    exact = np.where(t < t0, 0.0, 1.0 - np.exp(-(t - t0) / (rs * cap)))
    assert np.abs(v - exact).max() < 0.03
    # one time constant after the step
    q = int(np.argmin(np.abs(t - (t0 + rs * cap))))
    assert abs(v[q] - (1 - np.exp(-1))) < 0.02
    assert abs(v[-200:].mean() - exact[-200:].mean()) < 2e-3
    # twice the capacitance doubles the time constant
    t, v2, _ = _capacitor_circuit(2000, rs, 2 * cap)
    q = int(np.argmin(np.abs(t - (t0 + 2 * rs * cap))))
    assert abs(v2[q] - (1 - np.exp(-1))) < 0.02


# ------------------------------------------------------------------ CPML --
def _dipole_box(half, steps, pml=8, probe_half=10):
    s = S.Simulation(MM, MM, MM, 0.9)
    h = half * MM
    s.addObject(S.Object("domain", S.Cube(S.Vector(-h, -h, -h), S.Vector(2 * h, 2 * h, 2 * h)),
                         S.Material.createAir()))
    tau = 4 * MM / em.C0
    s.addObject(S.CurrentSource("src", S.Cube(S.Vector(0, 0, 0), S.Vector(0, 0, MM)), "ZP", 1e20,
                                S.Waveform.gaussianDerivative(tau, 4.5 * tau)))
    if pml:
        for d in ("XN", "XP", "YN", "YP", "ZN", "ZP"):
            s.addBoundary(S.PML(pml, d))
    p = probe_half * MM
    s.addMonitor(S.MovieMonitor(S.FieldMonitor(S.Cube(S.Vector(-p, -p, -p), S.Vector(2 * p, 2 * p, 2 * p)), "EZ"),
                                1, "m", ""))
    ref = RefSim(s.spec(), threads=(1, 1, 1))
    ref.run(steps)
    return np.array(ref.monitor_records(0))


def test_cpml_reflection_against_larger_box():
    steps = 160  # the large box's own boundary is 86 cells of round trip away
    small = _dipole_box(12, steps)
    large = _dipole_box(48, steps)
    pec = _dipole_box(12, steps, pml=0)
    mask = np.ones(small.shape[1:], bool)
    c = small.shape[1] // 2
    mask[c - 4:c + 4, c - 4:c + 4, c - 4:c + 4] = False  # away from the source's near field
    peak = np.abs(large[:, mask]).max()
    refl = np.abs(small[:, mask] - large[:, mask]).max() / peak
    refl_pec = np.abs(pec[:, mask] - large[:, mask]).max() / peak
    assert refl < 0.015, refl
    assert refl_pec > 0.5, refl_pec


def test_cpml_late_time_decay():
    s = S.Simulation(MM, MM, MM, 0.9)
    h = 12 * MM
    s.addObject(S.Object("domain", S.Cube(S.Vector(-h, -h, -h), S.Vector(2 * h, 2 * h, 2 * h)),
                         S.Material.createAir()))
    tau = 4 * MM / em.C0
    s.addObject(S.CurrentSource("src", S.Cube(S.Vector(0, 0, 0), S.Vector(0, 0, MM)), "ZP", 1e20,
                                S.Waveform.gaussianDerivative(tau, 4.5 * tau)))
    for d in ("XN", "XP", "YN", "YP", "ZN", "ZP"):
        s.addBoundary(S.PML(8, d))
    s.addMonitor(S.FieldTimeMonitor("p", S.Cube(S.Vector(6 * MM, 3 * MM, 2 * MM), S.Vector(MM, MM, MM)), "", "EZ"))
    ref = RefSim(s.spec(), threads=(1, 1, 1))
    ref.run(1200)
    v = np.abs(ref.monitor_records(0)[0][1])
    peaks = [v[i:i + 200].max() for i in range(0, 1200, 200)]
    assert all(b < a for a, b in zip(peaks, peaks[1:])), peaks
    assert peaks[-1] < 1e-6 * peaks[0]


def test_cpml_leaves_slab_uniform_mode_unchanged():
    """An Ex(y, z) cavity mode is uniform along x: the x slabs' dual
    differences vanish, psi stays zero, and the mode evolves by the PEC
    closed form even though the grid now extends through two x slabs."""
    n, dl = (24, 20, 16), (1e-3, 1.2e-3, 0.9e-3)
    steps = 120
    s = S.Simulation(dl[0], dl[1], dl[2], 0.9)
    size = [n[a] * dl[a] for a in range(3)]
    s.addObject(S.Object("cavity", S.Cube(S.Vector(0, 0, 0), S.Vector(*size)), S.Material.createAir()))
    s.addBoundary(S.PML(6, "XN"))
    s.addBoundary(S.PML(6, "XP"))
    ref = RefSim(s.spec())
    ref.init(steps)
    grid = ref.grid()
    assert grid == (n[0] + 12, n[1], n[2])
    e0, _, _ = em.mode("EX", grid)
    ref.write_field("EX", e0)
    ref.loop()
    g = em.growth("EX", grid, dl, ref.dt, steps)
    assert abs(g) > 0.05
    np.testing.assert_allclose(ref.field("EX"), g * e0, rtol=0, atol=2e-12 * np.abs(e0).max())
    for other in ("EY", "EZ", "HX"):
        assert np.abs(ref.field(other)).max() == 0.0


# ------------------------------------------------------------ oblique TFSF --
def _oblique_tfsf(theta, phi, steps, monitors=()):
    dl = 5e-3
    s = S.Simulation(dl, dl, dl, 0.9)
    half = 20 * dl
    s.addObject(S.Object("domain", S.Cube(S.Vector(-half, -half, -half), S.Vector(2 * half, 2 * half, 2 * half)),
                         S.Material.createAir()))
    tau = 20 * dl / 6e8
    s.addWaveformSource(S.TFSF3D(14, 14, 14, theta, phi, 0.0, S.Waveform.gaussian(tau, 4.5 * tau)))
    for d in ("XN", "XP", "YN", "YP", "ZN", "ZP"):
        s.addBoundary(S.PML(8, d))
    for q, (a, b, c) in enumerate(monitors):
        s.addMonitor(S.FieldTimeMonitor(f"p{q}", S.Cube(S.Vector(a * dl + dl / 2, b * dl, c * dl),
                                                        S.Vector(dl, dl, dl)), "", "EX"))
    ref = RefSim(s.spec())
    ref.run(steps)
    return ref, dl


@pytest.mark.parametrize("theta,phi", [(np.pi / 6, np.pi / 4), (np.pi / 3, np.pi / 5), (0.4, 1.9)])
def test_oblique_tfsf_leakage_and_polarisation(theta, phi):
    ref, _ = _oblique_tfsf(theta, phi, 160)
    n = ref.grid()[0]
    lo, hi = 14, n - 14
    total = scat = 0.0
    centre = []
    for f in ("EX", "EY", "EZ"):
        a = ref.field(f)
        total = max(total, np.abs(a[lo + 1:hi - 1, lo + 1:hi - 1, lo + 1:hi - 1]).max())
        outside = a.copy()
        outside[lo - 1:hi + 1, lo - 1:hi + 1, lo - 1:hi + 1] = 0
        scat = max(scat, np.abs(outside).max())
        centre.append(a[n // 2, n // 2, n // 2])
    assert total > 0.5
    assert scat < 1e-3 * total, scat / total
    theta_hat = np.array([np.cos(theta) * np.cos(phi), np.cos(theta) * np.sin(phi), -np.sin(theta)])
    v = np.array(centre)
    assert abs(v @ theta_hat) / np.linalg.norm(v) > 0.998


def test_oblique_tfsf_arrival_delays():
    theta, phi = np.pi / 6, np.pi / 4
    pts = [(0, 0, 0), (0, 0, 6), (6, 0, 0), (0, 6, 0), (-5, -5, -5)]
    ref, dl = _oblique_tfsf(theta, phi, 220, pts)
    k = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    peak = []
    for q in range(len(pts)):
        v = np.abs(ref.monitor_records(q)[0][1])
        m = int(np.argmax(v))
        y0, y1, y2 = v[m - 1:m + 2]
        peak.append(m + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2))
    for q in range(1, len(pts)):
        want = (np.array(pts[q]) * dl) @ k / em.C0 / ref.dt
        assert abs((peak[q] - peak[0]) - want) < 0.05, (pts[q], peak[q] - peak[0], want)


# --------------------------------------------------------- monitor rules --
N = (24, 20, 16)
DL = (1e-3, 1.2e-3, 0.9e-3)


def _node(x, axis):
    """Nearest e-node of coordinate x (cavity nodes at 0, dl, 2 dl, ...)."""
    nodes = np.arange(N[axis] + 1) * DL[axis]
    return int(np.argmin(np.abs(nodes - x)))


def _box(lo, hi):
    """[origin, end) node box of the cube [lo, hi] with the +1 shift of a box
    starting at node 0."""
    o = [_node(lo[a], a) for a in range(3)]
    e = [_node(hi[a], a) for a in range(3)]
    o = [1 if v == 0 else v for v in o]
    return o, e


# cube corners in units of the cell size: interior, rounding both ways,
# touching node 0 on one or two axes, and one whose shifted box is empty
CUBES = [
    ((5.3, 4.6, 2.2), (7.6, 9.4, 6.7)),
    ((0.0, 3.0, 4.0), (6.0, 7.0, 9.0)),
    ((0.2, 0.4, 0.0), (3.0, 2.0, 5.0)),
    ((8.0, 0.0, 1.0), (9.0, 1.0, 2.0)),
]


def _mode_run(cube, kind):
    s_cube = S.Cube(S.Vector(*[cube[0][a] * DL[a] for a in range(3)]),
                    S.Vector(*[(cube[1][a] - cube[0][a]) * DL[a] for a in range(3)]))
    sp = em.cavity_spec(N, DL)
    if kind == "time":
        sp += S.FieldTimeMonitor("p", s_cube, "", "EZ").line() + "\n"
    else:
        sp += S.MovieMonitor(S.FieldMonitor(s_cube, "EZ"), 5, "m", "").line() + "\n"
    steps = 30
    ref = RefSim(sp)
    ref.init(steps)
    e0, _, _ = em.mode("EZ", N)
    ref.write_field("EZ", e0)
    ref.loop()
    g = np.array([em.growth("EZ", N, DL, ref.dt, t) for t in range(steps + 1)])
    return ref, e0, g


@pytest.mark.parametrize("cube", CUBES)
def test_field_time_monitor_records_last_cell_of_box(cube):
    lo = [cube[0][a] * DL[a] for a in range(3)]
    hi = [cube[1][a] * DL[a] for a in range(3)]
    o, e = _box(lo, hi)
    ref, e0, g = _mode_run(cube, "time")
    rec = ref.monitor_records(0)[0][1]
    if any(e[a] <= o[a] for a in range(3)):
        assert np.all(rec == 0.0)  # empty box: nothing recorded
        return
    cell = tuple(v - 1 for v in e)
    want = e0[cell] * g[1:]
    assert abs(e0[cell]) > 0.05
    np.testing.assert_allclose(rec, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("cube", CUBES[:3])
def test_field_monitor_frame_is_shifted_box(cube):
    lo = [cube[0][a] * DL[a] for a in range(3)]
    hi = [cube[1][a] * DL[a] for a in range(3)]
    o, e = _box(lo, hi)
    ref, e0, g = _mode_run(cube, "movie")
    frames = ref.monitor_records(0)
    assert len(frames) == 6  # steps 0, 5, ..., 25 of 30
    box = e0[o[0]:e[0], o[1]:e[1], o[2]:e[2]]
    for r, fr in enumerate(frames):
        assert fr.shape == box.shape
        np.testing.assert_allclose(fr, box * g[5 * r + 1], rtol=0, atol=1e-12)
