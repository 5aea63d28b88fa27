"""Pins the oracle's 1D (TEM) and 2D (TE) paths against closed forms.

The reference builds a 1D grid when every object is infinite along x and y,
and a 2D grid when they are infinite along z (grid_space_generator.cpp:
16-181).  These tests fix what the mathematics gives for each piece the
restatement adds:

* 2D grid and TE update (grid_space_2d.cpp, basic_updator_te.cpp:23-38,
  time_param.cpp:15-17): a PEC-cavity Ez mode evolves by the discrete
  eigenmode closed form with dt from the two-dimensional CFL formula.
* 1D grid and TEM update (grid_space_1d.cpp, basic_updator.cpp:90-109,
  time_param.cpp:19-21): an Ex standing wave on the line evolves by the 1D
  closed form sin(theta/2) = c dt sin(p/2) / dz.
* TFSF1D / TFSF2D (tfsf_1d.cpp, tfsf_2d.cpp, tfsf_corrector.cpp:124-175,
  tfsf.cpp:153-161): along an axis the auxiliary line is the grid's own
  1D restriction, so the scattered-field side stays at roundoff and the
  total field equals the incident table; oblique 2D incidence leaks below
  1e-3 and reaches points at the delays k-hat . r / c; a dielectric
  half-space reflects (1 - n)/(1 + n) and transmits 2/(1 + n) of the pulse.
* Monitor index rules per dimension (field_monitor.cpp:29-38,
  field_time_monitor.cpp:43-52, grid_space_1d/2d handleTransform): the
  one-cell axes map any finite coordinate to node 0, and the +1 shift of a
  box starting at node 0 applies to x/y unless the grid is 1D and to z only
  in 3D.
"""

import numpy as np
import pytest

import eigenmode as em
from oracle import RefSim
from xfdtd_core_amd import spec as S

INF = float("inf")
MM = 1e-3


# ------------------------------------------------------------- 2D (TE) ---
def _plane(n, dl, cfl):
    s = S.Simulation(dl[0], dl[1], 1.0, cfl)
    s.addObject(S.Object("cavity", S.Cube(S.Vector(0, 0, -INF), S.Vector(n[0] * dl[0], n[1] * dl[1], INF)),
                         S.Material.createAir()))
    return s


def test_te_cavity_eigenmode():
    n, dl, cfl = (30, 24), (1e-3, 1.2e-3), 0.8
    steps = 150
    ref = RefSim(_plane(n, dl, cfl).spec())
    ref.init(steps)
    assert ref.grid() == (30, 24, 1)
    # two-dimensional CFL (the 3D formula with the z cell size 1 differs by 3e-7;
    # the grid's cell sizes carry the node arithmetic's last bits)
    dt2 = cfl / (em.C0 * np.sqrt(1.0 / (dl[0] * dl[0]) + 1.0 / (dl[1] * dl[1])))
    dt3 = cfl / (em.C0 * np.sqrt(1.0 / (dl[0] * dl[0]) + 1.0 / (dl[1] * dl[1]) + 1.0))
    assert abs(ref.dt / dt2 - 1) < 1e-12 and abs(ref.dt / dt3 - 1) > 1e-7
    # GridSpace2D passes based_dx for both in-plane sizes (grid_space_2d.cpp:
    # 14-23): the first y h-size is dy/2 + dx/2
    assert ref.grid_array(3, 1)[0] == pytest.approx(0.5 * (dl[0] + dl[1]), rel=1e-12)
    n3, dl3 = (n[0], n[1], 1), (dl[0], dl[1], 1.0)
    e0, _, _ = em.mode("EZ", n3)
    ref.write_field("EZ", e0)
    ref.loop()
    g = em.growth("EZ", n3, dl3, ref.dt, steps)
    assert abs(g) > 0.05
    np.testing.assert_allclose(ref.field("EZ"), g * e0, rtol=0, atol=2e-12)
    for f in ("EX", "EY", "HZ"):
        assert np.abs(ref.field(f)).max() == 0.0


def _te_tfsf(phi, steps, monitors=(), pml=8, half=30):
    dl = 5e-3
    s = S.Simulation(dl, dl, 1.0, 0.9)
    s.addObject(S.Object("domain", S.Cube(S.Vector(-half * dl, -half * dl, -INF),
                                          S.Vector(2 * half * dl, 2 * half * dl, INF)), S.Material.createAir()))
    tau = 20 * dl / 6e8
    s.addWaveformSource(S.TFSF2D(14, 14, phi, S.Waveform.gaussian(tau, 4.5 * tau)))
    for d in ("XN", "XP", "YN", "YP"):
        s.addBoundary(S.PML(pml, d))
    for q, (a, b) in enumerate(monitors):
        s.addMonitor(S.FieldTimeMonitor(f"p{q}", S.Cube(S.Vector(a * dl + dl / 2, b * dl + dl / 2, -INF),
                                                        S.Vector(dl, dl, INF)), "", "EZ"))
    ref = RefSim(s.spec())
    ref.run(steps)
    return ref, dl


def _leakage_2d(ref, z0=14):
    ez = ref.field("EZ")[:, :, 0]
    n = ref.grid()[0]
    lo, hi = z0, n - z0
    total = np.abs(ez[lo + 1:hi - 1, lo + 1:hi - 1]).max()
    outside = ez.copy()
    outside[lo - 1:hi + 2, lo - 1:hi + 2] = 0
    return total, np.abs(outside).max()


def test_te_tfsf_axis_incidence_is_exact():
    """phi = 0: the wave runs along x, the auxiliary line (dl / ratio_delta =
    dx) is the grid's own row, so nothing leaks out of the box."""
    ref, _ = _te_tfsf(0.0, 140)
    total, scat = _leakage_2d(ref)
    assert total > 0.5
    assert scat < 1e-12 * total
    # TFSF2D corrects Ez and Hx/Hy only; the polarisation is -z (theta = pi/2)
    ez = ref.field("EZ")
    n = ref.grid()[0]
    col = ez[:, n // 2, 0]
    assert col.min() < -0.5 and col.max() < 1e-12
    for f in ("EX", "EY", "HZ"):
        assert np.abs(ref.field(f)).max() == 0.0


@pytest.mark.parametrize("phi", [np.pi / 4, np.pi / 6, 2.0])
def test_te_tfsf_oblique_leakage(phi):
    ref, _ = _te_tfsf(phi, 160)
    total, scat = _leakage_2d(ref)
    assert total > 0.5
    assert scat < 1e-3 * total, scat / total


def test_te_tfsf_arrival_delays():
    phi = np.pi / 5
    pts = [(0, 0), (8, 0), (0, 8), (-6, -6), (7, -4)]
    ref, dl = _te_tfsf(phi, 240, pts)
    k = np.array([np.cos(phi), np.sin(phi)])
    peak = []
    for q in range(len(pts)):
        v = np.abs(ref.monitor_records(q)[0][1])
        m = int(np.argmax(v))
        y0, y1, y2 = v[m - 1:m + 2]
        peak.append(m + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2))
    for q in range(1, len(pts)):
        want = (np.array(pts[q]) * dl) @ k / em.C0 / ref.dt
        assert abs((peak[q] - peak[0]) - want) < 0.05, (pts[q], peak[q] - peak[0], want)


# ------------------------------------------------------------ 1D (TEM) ---
def _line(nz, dz, cfl, pml=0):
    s = S.Simulation(dz, dz, dz, cfl)
    s.addObject(S.Object("line", S.Cube(S.Vector(-INF, -INF, 0), S.Vector(INF, INF, nz * dz)),
                         S.Material.createAir()))
    for d in ("ZN", "ZP")[: 2 if pml else 0]:
        s.addBoundary(S.PML(pml, d))
    return s


def test_tem_line_standing_wave():
    nz, dz, cfl, m = 64, 0.7e-3, 0.9, 3
    steps = 200
    ref = RefSim(_line(nz, dz, cfl).spec())
    ref.init(steps)
    assert ref.grid() == (1, 1, nz)
    assert abs(ref.dt / (cfl * dz / em.C0) - 1) < 1e-12
    p = m * np.pi / nz
    e0 = np.zeros((1, 2, nz + 1))
    e0[0, 0, :] = np.sin(p * np.arange(nz + 1))
    ref.write_field("EX", e0)
    ref.loop()
    theta = 2 * np.arcsin(em.C0 * ref.dt * np.sin(p / 2) / dz)
    g = np.cos((steps + 0.5) * theta) / np.cos(theta / 2)
    assert abs(g) > 0.05
    ex = ref.field("EX")
    np.testing.assert_allclose(ex[0, 0], g * e0[0, 0], rtol=0, atol=2e-12)
    assert np.abs(ex[0, 1]).max() == 0.0  # the j = 1 edge is never updated
    for f in ("EY", "EZ", "HX"):
        assert np.abs(ref.field(f)).max() == 0.0


def _tem_tfsf(forward, steps, slab=None, nz=400, z0=20, tau_cells=20, monitors=()):
    dz = 1e-3
    s = _line(nz, dz, 0.9, pml=10)
    if slab is not None:
        k0, eps_r = slab
        s.addObject(S.Object("slab", S.Cube(S.Vector(-INF, -INF, k0 * dz), S.Vector(INF, INF, (nz - k0) * dz)),
                             S.Material("d", S.ElectroMagneticProperty(eps_r, 1.0, 0.0, 0.0))))
    tau = tau_cells * dz / 6e8
    s.addWaveformSource(S.TFSF1D(z0, forward, S.Waveform.gaussian(tau, 4.5 * tau)))
    for q, k in enumerate(monitors):  # grid node k: the grid starts 10 PML cells below z = 0
        s.addMonitor(S.FieldTimeMonitor(f"p{q}", S.Cube(S.Vector(-INF, -INF, (k - 10) * dz),
                                                        S.Vector(INF, INF, dz)), "", "EX"))
    ref = RefSim(s.spec())
    ref.run(steps)
    return ref


@pytest.mark.parametrize("forward", [True, False])
def test_tem_tfsf_is_exact(forward):
    """The auxiliary line has the grid's own dz and dt, so the scattered-field
    side of the injection face holds roundoff only and the total field equals
    the incident table row at the projected index."""
    z0, steps = 20, 150
    spec_ref = _tem_tfsf(forward, steps + 1)
    e_inc = spec_ref.tfsf_table(0, 0)
    ref = _tem_tfsf(forward, steps)
    ex = ref.field("EX")[0, 0]
    n = ref.grid()[2]
    if forward:
        scat, total_ks = ex[:z0 - 1], np.arange(z0 + 1, n - z0 - 1)
        proj = total_ks - z0 + 2
    else:
        scat, total_ks = ex[n - z0 + 2:], np.arange(z0 + 1, n - z0 - 1)
        proj = (n - z0) - total_ks + 2
    total = ex[total_ks]
    assert np.abs(total).max() > 0.5
    assert np.abs(scat).max() < 1e-12
    # E along theta-hat x psi: +x for theta = 0, -x for theta = pi
    sign = 1.0 if forward else -1.0
    np.testing.assert_allclose(total, sign * e_inc[steps, proj], rtol=0, atol=1e-12)


def test_tem_dielectric_interface_fresnel():
    """A pulse from vacuum onto a half-space of eps_r = 4 (n = 2): the
    reflected pulse in the scattered-field region has r = (1 - n)/(1 + n) =
    -1/3 of the incident peak, the transmitted one 2/(1 + n) = 2/3."""
    k0, eps_r = 120, 4.0
    nrm = np.sqrt(eps_r)
    ref = _tem_tfsf(True, 420, slab=(k0, eps_r), tau_cells=40, monitors=(15, 60, 200))
    refl = ref.monitor_records(0)[0][1]
    inc = ref.monitor_records(1)[0][1]
    trans = ref.monitor_records(2)[0][1]
    peak_inc = np.abs(inc).max()
    assert peak_inc == pytest.approx(1.0, abs=2e-3)
    r = refl[np.argmax(np.abs(refl))] / peak_inc
    t = trans[np.argmax(np.abs(trans))] / peak_inc
    assert r == pytest.approx((1 - nrm) / (1 + nrm), abs=5e-3), r
    assert t == pytest.approx(2 / (1 + nrm), abs=5e-3), t


def test_tem_cpml_absorbs():
    """After the pulse has left through the z slabs, the line between them
    holds less than 1e-3 of the incident peak.  The TFSF face's Ex range
    runs over j in [0, ny + 1) (tfsf_corrector.h:89-229), so the j = 1 edge
    at the injection node, which the TEM update never touches, keeps the sum
    of its corrections: nonzero there and nowhere else on that edge row."""
    ref = _tem_tfsf(True, 1400)
    ex = ref.field("EX")
    assert np.abs(ex[0, 0, 10:411]).max() < 1e-3
    assert list(np.nonzero(ex[0, 1])[0]) == [20]


# ------------------------------------------------------- monitor rules ---
def test_monitor_rules_per_dimension():
    """1D: a box from node 0 is not shifted (x, y and z); 2D: x and y shift
    from node 0, the one-cell z axis does not."""
    nz, dz = 40, 1e-3
    s = _line(nz, dz, 0.9)
    s.addMonitor(S.FieldTimeMonitor("p0", S.Cube(S.Vector(-INF, -INF, 0), S.Vector(INF, INF, dz)), "", "EX"))
    s.addMonitor(S.MovieMonitor(S.FieldMonitor(S.Cube(S.Vector(-INF, -INF, 0), S.Vector(INF, INF, 10 * dz)),
                                               "EX"), 4, "m", ""))
    ref = RefSim(s.spec())
    steps = 12
    ref.init(steps)
    p = np.pi / nz
    e0 = np.zeros((1, 2, nz + 1))
    e0[0, 0, :] = np.cos(p * np.arange(nz + 1))  # nonzero at node 0 until its first update
    ref.write_field("EX", e0)
    ref.loop()
    rec = ref.monitor_records(0)[0][1]
    # node 0 is a PEC end: never updated, so it keeps its initial value
    np.testing.assert_array_equal(rec, np.full(steps, e0[0, 0, 0]))
    frames = ref.monitor_records(1)
    assert len(frames) == 3 and frames[0].shape == (1, 1, 10)
    assert frames[0][0, 0, 0] == e0[0, 0, 0]

    n2, dl2 = (20, 16), (1e-3, 1e-3)
    s = _plane(n2, dl2, 0.8)
    s.addMonitor(S.MovieMonitor(S.FieldMonitor(S.Cube(S.Vector(0, 0, -INF), S.Vector(6e-3, 5e-3, INF)), "EZ"),
                                5, "m", ""))
    s.addMonitor(S.FieldTimeMonitor("p", S.Cube(S.Vector(2.2e-3, 3.6e-3, -INF), S.Vector(2e-3, 1e-3, INF)),
                                    "", "EZ"))
    ref = RefSim(s.spec())
    steps = 20
    ref.init(steps)
    n3, dl3 = (n2[0], n2[1], 1), (dl2[0], dl2[1], 1.0)
    e0, _, _ = em.mode("EZ", n3)
    ref.write_field("EZ", e0)
    ref.loop()
    g = np.array([em.growth("EZ", n3, dl3, ref.dt, t) for t in range(steps + 1)])
    frames = ref.monitor_records(0)
    assert len(frames) == 4
    for r, fr in enumerate(frames):
        assert fr.shape == (5, 4, 1)  # nodes [1, 6) x [1, 5) x [0, 1)
        np.testing.assert_allclose(fr, e0[1:6, 1:5, 0:1] * g[5 * r + 1], rtol=0, atol=1e-12)
    rec = ref.monitor_records(1)[0][1]
    # box nodes [2, 4) x [4, 5) x [0, 1): its last cell (3, 4, 0)
    np.testing.assert_allclose(rec, e0[3, 4, 0] * g[1:], rtol=0, atol=1e-12)


def test_dimension_errors():
    """The reference's refusals: PML along a one-cell axis, a grid that is
    neither 1D, 2D nor 3D, and dispersive material outside 3D."""
    s = _plane((20, 16), (1e-3, 1e-3), 0.8)
    s.addBoundary(S.PML(4, "ZN"))
    with pytest.raises(Exception, match="X or Y"):
        RefSim(s.spec()).init(1)
    s = _line(30, 1e-3, 0.9)
    s.addBoundary(S.PML(4, "XN"))
    with pytest.raises(Exception, match="Z direction"):
        RefSim(s.spec()).init(1)
    s = S.Simulation(1e-3, 1e-3, 1e-3, 0.9)
    s.addObject(S.Object("d", S.Cube(S.Vector(-INF, 0, 0), S.Vector(INF, 0.02, 0.02)), S.Material.createAir()))
    with pytest.raises(Exception, match="undefined"):
        RefSim(s.spec()).init(1)
    s = _line(30, 1e-3, 0.9)
    s.addObject(S.Object("dr", S.Cube(S.Vector(-INF, -INF, 0.01), S.Vector(INF, INF, 0.01)),
                         S.DrudeMedium.makeDrudeMedium("dr", 1.0, [1e11], [1e9])))
    with pytest.raises(Exception, match="updator"):
        RefSim(s.spec()).init(1)
