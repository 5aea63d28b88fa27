"""GPU parity of the 2D (TE) and 1D (TEM) paths against the oracle and
against closed forms.

A 2D grid (objects infinite along z) runs BasicUpdatorTE's Ez update and
the generic H update (src/updator/basic_updator_te.cpp:23-38,
basic_updator.cpp:23-41) with TFSF2D (tfsf_2d.cpp, tfsf_corrector.cpp:
153-175) and x/y CPML; a 1D grid (objects infinite along x and y) runs
BasicUpdatorTEM (basic_updator.cpp:90-109) with TFSF1D (tfsf_1d.cpp,
tfsf_corrector.cpp:124-151) and z CPML.  FP64 results must equal the
oracle's (relative L2 <= 1e-12, every value bitwise in practice); FP32 runs
state their error against the FP64 oracle.
"""

import sys

import numpy as np
import pytest

import eigenmode as em
from oracle import RefSim, rel_l2
from xfdtd_core_amd import native, scenarios
from xfdtd_core_amd import spec as S

pytestmark = pytest.mark.gpu

FIELDS = ("EX", "EY", "EZ", "HX", "HY", "HZ")
INF = float("inf")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare(spec, steps, label, precision="f64"):
    say(f"[gpu] {label}: oracle {steps} steps")
    ref = RefSim(spec)
    ref.run(steps)
    say(f"[gpu] {label}: device {precision}")
    dev = native.Simulation(spec, precision)
    dev.run(steps)
    assert dev.grid() == ref.grid()
    out = {}
    for f in FIELDS:
        a, b = dev.field(f), ref.field(f)
        assert a.shape == b.shape, f
        out[f] = (rel_l2(a, b), float(np.mean(a == b)), float(np.abs(b).max()))
    for m in range(dev.monitor_count()):
        ra, rb = dev.monitor_records(m), ref.monitor_records(m)
        assert len(ra) == len(rb), (m, len(ra), len(rb))
        errs = [rel_l2(x, y) for x, y in zip(ra, rb)]
        eq = [float(np.mean(x == y)) for x, y in zip(ra, rb)]
        out[f"monitor{m}"] = (max(errs) if errs else 0.0, min(eq) if eq else 1.0, 1.0)
    for k, (e, q, _) in out.items():
        say(f"  {label} {k}: rel_l2={e:.3e} bitwise_fraction={q:.6f}")
    return out, ref, dev


@pytest.mark.parametrize("eps_r", [0.0, 3.5])
def test_te_cylinder_fp64_matches_oracle(eps_r):
    """The reference's 2D example at reduced size (PEC or dielectric
    cylinder, TFSF2D at phi = pi/4, x/y CPML, Ez movie, Ez probes): every
    field, frame and probe series equal to the oracle."""
    sp = scenarios.cylinder_scatter_2d(nx=120, ny=132, movie_interval=25, probes=4, eps_r=eps_r, tfsf=20,
                                       radius=0.015).spec()
    out, ref, dev = _compare(sp, 500, f"te cylinder eps_r={eps_r}")
    assert out["EZ"][2] > 1e-3  # the wave reached the box
    for k, (e, q, _) in out.items():
        assert e <= 1e-12, (k, e)
        assert q == 1.0, (k, q)
    assert len(dev.monitor_records(0)) == 20


def test_tem_slab_line_fp64_matches_oracle():
    """The reference's 1D slab example (TFSF1D forward, z CPML, lossy
    dielectric slab, Ex probes and movie): equal to the oracle."""
    sp = scenarios.slab_line_1d(nz=300, slab=(50, 260), sigma=0.01).spec()
    out, ref, dev = _compare(sp, 700, "tem slab line")
    assert out["EX"][2] > 1e-3
    for k, (e, q, _) in out.items():
        assert e <= 1e-12, (k, e)
        assert q == 1.0, (k, q)


def test_tem_backward_tfsf_fp64_matches_oracle():
    """TFSF1D travelling towards -z (theta = pi, the ZP face's corrections)."""
    s = S.Simulation(1e-3, 1e-3, 1e-3, 0.9)
    s.addObject(S.Object("line", S.Cube(S.Vector(-INF, -INF, 0), S.Vector(INF, INF, 0.2)), S.Material.createAir()))
    s.addObject(S.Object("slab", S.Cube(S.Vector(-INF, -INF, 0.03), S.Vector(INF, INF, 0.05)),
                         S.Material("d", S.ElectroMagneticProperty(2.5, 1.2, 0.0, 0.0))))
    tau = 20e-3 / 6e8
    s.addWaveformSource(S.TFSF1D(30, False, S.Waveform.gaussian(tau, 4.5 * tau)))
    for d in ("ZN", "ZP"):
        s.addBoundary(S.PML(8, d))
    out, _, _ = _compare(s.spec(), 500, "tem backward")
    assert out["EX"][2] > 1e-3
    for k, (e, q, _) in out.items():
        assert e <= 1e-12, (k, e)


def test_te_cavity_eigenmode_through_capi():
    """xfd_create with dim = 2 and the oracle's coefficients: a PEC-cavity
    Ez mode evolves by the closed form, bitwise equal to the oracle."""
    n, dl, steps = (30, 24), (1e-3, 1.2e-3), 150
    s = S.Simulation(dl[0], dl[1], 1.0, 0.8)
    s.addObject(S.Object("cavity", S.Cube(S.Vector(0, 0, -INF), S.Vector(n[0] * dl[0], n[1] * dl[1], INF)),
                         S.Material.createAir()))
    ref = RefSim(s.spec())
    ref.init(steps)
    coeffs = {c: ref.coeff(c) for c in native.COEFF_NAMES}
    n3, dl3 = (n[0], n[1], 1), (dl[0], dl[1], 1.0)
    e0, _, _ = em.mode("EZ", n3)
    ctx = native.Context(coeffs, "f64", dim=2)
    assert ctx.shape == n3
    ctx.write("EZ", e0)
    ctx.step(0, steps)
    ref.write_field("EZ", e0)
    ref.loop()
    for f in FIELDS:
        np.testing.assert_array_equal(ctx.read(f), ref.field(f), err_msg=f)
    g = em.growth("EZ", n3, dl3, ref.dt, steps)
    np.testing.assert_allclose(ctx.read("EZ"), g * e0, rtol=0, atol=2e-12)


def test_te_grid_with_nonzero_ex_ey_hz_through_capi():
    """A 2D grid whose Ex, Ey and Hz are written nonzero: the H half-step
    must then evaluate every term (the reference's updateH on all three
    components) rather than the TE shortcut that assumes them zero; bitwise
    equal to the oracle."""
    n, dl, steps = (30, 24), (1e-3, 1.2e-3), 120
    s = S.Simulation(dl[0], dl[1], 1.0, 0.8)
    s.addObject(S.Object("cavity", S.Cube(S.Vector(0, 0, -INF), S.Vector(n[0] * dl[0], n[1] * dl[1], INF)),
                         S.Material.createAir()))
    ref = RefSim(s.spec())
    ref.init(steps)
    coeffs = {c: ref.coeff(c) for c in native.COEFF_NAMES}
    ctx = native.Context(coeffs, "f64", dim=2)
    rng = np.random.default_rng(3)
    for f in ("EZ", "EX", "EY", "HZ"):
        v = rng.standard_normal(ctx.read(f).shape)
        ctx.write(f, v)
        ref.write_field(f, v)
    ctx.step(0, steps)
    ref.loop()
    for f in FIELDS:
        np.testing.assert_array_equal(ctx.read(f), ref.field(f), err_msg=f)
    assert np.abs(ctx.read("HZ")).max() > 0


def test_tem_standing_wave_through_capi():
    """xfd_create with dim = 1: an Ex standing wave on the line evolves by
    sin(theta/2) = c dt sin(p/2) / dz, bitwise equal to the oracle."""
    nz, dz, cfl, m, steps = 64, 0.7e-3, 0.9, 3, 200
    s = S.Simulation(dz, dz, dz, cfl)
    s.addObject(S.Object("line", S.Cube(S.Vector(-INF, -INF, 0), S.Vector(INF, INF, nz * dz)),
                         S.Material.createAir()))
    ref = RefSim(s.spec())
    ref.init(steps)
    coeffs = {c: ref.coeff(c) for c in native.COEFF_NAMES}
    p = m * np.pi / nz
    e0 = np.zeros((1, 2, nz + 1))
    e0[0, 0, :] = np.sin(p * np.arange(nz + 1))
    ctx = native.Context(coeffs, "f64", dim=1)
    ctx.write("EX", e0)
    ctx.step(0, steps)
    ref.write_field("EX", e0)
    ref.loop()
    for f in FIELDS:
        np.testing.assert_array_equal(ctx.read(f), ref.field(f), err_msg=f)
    theta = 2 * np.arcsin(em.C0 * ref.dt * np.sin(p / 2) / dz)
    g = np.cos((steps + 0.5) * theta) / np.cos(theta / 2)
    np.testing.assert_allclose(ctx.read("EX")[0, 0], g * e0[0, 0], rtol=0, atol=2e-12)


def test_te_fp32_error_against_fp64_oracle():
    """FP32 fields on the 2D cylinder case: relative L2 against the FP64
    oracle, stated, within 1e-4 on the in-pulse Ez movie frames."""
    sp = scenarios.cylinder_scatter_2d(nx=120, ny=132, movie_interval=25, probes=4, eps_r=3.5, tfsf=20,
                                       radius=0.015).spec()
    steps = 300
    ref = RefSim(sp)
    ref.run(steps)
    dev = native.Simulation(sp, "f32")
    dev.run(steps)
    fr_ref, fr_dev = ref.monitor_records(0), dev.monitor_records(0)
    errs = [rel_l2(a, b) for a, b in zip(fr_dev, fr_ref) if np.abs(b).max() > 0]
    say(f"  te fp32: Ez frames rel_l2 max {max(errs):.3e} over {len(errs)} frames, final Ez "
        f"{rel_l2(dev.field('EZ'), ref.field('EZ')):.3e}")
    assert max(errs) <= 1e-4
