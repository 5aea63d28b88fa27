"""Pins the oracle's Capacitor, Inductor and PecPlane restatements.

Each element is placed so that a PEC-cavity eigenmode stays an eigenmode
and the discrete result has a closed form:

* a Capacitor on every Ez edge is a uniform extra permittivity for Ez, so the
  TM mode runs at the vacuum closed form with c / sqrt(eps_eff / eps0);
* an Inductor on every Ez edge adds a current state per edge; the mode's
  (H, E, J) amplitudes obey a 3x3 linear recurrence written from the
  reference's update order (Yee H, Yee E, then E += cecjc J, J += cjcec E);
* a PecPlane across the cavity splits it: a mode of one part runs at its own
  closed form and nothing reaches the other part.
"""

import numpy as np
import pytest

import eigenmode as em
from oracle import RefSim
from xfdtd_core_amd import spec as S

EPS0 = em.EPS0
DL = 1e-3


def _cavity(n, extra=()):
    s = S.Simulation(DL, DL, DL, 0.9)
    size = [n[a] * DL for a in range(3)]
    s.addObject(S.Object("cavity", S.Cube(S.Vector(0, 0, 0), S.Vector(*size)), S.Material.createAir()))
    for e in extra:
        s.addObject(e)
    return s.spec()


def _run_mode(spec, field, n_mode, steps, embed_shape=None):
    ref = RefSim(spec)
    ref.init(steps)
    e0, _, _ = em.mode(field, n_mode)
    if embed_shape is not None:
        full = np.zeros(embed_shape)
        full[tuple(slice(0, s) for s in e0.shape)] = e0
        e0 = full
    ref.write_field(field, e0)
    ref.loop()
    return ref, e0


def test_capacitor_is_a_uniform_permittivity():
    """Capacitor on all Ez edges of the box: per-edge capacitance C / nb
    (capacitor.cpp:24-25, with nb = edges along the second tangential axis)
    and beta = 2 C' dc / (da db) give eps_eff = eps0 + C' dc / (da db)."""
    n = (12, 10, 8)
    steps = 150
    extra_eps = 1.5 * EPS0
    nb = n[1] + 1  # tangential B of z is y
    c_total = extra_eps * DL * nb
    cap = S.Capacitor("cap", S.Cube(S.Vector(0, 0, 0), S.Vector(n[0] * DL, n[1] * DL, n[2] * DL)), "Z", c_total)
    ref, e0 = _run_mode(_cavity(n, [cap]), "EZ", n, steps)
    eps_eff = 1 + extra_eps / EPS0
    g = em.growth("EZ", n, (DL, DL, DL), ref.dt / np.sqrt(eps_eff), steps)
    assert abs(g) > 0.05
    np.testing.assert_allclose(ref.field("EZ"), g * e0, rtol=0, atol=1e-11 * np.abs(e0).max())


def _inductor_recurrence(omega2, cd, steps):
    """Ez amplitude after `steps` steps for the modal state (H~, E, J~) with
    H~ = b H, J~ = cecjc J: H~ += omega2 E; E -= H~; E += J~; J~ += cd E."""
    h, e, j = 0.0, 1.0, 0.0
    for _ in range(steps):
        h = h + omega2 * e
        e = e - h
        e = e + j
        j = j + cd * e
    return e


def test_inductor_mode_recurrence():
    """Inductor on all Ez edges: J per edge, E += cecjc J then
    J += cjcec E after the E update (lumped_element_corrrector.cpp:88-93);
    cecjc = -2dt/(2eps+dt sigma), cjcec = dt dc/(L' da db), L' = L na nb / nc
    (inductor.cpp:29-31, 69, 84)."""
    n = (12, 10, 8)
    steps = 200
    wp = 3e11  # plasma-like frequency of the inductive loading
    lprime = 1 / (EPS0 * DL * wp ** 2)
    na, nb, nc = n[0] + 1, n[1] + 1, n[2]
    ind_l = lprime * nc / (na * nb)
    ind = S.Inductor("ind", S.Cube(S.Vector(0, 0, 0), S.Vector(n[0] * DL, n[1] * DL, n[2] * DL)), "Z", ind_l)
    ref, e0 = _run_mode(_cavity(n, [ind]), "EZ", n, steps)
    dt = ref.dt
    # vacuum modal factor: 1 - omega2 = E^1/E^0 of the plain leapfrog
    omega2 = 1 - em.growth("EZ", n, (DL, DL, DL), dt, 1)
    lf = ind_l * na * nb / nc
    cd = (-(2 * dt) / (2 * EPS0 + dt * 1e-20)) * (dt * DL / (lf * DL * DL))
    amp = _inductor_recurrence(omega2, cd, steps)
    np.testing.assert_allclose(ref.field("EZ"), amp * e0, rtol=0, atol=1e-10 * np.abs(e0).max())
    # the loading shifts the mode as w^2 = w_mode^2 + wp^2 (continuous limit)
    w_mode = np.arccos(1 - omega2 / 2) / dt  # from the vacuum recurrence
    ev = np.linalg.eigvals(_inductor_step(omega2, cd))
    w_disc = np.max(np.abs(np.angle(ev))) / dt
    w_cont = np.sqrt(w_mode ** 2 + wp ** 2)
    assert abs(w_disc / w_cont - 1) < 5e-2, (w_disc, w_cont)


def _inductor_step(omega2, cd):
    """One step of _inductor_recurrence as a matrix on (H~, E, J~)."""
    return np.array([[1, omega2, 0],
                     [-1, 1 - omega2, 1],
                     [-cd, cd * (1 - omega2), 1 + cd]])


def test_inductor_recurrence_matrix_matches_loop():
    omega2, cd = 0.3, -0.2
    v = np.array([0.0, 1.0, 0.0])
    for _ in range(7):
        v = _inductor_step(omega2, cd) @ v
    assert abs(v[1] - _inductor_recurrence(omega2, cd, 7)) < 1e-14


CUTS = [("X", "EZ"), ("X", "EY"), ("Y", "EZ"), ("Y", "EX"), ("Z", "EX"), ("Z", "EY")]


@pytest.mark.parametrize("axis,field", CUTS)
def test_pec_plane_splits_cavity(axis, field):
    """PecPlane at 10 cells along `axis`: sigma_e = 1e10 on the plane's
    tangential edges (pec_plane.cpp:22-63).  A mode of the 10-cell part runs
    at its closed form and the far part stays at round-off."""
    n = (16, 14, 13)
    a = "XYZ".index(axis)
    cut = 10
    org = [0.0, 0.0, 0.0]
    size = [n[0] * DL, n[1] * DL, n[2] * DL]
    org[a] = cut * DL
    size[a] = 0.0
    plane = S.PecPlane("plane", S.Cube(S.Vector(*org), S.Vector(*size)))
    n_sub = list(n)
    n_sub[a] = cut
    steps = 120
    comp = "XYZ".index(field[1])
    ext = [n[b] + 1 for b in range(3)]
    ext[comp] = n[comp]
    ref, e0 = _run_mode(_cavity(n, [plane]), field, tuple(n_sub), steps, embed_shape=tuple(ext))
    g = em.growth(field, tuple(n_sub), (DL, DL, DL), ref.dt, steps)
    assert abs(g) > 0.05
    got = ref.field(field)
    near = [slice(None)] * 3
    near[a] = slice(0, cut + 1)
    far = [slice(None)] * 3
    far[a] = slice(cut + 1, None)
    scale = np.abs(e0).max()
    np.testing.assert_allclose(got[tuple(near)], g * e0[tuple(near)], rtol=0, atol=1e-6 * scale)
    assert np.abs(got[tuple(far)]).max() < 1e-6 * scale
    # without the plane the same start leaks across
    ref2, _ = _run_mode(_cavity(n), field, tuple(n_sub), steps, embed_shape=tuple(ext))
    assert np.abs(ref2.field(field)[tuple(far)]).max() > 1e-2 * scale
