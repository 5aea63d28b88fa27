"""Reverse-time discrete adjoint on the GPU.

A small version of the inverse-design workload (x-directed TFSF plane wave,
random design block, DFT objective on an Ez patch, CPML on all faces):

* the forward objective equals the oracle's (FP64, <= 1e-12);
* the FP64 adjoint gradient equals central finite differences of the
  oracle's objective (h = 1e-3 on eps_r) at single cells and along a random
  direction through all design cells, to <= 1e-3 relative: a wrong sign,
  index or missing transposed term (CPML, boundary faces) fails this;
* the gradient does not depend on the checkpoint interval (the segments are
  recomputed exactly);
* the FP32 adjoint states its error against the FP64 adjoint;
* the torch.autograd.Function returns the same gradient;
* with a dispersive (Drude, Debye or Lorentz ADE) block in the wave's path
  and an inductor between the design block and the objective, the FP64
  gradient still equals central differences of the oracle's objective
  (the transposed ADE update and inductor records);
* the i-marching adjoint sweeps equal the one-point-per-thread transposes
  bitwise in FP64.
"""

import copy
import sys

import numpy as np
import pytest

from oracle import RefSim, rel_l2
from xfdtd_core_amd import native, scenarios
from xfdtd_core_amd import spec as S

pytestmark = pytest.mark.gpu

STEPS = 400
H = 1e-3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(seed=4):
    return scenarios.adjoint_design(n_total=32, pml=6, design=6, patch_gap=3, tfsf_margin=2, seed=seed)


def _oracle_J(s, eps, origin):
    s = copy.deepcopy(s)  # the shared scenario keeps its own permittivity
    s.setDesignRegion(origin, eps.shape, eps)
    r = RefSim(s.spec())
    r.run(STEPS)
    return r.objective()[0]


def _device(s, precision, interval):
    d = native.Simulation(s.spec(), precision)
    J = d.forward(STEPS, interval)
    return d, J


@pytest.fixture(scope="module")
def fp64_grad():
    s, origin, eps = _problem()
    d, J = _device(s, "f64", 50)
    g = d.backward(eps.shape)
    return s, origin, eps, J, g


def test_forward_objective_matches_oracle(fp64_grad):
    s, origin, eps, J, g = fp64_grad
    r = RefSim(s.spec())
    r.run(STEPS)
    Jr, Dr = r.objective()
    say(f"[gpu] adjoint forward J={J:.6e} oracle={Jr:.6e}")
    assert abs(J / Jr - 1) <= 1e-12


def test_gradient_against_central_differences(fp64_grad):
    s, origin, eps, J, g = fp64_grad
    rng = np.random.default_rng(1)
    cells = [tuple(int(v) for v in rng.integers(0, 6, size=3)) for _ in range(6)]
    worst = 0.0
    for c in cells:
        ep, em = eps.copy(), eps.copy()
        ep[c] += H
        em[c] -= H
        fd = (_oracle_J(s, ep, origin) - _oracle_J(s, em, origin)) / (2 * H)
        err = abs(g[c] - fd) / abs(fd)
        worst = max(worst, err)
        say(f"  cell {c}: adjoint {g[c]:.6e} fd {fd:.6e} rel {err:.2e}")
        assert err <= 1e-3
    say(f"  worst single-cell relative error {worst:.2e}")


def test_directional_derivative(fp64_grad):
    s, origin, eps, J, g = fp64_grad
    rng = np.random.default_rng(2)
    v = rng.uniform(-1, 1, size=eps.shape)
    fd = (_oracle_J(s, eps + H * v, origin) - _oracle_J(s, eps - H * v, origin)) / (2 * H)
    ad = float(np.sum(g * v))
    say(f"  directional: adjoint {ad:.6e} fd {fd:.6e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(ad - fd) <= 1e-3 * abs(fd)


@pytest.mark.parametrize("interval", [7, 400])
def test_checkpoint_interval_does_not_matter(fp64_grad, interval):
    s, origin, eps, J, g = fp64_grad
    d, J2 = _device(s, "f64", interval)
    assert J2 == J
    np.testing.assert_array_equal(d.backward(eps.shape), g)


def test_fp32_adjoint_against_fp64(fp64_grad):
    s, origin, eps, J, g = fp64_grad
    d, J32 = _device(s, "f32", 50)
    g32 = d.backward(eps.shape)
    e = rel_l2(g32, g)
    say(f"  fp32 adjoint: J rel {abs(J32 / J - 1):.2e}, gradient rel L2 {e:.2e}")
    assert abs(J32 / J - 1) < 1e-4
    assert e < 1e-3


def test_torch_autograd_function(fp64_grad):
    import torch

    from xfdtd_core_amd.adjoint import DesignProblem

    s, origin, eps, J, g = fp64_grad
    prob = DesignProblem(s.spec(), origin, eps.shape, STEPS, checkpoint_interval=50, precision="f64")
    x = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    val = prob.objective(x)
    val.backward()
    assert abs(val.item() / J - 1) <= 1e-12
    np.testing.assert_allclose(x.grad.numpy(), g, rtol=1e-12, atol=0)


def test_torch_backward_of_an_earlier_forward(fp64_grad):
    """Two objective() calls on one problem before the first backward (a line
    search does this): the first graph's gradient still belongs to its own
    eps_r, recomputed from the saved parameters, not to the second's."""
    import torch

    from xfdtd_core_amd.adjoint import DesignProblem

    s, origin, eps, J, g = fp64_grad
    prob = DesignProblem(s.spec(), origin, eps.shape, STEPS, checkpoint_interval=50, precision="f64")
    x = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    y = torch.tensor(eps * 1.25, dtype=torch.float64, requires_grad=True)
    vx = prob.objective(x)
    vy = prob.objective(y)
    assert vy.item() != vx.item()
    vx.backward()
    np.testing.assert_allclose(x.grad.numpy(), g, rtol=1e-12, atol=0)
    vy.backward()
    assert np.abs(y.grad.numpy() - g).max() > 0


@pytest.mark.parametrize("n_total,layout,ichunk", [(32, "classed", None), (70, "classed", "3"), (70, "dense", None)])
def test_marching_sweeps_equal_point_kernels_fp64(monkeypatch, n_total, layout, ichunk):
    """The i-marching adjoint sweeps (products carried along i, shuffled along
    k) against the one-point-per-thread transposes: in FP64 (no contraction)
    both evaluate every term in the same order, so the gradients agree
    bitwise.  70 cells per axis put tile edges inside the grid (k = 63 / 64)
    and an i-chunk of 3 leaves a partial last chunk."""
    s, origin, eps = scenarios.adjoint_design(n_total=n_total, pml=6, design=6, patch_gap=3, tfsf_margin=2, seed=5)
    if layout == "dense":
        monkeypatch.setenv("XFD_COEF_MODE", "dense")
    if ichunk:
        monkeypatch.setenv("XFD_ICHUNK", ichunk)
    d = native.Simulation(s.spec(), "f64")
    d.forward(STEPS, 100)
    g_sweep = d.backward(eps.shape)
    monkeypatch.setenv("XFD_ADJ_POINT", "1")
    d.forward(STEPS, 100)
    g_point = d.backward(eps.shape)
    say(f"[gpu] adjoint sweeps vs point kernels, {n_total}^3 {layout}: |g| {np.linalg.norm(g_sweep):.3e}")
    assert np.linalg.norm(g_sweep) > 0
    np.testing.assert_array_equal(g_sweep, g_point)


DL = 1e-3
DISPERSIVE = {
    "drude": lambda: S.DrudeMedium.makeDrudeMedium("drude", 1.0, [2e11], [1e10]),
    "debye": lambda: S.DebyeMedium.makeDebyeMedium("debye", 2.0, [6.0], [1e-11]),
    "lorentz": lambda: S.LorentzMedium.makeMLorentz("lorentz", 2.0, [1.5 * 2e10 ** 2, 0.8 * 6e10 ** 2], [0.0, 0.0],
                                                    [2e10 ** 2, 6e10 ** 2], [2e9, 6e9], [1.0, 1.0]),
}


def _state_problem(kind, seed=4):
    """The 32^3 problem with a dispersive block in the plane wave's path
    (cells [9, 12) along x, [10, 22) along y and z: it changes J by 20-60 %)
    and a 5 nH inductor along z between the design block and the objective
    patch (cells (20, 15, 14..16): it changes J by 5 %)."""
    s, origin, eps = _problem(seed)
    s.addObject(S.Object("disp", S.Cube(S.Vector(-7 * DL, -6 * DL, -6 * DL), S.Vector(3 * DL, 12 * DL, 12 * DL)),
                         DISPERSIVE[kind]()))
    s.addObject(S.Inductor("l1", S.Cube(S.Vector(4 * DL, -1 * DL, -2 * DL), S.Vector(0.0, 0.0, 3 * DL)), "Z", 5e-9))
    return s, origin, eps


@pytest.mark.parametrize("kind", ["drude", "debye", "lorentz"])
def test_gradient_through_dispersive_materials_and_inductors(kind):
    s, origin, eps = _state_problem(kind)
    d, J = _device(s, "f64", 50)
    Jr = _oracle_J(s, eps, origin)
    say(f"[gpu] adjoint with {kind} block + inductor: J={J:.6e} oracle={Jr:.6e}")
    assert abs(J / Jr - 1) <= 1e-12
    g = d.backward(eps.shape)
    rng = np.random.default_rng(3)
    for c in [tuple(int(v) for v in rng.integers(0, 6, size=3)) for _ in range(3)]:
        ep, em = eps.copy(), eps.copy()
        ep[c] += H
        em[c] -= H
        fd = (_oracle_J(s, ep, origin) - _oracle_J(s, em, origin)) / (2 * H)
        err = abs(g[c] - fd) / abs(fd)
        say(f"  cell {c}: adjoint {g[c]:.6e} fd {fd:.6e} rel {err:.2e}")
        assert err <= 1e-3
    v = rng.uniform(-1, 1, size=eps.shape)
    fd = (_oracle_J(s, eps + H * v, origin) - _oracle_J(s, eps - H * v, origin)) / (2 * H)
    ad = float(np.sum(g * v))
    say(f"  directional: adjoint {ad:.6e} fd {fd:.6e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(ad - fd) <= 1e-3 * abs(fd)
    d32, J32 = _device(s, "f32", 50)
    e = rel_l2(d32.backward(eps.shape), g)
    say(f"  fp32 adjoint gradient rel L2 {e:.2e}")
    assert e < 1e-3


def test_design_edge_under_a_record_is_refused():
    """The gradient reads each design edge's E after the Yee update, so an
    inductor on a design edge is refused rather than differentiated wrongly."""
    s, origin, eps = _problem()
    o = origin[0] - 16  # design cell 0 along each axis, in cells from the centre
    s.addObject(S.Inductor("bad", S.Cube(S.Vector(o * DL, o * DL, o * DL), S.Vector(0.0, 0.0, 2 * DL)), "Z", 5e-9))
    d, J = _device(s, "f64", 50)
    with pytest.raises(RuntimeError, match="design edge"):
        d.backward(eps.shape)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_wide_class_ids_equal_dense(monkeypatch, precision):
    """A 40^3 design block of random permittivities gives each component
    ~67k distinct coefficient triples, more than 16-bit class ids hold: the
    context takes 32-bit class ids.  Forward objective and gradient equal
    the dense layout's bitwise (the same coefficient values, read either
    way)."""
    s, origin, eps = scenarios.adjoint_design(n_total=64, pml=6, design=40, patch_gap=2, tfsf_margin=3, seed=6)
    runs = {}
    for mode in ("wide", "dense"):
        if mode == "dense":
            monkeypatch.setenv("XFD_COEF_MODE", "dense")
        d = native.Simulation(s.spec(), precision)
        J = d.forward(200, 50)
        layout = d.coef_layout()
        g = d.backward(eps.shape)
        runs[mode] = (J, g, layout)
        d.close()
    say(f"[gpu] 32-bit class ids vs dense coefficients ({precision}): J {runs['wide'][0]:.6e}, layouts "
        f"{runs['wide'][2]} / {runs['dense'][2]}")
    assert runs["wide"][2] == 2 and runs["dense"][2] == 0
    assert runs["wide"][0] == runs["dense"][0] and runs["wide"][0] > 0
    np.testing.assert_array_equal(runs["wide"][1], runs["dense"][1])
