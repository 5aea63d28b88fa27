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
* the torch.autograd.Function returns the same gradient.
"""

import copy
import sys

import numpy as np
import pytest

from oracle import RefSim, rel_l2
from xfdtd_core_amd import native, scenarios

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
