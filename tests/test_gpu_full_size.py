"""The batched design sweep and the adjoint workload at their full sizes.

Parity at these sizes cannot run through the oracle (its cost), so these
tests check what holds at any size:

* design sweep, 96 cases of 48-160 cells per axis, 800 HIP-graph steps:
  the FP64 batch equals each of three cases (the smallest, the largest and a
  random one) stepped alone, bitwise, in the fields and every DFT monitor
  current (the single-context path equals the oracle bitwise on every
  parity case); the FP32 batch's DFT currents are within 1e-4 of the FP64
  batch's for every case;
* adjoint, 160^3 with a 48^3 design block, 2000 steps, checkpoints every
  100: the FP64 gradient equals central differences of FP64 forward runs
  (h = 1e-3 on eps_r) at four random design cells within 1e-3, and the FP32
  gradient is within 1e-3 (relative L2) of the FP64 one.
"""

import sys

import numpy as np
import pytest

from xfdtd_core_amd import native, scenarios

pytestmark = pytest.mark.gpu

FIELDS = ("EX", "EY", "EZ", "HX", "HY", "HZ")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dft(sim, nf=4):
    return {(f, d, w): sim.nf2ff_current(0, f, d, w) for f in range(nf) for d in range(6) for w in range(4)}


def _batch(cases, precision, steps):
    sims = [native.Simulation(c.spec(), precision) for c in cases]
    b = native.SimulationBatch(sims)
    b.init(steps)
    b.set_graph(True)
    b.advance(0, steps)
    b.finish()
    return sims, b


def test_full_size_design_sweep():
    steps = 800
    cases = scenarios.design_sweep(n_cases=96)
    say(f"[gpu] design sweep 96 cases, {steps} FP64 graph steps")
    sims64, b64 = _batch(cases, "f64", steps)
    sizes = [int(np.prod(s.grid())) for s in sims64]
    say(f"  {sum(sizes) / 1e6:.1f} M cells; shared launches {b64.info()['shared']}")
    rng = np.random.default_rng(5)
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), int(rng.integers(0, len(cases)))})
    for q in picks:
        say(f"  case {q} ({sims64[q].grid()}) alone, FP64")
        one = native.Simulation(cases[q].spec(), "f64")
        one.run(steps)
        for f in FIELDS:
            np.testing.assert_array_equal(sims64[q].field(f), one.field(f), err_msg=f"case {q} {f}")
        a, b = _dft(sims64[q]), _dft(one)
        for k in b:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"case {q} dft {k}")
        one.close()
    ref = [_dft(s) for s in sims64]
    for s in sims64:
        s.close()
    b64.close()
    say(f"[gpu] design sweep 96 cases, {steps} FP32 graph steps")
    sims32, b32 = _batch(cases, "f32", steps)
    worst = 0.0
    for q, s in enumerate(sims32):
        got = _dft(s)
        num = sum(float(np.sum(np.abs(got[k] - ref[q][k]) ** 2)) for k in got)
        den = sum(float(np.sum(np.abs(ref[q][k]) ** 2)) for k in got)
        if den > 0:
            worst = max(worst, float(np.sqrt(num / den)))
    say(f"  FP32 DFT currents vs FP64: worst case relative L2 {worst:.2e}")
    assert worst <= 1e-4
    for s in sims32:
        s.close()
    b32.close()


def test_full_size_adjoint():
    steps, K, h = 2000, 100, 1e-3
    s, origin, eps = scenarios.adjoint_design()
    grads = {}
    for prec in ("f64", "f32"):
        say(f"[gpu] adjoint 160^3, {steps} {prec} steps, checkpoints every {K}")
        sim = native.Simulation(s.spec(), prec)
        J = sim.forward(steps, K)
        grads[prec] = sim.backward(eps.shape)
        say(f"  J = {J:.6e}, |grad| = {np.linalg.norm(grads[prec]):.3e}")
        sim.close()
    g64 = grads["f64"]
    d = float(np.linalg.norm(grads["f32"] - g64) / np.linalg.norm(g64))
    say(f"  FP32 vs FP64 gradient relative L2 {d:.2e}")
    assert d <= 1e-3
    rng = np.random.default_rng(9)
    sim = native.Simulation(s.spec(), "f64")
    for _ in range(4):
        c = tuple(int(v) for v in rng.integers(0, eps.shape[0], size=3))
        vals = []
        for sgn in (1, -1):
            e = eps.copy()
            e[c] += sgn * h
            sim.set_design(origin, e)
            vals.append(sim.forward(steps, steps))
        fd = (vals[0] - vals[1]) / (2 * h)
        rel = abs(g64[c] - fd) / abs(fd)
        say(f"  cell {c}: adjoint {g64[c]:.6e}, central difference {fd:.6e}, relative {rel:.2e}")
        assert rel <= 1e-3
    sim.close()
