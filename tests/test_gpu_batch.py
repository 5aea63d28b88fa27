"""Batched stepping (SimulationBatch / xfd_batch_*): many uneven cases in one
set of persistent launches, optionally replayed as a HIP graph.

The batched launches run the same per-cell code as a single context, so a
case stepped in a batch must equal the same case stepped alone: bitwise in
FP64 (no FMA contraction) and in FP32 within 1e-6 relative L2 (the compiler
may contract the two instantiations differently).  A small FP64 batch is also
compared case by case with the oracle (<= 1e-12), which pins the batched
result independently of the single-context path.
"""

import sys

import numpy as np
import pytest

from oracle import RefSim, rel_l2
from xfdtd_core_amd import native, scenarios

pytestmark = pytest.mark.gpu

FIELDS = ("EX", "EY", "EZ", "HX", "HY", "HZ")
STEPS = 120


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cases(n=5, seed=7, sizes=(24, 32, 40), pml=6):
    return scenarios.design_sweep(n_cases=n, sizes=sizes, pml=pml, freqs=(1e10, 2e10), seed=seed)


def _outputs(sim, nf=2):
    out = {f: sim.field(f) for f in FIELDS}
    for f in range(nf):
        for d in range(6):
            for w in range(4):
                out[f"nf{f}{d}{w}"] = sim.nf2ff_current(0, f, d, w)
    return out


def _alone(specs, precision, steps=STEPS):
    res = []
    for sp in specs:
        s = native.Simulation(sp.spec(), precision)
        s.run(steps)
        res.append(_outputs(s))
        s.close()
    return res


def _batched(specs, precision, steps=STEPS, graph=False, tune=0, chunks=None):
    sims = [native.Simulation(sp.spec(), precision) for sp in specs]
    b = native.SimulationBatch(sims)
    b.init(steps)
    b.set_graph(graph)
    if tune:
        b.tune(tune)
    t = 0
    for c in chunks or [steps]:
        b.advance(t, c)
        t += c
    b.finish()
    info = b.info()
    return sims, b, [_outputs(s) for s in sims], info


def _compare(got, want, exact, label):
    worst = 0.0
    for i, (g, w) in enumerate(zip(got, want)):
        for k in w:
            assert g[k].shape == w[k].shape, (label, i, k)
            if exact:
                np.testing.assert_array_equal(g[k], w[k], err_msg=f"{label} case {i} {k}")
            elif w[k].size:
                worst = max(worst, rel_l2(g[k], w[k]))
    say(f"  {label}: worst relative L2 {worst:.3e}")
    return worst


ETA0 = 376.730313668


def _state_err(a, b):
    """Relative L2 of the whole field state (E, eta0 H): the Gaussian dipole
    leaves a static E field while H decays to roundoff, so a component's own
    norm is no scale for H."""
    da = np.concatenate([a[f].ravel() * (ETA0 if f[0] == "H" else 1.0) for f in FIELDS])
    db = np.concatenate([b[f].ravel() * (ETA0 if f[0] == "H" else 1.0) for f in FIELDS])
    return float(np.linalg.norm(da - db) / np.linalg.norm(db))


@pytest.mark.parametrize("fused", ["0", "1"])
def test_batch_matches_single_cases_fp64(monkeypatch, fused):
    """Also with the x/y CPML slabs applied inside the batched sweeps."""
    monkeypatch.setenv("XFD_CPML_FUSED", fused)
    specs = _cases()
    say(f"[gpu] batch vs single cases f64 (CPML in sweeps: {fused})")
    want = _alone(specs, "f64")
    sims, b, got, info = _batched(specs, "f64")
    assert info["shared"], info
    _compare(got, want, True, "batch f64")


@pytest.mark.parametrize("zfused", ["auto", "1"])
def test_batch_fp32_accuracy_matches_single_cases(monkeypatch, zfused):
    """FP32: the batched and the single-context sweeps differ only in the
    compiler's contraction choices, so the batch is as close to the FP64 run
    as each case run alone.  Batched cases keep their z CPML slabs separate
    even when XFD_CPML_ZFUSED=1 asks single contexts to fuse them."""
    if zfused == "1":
        monkeypatch.setenv("XFD_CPML_ZFUSED", "1")
    else:
        monkeypatch.delenv("XFD_CPML_ZFUSED", raising=False)
    specs = _cases()
    say(f"[gpu] batch vs single cases f32 (XFD_CPML_ZFUSED: {zfused})")
    ref = _alone(specs, "f64")
    single = _alone(specs, "f32")
    sims, b, got, info = _batched(specs, "f32")
    assert info["shared"], info
    for i in range(len(specs)):
        eb, es, ebs = _state_err(got[i], ref[i]), _state_err(single[i], ref[i]), _state_err(got[i], single[i])
        say(f"  case {i}: batch-vs-f64 {eb:.2e}, single-vs-f64 {es:.2e}, batch-vs-single {ebs:.2e}")
        assert eb <= 2 * es + 1e-6
        assert ebs <= 1e-4
        assert es <= 1e-4


def test_batch_fp64_against_oracle():
    specs = _cases(n=3, seed=11)
    say("[gpu] batch fp64 vs oracle")
    sims, b, got, info = _batched(specs, "f64")
    for i, sp in enumerate(specs):
        ref = RefSim(sp.spec())
        ref.run(STEPS)
        want = {f: ref.field(f) for f in FIELDS}
        for f in range(2):
            for d in range(6):
                for w in range(4):
                    want[f"nf{f}{d}{w}"] = ref.nf2ff_current(0, f, d, w)
        for k, v in want.items():
            assert got[i][k].shape == v.shape, (i, k)
            if v.size:
                assert rel_l2(got[i][k], v) <= 1e-12, (i, k, rel_l2(got[i][k], v))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_batch_graph_replays_match_eager(precision):
    """One captured step replayed many times, with the host reading every
    case's outputs between replays, equals the eager batch after each chunk."""
    specs = _cases(n=4, seed=5)
    say(f"[gpu] batch graph replays {precision}")
    chunks = [1, 1, 1, 3, 10, 24, 80]
    sims_g = [native.Simulation(sp.spec(), precision) for sp in specs]
    sims_e = [native.Simulation(sp.spec(), precision) for sp in specs]
    bg, be = native.SimulationBatch(sims_g), native.SimulationBatch(sims_e)
    bg.init(STEPS)
    be.init(STEPS)
    bg.set_graph(True)
    t = 0
    for c in chunks:
        bg.advance(t, c)
        be.advance(t, c)
        t += c
        bg.finish()
        be.finish()
        for i in range(len(specs)):
            og = _outputs(sims_g[i])
            for k, v in _outputs(sims_e[i]).items():
                np.testing.assert_array_equal(og[k], v, err_msg=f"t={t} case {i} {k}")
        say(f"  replay chunk ending at step {t}: equal")
    assert t == STEPS


def test_batch_tune_resets_state():
    """The autotuner steps the sweeps to time them, then clears every case:
    a tuned batch run equals the single cases."""
    specs = _cases(n=4, seed=3)
    say("[gpu] batch tune")
    want = _alone(specs, "f64")
    sims, b, got, info = _batched(specs, "f64", tune=2)
    say(f"  tuned: ichunk={info['ichunk']} grid={info['grid']}")
    assert info["ichunk"] in (4, 8, 16, 32)
    _compare(got, want, True, "tuned batch")


@pytest.mark.parametrize("tile", [0, 2])
def test_batch_with_dispersive_cases_fp64(monkeypatch, tile):
    """Cases with ADE materials share the batched launches: their E sweep is
    one launch over the dispersive cases (k_batch_sweep_ade, each case's own
    scheme: Drude only, Debye only, and the modified-Lorentz scheme a Drude +
    Lorentz mix takes), beside the plain cases' E sweep.  Bitwise the single
    contexts' FP64 results, with a HIP graph, for two sweep tiles."""
    monkeypatch.setenv("XFD_BATCH_TILE", str(tile))
    specs = _cases(n=2, seed=9)
    kw = dict(n_total=40, pml=6, slab=6, gap=2, tfsf_offset=9, lumped=False, nf2ff_distance=8,
              nf2ff_freqs=[1e10, 2e10])
    disp = [scenarios.dispersive_slabs(kinds=("drude", "lorentz"), **kw),
            scenarios.dispersive_slabs(kinds=("drude",), drude_wp=2e11, **kw),
            scenarios.dispersive_slabs(kinds=("debye",), **kw)]
    specs = specs + disp
    say(f"[gpu] batch with dispersive cases, tile {tile}")
    want = _alone(specs, "f64", steps=60)
    sims, b, got, info = _batched(specs, "f64", steps=60, graph=True)
    assert info["shared"]
    _compare(got, want, True, "batch with dispersive cases")


def test_batch_tuner_with_dispersive_cases_fp64():
    """The autotuner times the dispersive cases' E sweep launch too, then
    zeroes every case: the tuned batch still equals the single contexts."""
    kw = dict(n_total=40, pml=6, slab=6, gap=2, tfsf_offset=9, lumped=False, nf2ff_distance=8,
              nf2ff_freqs=[1e10, 2e10])
    specs = _cases(n=2, seed=13) + [scenarios.dispersive_slabs(kinds=("drude", "lorentz"), **kw)]
    say("[gpu] tuned batch with a dispersive case")
    want = _alone(specs, "f64", steps=40)
    sims, b, got, info = _batched(specs, "f64", steps=40, graph=True, tune=1)
    assert info["shared"]
    _compare(got, want, True, "tuned batch with a dispersive case")


def test_batch_with_32bit_class_ids_steps_one_by_one():
    """A case whose coefficients need 32-bit class ids is outside the batched
    launches (they read 16-bit ids): the batch steps its contexts in turn,
    with the same results."""
    s, origin, eps = scenarios.adjoint_design(n_total=64, pml=6, design=40, patch_gap=2, tfsf_margin=3, seed=6)
    specs = _cases(n=1, seed=10) + [s]
    say("[gpu] batch with a 32-bit class-id case")
    want = []
    for sp in specs:
        one = native.Simulation(sp.spec(), "f64")
        one.run(40)
        want.append({f: one.field(f) for f in FIELDS})
        one.close()
    sims = [native.Simulation(sp.spec(), "f64") for sp in specs]
    b = native.SimulationBatch(sims)
    b.init(40)
    b.advance(0, 40)
    b.finish()
    assert sims[1].coef_layout() == 2
    assert not b.info()["shared"]
    _compare([{f: x.field(f) for f in FIELDS} for x in sims], want, True, "fallback batch")


@pytest.mark.parametrize("tile", [1, 2, 3, 4])
def test_batch_sweep_tiles_match_single_cases_fp64(monkeypatch, tile):
    """Every compiled sweep tile of the batched launches (64x8, 128x2, 128x4,
    64x2 k-lanes by j-rows; 64x4 is the default the other tests use) gives
    the single cases' FP64 results bitwise."""
    monkeypatch.setenv("XFD_BATCH_TILE", str(tile))
    specs = _cases(n=4, seed=11, sizes=(24, 40, 72))
    say(f"[gpu] batch tile {tile} vs single cases f64")
    sims, b, got, info = _batched(specs, "f64", graph=True)
    assert info["shared"]
    assert info["tile"] == [(64, 4), (64, 8), (128, 2), (128, 4), (64, 2)][tile]
    _compare(got, _alone(specs, "f64"), True, f"tile {tile}")


def test_batch_tuner_picks_a_compiled_tile():
    specs = _cases(n=4, seed=12, sizes=(32, 48))
    sims = [native.Simulation(sp.spec(), "f32") for sp in specs]
    b = native.SimulationBatch(sims)
    b.init(20)
    b.tune(2)
    info = b.info()
    say(f"[gpu] tuned: tile {info['tile']}, ichunk {info['ichunk']}, grid {info['grid']}")
    assert info["tile"] in [(64, 4), (64, 8), (128, 2), (128, 4), (64, 2)]
    assert info["ichunk"] in (4, 8, 16, 32)

