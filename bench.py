#!/usr/bin/env python3
"""Throughput of the FDTD time-stepping loop on one MI355X.

Workload: the 384^3 inhomogeneous box (12-cell CPML on all faces, 64 random
dielectric spheres, centred Ez current-source dipole, CFL 0.95, full-field
snapshots every 500 steps and 32 Ez probes), synthetic inputs drawn from a
fixed seed.  A "step" is one full time step of the hot path (Yee H, CPML H,
Yee E, source and CPML E corrections, recordings).  Inputs are resident in
HBM when the timed region starts; the host precompute and the upload are
outside it, as Simulation::init is outside the reference's update timer.

Prints ONE JSON line: `value` is FP32 Mcells/s (the headline form); the FP64
form, measured in the same run on the same scenario, is reported under
"fp64".  `roofline` is for the dominant kernel, from HIP-event launch
durations measured on the library's stream over the timed region;
`cpu_baseline` times the oracle (the FP64 CPU restatement of the reference
path) on this host.
"""

from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0
FIELDS = ("EX", "EY", "EZ", "HX", "HY", "HZ")
METRIC = "Mcells/s + HBM GB/s, FP32 vs FP64 on 1 MI355X at stated field error"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def alg_bytes_per_launch(kernel: str, n: int, pml: int, s: int, fused: int = 0,
                         cpml_launches: float = 6.0) -> float:
    """Algorithmic HBM bytes of one launch of `kernel` on an n^3 grid, the
    per-cell-step figures of the leapfrog step split over its launches:
    B_yee = (6 field reads + 6 field writes + 18 coefficient reads) * s per
    cell and step, so each of the two Yee sweeps owns 15 * s per cell (its
    own 3 reads, 3 writes and 9 coefficients; the 3 dual-field reads a
    separate sweep adds are actual, not algorithmic, traffic); CPML = 12 * s
    per slab cell and step (4 components x psi read + write + c_psi read), so
    one slab launch (2 components) owns 6 * s per slab cell.  Slabs applied
    inside the sweeps (`fused` 1: the four x/y slabs, 2: the two z slabs)
    move their share to the sweep; their separate launches only finish the
    far grid planes (not counted), and the CPML figure is the mean over the
    `cpml_launches` launches of a half-step (six, one per slab; fewer where
    the two slabs of one axis share a launch)."""
    if kernel == "step":  # the single-pass step: the whole step's B_yee and CPML
        return step_alg_bytes(n, pml, s)
    cells = n ** 3
    slab_cells = pml * (n + 1) * (n + 1)
    inside = {0: 0, 1: 4, 2: 2, 3: 6, 4: 6}[int(fused)]
    if kernel in ("update_h", "update_e"):
        return 15.0 * s * cells + 6.0 * s * inside * slab_cells
    if kernel in ("cpml_h", "cpml_e"):
        return 6.0 * s * slab_cells * (6 - inside) / cpml_launches
    return 0.0


TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_inhomogeneous_box_384.json"


def measured_traffic(precision, n, pml):
    """HBM bytes per launch of each kernel, from the calibrated rocprofv3
    FETCH_SIZE / WRITE_SIZE runs recorded for the 384^3 box (None for other
    sizes, which were not profiled)."""
    if n != 384 or pml != 12 or not TRAFFIC_FILE.exists():
        return None
    d = json.loads(TRAFFIC_FILE.read_text())
    return {k: v["hbm_bytes"] for k, v in d["kernels"].get(precision, {}).items()}


def step_alg_bytes(n: int, pml: int, s: int) -> float:
    """Algorithmic bytes of one whole step: B_yee = 30 s per cell (6 field
    reads, 6 writes, 18 coefficient reads) plus CPML 12 s per slab cell and
    slab (4 components x psi read/write + c_psi read)."""
    return 30.0 * s * n ** 3 + 12.0 * s * 6 * pml * n * n


BREAKDOWN_STEPS = 20


def run_device(sim, precision, warmup, steps, n, pml):
    from xfdtd_core_amd import native

    log(f"{precision}: init (precompute + upload)")
    sim.set_precision(precision)
    t0 = time.perf_counter()
    sim.init(warmup + BREAKDOWN_STEPS + steps)
    init_s = time.perf_counter() - t0
    log(f"{precision}: init {init_s:.1f} s, device bytes {sim.device_bytes() / 2**30:.2f} GiB; warmup {warmup}")
    sim.set_profiling(False)
    sim.advance(0, warmup)
    # Per-kernel breakdown: every kernel class timed with HIP events.  An event
    # pair costs about 10 us of idle GPU time per launch (~16 launches per
    # step), so this short run stays outside the timed region.
    sim.set_profiling(True)
    sim.advance(warmup, BREAKDOWN_STEPS)
    stats = sim.kernel_stats()
    s = 4 if precision == "f32" else 8
    fused = sim.cpml_fused()
    cells = n ** 3
    kernels = {}
    for name, (launches, ms) in stats.items():
        if launches == 0:
            continue
        avg_ms = ms / launches
        b = alg_bytes_per_launch(name, n, pml, s, fused, launches / BREAKDOWN_STEPS)
        kernels[name] = {"launches": launches, "avg_ms": round(avg_ms, 5), "total_ms": round(ms, 3),
                         "alg_GBps": round(b / (avg_ms * 1e-3) / 1e9, 1) if b else None}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    # Timed region: HIP events around the dominant kernel's launches only.
    # The library's step call returns after its stream drains; the device
    # barrier on both sides brackets the timed region.
    sim.set_profiling(True, classes=[dom])
    native.device_synchronize()
    t0 = time.perf_counter()
    sim.advance(warmup + BREAKDOWN_STEPS, steps)
    native.device_synchronize()
    wall = time.perf_counter() - t0
    dl, dms = sim.kernel_stats()[dom]
    dom_avg = dms / dl
    kernels[dom]["timed_region"] = {"launches": dl, "avg_ms": round(dom_avg, 5)}
    ach = round(alg_bytes_per_launch(dom, n, pml, s, fused, kernels[dom]["launches"] / BREAKDOWN_STEPS)
                / (dom_avg * 1e-3) / 1e9, 1)
    traffic = measured_traffic(precision, n, pml)
    for name, k in kernels.items():
        if traffic and name in traffic:
            k["hbm_bytes_pmc"] = traffic[name]
            avg = dom_avg if name == dom else k["avg_ms"]
            k["hbm_GBps_pmc"] = round(traffic[name] / (avg * 1e-3) / 1e9, 1)
    step_b = step_alg_bytes(n, pml, s)
    step_gbps = step_b * steps / wall / 1e9
    res = {
        "mcells_per_s": cells * steps / wall / 1e6,
        "ms_per_step": wall / steps * 1e3,
        "init_s": round(init_s, 2),
        "step_alg_GBps": round(step_gbps, 1),
        "step_roofline": {"alg_bytes_per_step": step_b, "achieved_GBps": round(step_gbps, 1),
                          "frac": round(step_gbps / PEAK_HBM_GBS, 4),
                          "mcells_at_peak": round(cells / (step_b / (PEAK_HBM_GBS * 1e9)) / 1e6, 1)},
        "kernels": kernels,
        "kernel_breakdown_steps": BREAKDOWN_STEPS,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(ach / PEAK_HBM_GBS, 4) if ach else None,
                     "traffic": traffic.get(dom) if traffic else None,
                     "traffic_source": (f"stored rocprofv3 FETCH_SIZE/WRITE_SIZE profile "
                                        f"({TRAFFIC_FILE.relative_to(ROOT)}), not measured in this run"
                                        if traffic else None),
                     "hbm_GBps_measured": kernels[dom].get("hbm_GBps_pmc")},
    }
    log(f"{precision}: {res['mcells_per_s']:.0f} Mcells/s, {res['ms_per_step']:.3f} ms/step, dominant {dom} "
        f"{ach} GB/s")
    if sim.coef_classed():
        # bytes the class-id layout must move per sweep: 3 own-field reads and
        # writes, 3 dual-field reads, 3 16-bit class ids per cell
        ib = (9.0 * s + 6.0) * cells
        for k in ("update_h", "update_e"):
            if k in kernels:
                kernels[k]["layout_bytes"] = ib
                kernels[k]["layout_GBps"] = round(ib / (kernels[k]["avg_ms"] * 1e-3) / 1e9, 1)
        if "step" in kernels:
            # the single pass: 6 field reads and 6 writes, 6 class ids per
            # cell, and psi read/write + c_psi read of 4 components per slab
            # cell (six slabs)
            sb = (12.0 * s + 12.0) * cells + 12.0 * s * 6 * pml * n * n
            avg = kernels["step"].get("timed_region", {}).get("avg_ms", kernels["step"]["avg_ms"])
            kernels["step"]["layout_bytes"] = sb
            kernels["step"]["layout_GBps"] = round(sb / (avg * 1e-3) / 1e9, 1)
    res["coefficients"] = {0: "dense", 1: "class ids", 2: "32-bit class ids"}[sim.coef_layout()]
    res["cpml"] = {0: "separate slab launches", 1: "x/y slabs inside the Yee sweeps",
                   2: "z slabs inside the Yee sweeps, x/y slabs separate",
                   3: "every slab inside the single-pass step",
                   4: "every slab inside the two Yee sweeps"}[int(fused)]
    sim.finish()
    return res


ACCURACY_STEPS = 1000


def accuracy_leg(sim, steps, interval=500):
    """FP32 against FP64 on the headline box, outside the timed region and
    independent of --steps: one untimed run per precision of `steps` + 1
    steps, so the scenario's snapshots land on steps 0, 500, ..., `steps`,
    plus the 32 probe series.  The FP64 device run is the error reference:
    it equals the oracle bitwise on every parity case, and at this full size
    the GPU test test_full_size_fp64_matches_oracle checks it against the
    oracle over 100 steps."""
    from xfdtd_core_amd import accuracy

    runs = {}
    for prec in ("f64", "f32"):
        t0 = time.perf_counter()
        sim.set_precision(prec)
        sim.init(steps + 1)
        sim.set_profiling(False)
        sim.advance(0, steps + 1)
        sim.finish()
        runs[prec] = {"snapshots": {f: sim.monitor_records(i) for i, f in enumerate(FIELDS)},
                      "probes": [sim.monitor_records(m)[0][1] for m in range(len(FIELDS), sim.monitor_count())]}
        log(f"accuracy leg: {prec} {steps + 1} steps in {time.perf_counter() - t0:.1f} s")
    f32, f64 = runs["f32"], runs["f64"]
    n = len(f64["snapshots"]["EX"])
    rows = accuracy.snapshot_errors(f32["snapshots"], f64["snapshots"], [k * interval for k in range(n)])
    probes = [accuracy.series_error(a, b) for a, b in zip(f32["probes"], f64["probes"])]
    seen = [p for p in probes if p is not None]
    out = {"reference": "FP64 run of the same scenario on the GPU (equal to the oracle; see the docstring)",
           "steps": steps,
           "snapshots": [{"step": r["step"], "E_rel_l2": r["E"]["own"], "H_rel_l2": r["H"]["own"],
                          "E_rel_peak": r["E"]["peak"], "H_rel_peak": r["H"]["peak"]} for r in rows if r["step"] > 0],
           "probes_rel_l2_max": max(seen) if seen else None,
           "probes_compared": len(seen), "probes": len(probes)}
    for r in out["snapshots"]:
        log(f"fp32 error step {r['step']}: E {r['E_rel_l2']:.2e} (peak-normalised {r['E_rel_peak']:.2e}), "
            f"H {r['H_rel_l2']:.2e} ({r['H_rel_peak']:.2e})")
    if seen:
        log(f"fp32 error probes: max relative L2 {out['probes_rel_l2_max']:.2e} over {len(seen)} probes")
    del runs, f32, f64
    gc.collect()
    return out


def physical_cores():
    """Physical cores of this host from lscpu (cores per socket x sockets),
    or None when lscpu is unavailable."""
    import subprocess

    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    kv = {}
    for line in txt.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    try:
        return int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (KeyError, ValueError):
        return None


def cpu_baseline(spec, n, steps):
    """The oracle's std::thread loop on this host's cores: one warm-up step
    (it takes the step-0 snapshots), then `steps` timed steps, for the two
    ThreadConfig layouts {1,1,T} and {T,1,1}; the faster one is reported."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle import RefSim  # CPU restatement of the reference path

    # Threads: this process's CPU share (its affinity set, capped by the
    # OMP_NUM_THREADS the host sets for one GPU), never more than the host's
    # physical cores.
    phys = physical_cores()
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        share = min(share, int(omp))
    cores = max(1, min(share, phys or share))
    log(f"cpu baseline: host has {phys} physical cores (lscpu), CPU share {share}; oracle init on {n}^3 "
        f"(untimed), then 1 + {steps} steps per layout on {cores} threads")
    t0 = time.perf_counter()
    ref = RefSim(spec)
    ref.init(steps + 1)
    log(f"cpu baseline: init {time.perf_counter() - t0:.1f} s")
    best = None
    for layout in ((1, 1, cores), (cores, 1, 1)):
        ref.set_threads(*layout)
        ref.loop()
        t = ref.step_end_seconds()
        sec = t[-1] - t[0]
        rate = n ** 3 * (len(t) - 1) / sec / 1e6
        log(f"cpu baseline: ThreadConfig{layout} {rate:.1f} Mcells/s ({sec:.1f} s for {len(t) - 1} steps)")
        if best is None or rate > best[0]:
            best = (rate, layout, sec)
    del ref
    gc.collect()
    rate, layout, sec = best
    return {"value": round(rate, 2), "unit": "Mcells/s", "cores": cores, "kind": "port",
            "host_physical_cores_lscpu": phys, "process_cpu_share": share,
            "sample": f"same {n}^3 scenario, FP64 oracle, {steps} steps timed after 1 warm-up step "
                      f"({sec:.1f} s), ThreadConfig{{{layout[0]},{layout[1]},{layout[2]}}} "
                      f"(faster of {{1,1,{cores}}} and {{{cores},1,1}})"}


def bench_design_sweep(args):
    """Batched design sweep: 96 independent FP32 cases of uneven size in one
    set of persistent launches per step, one step captured as a HIP graph and
    replayed, tile shape from the autotuner.  value = total cell updates / s
    over all cases."""
    from xfdtd_core_amd import native, scenarios

    cases = scenarios.design_sweep(n_cases=args.cases)
    log(f"design sweep: {len(cases)} cases, precompute + upload")
    t0 = time.perf_counter()
    sims = [native.Simulation(c.spec(), "f32") for c in cases]
    batch = native.SimulationBatch(sims)
    batch.init(args.warmup + args.steps)
    cells = sum(int(np.prod(s.grid())) for s in sims)
    log(f"design sweep: init {time.perf_counter() - t0:.1f} s, {cells / 1e6:.1f} M cells")
    batch.set_graph(True)
    t0 = time.perf_counter()
    batch.tune(3)
    info = batch.info()
    log(f"design sweep: tuned in {time.perf_counter() - t0:.1f} s: tile {info['tile'][0]}x{info['tile'][1]} "
        f"(k-lanes x j-rows), ichunk {info['ichunk']}, grid {info['grid']}")
    batch.advance(0, args.warmup)
    native.device_synchronize()
    t0 = time.perf_counter()
    batch.advance(args.warmup, args.steps)
    native.device_synchronize()
    wall = time.perf_counter() - t0
    info = batch.info()
    value = cells * args.steps / wall / 1e6
    log(f"design sweep: {value:.0f} Mcells/s ({wall / args.steps * 1e3:.3f} ms/step; device "
        f"{info['device_ms'] / args.steps:.3f} ms/step)")
    batch.finish()
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mcells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random sizes, cubes and permittivities; no dataset)",
        "config": {"workload": f"batched design sweep: {len(cases)} cases, per-axis size from "
                               "{48,64,96,128,160} incl. 8-cell CPML, one dielectric cube each (eps_r U[1,8]), "
                               "Ez Gaussian dipole, 4-frequency FP64 DFT on the cube faces, HIP-graph steps",
                   "cells": cells, "parallelism": "one persistent launch per phase over all cases, 1 GPU"},
        "batch": {"shared_launches": info["shared"], "tile": list(info["tile"]), "ichunk": info["ichunk"],
                  "grid": info["grid"],
                  "device_ms_per_step": round(info["device_ms"] / args.steps, 4)},
    }
    batch.close()
    del batch
    if args.sequential:
        # the same cases stepped one context at a time (each its own launches)
        sub = sims[: args.sequential]
        cells_sub = sum(int(np.prod(s_.grid())) for s_ in sub)
        n_seq = min(args.steps, 100)
        t_seq = 0.0
        for s_ in sub:
            s_.init(n_seq)
            native.device_synchronize()
            t0 = time.perf_counter()
            s_.advance(0, n_seq)
            native.device_synchronize()
            t_seq += time.perf_counter() - t0
        seq = cells_sub * n_seq / t_seq / 1e6
        log(f"design sweep: the first {len(sub)} cases one context at a time: {seq:.0f} Mcells/s")
        line["sequential_single_contexts"] = {"cases": len(sub), "steps": n_seq, "value": round(seq, 1),
                                              "unit": "Mcells/s", "batch_speedup": round(value / seq, 2)}
    for s_ in sims:
        s_.close()
    return line


def bench_dispersive(args):
    """Dispersive + lumped box at full size: 256^3 (10-cell CPML), ADE slabs,
    oblique TFSF, 50-ohm resistor and voltage source, 8-frequency NF2FF 15
    cells inside the grid.

    Throughput (FP64 and FP32, `--dispersive-steps` steps each) on two
    material sets: the four slabs as specified (Drude wp = 1.37e16 rad/s,
    Debye, two-pole Lorentz, modified Lorentz) and the same box without the
    Debye slab.  With the Debye slab every dispersive cell runs the
    modified-Lorentz scheme with b2 = 0 for Debye, outside that scheme's
    stability condition 4 b2 - b0 dt^2 >= 0 (MLorentzEqDecision::
    stabilityCondition, which the reference defines but never calls): the
    fields grow without bound in FP64 exactly as in the oracle.  The FP32
    error is stated on the box without Debye, with the specified stiff Drude
    slab (wp dt ~ 1.25e4): FP64 and FP32 runs advance side by side and are
    compared every `--dispersive-check` steps (field state, far-field
    potentials, radiated power), together with the FP64 run's largest |E|,
    which shows where the reference's own scheme stops being bounded."""
    from xfdtd_core_amd import native, scenarios

    steps = args.dispersive_steps
    res = {"workload": "dispersive + lumped 256^3 (10-cell CPML), Drude/Debye/Lorentz/M-Lorentz slabs, oblique TFSF, "
                       "50-ohm resistor + voltage source, 8-frequency NF2FF", "steps": steps}
    sets = {"as_specified": ("drude", "debye", "lorentz", "mlorentz"), "without_debye": ("drude", "lorentz", "mlorentz")}
    for label, kinds in sets.items():
        sp = scenarios.dispersive_slabs(nf2ff_distance=15, kinds=kinds).spec()
        res[label] = {"kinds": list(kinds), "drude_wp": 1.37e16, **_dispersive_throughput(sp, steps, label)}
    sp = scenarios.dispersive_slabs(nf2ff_distance=15, kinds=sets["without_debye"]).spec()
    res["fp32_error"] = _dispersive_error(sp, steps, args.dispersive_check)
    return res


def _dispersive_throughput(spec, steps, label):
    from xfdtd_core_amd import native

    out = {}
    for prec in ("f64", "f32"):
        sim = native.Simulation(spec, prec)
        t0 = time.perf_counter()
        sim.init(steps)
        init_s = time.perf_counter() - t0
        native.device_synchronize()
        t0 = time.perf_counter()
        sim.advance(0, steps)
        native.device_synchronize()
        wall = time.perf_counter() - t0
        sim.finish()
        cells = int(np.prod(sim.grid()))
        finite = all(bool(np.isfinite(sim.field(f)).all()) for f in FIELDS)
        out[prec] = {"mcells_per_s": round(cells * steps / wall / 1e6, 1), "ms_per_step": round(wall / steps * 1e3, 4),
                     "init_s": round(init_s, 2), "fields_finite": finite,
                     "max_abs_E": float(max(np.abs(sim.field(f)).max() for f in FIELDS[:3]))}
        log(f"dispersive {label} {prec}: {out[prec]['mcells_per_s']:.0f} Mcells/s ({out[prec]['ms_per_step']:.3f} "
            f"ms/step), max |E| at the end {out[prec]['max_abs_E']:.3e}")
        sim.close()
    out["fp32_over_fp64"] = round(out["f32"]["mcells_per_s"] / out["f64"]["mcells_per_s"], 3)
    return out


def _dispersive_error(spec, steps, check):
    """FP64 and FP32 contexts advanced side by side; every `check` steps the
    FP32 field state (E and eta0 H together), far-field potentials (each of
    the four against the largest at its frequency, 181 directions) and
    radiated power are compared with FP64."""
    from xfdtd_core_amd import native

    eta = 376.730313668
    sims = {p: native.Simulation(spec, p) for p in ("f64", "f32")}
    for sim in sims.values():
        sim.init(steps)
    theta = np.linspace(0.0, np.pi, 181)
    rows = []
    t = 0
    while t < steps:
        n = min(check, steps - t)
        for sim in sims.values():
            sim.advance(t, n)
            sim.finish()  # fields and NF2FF currents back to the host
        t += n
        a, b = sims["f32"], sims["f64"]
        num = den = 0.0
        emax = 0.0
        for f in FIELDS:
            w = eta if f[0] == "H" else 1.0
            x, y = a.field(f) * w, b.field(f) * w
            num += float(np.sum((x - y) ** 2))
            den += float(np.sum(y ** 2))
            if f[0] == "E":
                emax = max(emax, float(np.abs(y).max()))
        row = {"step": t, "fp64_max_abs_E": emax,
               "field_state_rel_l2": float(np.sqrt(num / den)) if den > 0 and np.isfinite(num) else None}
        if np.isfinite(emax) and row["field_state_rel_l2"] is not None:
            pot_err, pow_err = 0.0, 0.0
            for f in range(8):
                pa = [a.nf2ff_potential(0, f, ap, tc, theta, 0.0) for ap in (True, False) for tc in (True, False)]
                pb = [b.nf2ff_potential(0, f, ap, tc, theta, 0.0) for ap in (True, False) for tc in (True, False)]
                scale = max(float(np.abs(x).max()) for x in pb)
                if scale > 0:  # nothing has reached the NF2FF surfaces yet: no figure
                    pot_err = max([pot_err] + [float(np.abs(x - y).max()) / scale for x, y in zip(pa, pb)])
                pw = b.nf2ff_power(0, f)
                if pw > 0:
                    pow_err = max(pow_err, abs(a.nf2ff_power(0, f) / pw - 1))
            row["far_field_potential_rel_max"] = pot_err
            row["radiated_power_rel_max"] = pow_err
        rows.append(row)
        log(f"dispersive error step {t}: fields {row['field_state_rel_l2']}, far field "
            f"{row.get('far_field_potential_rel_max')}, power {row.get('radiated_power_rel_max')}, "
            f"FP64 max |E| {emax:.3e}")
    for sim in sims.values():
        sim.close()
    # bounded: the FP64 field's maximum stays within twice the largest value
    # it took during the first third of the run (the pulse and the sources)
    early = max((r["fp64_max_abs_E"] for r in rows if r["step"] <= max(steps // 3, check)), default=0.0)
    bounded = [r for r in rows if r["fp64_max_abs_E"] <= 2 * early and r["field_state_rel_l2"] is not None]
    last = bounded[-1] if bounded else None
    return {"reference": "FP64 run of the same scenario on the GPU (equal to the oracle)",
            "kinds": ["drude", "lorentz", "mlorentz"], "drude_wp": 1.37e16, "check_interval": check,
            "timeline": rows, "last_bounded_step": last["step"] if last else None,
            "at_last_bounded_step": last}


def bench_nf2ff_td(args):
    """Time-domain NF2FF on a 256^3 box (10-cell CPML, radiating dipole),
    surfaces 15 cells inside the grid, one oblique direction: the step with
    and without it (FP64 and FP32), the NF2FF kernels' share from the HIP
    event timers, and the FP32 far field against FP64."""
    from xfdtd_core_amd import native, scenarios

    steps = args.td_steps
    res = {"workload": "time-domain NF2FF: 256^3 vacuum box, 10-cell CPML, Ez dipole, surfaces 15 cells inside, "
                       "theta 60 deg phi 45 deg", "steps": steps}
    far = {}
    for prec in ("f64", "f32"):
        row = {}
        for label, td in (("with", (15, np.pi / 3, np.pi / 4)), ("without", None)):
            sim = native.Simulation(scenarios.radiating_dipole(td_nf2ff=td).spec(), prec)
            sim.init(2 * steps)
            sim.set_profiling(False)  # the step unprofiled, then the NF2FF kernels timed
            native.device_synchronize()
            t0 = time.perf_counter()
            sim.advance(0, steps)
            native.device_synchronize()
            wall = time.perf_counter() - t0
            sim.set_profiling(True, classes=["nf2ff"])
            sim.advance(steps, steps)
            sim.finish()
            stats = sim.kernel_stats()
            launches, ms = stats.get("nf2ff", (0, 0.0))
            row[label] = {"ms_per_step": round(wall / steps * 1e3, 4),
                          "nf2ff_ms_per_step": round(ms / steps, 5) if launches else 0.0}
            if td is not None:
                far[prec] = sim.nf2ff_td_far_field(0)
            sim.close()
        row["surface_cells"] = int(6 * (256 - 30) ** 2)
        res[prec] = row
        log(f"nf2ff_td {prec}: {row['with']['ms_per_step']:.3f} ms/step with, {row['without']['ms_per_step']:.3f} "
            f"without; NF2FF kernels {row['with']['nf2ff_ms_per_step'] * 1e3:.1f} us/step")
    eta = 376.730313668
    a, b = far["f32"], far["f64"]
    peak = float(np.abs(eta * b[0] + b[3]).max())
    err = max(float(np.abs(a[w] - b[w]).max()) * (eta if w < 2 else 1.0) for w in range(4)) / peak
    res["fp32_far_field_error_of_copolar_peak"] = err
    log(f"nf2ff_td fp32 far-field error {err:.2e} of the co-polarised peak")
    return res


def bench_low_dim(args):
    """The 2D (TE) and 1D (TEM) paths: the reference's cylinder-scatter
    example (TFSF2D at phi = pi/4, PEC cylinder, 10-cell x/y PML, Ez probes)
    on an n2d x n2d grid, and its slab reflection example (TFSF1D, z PML,
    Ex probes) at its own size (300 cells, launch-bound), FP64 and FP32.
    Algorithmic bytes of a TE cell-step: Ez, Hx, Hy read and written and the
    seven coefficients their updates use (chxh, chxez, chyh, chyez, ceze,
    cezhx, cezhy), 13 s.  The CPU baseline is the oracle on the 2D case."""
    from xfdtd_core_amd import native, scenarios

    n2d, steps = args.n2d, args.low_steps
    sp2 = scenarios.cylinder_scatter_2d(nx=n2d - 20, ny=n2d - 20, movie_interval=0, probes=8).spec()
    sp1 = scenarios.slab_line_1d(movie_interval=0).spec()
    res = {"workload": f"2D TE cylinder scatter {n2d}x{n2d} incl. 10-cell PML (reference example "
                       f"cylinder_scatter_2d at a larger size), 1D TEM slab line 320 cells", "steps": steps}
    cells2 = n2d * n2d
    for prec in ("f64", "f32"):
        s = 8 if prec == "f64" else 4
        row = {}
        for label, sp, cells in (("te2d", sp2, cells2), ("tem1d", sp1, None)):
            sim = native.Simulation(sp, prec)
            sim.init(args.warmup + 20 + steps)
            cells = cells or int(np.prod(sim.grid()))
            sim.set_profiling(False)
            sim.advance(0, args.warmup)
            sim.set_profiling(True)
            sim.advance(args.warmup, 20)
            stats = sim.kernel_stats()
            sim.set_profiling(False)
            native.device_synchronize()
            t0 = time.perf_counter()
            sim.advance(args.warmup + 20, steps)
            native.device_synchronize()
            wall = time.perf_counter() - t0
            sim.finish()
            kern = {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 5)} for k, v in stats.items() if v[0]}
            r = {"cells": cells, "grid": list(sim.grid()), "ms_per_step": round(wall / steps * 1e3, 4),
                 "mcells_per_s": round(cells * steps / wall / 1e6, 1), "kernels": kern}
            if label == "te2d":
                yee_ms = sum(kern[k]["avg_ms"] for k in ("update_h", "update_e") if k in kern)
                alg = 13.0 * s * cells
                r["roofline"] = {"bound": "hbm", "kernel": "update_h + update_e (k_te_h, k_te_e)",
                                 "achieved": round(alg / (yee_ms * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS,
                                 "unit": "GB/s", "frac": round(alg / (yee_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                                 "alg_bytes_per_step": alg}
            row[label] = r
            log(f"{label} {prec}: {r['mcells_per_s']:.0f} Mcells/s, {r['ms_per_step']:.4f} ms/step")
            sim.close()
        res[prec] = row
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "tests"))
        from oracle import RefSim  # CPU restatement of the reference path

        cs = min(steps, 40)
        ref = RefSim(sp2)
        ref.init(cs + 1)
        cores = max(1, min(len(os.sched_getaffinity(0)), physical_cores() or 16))
        ref.set_threads(cores, 1, 1)
        ref.loop()
        t = ref.step_end_seconds()
        sec = t[-1] - t[0]
        res["cpu_baseline"] = {"value": round(cells2 * (len(t) - 1) / sec / 1e6, 2), "unit": "Mcells/s",
                               "cores": cores, "kind": "port",
                               "sample": f"the same 2D scenario, FP64 oracle, {len(t) - 1} steps after 1 warm-up "
                                         f"({sec:.1f} s), ThreadConfig{{{cores},1,1}}"}
    return res


def bench_adjoint(args):
    """Adjoint inverse design at full size: FP32 forward with checkpoints
    every 100 steps and the reverse-time adjoint; optionally the FP64 adjoint
    and central differences (FP64 forward runs) at --fd-cells random design
    cells."""
    from xfdtd_core_amd import native, scenarios

    s, origin, eps = scenarios.adjoint_design()
    steps, K = args.adjoint_steps, 100
    out = {}
    grads = {}
    for prec in (["f32", "f64"] if args.fd_cells or args.adjoint_f64 else ["f32"]):
        sim = native.Simulation(s.spec(), prec)
        t0 = time.perf_counter()
        J = sim.forward(steps, K)
        tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        g = sim.backward(eps.shape)
        tb = time.perf_counter() - t0
        grads[prec] = g
        out[prec] = {"J": J, "forward_s": round(tf, 3), "backward_s": round(tb, 3),
                     "coefficients": {0: "dense", 1: "class ids", 2: "32-bit class ids"}[sim.coef_layout()]}
        log(f"adjoint {prec}: J {J:.6e}, forward {tf:.2f} s (incl. init), backward {tb:.2f} s")
        sim.close()
    res = {"steps": steps, "checkpoint_interval": K, "design_cells": int(eps.size), "runs": out}
    if "f64" in grads:
        d = grads["f32"] - grads["f64"]
        res["fp32_vs_fp64_gradient_rel_l2"] = float(np.linalg.norm(d) / np.linalg.norm(grads["f64"]))
        log(f"adjoint: FP32 vs FP64 gradient relative L2 {res['fp32_vs_fp64_gradient_rel_l2']:.2e}")
    if args.fd_cells:
        rng = np.random.default_rng(7)
        h = 1e-3
        errs = []
        sim = native.Simulation(s.spec(), "f64")
        for q in range(args.fd_cells):
            c = tuple(int(v) for v in rng.integers(0, eps.shape[0], size=3))
            vals = []
            for sgn in (1, -1):
                e = eps.copy()
                e[c] += sgn * h
                sim.set_design(origin, e)
                vals.append(sim.forward(steps, steps))
            fd = (vals[0] - vals[1]) / (2 * h)
            errs.append(abs(grads["f64"][c] - fd) / abs(fd))
            log(f"adjoint fd cell {c}: adjoint {grads['f64'][c]:.6e} fd {fd:.6e} rel {errs[-1]:.2e}")
        res["fd_cells"] = args.fd_cells
        res["fd_rel_err_max"] = max(errs)
        res["fd_rel_err_median"] = float(np.median(errs))
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=384, help="cells per axis including the CPML")
    ap.add_argument("--pml", type=int, default=12)
    ap.add_argument("--precision", choices=["both", "f32", "f64"], default="both")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["box", "design_sweep", "adjoint", "dispersive", "nf2ff_td", "low_dim"],
                    default="box",
                    help="box: the 384^3 headline; design_sweep: the batched form; adjoint: the gradient form; "
                         "dispersive: the 256^3 ADE + lumped + NF2FF box; nf2ff_td: the time-domain NF2FF cost; "
                         "low_dim: the 2D (TE) and 1D (TEM) paths")
    ap.add_argument("--n2d", type=int, default=4096, help="low_dim: 2D grid cells per axis including the PML")
    ap.add_argument("--low-steps", type=int, default=1000)
    ap.add_argument("--td-steps", type=int, default=1000)
    ap.add_argument("--dispersive-steps", type=int, default=3000)
    ap.add_argument("--dispersive-check", type=int, default=250)
    ap.add_argument("--cases", type=int, default=96)
    ap.add_argument("--sequential", type=int, default=24,
                    help="design_sweep: also time this many cases stepped one context at a time")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-f64", action="store_true")
    ap.add_argument("--fd-cells", type=int, default=0)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs a single domain on one MI355X "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    from xfdtd_core_amd import native, scenarios

    if args.workload == "design_sweep":
        print(json.dumps(bench_design_sweep(args)), flush=True)
        return
    if args.workload == "dispersive":
        print(json.dumps(bench_dispersive(args)), flush=True)
        return
    if args.workload == "nf2ff_td":
        print(json.dumps(bench_nf2ff_td(args)), flush=True)
        return
    if args.workload == "low_dim":
        print(json.dumps(bench_low_dim(args)), flush=True)
        return
    if args.workload == "adjoint":
        print(json.dumps({"workload": "adjoint inverse design 160^3", **bench_adjoint(args)}), flush=True)
        return

    spec = scenarios.inhomogeneous_box(n_total=args.n, pml=args.pml, snapshot_interval=500).spec()
    dev_name, dev_bytes = native.device_info()
    log(f"device: {dev_name}, {dev_bytes / 2**30:.0f} GiB")
    sim = native.Simulation(spec)
    out = {}
    order = ["f64", "f32"] if args.precision == "both" else [args.precision]
    for p in order:
        out[p] = run_device(sim, p, args.warmup, args.steps, args.n, args.pml)

    head = out.get("f32") or out["f64"]
    line = {
        "metric": METRIC,
        "value": round(head["mcells_per_s"], 1),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if "f32" in out else "f64",
        "data": "synthetic (seeded random spheres; no dataset)",
        "config": {"workload": f"inhomogeneous box {args.n}^3 incl. {args.pml}-cell CPML, 64 dielectric spheres, "
                               f"Ez current-source dipole, CFL 0.95, snapshots every 500 steps + 32 Ez probes",
                   "cells": args.n ** 3, "parallelism": "single domain, 1 GPU", "device": dev_name},
        "hbm_GBps_alg_step": head["step_alg_GBps"],
        "step_roofline": head["step_roofline"],
        "roofline": head["roofline"],
        "coefficients": head["coefficients"],
        "kernels": head["kernels"],
    }
    if "f64" in out and "f32" in out:
        line["fp32_error"] = accuracy_leg(sim, ACCURACY_STEPS)
        f64 = out["f64"]
        line["fp64"] = {"value": round(f64["mcells_per_s"], 1), "unit": "Mcells/s",
                        "ms_per_step": round(f64["ms_per_step"], 4), "hbm_GBps_alg_step": f64["step_alg_GBps"],
                        "step_roofline": f64["step_roofline"],
                        "roofline": f64["roofline"], "kernels": f64["kernels"]}
        line["fp32_over_fp64"] = round(out["f32"]["mcells_per_s"] / f64["mcells_per_s"], 3)
    sim.close()
    del sim
    gc.collect()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(spec, args.n, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
