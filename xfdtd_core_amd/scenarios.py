"""The workloads the hot path is measured and checked on, built with the
reference-shaped scenario API.

``dielectric_sphere`` is the reference README example (README.md:73-122):
63^3 cells including an 8-cell CPML, TFSF plane wave, Ex XZ movie.
``inhomogeneous_box`` is the 384^3 box with 64 random dielectric spheres and a
centred Ez dipole (a one-edge CurrentSource), in full size or scaled down for
parity runs.  Random draws use numpy's PCG64 seeded per workload; the spec text
carries the drawn numbers, so every consumer sees identical inputs.
"""

from __future__ import annotations

import math

import numpy as np

from . import spec as S

C0 = 1.0 / math.sqrt(8.854187817e-12 * 4 * math.pi * 1e-7)


def dielectric_sphere(movie_dir: str = "", threads=(1, 1, 1), movie_interval: int = 10):
    """README.md:73-122: dl=7.5 mm, 0.35 m air cube, sphere r=0.1 m (eps_r=3,
    mu_r=2), TFSF at index 15, 6 x PML(8), CFL 0.9, Ex XZ movie every 10 steps."""
    dl = 7.5e-3
    s = S.Simulation(dl, dl, dl, 0.9, threads)
    s.addObject(S.Object("domain", S.Cube(S.Vector(-0.175, -0.175, -0.175), S.Vector(0.35, 0.35, 0.35)),
                         S.Material.createAir()))
    s.addObject(S.Object("dielectric_sphere", S.Sphere(S.Vector(0, 0, 0), 0.1),
                         S.Material("a", S.ElectroMagneticProperty(3, 2, 0, 0))))
    l_min = dl * 20
    tau = l_min / 6e8
    t_0 = 4.5 * tau
    s.addWaveformSource(S.TFSF3D(15, 15, 15, 0.0, 0.0, 0.0, S.Waveform.gaussian(tau, t_0)))
    for d in ("XN", "XP", "YN", "YP", "ZN", "ZP"):
        s.addBoundary(S.PML(8, d))
    s.addMonitor(S.MovieMonitor(
        S.FieldMonitor(S.Cube(S.Vector(-0.175, 0, -0.175), S.Vector(0.35, dl, 0.35)), "EX", "", ""),
        movie_interval, "movie", movie_dir))
    return s


def inhomogeneous_box(n_total: int = 384, pml: int = 12, n_spheres: int = 64,
                      r_cells=(10, 40), dl: float = 1e-3, seed: int = 2028,
                      snapshot_interval: int = 500, probes=(4, 4, 2),
                      threads=(1, 1, 1), monitors: bool = True):
    """Inhomogeneous box: n_total^3 cells including a `pml`-cell CPML, random
    non-overlapping dielectric spheres (eps_r ~ U[1,6], sigma ~ U[0,0.05] S/m,
    at least one cell clear of the CPML), a centred one-edge Ez current source
    driven by a derivative-of-Gaussian, CFL 0.95.  Monitors: all six fields
    every `snapshot_interval` steps and a lattice of Ez probes."""
    rng = np.random.default_rng(seed)
    n_in = n_total - 2 * pml
    half = n_in * dl / 2
    s = S.Simulation(dl, dl, dl, 0.95, threads)
    s.addObject(S.Object("domain", S.Cube(S.Vector(-half, -half, -half), S.Vector(2 * half, 2 * half, 2 * half)),
                         S.Material.createAir()))
    placed = []
    tries = 0
    while len(placed) < n_spheres and tries < 200000:
        tries += 1
        r = int(rng.integers(r_cells[0], r_cells[1] + 1))
        lo, hi = -half + (r + 1) * dl, half - (r + 1) * dl
        if hi <= lo:
            continue
        c = rng.uniform(lo, hi, size=3)
        if any(np.linalg.norm(c - p) < (r + q) * dl for p, q in placed):
            continue
        placed.append((c, r))
    for idx, (c, r) in enumerate(placed):
        eps_r = float(rng.uniform(1.0, 6.0))
        sigma = float(rng.uniform(0.0, 0.05))
        s.addObject(S.Object(f"sphere{idx}", S.Sphere(S.Vector(*map(float, c)), r * dl),
                             S.Material(f"m{idx}", S.ElectroMagneticProperty(eps_r, 1.0, sigma, 0.0))))
    tau = 20 * dl / (2 * C0)
    t0 = 4.5 * tau
    s.addObject(S.CurrentSource("dipole", S.Cube(S.Vector(0.0, 0.0, 0.0), S.Vector(0.0, 0.0, dl)), "ZP",
                                1e20, S.Waveform.gaussianDerivative(tau, t0)))
    for d in ("XN", "XP", "YN", "YP", "ZN", "ZP"):
        s.addBoundary(S.PML(pml, d))
    if monitors:
        whole = S.Cube(S.Vector(-half - pml * dl, -half - pml * dl, -half - pml * dl),
                       S.Vector(n_total * dl, n_total * dl, n_total * dl))
        for f in S.FIELDS:
            s.addMonitor(S.MovieMonitor(S.FieldMonitor(whole, f), snapshot_interval, f"snap_{f}", ""))
        px, py, pz = probes
        for a in range(px):
            for b in range(py):
                for c in range(pz):
                    x = -half + (a + 1) * 2 * half / (px + 1)
                    y = -half + (b + 1) * 2 * half / (py + 1)
                    z = -half + (c + 1) * 2 * half / (pz + 1)
                    s.addMonitor(S.FieldTimeMonitor(f"probe_{a}{b}{c}", S.Cube(S.Vector(x, y, z), S.Vector(dl, dl, dl)),
                                                    "", "EZ"))
    return s


def dispersive_slabs(n_total: int = 256, pml: int = 10, dl: float = 0.5e-3, slab: int = 40, gap: int = 8,
                     tfsf_offset: int = 14, lumped: bool = True, kinds=("drude", "debye", "lorentz", "mlorentz"),
                     snapshot_interval: int = 0, nf2ff_distance: int = 0, nf2ff_freqs=None,
                     drude_wp: float = 1.37e16):
    """Dispersive + lumped box: four z-slabs (`slab` cells thick, `gap`-cell
    air gaps, centred, full transverse extent inside the CPML) of Drude
    (eps_inf=1, wp=1.37e16, gamma=1e14), Debye (eps_inf=4, eps_s=50,
    tau=1e-11 s), two-pole Lorentz (eps_inf=2, d_eps={1.5, 0.8},
    w0={2e10, 6e10}, delta={1e9, 3e9}) and a modified Lorentz with a1 != 0
    (eps_inf=2, a0=1.35e21, a1=5e9, b0=9e20, b1=1e10, b2=1; a0 b1 - a1 b0 > 0
    and 4 b2 - b0 dt^2 > 0, the stability conditions of
    dispersive_material_equation.cpp:62-99).  TFSF plane wave (theta=30 deg,
    phi=45 deg) with a Gaussian of tau = 20 dl / 6e8; a 50-ohm Resistor and a
    50-ohm, 1 V VoltageSource along z, each a 6-cell gap in the air below the
    slabs.  CFL 0.95.  `kinds` selects and orders the slab materials.  With
    `nf2ff_distance`, a frequency-domain NF2FF on the box that many cells
    inside the grid (default frequencies: 8 from 5 to 40 GHz).  With the
    default wp = 1.37e16 rad/s (wp dt ~ 1e4 at dl = 0.5 mm) the reference's
    first-order Drude scheme diverges (the oracle shows the same growth);
    `drude_wp` sets another plasma frequency."""
    s = S.Simulation(dl, dl, dl, 0.95)
    n = n_total - 2 * pml
    half = n * dl / 2
    s.addObject(S.Object("domain", S.Cube(S.Vector(-half, -half, -half), S.Vector(2 * half, 2 * half, 2 * half)),
                         S.Material.createAir()))
    mats = {
        "drude": S.DrudeMedium.makeDrudeMedium("drude", 1.0, [drude_wp], [1e14]),
        "debye": S.DebyeMedium.makeDebyeMedium("debye", 4.0, [50.0], [1e-11]),
        "lorentz": S.LorentzMedium.makeMLorentz("lorentz", 2.0, [1.5 * 2e10 ** 2, 0.8 * 6e10 ** 2], [0.0, 0.0],
                                                [2e10 ** 2, 6e10 ** 2], [2e9, 6e9], [1.0, 1.0]),
        "mlorentz": S.MLorentzMaterial.makeMLorentz("mlorentz", 2.0, [1.35e21], [5e9], [9e20], [1e10], [1.0]),
    }
    total = len(kinds) * slab + (len(kinds) - 1) * gap
    z0 = (n - total) // 2
    for q, kind in enumerate(kinds):
        lo = z0 + q * (slab + gap)
        s.addObject(S.Object(f"slab_{kind}", S.Cube(S.Vector(-half, -half, -half + lo * dl),
                                                    S.Vector(2 * half, 2 * half, slab * dl)), mats[kind]))
    if lumped:
        zl = -half + max(1, z0 // 2 - 3) * dl
        s.addObject(S.Resistor("r50", S.Cube(S.Vector(-4 * dl, 0.0, zl), S.Vector(0.0, 0.0, 6 * dl)), "Z", 50.0))
        tau_v = 20 * dl / 6e8
        s.addObject(S.VoltageSource("v1", S.Cube(S.Vector(4 * dl, 0.0, zl), S.Vector(0.0, 0.0, 6 * dl)), "ZP", 50.0,
                                    S.Waveform.gaussian(tau_v, 4.5 * tau_v, 1.0)))
    tau = 20 * dl / 6e8
    s.addWaveformSource(S.TFSF3D(tfsf_offset, tfsf_offset, tfsf_offset, math.radians(30), math.radians(45), 0.0,
                                 S.Waveform.gaussian(tau, 4.5 * tau)))
    for d in ("XN", "XP", "YN", "YP", "ZN", "ZP"):
        s.addBoundary(S.PML(pml, d))
    if snapshot_interval:
        whole = S.Cube(S.Vector(-half, -half, -half), S.Vector(2 * half, 2 * half, 2 * half))
        for f in S.FIELDS:
            s.addMonitor(S.MovieMonitor(S.FieldMonitor(whole, f), snapshot_interval, f"snap_{f}", ""))
    if nf2ff_distance:
        freqs = list(nf2ff_freqs) if nf2ff_freqs is not None else [float(f) for f in np.linspace(5e9, 40e9, 8)]
        s.addNF2FF(S.NFFFTFrequencyDomain(nf2ff_distance, nf2ff_distance, nf2ff_distance, freqs))
    return s


def radiating_dipole(n_total: int = 256, pml: int = 10, dl: float = 1e-3, td_nf2ff=(15, np.pi / 3, np.pi / 4),
                     out_dir: str = ""):
    """A one-edge Ez current source (derivative-of-Gaussian, tau = 20 dl /
    (2 c)) at the centre of an n_total^3 vacuum box including a `pml`-cell
    CPML, CFL 0.9; with `td_nf2ff` = (distance, theta, phi) a time-domain
    NF2FF `distance` cells inside the grid along that direction."""
    n_in = n_total - 2 * pml
    half = n_in * dl / 2
    s = S.Simulation(dl, dl, dl, 0.9)
    s.addObject(S.Object("domain", S.Cube(S.Vector(-half, -half, -half), S.Vector(2 * half, 2 * half, 2 * half)),
                         S.Material.createAir()))
    c0 = 299792458.0
    tau = 20 * dl / (2 * c0)
    s.addObject(S.CurrentSource("src", S.Cube(S.Vector(0.0, 0.0, 0.0), S.Vector(0.0, 0.0, dl)), "ZP", 1e20,
                                S.Waveform.gaussianDerivative(tau, 4.5 * tau, 1.0)))
    for d in ("XN", "XP", "YN", "YP", "ZN", "ZP"):
        s.addBoundary(S.PML(pml, d))
    if td_nf2ff is not None:
        dist, theta, phi = td_nf2ff
        s.addNF2FF(S.NFFFTTimeDomain(dist, dist, dist, out_dir, theta, phi))
    return s


def design_sweep(n_cases: int = 96, sizes=(48, 64, 96, 128, 160), pml: int = 8, dl: float = 1e-3,
                 side=(0.2, 0.5), eps_r=(1.0, 8.0), freqs=(2.5e9, 5e9, 7.5e9, 1e10), seed: int = 2030,
                 cfl: float = 0.9):
    """Batched design sweep: `n_cases` independent boxes.  Per case and axis
    the total cell count (CPML included) is drawn from `sizes`; one centred
    dielectric cuboid spans a fraction U[side] of the interior along each axis
    (eps_r ~ U[eps_r]); a one-edge Ez current source at the centre carries a
    Gaussian (tau = 20 dl / 6e8, t0 = 4.5 tau); `pml`-cell CPML on all faces;
    a plain (one accumulation per step) running DFT at `freqs` on the
    cuboid's bounding faces.  Returns the list of scenarios."""
    rng = np.random.default_rng(seed)
    tau = 20 * dl / 6e8
    cases = []
    for c in range(n_cases):
        n = [int(rng.choice(sizes)) for _ in range(3)]
        n_in = [a - 2 * pml for a in n]
        frac = float(rng.uniform(*side))
        eps = float(rng.uniform(*eps_r))
        # cuboid sides keep the parity of the interior, so its faces sit on
        # grid planes `pml + d` cells from each end
        cube = [max(2, 2 * int(round(frac * a / 2))) for a in n_in]
        s = S.Simulation(dl, dl, dl, cfl)
        half = [a * dl / 2 for a in n_in]
        s.addObject(S.Object("domain", S.Cube(S.Vector(*[-h for h in half]), S.Vector(*[2 * h for h in half])),
                             S.Material.createAir()))
        s.addObject(S.Object("cube", S.Cube(S.Vector(*[-q * dl / 2 for q in cube]), S.Vector(*[q * dl for q in cube])),
                             S.Material(f"m{c}", S.ElectroMagneticProperty(eps, 1.0, 0.0, 0.0))))
        s.addObject(S.CurrentSource("dipole", S.Cube(S.Vector(0.0, 0.0, 0.0), S.Vector(0.0, 0.0, dl)), "ZP",
                                    1e20, S.Waveform.gaussian(tau, 4.5 * tau)))
        for d in ("XN", "XP", "YN", "YP", "ZN", "ZP"):
            s.addBoundary(S.PML(pml, d))
        dist = [pml + (a - q) // 2 for a, q in zip(n_in, cube)]
        s.addNF2FF(S.NFFFTFrequencyDomain(dist[0], dist[1], dist[2], [float(f) for f in freqs],
                                          reference_repeat=False))
        cases.append(s)
    return cases


def adjoint_design(n_total: int = 160, pml: int = 10, dl: float = 1e-3, design: int = 48, f0: float = 1e10,
                   eps_r=(1.0, 4.0), patch=(4, 4, 1), patch_gap: int = 8, tfsf_margin: int = 3, seed: int = 2031,
                   cfl: float = 0.9):
    """Adjoint inverse-design workload: an `n_total`^3 grid (CPML included) at
    `dl`, an x-directed TFSF plane wave (theta = 90 deg, phi = 0) carrying a
    cosine-modulated Gaussian at f0 (tau = 1 / (2 f0 0.5), t0 = 4.5 tau),
    a central `design`^3 block of cells with eps_r ~ U[eps_r] (the
    differentiable parameter) and the objective sum |DFT_f0(Ez)|^2 over a
    `patch` of Ez points `patch_gap` cells beyond the block along +x.
    Returns (scenario, design origin, design eps_r array)."""
    rng = np.random.default_rng(seed)
    n_in = n_total - 2 * pml
    half = n_in * dl / 2
    s = S.Simulation(dl, dl, dl, cfl)
    s.addObject(S.Object("domain", S.Cube(S.Vector(-half, -half, -half), S.Vector(2 * half, 2 * half, 2 * half)),
                         S.Material.createAir()))
    tau = 1.0 / (2 * f0 * 0.5)
    s.addWaveformSource(S.TFSF3D(pml + tfsf_margin, pml + tfsf_margin, pml + tfsf_margin, math.pi / 2, 0.0, 0.0,
                                 S.Waveform.cosineModulatedGaussian(tau, 4.5 * tau, f0)))
    for d in ("XN", "XP", "YN", "YP", "ZN", "ZP"):
        s.addBoundary(S.PML(pml, d))
    o = (n_total - design) // 2
    origin = (o, o, o)
    eps = rng.uniform(eps_r[0], eps_r[1], size=(design, design, design))
    s.setDesignRegion(origin, (design,) * 3, eps)
    po = (o + design + patch_gap, n_total // 2 - patch[1] // 2, n_total // 2 - patch[2] // 2)
    s.setDftObjective("EZ", po, patch, f0)
    return s, origin, eps


def cylinder_scatter_2d(nx: int = 330, ny: int = 350, pml: int = 10, movie_interval: int = 20,
                        probes: int = 8, eps_r: float = 0.0, tfsf: int = 50, radius: float = 0.03):
    """2D (TE) workload after the reference's own example
    (test/src/examples/cylinder_scatter_2d.cpp): dx = dy = c/(20 f_max) with
    f_max = 20 GHz, an `nx` x `ny` air domain infinite along z, a cylinder of
    radius 0.03 m at the centre (PEC, or a dielectric of `eps_r` when given),
    TFSF2D at `tfsf` = 50 cells with phi = pi/4 and a cosine-modulated Gaussian
    (f_c = 12 GHz), 10-cell PML on x and y, CFL 0.8, an Ez movie of the whole
    plane every `movie_interval` steps and `probes` Ez time monitors on the
    diagonal."""
    inf = float("inf")
    fc, fmax = 12e9, 20e9
    dx = 3e8 / fmax / 20
    tau = 1.7 / (fmax - fc)
    t0 = 0.8 * tau
    s = S.Simulation(dx, dx, 1.0, 0.8)
    ox, oy = -(nx // 2) * dx, -(ny // 2) * dx
    s.addObject(S.Object("domain", S.Cube(S.Vector(ox, oy, -inf), S.Vector(nx * dx, ny * dx, inf)),
                         S.Material.createAir()))
    mat = (S.Material("pec", S.ElectroMagneticProperty(1, 1, 1e10, 0)) if eps_r <= 0
           else S.Material("dielectric", S.ElectroMagneticProperty(eps_r, 1, 0, 0)))
    s.addObject(S.Object("cylinder", S.Cylinder(S.Vector(0.0, 0.0, -inf), radius, inf, "Z"), mat))
    s.addWaveformSource(S.TFSF2D(tfsf, tfsf, math.pi * 0.25, S.Waveform.cosineModulatedGaussian(tau, t0, fc)))
    for d in ("XN", "XP", "YN", "YP"):
        s.addBoundary(S.PML(pml, d))
    if movie_interval > 0:
        s.addMonitor(S.MovieMonitor(S.FieldMonitor(S.Cube(S.Vector(ox, oy, -inf), S.Vector(nx * dx, ny * dx, inf)),
                                                   "EZ"), movie_interval, "movie", ""))
    for q in range(probes):
        f = (q + 1) / (probes + 1)
        x, y = ox + f * nx * dx, oy + f * ny * dx
        s.addMonitor(S.FieldTimeMonitor(f"probe{q}", S.Cube(S.Vector(x, y, -inf), S.Vector(dx, dx, inf)), "", "EZ"))
    return s


def slab_line_1d(nz: int = 300, pml: int = 10, dl: float = 0.5e-3, slab=(50, 260), eps_r: float = 4.0,
                 sigma: float = 0.0, movie_interval: int = 5):
    """1D (TEM) workload after the reference's own example
    (test/src/examples/slab_refection_transmission.cpp): a line of `nz`
    cells at dl = 0.5 mm along z (infinite along x and y), a slab over cells
    [50, 310) of eps_r = 4 (the example's modified-Lorentz blood model needs
    the 3D ADE updator; the reference refuses it on a 1D grid), TFSF1D at 25
    cells going forward, a Gaussian of tau = 40 dl / 6e8, 10-cell PML on z,
    CFL 0.9, Ex time monitors at cells 10 and 50 and an Ex movie."""
    inf = float("inf")
    s = S.Simulation(dl, dl, dl, 0.9)
    s.addObject(S.Object("domain", S.Cube(S.Vector(-inf, -inf, 0.0), S.Vector(inf, inf, nz * dl)),
                         S.Material.createAir()))
    s.addObject(S.Object("slab", S.Cube(S.Vector(-inf, -inf, slab[0] * dl), S.Vector(inf, inf, slab[1] * dl)),
                         S.Material("slab", S.ElectroMagneticProperty(eps_r, 1, sigma, 0))))
    tau = dl * 40 / 6e8
    s.addWaveformSource(S.TFSF1D(25, True, S.Waveform.gaussian(tau, 4.5 * tau)))
    s.addMonitor(S.FieldTimeMonitor("reflect_ex", S.Cube(S.Vector(-inf, -inf, 10 * dl), S.Vector(inf, inf, dl)),
                                    "", "EX"))
    s.addMonitor(S.FieldTimeMonitor("transmit_ex", S.Cube(S.Vector(-inf, -inf, 50 * dl), S.Vector(inf, inf, dl)),
                                    "", "EX"))
    if movie_interval > 0:
        s.addMonitor(S.MovieMonitor(S.FieldMonitor(S.Cube(S.Vector(-inf, -inf, 0.0), S.Vector(inf, inf, 200 * dl)),
                                                   "EX"), movie_interval, "movie", ""))
    for d in ("ZP", "ZN"):
        s.addBoundary(S.PML(pml, d))
    return s
