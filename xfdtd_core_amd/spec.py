"""Scenario description mirroring the XFDTD_CORE C++ API, serialised as spec text.

The classes carry the reference's names and argument meaning
(``include/xfdtd/simulation/simulation.h:37-155``, ``object/object.h``,
``shape/*.h``, ``material/material.h``, ``boundary/pml.h:19-21``,
``waveform_source/tfsf_3d.h``, ``waveform/waveform.h``,
``object/lumped_element/*.h``, ``monitor/*.h``).  ``Simulation.spec()``
renders the scenario as line-oriented text, one directive per line:

    sim DX DY DZ CFL
    threads TX TY TZ
    object NAME SHAPE... MATERIAL...
        SHAPE    = cube OX OY OZ SX SY SZ | sphere CX CY CZ R
                 | cylinder CX CY CZ R H AXIS
        MATERIAL = plain EPS_R MU_R SIGMA_E SIGMA_M
                 | lorentz EPS_INF P A0[P] A1[P] B0[P] B1[P] B2[P]
                 | drude EPS_INF P OMEGA_P[P] GAMMA[P]
                 | debye EPS_INF P EPS_STATIC[P] TAU[P]
    pml THICKNESS DIR [ORDER SIGMA_RATIO ALPHA_MIN ALPHA_MAX KAPPA_MAX]
    tfsf3d X Y Z THETA PHI PSI WAVEFORM...
    tfsf2d X Y PHI WAVEFORM...
    tfsf1d Z FORWARD(0|1) WAVEFORM...
    current_source NAME OX OY OZ SX SY SZ DIR R WAVEFORM...
    voltage_source NAME OX OY OZ SX SY SZ DIR R WAVEFORM...
    resistor NAME OX OY OZ SX SY SZ AXIS R
    capacitor NAME OX OY OZ SX SY SZ AXIS C
    inductor NAME OX OY OZ SX SY SZ AXIS L
    pec_plane NAME OX OY OZ SX SY SZ
    movie FIELD OX OY OZ SX SY SZ INTERVAL NAME OUT_DIR
    nf2ff DX DY DZ NF F1 .. FNF OUT_DIR [plain_dft]
    nf2ff_td DX DY DZ THETA PHI OUT_DIR
    field_time NAME FIELD OX OY OZ SX SY SZ OUT_DIR
    voltage_monitor NAME OX OY OZ SX SY SZ DIR OUT_DIR
    current_monitor NAME OX OY OZ SX SY SZ DIR OUT_DIR
    network NAME OUT_DIR NF F1 .. FNF NP (INDEX SOURCE Z_RE Z_IM CURRENT VOLTAGE)...
    design_region OI OJ OK NI NJ NK EPS_R...      (cells; adjoint form)
    dft_objective FIELD OI OJ OK NI NJ NK FREQ    (grid points; adjoint form)
    WAVEFORM = gaussian TAU T0 AMP | cosine_modulated_gaussian TAU T0 F AMP
             | sine F AMP | cosine F AMP | step T0 AMP
             | gaussian_derivative TAU T0 AMP

Numbers are written with ``repr`` so they round-trip exactly through
``strtod``.  ``OUT_DIR`` of ``-`` means "keep in memory, write no file".
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


def _n(x: float) -> str:
    return repr(float(x))


@dataclass
class Vector:
    x: float
    y: float
    z: float

    def tokens(self):
        return [_n(self.x), _n(self.y), _n(self.z)]


class Shape:
    def tokens(self):  # pragma: no cover - interface
        raise NotImplementedError


@dataclass
class Cube(Shape):
    """``xfdtd::Cube(origin, size)`` (shape/cube.h)."""

    origin: Vector
    size: Vector

    def tokens(self):
        return ["cube"] + self.origin.tokens() + self.size.tokens()

    def box_tokens(self):
        return self.origin.tokens() + self.size.tokens()


@dataclass
class Sphere(Shape):
    """``xfdtd::Sphere(center, radius)`` (shape/sphere.h)."""

    center: Vector
    radius: float

    def tokens(self):
        return ["sphere"] + self.center.tokens() + [_n(self.radius)]


@dataclass
class Cylinder(Shape):
    """``xfdtd::Cylinder(center, radius, height, axis)`` (shape/cylinder.h)."""

    center: Vector
    radius: float
    height: float
    axis: str = "Z"

    def tokens(self):
        return ["cylinder"] + self.center.tokens() + [_n(self.radius), _n(self.height), self.axis]


@dataclass
class ElectroMagneticProperty:
    """``ElectroMagneticProperty(epsilon_r, mu_r, sigma_e, sigma_m)`` (material.cpp:11-26)."""

    epsilon_r: float = 1.0
    mu_r: float = 1.0
    sigma_e: float = 0.0
    sigma_m: float = 0.0


@dataclass
class Material:
    name: str
    em_property: ElectroMagneticProperty = field(default_factory=ElectroMagneticProperty)

    @staticmethod
    def createAir(name: str = "air") -> "Material":
        return Material(name, ElectroMagneticProperty(1, 1, 0, 0))

    def tokens(self):
        p = self.em_property
        return ["plain", _n(p.epsilon_r), _n(p.mu_r), _n(p.sigma_e), _n(p.sigma_m)]


@dataclass
class LinearDispersiveMaterial(Material):
    """Modified-Lorentz form of a linear dispersive material
    (dispersive_material.h): eps_r(w) = eps_inf + sum_p (a1 jw + a0) /
    (b2 (jw)^2 + b1 jw + b0).  `kind` is the reference class the material
    was made with ("lorentz", "drude" or "debye"); `params` are that
    factory's per-pole arrays."""

    kind: str = "lorentz"
    epsilon_inf: float = 1.0
    params: List[List[float]] = field(default_factory=list)

    def tokens(self):
        n = len(self.params[0])
        out = [self.kind, _n(self.epsilon_inf), str(n)]
        for arr in self.params:
            assert len(arr) == n, "every per-pole array has the pole count"
            out += [_n(v) for v in arr]
        return out

    def susceptibility(self, omega, p):
        """(a1 jw + a0) / (b2 (jw)^2 + b1 jw + b0) of pole p at angular
        frequency `omega` (dispersive_material_equation.cpp:35-41)."""
        a0, a1, b0, b1, b2 = self.lorentz_params()
        jw = 1j * omega
        return (a1[p] * jw + a0[p]) / (b2[p] * jw * jw + b1[p] * jw + b0[p])

    def relative_permittivity(self, omega):
        return self.epsilon_inf + sum(self.susceptibility(omega, p) for p in range(len(self.params[0])))

    def lorentz_params(self):
        """(a0, a1, b0, b1, b2) per pole, as the equation classes set them."""
        n = len(self.params[0])
        if self.kind == "lorentz":
            return tuple(list(a) for a in self.params)
        if self.kind == "drude":
            wp, g = self.params
            return ([w * w for w in wp], [0.0] * n, [0.0] * n, list(g), [1.0] * n)
        es, tau = self.params
        return ([e - self.epsilon_inf for e in es], [0.0] * n, [1.0] * n, list(tau), [0.0] * n)


class MLorentzMaterial:
    @staticmethod
    def makeMLorentz(name, epsilon_inf, a_0, a_1, b_0, b_1, b_2):
        """MLorentzMaterial::makeMLorentz (dipersive_material.cpp:47-56)."""
        return LinearDispersiveMaterial(name, kind="lorentz", epsilon_inf=epsilon_inf,
                                        params=[list(a_0), list(a_1), list(b_0), list(b_1), list(b_2)])


LorentzMedium = MLorentzMaterial


class DrudeMedium:
    @staticmethod
    def makeDrudeMedium(name, eps_inf, omega_p, gamma):
        """DrudeMedium::makeDrudeMedium (drude_medium.cpp:9-14)."""
        return LinearDispersiveMaterial(name, kind="drude", epsilon_inf=eps_inf,
                                        params=[list(omega_p), list(gamma)])


class DebyeMedium:
    @staticmethod
    def makeDebyeMedium(name, epsilon_inf, epsilon_static, tau):
        """DebyeMedium::makeDebyeMedium (debye_medium.cpp:10-17)."""
        return LinearDispersiveMaterial(name, kind="debye", epsilon_inf=epsilon_inf,
                                        params=[list(epsilon_static), list(tau)])


@dataclass
class Waveform:
    """Source time function (waveform.cpp:80-101); sampled by the host."""

    kind: str
    params: List[float]
    amplitude: float = 1.0

    @staticmethod
    def gaussian(tau, t_0, amplitude=1.0):
        return Waveform("gaussian", [tau, t_0], amplitude)

    @staticmethod
    def cosineModulatedGaussian(tau, t_0, frequency, amplitude=1.0):
        return Waveform("cosine_modulated_gaussian", [tau, t_0, frequency], amplitude)

    @staticmethod
    def sine(frequency, amplitude=1.0):
        return Waveform("sine", [frequency], amplitude)

    @staticmethod
    def cosine(frequency, amplitude=1.0):
        return Waveform("cosine", [frequency], amplitude)

    @staticmethod
    def step(t_0, amplitude=1.0):
        return Waveform("step", [t_0], amplitude)

    @staticmethod
    def gaussianDerivative(tau, t_0, amplitude=1.0):
        """-(t-t0)/tau * exp(-((t-t0)/tau)^2): a user functor (waveform.h:13)."""
        return Waveform("gaussian_derivative", [tau, t_0], amplitude)

    def tokens(self):
        return [self.kind] + [_n(p) for p in self.params] + [_n(self.amplitude)]


@dataclass
class Object:
    """``xfdtd::Object(name, shape, material)`` (object/object.h)."""

    name: str
    shape: Shape
    material: Material

    def line(self):
        return " ".join(["object", self.name] + self.shape.tokens() + self.material.tokens())


@dataclass
class PML:
    """``xfdtd::PML(thickness, direction, order=4, sigma_ratio=1, alpha_min=0,
    alpha_max=0.05, kappa_max=10)`` (boundary/pml.h:19-21)."""

    thickness: int
    direction: str
    order: int = 4
    sigma_ratio: float = 1.0
    alpha_min: float = 0.0
    alpha_max: float = 0.05
    kappa_max: float = 10.0

    def line(self):
        return " ".join(["pml", str(int(self.thickness)), self.direction, str(int(self.order)),
                         _n(self.sigma_ratio), _n(self.alpha_min), _n(self.alpha_max),
                         _n(self.kappa_max)])


@dataclass
class TFSF3D:
    """``xfdtd::TFSF3D(x, y, z, theta, phi, psi, waveform)`` (waveform_source/tfsf_3d.h)."""

    x: int
    y: int
    z: int
    theta: float
    phi: float
    psi: float
    waveform: Waveform

    def line(self):
        return " ".join(["tfsf3d", str(int(self.x)), str(int(self.y)), str(int(self.z)),
                         _n(self.theta), _n(self.phi), _n(self.psi)] + self.waveform.tokens())


@dataclass
class TFSF2D:
    """``xfdtd::TFSF2D(distance_x, distance_y, phi, waveform)`` (waveform_source/tfsf_2d.h): the
    plane wave of a 2D (TE) grid, TFSF with theta = pi/2, psi = 0 (tfsf_2d.cpp:10-13)."""

    x: int
    y: int
    phi: float
    waveform: Waveform

    def line(self):
        return " ".join(["tfsf2d", str(int(self.x)), str(int(self.y)), _n(self.phi)] + self.waveform.tokens())


@dataclass
class TFSF1D:
    """``xfdtd::TFSF1D(z, forward, waveform)`` (waveform_source/tfsf_1d.h): the plane wave of a
    1D (TEM) grid travelling towards +z (forward) or -z (tfsf_1d.cpp:14-23)."""

    z: int
    forward: bool
    waveform: Waveform

    def line(self):
        return " ".join(["tfsf1d", str(int(self.z)), "1" if self.forward else "0"] + self.waveform.tokens())


@dataclass
class CurrentSource:
    """``CurrentSource(name, cube, direction, resistance, waveform)`` (current_source.h:11-14)."""

    name: str
    cube: Cube
    direction: str
    resistance: float
    waveform: Waveform

    def line(self):
        return " ".join(["current_source", self.name] + self.cube.box_tokens() +
                        [self.direction, _n(self.resistance)] + self.waveform.tokens())


@dataclass
class VoltageSource:
    """``VoltageSource(name, cube, direction, resistance, waveform)`` (voltage_source.h)."""

    name: str
    cube: Cube
    direction: str
    resistance: float
    waveform: Waveform

    def line(self):
        return " ".join(["voltage_source", self.name] + self.cube.box_tokens() +
                        [self.direction, _n(self.resistance)] + self.waveform.tokens())


@dataclass
class Resistor:
    """``Resistor(name, cube, xyz, resistance)`` (resistor.h)."""

    name: str
    cube: Cube
    xyz: str
    resistance: float

    def line(self):
        return " ".join(["resistor", self.name] + self.cube.box_tokens() +
                        [self.xyz, _n(self.resistance)])


@dataclass
class Capacitor:
    """``Capacitor(name, cube, xyz, capacitance)`` (capacitor.h)."""

    name: str
    cube: Cube
    xyz: str
    capacitance: float

    def line(self):
        return " ".join(["capacitor", self.name] + self.cube.box_tokens() +
                        [self.xyz, _n(self.capacitance)])


@dataclass
class Inductor:
    """``Inductor(name, cube, xyz, inductance)`` (inductor.h)."""

    name: str
    cube: Cube
    xyz: str
    inductance: float

    def line(self):
        return " ".join(["inductor", self.name] + self.cube.box_tokens() +
                        [self.xyz, _n(self.inductance)])


@dataclass
class PecPlane:
    """``PecPlane(name, cube)``: a zero-thickness cube (pec_plane.h)."""

    name: str
    cube: Cube

    def line(self):
        return " ".join(["pec_plane", self.name] + self.cube.box_tokens())


@dataclass
class FieldMonitor:
    """``FieldMonitor(shape, field, name, output_dir)`` (monitor/field_monitor.h).
    Inside a MovieMonitor it is the frame; added on its own it records the box
    of the final fields (field_monitor.cpp:80-96)."""

    shape: Cube
    field: str
    name: str = "feild_monitor"
    output_dir: str = ""

    def line(self):
        return " ".join(["field_monitor", self.field] + self.shape.box_tokens() +
                        [self.name or "feild_monitor", self.output_dir or "-"])


@dataclass
class MovieMonitor:
    """``MovieMonitor(frame, frame_interval, name, output_dir)`` (monitor/movie_monitor.h)."""

    frame: FieldMonitor
    frame_interval: int
    name: str = "movie"
    output_dir: str = ""

    def line(self):
        return " ".join(["movie", self.frame.field] + self.frame.shape.box_tokens() +
                        [str(int(self.frame_interval)), self.name or "movie",
                         self.output_dir or "-"])


@dataclass
class FieldTimeMonitor:
    """``FieldTimeMonitor(name, cube, output_dir, field)`` (monitor/field_time_monitor.h)."""

    name: str
    cube: Cube
    output_dir: str
    field: str

    def line(self):
        return " ".join(["field_time", self.name, self.field] + self.cube.box_tokens() +
                        [self.output_dir or "-"])


@dataclass
class VoltageMonitor:
    """``VoltageMonitor(name, cube, direction, output_dir)``
    (monitor/voltage_monitor.h): every step, the line integral of E along
    `direction` through the cube, averaged over its cross-section."""

    name: str
    cube: Cube
    direction: str
    output_dir: str = ""

    def line(self):
        return " ".join(["voltage_monitor", self.name] + self.cube.box_tokens() +
                        [self.direction, self.output_dir or "-"])


@dataclass
class CurrentMonitor:
    """``CurrentMonitor(name, cube, direction, output_dir)``
    (monitor/current_monitor.h): every step, the loop integral of H around
    the cube's last cross-section along `direction`."""

    name: str
    cube: Cube
    direction: str
    output_dir: str = ""

    def line(self):
        return " ".join(["current_monitor", self.name] + self.cube.box_tokens() +
                        [self.direction, self.output_dir or "-"])


@dataclass
class NFFFTFrequencyDomain:
    """``NFFFTFrequencyDomain(distance_x, distance_y, distance_z, frequencies,
    output_dir)`` (nffft_frequency_domain.h): a running DFT of the tangential
    fields on the box `distance` cells inside the grid boundary."""

    distance_x: int
    distance_y: int
    distance_z: int
    frequencies: List[float]
    output_dir: str = ""
    reference_repeat: bool = True  # False: the plain DFT, one accumulation per step

    def line(self):
        return " ".join(["nf2ff", str(int(self.distance_x)), str(int(self.distance_y)), str(int(self.distance_z)),
                         str(len(self.frequencies))] + [_n(f) for f in self.frequencies] + [self.output_dir or "-"] +
                        ([] if self.reference_repeat else ["plain_dft"]))


@dataclass
class NFFFTTimeDomain:
    """``NFFFTTimeDomain(distance_x, distance_y, distance_z, output_dir,
    theta, phi)`` (nffft/nffft_time_domain.h): time-domain near-to-far-field
    potentials along one observation direction."""

    distance_x: int
    distance_y: int
    distance_z: int
    output_dir: str = ""
    theta: float = 0.0
    phi: float = 0.0

    def line(self):
        return " ".join(["nf2ff_td", str(int(self.distance_x)), str(int(self.distance_y)), str(int(self.distance_z)),
                         _n(self.theta), _n(self.phi), self.output_dir or "-"])


@dataclass
class Port:
    """``Port(index, is_source, impedance, current_monitor, voltage_monitor)``
    (network/port.h): a port of an S-parameter network."""

    index: int
    is_source: bool
    impedance: complex
    current_monitor: CurrentMonitor
    voltage_monitor: VoltageMonitor

    def tokens(self):
        z = complex(self.impedance)
        return [str(int(self.index)), "1" if self.is_source else "0", _n(z.real), _n(z.imag),
                self.current_monitor.name, self.voltage_monitor.name]


@dataclass
class Network:
    """``Network(ports, frequencies, output_dir)`` (network/network.h):
    ``output()`` gives S(j, i) = b_j / a_i for every source port i, keyed
    10 j + i (network.cpp:47-84).  The monitors must be added to the
    simulation before the network."""

    ports: List[Port]
    frequencies: List[float]
    output_dir: str = ""
    name: str = "network"

    def line(self):
        f = [float(v) for v in self.frequencies]
        return " ".join(["network", self.name, self.output_dir or "-", str(len(f))] + [_n(v) for v in f] +
                        [str(len(self.ports))] + [t for p in self.ports for t in p.tokens()])


class Simulation:
    """``xfdtd::Simulation(dx, dy, dz, cfl, ThreadConfig)`` (simulation.h:39-40)."""

    def __init__(self, dx, dy, dz, cfl, thread_config=(1, 1, 1)):
        self.dx, self.dy, self.dz, self.cfl = dx, dy, dz, cfl
        self.thread_config = tuple(thread_config)
        self._lines: List[str] = []
        self.monitors: List[object] = []
        self.nf2ffs: List[object] = []

    def addObject(self, obj):
        self._lines.append(obj.line())

    def addWaveformSource(self, src):
        self._lines.append(src.line())

    def addBoundary(self, b):
        self._lines.append(b.line())

    def addMonitor(self, m):
        self._lines.append(m.line())
        self.monitors.append(m)

    def addNF2FF(self, n):
        self._lines.append(n.line())
        self.nf2ffs.append(n)

    def addNetwork(self, n):
        self._lines.append(n.line())

    def setDesignRegion(self, origin, size, eps_r):
        """Per-cell relative permittivity of the design cells [origin,
        origin+size) (grid cell indices); `eps_r` has shape `size`."""
        e = np.asarray(eps_r, dtype=np.float64).reshape(tuple(int(n) for n in size))
        self._design = (tuple(int(o) for o in origin), tuple(int(n) for n in size), e)

    def setDftObjective(self, field, origin, size, frequency):
        """J = sum over the E-component grid points [origin, origin+size) of
        |sum_n dt exp(-j 2 pi f (n+1) dt) F^{n+1}|^2."""
        self._objective = (str(field), tuple(int(o) for o in origin), tuple(int(n) for n in size), float(frequency))

    def spec(self) -> str:
        extra = []
        design = getattr(self, "_design", None)
        if design is not None:
            o, n, e = design
            extra.append("design_region " + " ".join(str(v) for v in (*o, *n)) + " " +
                         " ".join(_n(v) for v in e.ravel()))
        obj = getattr(self, "_objective", None)
        if obj is not None:
            f, o, n, fr = obj
            extra.append(f"dft_objective {f} " + " ".join(str(v) for v in (*o, *n)) + f" {_n(fr)}")
        return self._spec_base() + "".join(line + "\n" for line in extra)

    def _spec_base(self) -> str:
        head = [f"sim {_n(self.dx)} {_n(self.dy)} {_n(self.dz)} {_n(self.cfl)}",
                "threads " + " ".join(str(int(t)) for t in self.thread_config)]
        return "\n".join(head + self._lines) + "\n"


FIELDS = ("EX", "EY", "EZ", "HX", "HY", "HZ")
