// xfdtd_core_amd.hpp — C++20 scenario API with the reference's class names.
//
// Code written against XFDTD_CORE's public headers (include/xfdtd/**) for the
// 3D time-stepping path builds the same scenario with these classes:
// Simulation, Object, Cube/Sphere/Cylinder, Material/ElectroMagneticProperty,
// PML, TFSF3D, Waveform, CurrentSource/VoltageSource/Resistor and the
// FieldMonitor/MovieMonitor/FieldTimeMonitor monitors keep their constructor
// arguments and meaning.  Simulation::run(n) performs the reference's FP64
// precompute on the host and runs the time loop on the GPU through the C-ABI
// in xfdtd_core_amd.h.  Arrays are row-major with k fastest, as in the
// reference; xtensor is not used, so array accessors return Array3D/Array1D
// views of this library.
#ifndef XFDTD_CORE_AMD_HPP_
#define XFDTD_CORE_AMD_HPP_

#include <array>
#include <cmath>
#include <complex>
#include <cstddef>
#include <functional>
#include <limits>
#include <memory>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>

namespace xfdtd {

using Real = double;
using Index = std::size_t;

// include/xfdtd/common/constant.h:11-23: the reference's constants, with
// eps0 and mu0 exactly as it writes them (not CODATA 2018).
namespace constant {
inline constexpr Real PI{std::numbers::pi_v<Real>};
inline constexpr Real EPSILON_0{8.854187817e-12};
inline constexpr Real MU_0{4 * PI * 1e-7};
inline const Real C_0{1 / std::sqrt(EPSILON_0 * MU_0)};
inline const Real Z_0{std::sqrt(MU_0 / EPSILON_0)};
inline constexpr Real SIGMA_E_ZERO_APPROX{1e-20};
inline constexpr Real SIGMA_M_ZERO_APPROX{1e-20};
inline constexpr Real INF{std::numeric_limits<Real>::infinity()};
inline constexpr Real NEG_INF{-1 * std::numeric_limits<Real>::infinity()};
constexpr auto II = std::complex<Real>(0.0, 1.0);
}  // namespace constant

// include/xfdtd/exception/exception.h
class XFDTDException : public std::runtime_error {
 public:
  explicit XFDTDException(const std::string& m = "XFDTD Exception") : std::runtime_error(m) {}
};
class XFDTDSimulationException : public XFDTDException {
 public:
  using XFDTDException::XFDTDException;
};
// include/xfdtd/grid_space/grid_space.h
class XFDTDGridSpaceException : public XFDTDException {
 public:
  using XFDTDException::XFDTDException;
};

template <class T>
struct Array1D {
  std::vector<T> data;
  Array1D() = default;
  explicit Array1D(std::size_t n, T v = T{}) : data(n, v) {}
  T& operator()(std::size_t i) { return data[i]; }
  const T& operator()(std::size_t i) const { return data[i]; }
  std::size_t size() const { return data.size(); }
};

template <class T>
struct Array3D {
  std::array<std::size_t, 3> shape{0, 0, 0};
  std::vector<T> data;
  Array3D() = default;
  Array3D(std::size_t a, std::size_t b, std::size_t c, T v = T{}) : shape{a, b, c}, data(a * b * c, v) {}
  T& operator()(std::size_t i, std::size_t j, std::size_t k) { return data[(i * shape[1] + j) * shape[2] + k]; }
  const T& operator()(std::size_t i, std::size_t j, std::size_t k) const {
    return data[(i * shape[1] + j) * shape[2] + k];
  }
  std::size_t size() const { return data.size(); }
};

struct Axis {
  enum class XYZ { X = 0, Y = 1, Z = 2 };
  enum class Direction { XN = 0, XP, YN, YP, ZN, ZP };
};

class Vector {
 public:
  Vector(Real x = 0, Real y = 0, Real z = 0) : _d{x, y, z} {}
  Real x() const { return _d[0]; }
  Real y() const { return _d[1]; }
  Real z() const { return _d[2]; }
  Real operator[](int i) const { return _d[i]; }

 private:
  std::array<Real, 3> _d;
};

class Cube;
class Shape {
 public:
  virtual ~Shape() = default;
  virtual bool isInside(const Vector& v, Real eps) const = 0;
  virtual std::unique_ptr<Cube> wrappedCube() const = 0;
  virtual std::unique_ptr<Shape> clone() const = 0;
};

class Cube : public Shape {
 public:
  Cube(Vector origin, Vector size);
  Vector origin() const { return _origin; }
  Vector size() const { return _size; }
  Vector end() const { return _end; }
  bool isInside(const Vector& v, Real eps) const override;
  std::unique_ptr<Cube> wrappedCube() const override;
  std::unique_ptr<Shape> clone() const override;

 private:
  Vector _origin, _size, _end;
};

class Sphere : public Shape {
 public:
  Sphere(Vector center, Real radius);
  bool isInside(const Vector& v, Real eps) const override;
  std::unique_ptr<Cube> wrappedCube() const override;
  std::unique_ptr<Shape> clone() const override;

 private:
  Vector _center;
  Real _radius;
};

class Cylinder : public Shape {
 public:
  Cylinder(Vector center, Real radius, Real height, Axis::XYZ axis);
  bool isInside(const Vector& v, Real eps) const override;
  std::unique_ptr<Cube> wrappedCube() const override;
  std::unique_ptr<Shape> clone() const override;

 private:
  Vector _center;
  Real _radius, _height;
  Axis::XYZ _axis;
};

class ElectroMagneticProperty {
 public:
  ElectroMagneticProperty(Real epsilon_r = 1, Real mu_r = 1, Real sigma_e = 0, Real sigma_m = 0);
  static ElectroMagneticProperty air() { return {1, 1, 0, 0}; }
  static ElectroMagneticProperty pec() { return {1, 1, 1e10, 0}; }  // material.cpp:32-34
  Real epsilonR() const { return _eps_r; }
  Real muR() const { return _mu_r; }
  Real sigmaE() const { return _sigma_e; }
  Real sigmaM() const { return _sigma_m; }
  Real epsilon() const { return _eps; }
  Real mu() const { return _mu; }

 private:
  Real _eps_r, _mu_r, _sigma_e, _sigma_m, _eps, _mu;
};

class Material {
 public:
  Material(std::string name, ElectroMagneticProperty p = {}) : _name(std::move(name)), _p(p) {}
  virtual ~Material() = default;
  static std::unique_ptr<Material> createAir(std::string name = "air") {
    return std::make_unique<Material>(std::move(name), ElectroMagneticProperty::air());
  }
  static std::unique_ptr<Material> createPec(std::string name = "pec") {
    return std::make_unique<Material>(std::move(name), ElectroMagneticProperty::pec());
  }
  const ElectroMagneticProperty& emProperty() const { return _p; }
  virtual bool dispersion() const { return false; }

 private:
  std::string _name;
  ElectroMagneticProperty _p;
};

// Linear dispersive material in the modified-Lorentz form
// (dispersive_material.h, dispersive_material_equation.h):
//   eps_r(w) = eps_inf + sum_p (a1_p jw + a0_p) / (b2_p (jw)^2 + b1_p jw + b0_p).
// The subclasses are the reference's factories; which of them a scenario
// uses decides the ADE scheme (Drude-only, Debye-only, else modified Lorentz).
class LinearDispersiveMaterial : public Material {
 public:
  enum class Kind { MLorentz, Drude, Debye };
  LinearDispersiveMaterial(std::string name, Real epsilon_inf, Kind kind, std::vector<Real> a0,
                           std::vector<Real> a1, std::vector<Real> b0, std::vector<Real> b1, std::vector<Real> b2,
                           ElectroMagneticProperty emp = ElectroMagneticProperty::air());
  bool dispersion() const override { return true; }
  Kind kind() const { return _kind; }
  Real epsilonInf() const { return _epsilon_inf; }
  std::size_t numPoles() const { return _a0.size(); }
  const std::vector<Real>& a0() const { return _a0; }
  const std::vector<Real>& a1() const { return _a1; }
  const std::vector<Real>& b0() const { return _b0; }
  const std::vector<Real>& b1() const { return _b1; }
  const std::vector<Real>& b2() const { return _b2; }

 private:
  Kind _kind;
  Real _epsilon_inf;
  std::vector<Real> _a0, _a1, _b0, _b1, _b2;
};

// MLorentzMaterial::makeMLorentz (dipersive_material.cpp:47-56).
class MLorentzMaterial : public LinearDispersiveMaterial {
 public:
  using LinearDispersiveMaterial::LinearDispersiveMaterial;
  static std::unique_ptr<MLorentzMaterial> makeMLorentz(std::string name, Real epsilon_inf, std::vector<Real> a_0,
                                                        std::vector<Real> a_1, std::vector<Real> b_0,
                                                        std::vector<Real> b_1, std::vector<Real> b_2);
};
using LorentzMedium = MLorentzMaterial;

// DrudeMedium::makeDrudeMedium: a0 = wp^2, a1 = 0, b0 = 0, b1 = gamma, b2 = 1
// (drude_medium.cpp:9-14, dispersive_material_equation.cpp:55-59).
class DrudeMedium : public LinearDispersiveMaterial {
 public:
  using LinearDispersiveMaterial::LinearDispersiveMaterial;
  static std::unique_ptr<DrudeMedium> makeDrudeMedium(std::string name, Real eps_inf, std::vector<Real> omega_p,
                                                      std::vector<Real> gamma);
};

// DebyeMedium::makeDebyeMedium: a0 = eps_s - eps_inf, a1 = 0, b0 = 1,
// b1 = tau, b2 = 0 (debye_medium.cpp:10-17, dispersive_material_equation.cpp:48-53).
class DebyeMedium : public LinearDispersiveMaterial {
 public:
  using LinearDispersiveMaterial::LinearDispersiveMaterial;
  static std::unique_ptr<DebyeMedium> makeDebyeMedium(std::string name, Real epsilon_inf,
                                                      std::vector<Real> epsilon_static, std::vector<Real> tau);
};

class Waveform {
 public:
  explicit Waveform(std::function<Real(Real)> func, Real amplitude = 1.0)
      : _func(std::move(func)), _amplitude(amplitude) {}
  static std::unique_ptr<Waveform> gaussian(Real tau, Real t_0, Real amplitude = 1.0);
  static std::unique_ptr<Waveform> cosineModulatedGaussian(Real tau, Real t_0, Real frequency,
                                                           Real amplitude = 1.0);
  static std::unique_ptr<Waveform> sine(Real frequency, Real amplitude = 1.0);
  static std::unique_ptr<Waveform> cosine(Real frequency, Real amplitude = 1.0);
  static std::unique_ptr<Waveform> step(Real t_0, Real amplitude = 1.0);
  Real amplitude() const { return _amplitude; }
  Real operator()(Real t) const { return _func(t); }
  void init(const std::vector<Real>& time);  // waveform.cpp:28-34
  const std::vector<Real>& value() const { return _value; }

 private:
  std::function<Real(Real)> _func;
  Real _amplitude;
  std::vector<Real> _value;
};

class Simulation;

class Object {
 public:
  Object(std::string name, std::unique_ptr<Shape> shape, std::unique_ptr<Material> material);
  virtual ~Object() = default;
  const std::string& name() const { return _name; }
  const Shape& shape() const { return *_shape; }
  const Material& material() const { return *_material; }

 private:
  std::string _name;
  std::unique_ptr<Shape> _shape;
  std::unique_ptr<Material> _material;
};

class LumpedElement : public Object {
 public:
  LumpedElement(std::string name, std::unique_ptr<Cube> cube, Axis::XYZ xyz,
                std::unique_ptr<Material> material = Material::createAir());
  Axis::XYZ xyz() const { return _xyz; }

 private:
  Axis::XYZ _xyz;
};

class Resistor : public LumpedElement {
 public:
  Resistor(std::string name, std::unique_ptr<Cube> cube, Axis::XYZ xyz, Real resistance,
           std::unique_ptr<Material> material = Material::createAir());
  Real resistance() const { return _resistance; }

 private:
  Real _resistance;
};

// capacitor.h: coefficient-only element (capacitor.cpp:61-112).
class Capacitor : public LumpedElement {
 public:
  Capacitor(std::string name, std::unique_ptr<Cube> cube, Axis::XYZ xyz, Real capacitance,
            std::unique_ptr<Material> material = Material::createAir());
  Real capacitance() const { return _capacitance; }

 private:
  Real _capacitance;
};

// inductor.h: a current state per edge updated after every E update
// (inductor.cpp:76-135).  An inductance of 0 is replaced by 1e-10 H.
class Inductor : public LumpedElement {
 public:
  Inductor(std::string name, std::unique_ptr<Cube> cube, Axis::XYZ xyz, Real inductance,
           std::unique_ptr<Material> material = Material::createAir());
  Real inductance() const { return _inductance; }

 private:
  Real _inductance;
};

// pec_plane.h: a zero-thickness cube whose tangential E edges get the PEC
// sigma_e after material averaging (pec_plane.cpp:12-67).
class PecPlane : public Object {
 public:
  PecPlane(std::string name, std::unique_ptr<Cube> cube);
};

class VoltageSource : public LumpedElement {
 public:
  VoltageSource(std::string name, std::unique_ptr<Cube> cube, Axis::Direction direction, Real resistance,
                std::unique_ptr<Waveform> waveform, std::unique_ptr<Material> material = Material::createAir());
  Axis::Direction direction() const { return _direction; }
  Real resistance() const { return _resistance; }
  Waveform& waveform() const { return *_waveform; }

 private:
  Axis::Direction _direction;
  Real _resistance;
  std::unique_ptr<Waveform> _waveform;
};

class CurrentSource : public LumpedElement {
 public:
  CurrentSource(std::string name, std::unique_ptr<Cube> cube, Axis::Direction direction, Real resistance,
                std::unique_ptr<Waveform> waveform, std::unique_ptr<Material> material = Material::createAir());
  Axis::Direction direction() const { return _direction; }
  Real resistance() const { return _resistance; }
  Waveform& waveform() const { return *_waveform; }

 private:
  Axis::Direction _direction;
  Real _resistance;
  std::unique_ptr<Waveform> _waveform;
};

class Boundary {
 public:
  virtual ~Boundary() = default;
};

class PML : public Boundary {
 public:
  PML(int thickness, Axis::Direction direction, int order = 4, Real sigma_ratio = 1, Real alpha_min = 0,
      Real alpha_max = 0.05, Real kappa_max = 10);
  int thickness() const { return _thickness; }
  Axis::Direction direction() const { return _direction; }
  int order() const { return _order; }
  Real sigmaRatio() const { return _sigma_ratio; }
  Real alphaMin() const { return _alpha_min; }
  Real alphaMax() const { return _alpha_max; }
  Real kappaMax() const { return _kappa_max; }

 private:
  int _thickness;
  Axis::Direction _direction;
  int _order;
  Real _sigma_ratio, _alpha_min, _alpha_max, _kappa_max;
};

class WaveformSource {
 public:
  explicit WaveformSource(std::unique_ptr<Waveform> w) : _waveform(std::move(w)) {}
  virtual ~WaveformSource() = default;
  Waveform& waveform() const { return *_waveform; }

 private:
  std::unique_ptr<Waveform> _waveform;
};

// include/xfdtd/waveform_source/tfsf.h: the plane-wave source; the box is
// the grid shrunk by (x, y, z) cells on each side.  TFSF3D / TFSF2D / TFSF1D
// pick the corrector calls of their grid's dimension.
class TFSF : public WaveformSource {
 public:
  enum class Kind { ONE = 1, TWO = 2, THREE = 3 };
  TFSF(std::size_t x, std::size_t y, std::size_t z, Real theta, Real phi, Real psi,
       std::unique_ptr<Waveform> waveform, Kind kind)
      : WaveformSource(std::move(waveform)), _x(x), _y(y), _z(z), _theta(theta), _phi(phi), _psi(psi),
        _kind(kind) {}
  std::size_t x() const { return _x; }
  std::size_t y() const { return _y; }
  std::size_t z() const { return _z; }
  Real theta() const { return _theta; }
  Real phi() const { return _phi; }
  Real psi() const { return _psi; }
  Kind kind() const { return _kind; }

 private:
  std::size_t _x, _y, _z;
  Real _theta, _phi, _psi;
  Kind _kind;
};

// include/xfdtd/waveform_source/tfsf_3d.h
class TFSF3D : public TFSF {
 public:
  TFSF3D(std::size_t x, std::size_t y, std::size_t z, Real theta, Real phi, Real psi,
         std::unique_ptr<Waveform> waveform);
};

// include/xfdtd/waveform_source/tfsf_2d.h: a 2D (TE) grid's plane wave,
// theta = pi/2 and psi = 0 (src/waveform_source/tfsf_2d.cpp:10-13)
class TFSF2D : public TFSF {
 public:
  TFSF2D(std::size_t distance_x, std::size_t distance_y, Real phi, std::unique_ptr<Waveform> waveform);
};

// include/xfdtd/waveform_source/tfsf_1d.h: a 1D (TEM) grid's plane wave
// towards +z (forward) or -z (src/waveform_source/tfsf_1d.cpp:14-23)
class TFSF1D : public TFSF {
 public:
  TFSF1D(Index z, bool forward, std::unique_ptr<Waveform> waveform);
  bool forward() const { return _forward; }
  Index start() const { return _start; }

 private:
  bool _forward;
  Index _start;
};

struct EMF {
  enum class Field { EX = 0, EY, EZ, EM, HX, HY, HZ, HM };
  Array3D<Real> ex, ey, ez, hx, hy, hz;
  Array3D<Real>& field(Field f);
};

// include/xfdtd/monitor/monitor.h.  output() writes the monitor's result to
// <outputDir()>/<name()>.npy (monitor.cpp:42-54): `<f8`, C order, as
// xt::dump_npy does; the data it writes is the one copied back by
// Simulation::run().
class Monitor {
 public:
  Monitor(std::string name, std::string output_dir = "xfdtd_output")
      : _name(std::move(name)), _output_dir(std::move(output_dir)) {}
  virtual ~Monitor() = default;
  const std::string& name() const { return _name; }
  const std::string& outputDir() const { return _output_dir; }
  void setName(std::string n) { _name = std::move(n); }
  void setOutputDir(std::string d) { _output_dir = std::move(d); }
  virtual void output() {}

 private:
  std::string _name, _output_dir;
};

// A box of one field (field_monitor.h).  Inside a MovieMonitor it is the
// frame; added on its own, output() writes the box of the final fields
// (field_monitor.cpp:82-96: update() records nothing, output() copies the
// current view).
class FieldMonitor : public Monitor {
 public:
  FieldMonitor(std::unique_ptr<Shape> shape, EMF::Field field, std::string name = "feild_monitor",
               std::string output_dir = "xfdtd_output");
  const Shape& shape() const { return *_shape; }
  EMF::Field field() const { return _field; }
  // Last frame written (field_monitor.cpp:82-96).
  Array3D<Real>& data() { return _data; }
  void output() override;

 private:
  std::unique_ptr<Shape> _shape;
  EMF::Field _field;
  Array3D<Real> _data;
};

// movie_monitor.h: frame files "%05d.npy" of the record count are written
// while the loop runs (movie_monitor.cpp:26-34); output() writes nothing.
class MovieMonitor : public Monitor {
 public:
  MovieMonitor(std::unique_ptr<Monitor> frame, std::size_t frame_interval, std::string name = "movie_monitor",
               std::string output_dir = "xfdtd_output");
  FieldMonitor& frame() const { return *_frame; }
  std::size_t frameInterval() const { return _interval; }
  // Frames in recording order, named "%05d" of the record count.
  std::vector<std::pair<std::string, Array3D<Real>>>& frames() { return _frames; }

 private:
  std::unique_ptr<FieldMonitor> _frame;
  std::size_t _interval;
  std::vector<std::pair<std::string, Array3D<Real>>> _frames;
};

class FieldTimeMonitor : public Monitor {
 public:
  FieldTimeMonitor(std::string name, std::unique_ptr<Cube> cube, std::string output_dir, EMF::Field field);
  const Cube& cube() const { return *_cube; }
  EMF::Field field() const { return _field; }
  std::vector<Real>& time() { return _time; }
  std::vector<Real>& data() { return _data; }
  // time_monitor.cpp:11-21: the (2, nt) stack of time and data.
  void output() override;

 private:
  std::unique_ptr<Cube> _cube;
  EMF::Field _field;
  std::vector<Real> _time, _data;
};

// VoltageMonitor (voltage_monitor.h): every step, -(line integral of E
// along `direction`) through the cube, averaged over its cross-section
// (voltage_monitor.cpp:21-227).  time() holds the E times.
class VoltageMonitor : public Monitor {
 public:
  VoltageMonitor(std::string name, std::unique_ptr<Cube> cube, Axis::Direction direction,
                 std::string output_dir = "");
  const Cube& cube() const { return *_cube; }
  Axis::Direction direction() const { return _direction; }
  std::vector<Real>& time() { return _time; }
  std::vector<Real>& data() { return _data; }
  // time_monitor.cpp:11-21: the (2, nt) stack of time and data.
  void output() override;

 private:
  std::unique_ptr<Cube> _cube;
  Axis::Direction _direction;
  std::vector<Real> _time, _data;
};

// CurrentMonitor (current_monitor.h): every step, the loop integral of H
// around the cube's last cross-section along `direction`
// (current_monitor.cpp:18-385).  time() holds the H times.
class CurrentMonitor : public Monitor {
 public:
  CurrentMonitor(std::string name, std::unique_ptr<Cube> cube, Axis::Direction direction,
                 std::string output_dir = "");
  const Cube& cube() const { return *_cube; }
  Axis::Direction direction() const { return _direction; }
  std::vector<Real>& time() { return _time; }
  std::vector<Real>& data() { return _data; }
  // time_monitor.cpp:11-21: the (2, nt) stack of time and data.
  void output() override;

 private:
  std::unique_ptr<Cube> _cube;
  Axis::Direction _direction;
  std::vector<Real> _time, _data;
};

// A port of an S-parameter network (network/port.h): a current and a
// voltage monitor and a reference impedance z.  calculateSParameters
// (port.cpp:44-130) transforms both recordings (util/dft.h:9-26, the current
// shifted by -1.5 dt) and forms the power waves
//   a = 0.5 (V + z I) / sqrt(Re z),   b = 0.5 (V - conj(z) I) / sqrt(Re z).
// The recordings come from the device loop; the transform runs on the host
// after the run.
class Port {
 public:
  Port(std::size_t index, bool is_source, std::complex<Real> impedance,
       std::shared_ptr<CurrentMonitor> current_monitor, std::shared_ptr<VoltageMonitor> voltage_monitor);
  std::size_t index() const { return _index; }
  bool isSource() const { return _is_source; }
  std::complex<Real> impedance() const { return _impedance; }
  const std::shared_ptr<CurrentMonitor>& currentMonitor() const { return _current; }
  const std::shared_ptr<VoltageMonitor>& voltageMonitor() const { return _voltage; }
  const std::vector<std::complex<Real>>& a() const { return _a; }
  const std::vector<std::complex<Real>>& b() const { return _b; }
  void init(Real dt) { _dt = dt; }  // Port::init keeps the time step (port.cpp:37-42)
  void calculateSParameters(const std::vector<Real>& frequencies);

 private:
  std::size_t _index;
  bool _is_source;
  std::complex<Real> _impedance;
  std::shared_ptr<CurrentMonitor> _current;
  std::shared_ptr<VoltageMonitor> _voltage;
  Real _dt = 0;
  std::vector<std::complex<Real>> _a, _b;
};

// An S-parameter network (network/network.h): output() computes every port's
// power waves and, for each source port i and each port j,
// S(j, i) = b_j / a_i, written as "s{10 j + i}.npy" (complex128) beside
// "frequencies.npy" (network.cpp:47-84).  Port indexes must be unique.
class Network {
 public:
  Network() = default;
  explicit Network(std::string output_dir) : _output_dir(std::move(output_dir)) {}
  Network(std::vector<std::shared_ptr<Port>> ports, std::vector<Real> frequencies, std::string output_dir)
      : _ports(std::move(ports)), _frequencies(std::move(frequencies)), _output_dir(std::move(output_dir)) {}
  void addPort(std::shared_ptr<Port> port) { _ports.emplace_back(std::move(port)); }
  void setFrequencies(std::vector<Real> frequencies) { _frequencies = std::move(frequencies); }
  void setOutputDir(std::string output_dir) { _output_dir = std::move(output_dir); }
  const std::vector<std::shared_ptr<Port>>& ports() const { return _ports; }
  const std::vector<Real>& frequencies() const { return _frequencies; }
  const std::string& outputDir() const { return _output_dir; }
  void init(Real dt);  // network.cpp:32-45
  // Computes the S-parameters; writes them when the output directory is set.
  void output();
  // Key 10 j + i -> S(j, i) over the frequencies, after output().
  const std::vector<std::pair<int, std::vector<std::complex<Real>>>>& sParameters() const { return _s; }

 private:
  std::vector<std::shared_ptr<Port>> _ports;
  std::vector<Real> _frequencies;
  std::string _output_dir;
  std::vector<std::pair<int, std::vector<std::complex<Real>>>> _s;
};

// Frequency-domain near-to-far-field transform (nffft_frequency_domain.h):
// a running DFT of the tangential fields on the box `distance` cells inside
// the grid boundary, run on the GPU every step; after the run the far-field
// potentials are summed on the GPU and the radiated power on the host.
class NFFFTFrequencyDomain {
 public:
  NFFFTFrequencyDomain(std::size_t distance_x, std::size_t distance_y, std::size_t distance_z,
                       std::vector<Real> frequencies, std::string output_dir = "");
  ~NFFFTFrequencyDomain();
  std::size_t freqCount() const { return _frequencies.size(); }
  const std::vector<Real>& frequencies() const { return _frequencies; }
  const std::string& outputDir() const { return _output_dir; }
  std::size_t distance(int axis) const { return _distance[axis]; }
  // The reference advances every frequency's DFT once per frequency each
  // step (nffft_frequency_domain.cpp:45-53), so its currents carry a factor
  // nf and its power nf^2.  true (the default) keeps that; false accumulates
  // each step once, the plain DFT.
  void setReferenceRepeat(bool on) { _reference_repeat = on; }
  bool referenceRepeat() const { return _reference_repeat; }

  // nffft_frequency_domain.cpp:56-160: writes <f>GHz_a_theta.npy,
  // _f_phi, _a_phi, _f_theta (complex) under outputDir()/sub_dir, and
  // power.npy (frequencies; radiated power).
  void processFarField(const std::vector<Real>& theta, Real phi, const std::string& sub_dir,
                       const Vector& origin = Vector{0.0, 0.0, 0.0}) const;
  void processFarField(Real theta, const std::vector<Real>& phi, const std::string& sub_dir,
                       const Vector& origin = Vector{0.0, 0.0, 0.0}) const;
  void outputRadiationPower() const;

  // FDPlaneData::aTheta / aPhi / fTheta / fPhi and power()
  // (nffft_fd_data_process.cpp) of frequency index f; one of theta and phi
  // has a single element.
  enum class Potential { A, F };
  std::vector<std::complex<Real>> potential(std::size_t f, Potential p, bool theta_component,
                                            const std::vector<Real>& theta, const std::vector<Real>& phi,
                                            const Vector& origin = Vector{0.0, 0.0, 0.0}) const;
  // A_theta, A_phi, F_theta, F_phi over the same directions in one pass.
  std::array<std::vector<std::complex<Real>>, 4> potentials(std::size_t f, const std::vector<Real>& theta,
                                                            const std::vector<Real>& phi,
                                                            const Vector& origin = Vector{0.0, 0.0, 0.0}) const;
  // True while the currents live on a device context: the potentials are
  // then summed there (xfd_nf2ff_far_field); XFD_FARFIELD=host forces the
  // host sums.
  bool deviceFarField() const;
  Real power(std::size_t f) const;
  // Surface d (XN XP YN YP ZN ZP): index box [lo, hi) (empty when unused)
  // and the accumulated currents, which = 0 ja, 1 jb, 2 ma, 3 mb.
  std::array<std::size_t, 6> surface(int d) const;
  const std::vector<std::complex<Real>>& current(std::size_t f, int d, int which) const;

  struct Data;
  Data& data() const { return *_data; }

 private:
  std::size_t _distance[3];
  std::vector<Real> _frequencies;
  std::string _output_dir;
  bool _reference_repeat = true;
  std::unique_ptr<Data> _data;
};

// Time-domain near-to-far-field transform (nffft/nffft_time_domain.h): for
// one observation direction (theta, phi), the potentials of the surface
// currents' time derivatives binned by retarded time (TDPlaneData,
// src/include/nffft/nffft_td_data.h), accumulated on the GPU every step.
class NFFFTTimeDomain {
 public:
  NFFFTTimeDomain(std::size_t distance_x, std::size_t distance_y, std::size_t distance_z, Real theta, Real phi);
  NFFFTTimeDomain(std::size_t distance_x, std::size_t distance_y, std::size_t distance_z, std::string output_dir,
                  Real theta, Real phi);
  ~NFFFTTimeDomain();
  Real theta() const { return _theta; }
  Real phi() const { return _phi; }
  std::size_t distance(int axis) const { return _distance[axis]; }
  const std::string& outputDir() const { return _output_dir; }
  void setOutputDir(std::string d) { _output_dir = std::move(d); }
  // processFarField (nfffft_time_domain.cpp:84-160): writes w_theta.npy,
  // w_phi.npy, u_theta.npy and u_phi.npy under outputDir().
  void processFarField() const;
  // The same four arrays (w_theta, w_phi, u_theta, u_phi) after run().
  std::array<std::vector<Real>, 4> farField() const;
  // Potential `which` (0 ua, 1 ub, 2 wa, 3 wb) of surface d after run().
  const std::vector<Real>& potential(int d, int which) const;
  // Length of the potentials, ceil((dmax - dmin) / (c dt)) + steps + 4 (after init).
  std::size_t length() const;

  struct Data;
  Data& data() const { return *_data; }

 private:
  std::size_t _distance[3];
  std::string _output_dir;
  Real _theta, _phi;
  std::unique_ptr<Data> _data;
};

struct ThreadConfig {  // accepted for API compatibility; the GPU path ignores it
  int x = 1, y = 1, z = 1;
};

enum class Precision { FP64, FP32 };

// include/xfdtd/simulation/simulation_flag.h:8-31.  Simulation::run(n) sends
// SimulationStart / SimulationEnd around itself and UpdateStart / UpdateEnd
// around the time loop (simulation.cpp:284-289, domain.cpp:80-137).  The
// loop runs on the GPU, so the per-phase iterator flags (UpdateHSStart ..
// UpdateEEnd) have no host-side moment and are not sent; NextStep is sent
// with cur = the number of steps done after each piece of the loop the host
// waits for (about 100 pieces per run when a visitor is registered, one
// otherwise), start = 0 and end = the run length.
enum class SimulationInitFlag { SimulationStart, SimulationEnd, InitStart, InitEnd, UpdateStart, UpdateEnd };
enum class SimulationIteratorFlag { UpdateHSStart, UpdateHSEnd, UpdateEStart, UpdateEEnd, NextStep };

class SimulationFlagVisitor {
 public:
  virtual ~SimulationFlagVisitor() = default;
  virtual auto initStep(SimulationInitFlag flag) -> void = 0;
  virtual auto iteratorStep(SimulationIteratorFlag flag, Index cur, Index start, Index end) -> void = 0;
};

// include/xfdtd/calculation_param/time_param.h: the run's time axis, valid
// after Simulation::init().  As time_param.cpp:63-71 writes them, eTime() is
// linspace((start + 1) dt, (start + size) dt, size) and hTime() is
// linspace((start + 0.5) dt, (start + size + 0.5) dt, size), whose spacing
// is size / (size - 1) dt.
class TimeParam {
 public:
  TimeParam(Real dt, Index size, Index start_time_step = 0);
  // time_param.cpp:10-21
  static Real calculateDt(Real cfl, Real dx, Real dy, Real dz);
  static Real calculateDt(Real cfl, Real dx, Real dy);
  static Real calculateDt(Real cfl, Real dz);
  Real dt() const { return _dt; }
  Index startTimeStep() const { return _start; }
  Index size() const { return _size; }
  Index endTimeStep() const { return _start + _size; }
  // The step the loop is at: endTimeStep() once run() has returned.
  Index currentTimeStep() const { return _current; }
  const Array1D<Real>& eTime() const { return _e_time; }
  const Array1D<Real>& hTime() const { return _h_time; }
  void setCurrentTimeStep(Index t) { _current = t; }

 private:
  Real _dt;
  Index _start, _size, _current;
  Array1D<Real> _e_time, _h_time;
};

// include/xfdtd/calculation_param/calculation_param.h (its time part).
class CalculationParam {
 public:
  explicit CalculationParam(std::shared_ptr<TimeParam> time_param) : _time(std::move(time_param)) {}
  const std::shared_ptr<TimeParam>& timeParam() const { return _time; }

 private:
  std::shared_ptr<TimeParam> _time;
};

class Simulation {
 public:
  Simulation(Real dx, Real dy, Real dz, Real cfl, ThreadConfig thread_config = ThreadConfig{});
  ~Simulation();
  Simulation(const Simulation&) = delete;
  Simulation& operator=(const Simulation&) = delete;

  // simulation.h:64-66, simulation.cpp:87-91, 148-268: flag visitors; the
  // default one prints the reference's start/end lines and progress bar to
  // stderr.
  auto addVisitor(std::shared_ptr<SimulationFlagVisitor> visitor) -> void;
  auto addDefaultVisitor() -> void;
  auto visitors() -> std::vector<std::shared_ptr<SimulationFlagVisitor>>;
  // simulation.h:78: dt and the run's E/H times (after init()).
  const std::shared_ptr<CalculationParam>& calculationParam() const;

  void addObject(std::shared_ptr<Object> object);
  void addWaveformSource(std::shared_ptr<WaveformSource> source);
  void addBoundary(std::shared_ptr<Boundary> boundary);
  void addMonitor(std::shared_ptr<Monitor> monitor);
  void addNF2FF(std::shared_ptr<NFFFTFrequencyDomain> nffft);
  void addNF2FF(std::shared_ptr<NFFFTTimeDomain> nffft);
  void addNetwork(std::shared_ptr<Network> network);  // simulation.cpp:79-81
  const std::vector<std::shared_ptr<Network>>& networks() const;

  // Field and update precision of the GPU loop (default FP64).  The
  // precompute is FP64 in both cases.
  void setPrecision(Precision p);

  void init(Index time_step);  // simulation.cpp:373-389
  // The host half of init(): grid, materials, coefficients, CPML, TFSF and
  // lumped precompute, without touching the GPU.
  void precompute(Index time_step);
  void run(Index time_step);   // simulation.cpp:284-289
  void run();                  // the time loop only, after init()
  // The loop in pieces: advance(t0, n) runs steps t0..t0+n-1 on the GPU and
  // returns when they are done; finish() then copies monitor recordings and
  // final fields back.  run() == advance(0, steps) + finish().
  void advance(Index t0, Index n);
  void finish();

  Real dt() const;
  std::array<Index, 3> gridSize() const;
  // Final fields, copied back after run() (simulation.h:76).
  const std::shared_ptr<EMF>& emf() const;
  // Precomputed Yee coefficient by reference name ("cexe" ... "chzey").
  const Array3D<Real>& coefficient(const std::string& name) const;
  const std::vector<Real>& eNode(int axis) const;
  const std::vector<Real>& hSize(int axis) const;
  // Time of the stepping loop in seconds (HIP events around the loop).
  double loopSeconds() const;
  struct KernelStat {
    std::string name;
    long long launches;
    double total_ms;
  };
  void setProfiling(bool on);
  // Times only the kernel classes whose bit is set (bit k: kernelStats()[k]).
  void setProfilingClasses(unsigned mask);
  std::vector<KernelStat> kernelStats() const;
  long long deviceBytes() const;
  // The device context (xfd_ctx* of include/xfdtd_core_amd.h) this
  // Simulation steps, for direct calls into the C-ABI; nullptr before init().
  // Borrowed: it lives until the next init() or the Simulation's end.
  void* deviceContext() const;
  // True when the device holds per-cell coefficient class ids instead of the
  // 18 dense coefficient arrays (same values; chosen at init).
  bool classedCoefficients() const;
  // 0: dense coefficient arrays, 1: class ids, 2: 32-bit class ids
  int coefficientLayout() const;
  // Which CPML slabs the Yee sweeps apply in the same pass: 0 none, 1 the
  // leading x/y slabs (XFD_CPML_FUSED=1), 2 the z slabs (FP32 when nz > 192,
  // either precision with XFD_CPML_ZFUSED=1).
  int fusedCpml() const;

  // ---- Adjoint (no reference counterpart) ----
  // Per-cell relative permittivity of the design cells [origin, origin+size)
  // (grid cell indices, k fastest).  It replaces the objects' permittivity
  // there before the edge averaging; conductivity and permeability stay the
  // objects'.
  void setDesignRegion(const std::array<Index, 3>& origin, const std::array<Index, 3>& size,
                       std::vector<Real> eps_r);
  // Objective J = sum over the points p of |D_p|^2, where D_p = sum_n dt
  // exp(-j 2 pi f (n+1) dt) F_p^{n+1} is the running DFT of E component
  // `field` at its grid points [origin, origin+size) over the whole run.
  void setDftObjective(EMF::Field field, const std::array<Index, 3>& origin, const std::array<Index, 3>& size,
                       Real frequency);
  // init(time_step), then the time loop with the state saved every
  // `checkpoint_interval` steps; returns J.  Monitors record as in run().
  Real forward(Index time_step, Index checkpoint_interval);
  // dJ/d eps_r of every design cell (k fastest), by the reverse-time
  // discrete adjoint from the checkpoints of the last forward().
  std::vector<Real> backward();
  std::vector<std::complex<Real>> objectiveDft() const;

  struct Impl;

 private:
  friend class SimulationBatch;
  std::unique_ptr<Impl> _impl;
};

// Many independent Simulations stepped together on one GPU.  No reference
// counterpart (the reference runs one Simulation per process).  Each phase of
// the time step runs as one launch over the work of every case, whatever
// their sizes, and one step can be captured once as a HIP graph and
// replayed.  Every Simulation keeps its own results (emf(), monitors, NF2FF)
// after finish(), bitwise those of running it alone.  The Simulations must
// share one precision and outlive the batch's use of them; their own
// advance() is disabled while they belong to an initialised batch.
class SimulationBatch {
 public:
  SimulationBatch();
  ~SimulationBatch();
  SimulationBatch(const SimulationBatch&) = delete;
  SimulationBatch& operator=(const SimulationBatch&) = delete;

  void add(Simulation& sim);
  std::size_t size() const;
  // Precomputes every case (as Simulation::precompute) and creates the batch.
  void init(Index time_step);
  // Capture one step as a HIP graph on the first advance() and replay it.
  void setGraph(bool on);
  // Picks the i-chunk and persistent grid size by timing the Yee sweeps,
  // then clears every case's state.
  void tune(int reps = 5);
  void advance(Index t0, Index n);
  void finish();  // monitor recordings, final fields and NF2FF currents of every case
  void run();
  void run(Index time_step);
  double loopSeconds() const;         // host clock around the last advance()
  double deviceMilliseconds() const;  // HIP events around the last advance()
  struct Layout {
    bool shared;  // false: cases with dense coefficients or ADE, stepped one by one
    int ichunk;
    int grid;
    int lanes_k, rows_j;  // the sweep tile
  };
  Layout layout() const;

  struct Impl;

 private:
  std::unique_ptr<Impl> _impl;
};

}  // namespace xfdtd

#endif  // XFDTD_CORE_AMD_HPP_
