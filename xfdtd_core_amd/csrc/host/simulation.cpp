// Host side of Simulation: the FP64 precompute of the reference's init()
// (src/simulation/simulation.cpp:319-389) and the hand-off to the device
// library.  The arithmetic of every precomputed value follows the reference
// expression for expression (cited per stage) so the coefficient arrays are
// bit-identical; the work itself is organised for large grids: each stage
// is a parallel sweep over i-planes, the staircase visits only the cells in
// a shape's bounding box, and the CPML/TFSF/lumped data are emitted directly
// in the form the C-ABI consumes.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <limits>
#include <map>
#include <thread>

#include "../../../include/xfdtd_core_amd.h"
#include "../../../include/xfdtd_core_amd.hpp"
#include "nf2ff_data.hpp"
#include "npy.hpp"

namespace xfdtd {

namespace {

constexpr Real kPi = 3.141592653589793238462643383279502884;
constexpr Real kEps0 = 8.854187817e-12;  // include/xfdtd/common/constant.h:12-13
constexpr Real kMu0 = 4 * kPi * 1e-7;
const Real kC0 = 1 / std::sqrt(kEps0 * kMu0);
constexpr Real kSigmaZero = 1e-20;  // material.h:40-41

// Runs fn(lo, hi) over [0, n) split across the host's cores.
template <class F>
void parallelFor(std::size_t n, F&& fn) {
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  std::size_t parts = std::min<std::size_t>({static_cast<std::size_t>(hw), n, 32});
  if (parts <= 1) {
    if (n) fn(std::size_t{0}, n);
    return;
  }
  std::vector<std::thread> th;
  std::size_t step = (n + parts - 1) / parts;
  for (std::size_t p = 0; p < parts; ++p) {
    std::size_t lo = p * step, hi = std::min(n, lo + step);
    if (lo >= hi) break;
    th.emplace_back([&fn, lo, hi] { fn(lo, hi); });
  }
  for (auto& t : th) t.join();
}

// xt::linspace into a container: running sum of the step, last element
// pinned to stop (xtensor 0.25 arange_generator::assign_to).  Used for the
// e-nodes (grid_space_generator.cpp:210-212) and eTime/hTime
// (time_param.cpp:63-71).
std::vector<Real> linspaceAssign(Real start, Real stop, std::size_t n) {
  std::vector<Real> v(n);
  const Real step = (stop - start) / std::fmax(Real(1), Real(n - 1));
  Real acc = start;
  for (std::size_t i = 0; i < n; ++i) {
    v[i] = acc;
    acc += step;
  }
  if (n > 1) v[n - 1] = stop;
  return v;
}

// float_compare.h:24-42, LessEqual
bool leq(Real a, Real b, Real eps) { return std::abs(a - b) <= eps ? true : a < b; }

int ax(Axis::XYZ a) { return static_cast<int>(a); }

}  // namespace

// ------------------------------------------------------------- shapes ----
Cube::Cube(Vector o, Vector s) : _origin(o), _size(s) {
  Real e[3];
  for (int a = 0; a < 3; ++a) {
    e[a] = o[a] + s[a];  // cube.cpp:128-139
    if (std::isnan(e[a])) e[a] = std::numeric_limits<Real>::infinity();
  }
  _end = Vector(e[0], e[1], e[2]);
}
bool Cube::isInside(const Vector& p, Real eps) const {  // cube.cpp:107-118
  bool in[3];
  for (int a = 0; a < 3; ++a) in[a] = leq(_origin[a], p[a], eps) && leq(p[a], _end[a], eps);
  return in[0] && in[1] && in[2];
}
std::unique_ptr<Cube> Cube::wrappedCube() const { return std::make_unique<Cube>(*this); }
std::unique_ptr<Shape> Cube::clone() const { return std::make_unique<Cube>(*this); }

Sphere::Sphere(Vector c, Real r) : _center(c), _radius(r) {
  if (r < 0) throw XFDTDException("Sphere radius must be positive");
}
bool Sphere::isInside(const Vector& p, Real eps) const {  // sphere.cpp:29-32
  const Real dx = p.x() - _center.x(), dy = p.y() - _center.y(), dz = p.z() - _center.z();
  return leq(std::sqrt((dx * dx + dy * dy) + dz * dz), _radius, eps);
}
std::unique_ptr<Cube> Sphere::wrappedCube() const {  // sphere.cpp:34-37
  return std::make_unique<Cube>(
      Vector(_center.x() - _radius, _center.y() - _radius, _center.z() - _radius),
      Vector(_radius * 2, _radius * 2, _radius * 2));
}
std::unique_ptr<Shape> Sphere::clone() const { return std::make_unique<Sphere>(*this); }

Cylinder::Cylinder(Vector c, Real r, Real h, Axis::XYZ a) : _center(c), _radius(r), _height(h), _axis(a) {
  if (r < 0 || h < 0) throw XFDTDException("Cylinder radius and height must be positive");
}
std::unique_ptr<Cube> Cylinder::wrappedCube() const {  // cylinder.cpp:73-98
  const Real s = _radius * 2;
  const int m = ax(_axis);
  Real o[3], z[3];
  for (int a = 0; a < 3; ++a) {
    o[a] = a == m ? _center[a] - _height / 2 : _center[a] - _radius;
    z[a] = a == m ? _height : s;
  }
  return std::make_unique<Cube>(Vector(o[0], o[1], o[2]), Vector(z[0], z[1], z[2]));
}
bool Cylinder::isInside(const Vector& p, Real eps) const {  // cylinder.cpp:41-66
  if (!wrappedCube()->isInside(p, eps)) return false;
  const int m = ax(_axis);
  const int u = m == 0 ? 1 : 0, w = m == 2 ? 1 : 2;
  const Real d = std::sqrt(std::pow(p[u] - _center[u], 2) + std::pow(p[w] - _center[w], 2));
  return leq(d, _radius, 1e-8);
}
std::unique_ptr<Shape> Cylinder::clone() const { return std::make_unique<Cylinder>(*this); }

ElectroMagneticProperty::ElectroMagneticProperty(Real er, Real mr, Real se, Real sm)
    : _eps_r(er), _mu_r(mr), _sigma_e(se), _sigma_m(sm), _eps(er * kEps0), _mu(mr * kMu0) {
  if (_sigma_e == 0) _sigma_e = kSigmaZero;  // material.cpp:20-25
  if (_sigma_m == 0) _sigma_m = kSigmaZero;
}

LinearDispersiveMaterial::LinearDispersiveMaterial(std::string name, Real epsilon_inf, Kind kind,
                                                   std::vector<Real> a0, std::vector<Real> a1,
                                                   std::vector<Real> b0, std::vector<Real> b1,
                                                   std::vector<Real> b2, ElectroMagneticProperty emp)
    : Material(std::move(name), emp),
      _kind(kind),
      _epsilon_inf(epsilon_inf),
      _a0(std::move(a0)),
      _a1(std::move(a1)),
      _b0(std::move(b0)),
      _b1(std::move(b1)),
      _b2(std::move(b2)) {
  const auto n = _a0.size();
  if (_a1.size() != n || _b0.size() != n || _b1.size() != n || _b2.size() != n)
    throw XFDTDException("Modified Lorentz Model paramater doesn't have same size.");
}
std::unique_ptr<MLorentzMaterial> MLorentzMaterial::makeMLorentz(std::string name, Real epsilon_inf,
                                                                 std::vector<Real> a_0, std::vector<Real> a_1,
                                                                 std::vector<Real> b_0, std::vector<Real> b_1,
                                                                 std::vector<Real> b_2) {
  return std::make_unique<MLorentzMaterial>(std::move(name), epsilon_inf, Kind::MLorentz, std::move(a_0),
                                            std::move(a_1), std::move(b_0), std::move(b_1), std::move(b_2));
}
std::unique_ptr<DrudeMedium> DrudeMedium::makeDrudeMedium(std::string name, Real eps_inf, std::vector<Real> omega_p,
                                                          std::vector<Real> gamma) {
  const auto n = omega_p.size();
  std::vector<Real> a0(n);
  for (std::size_t p = 0; p < n; ++p) a0[p] = omega_p[p] * omega_p[p];
  return std::make_unique<DrudeMedium>(std::move(name), eps_inf, Kind::Drude, std::move(a0),
                                       std::vector<Real>(n, 0.0), std::vector<Real>(n, 0.0), std::move(gamma),
                                       std::vector<Real>(n, 1.0));
}
std::unique_ptr<DebyeMedium> DebyeMedium::makeDebyeMedium(std::string name, Real epsilon_inf,
                                                          std::vector<Real> epsilon_static, std::vector<Real> tau) {
  const auto n = epsilon_static.size();
  std::vector<Real> a0(n);
  for (std::size_t p = 0; p < n; ++p) a0[p] = epsilon_static[p] - epsilon_inf;
  return std::make_unique<DebyeMedium>(std::move(name), epsilon_inf, Kind::Debye, std::move(a0),
                                       std::vector<Real>(n, 0.0), std::vector<Real>(n, 1.0), std::move(tau),
                                       std::vector<Real>(n, 0.0));
}

// ----------------------------------------------------------- waveforms ---
void Waveform::init(const std::vector<Real>& time) {
  _value.assign(time.size(), 0);
  for (std::size_t i = 0; i < time.size(); ++i) _value[i] = _amplitude * _func(time[i]);
}
std::unique_ptr<Waveform> Waveform::gaussian(Real tau, Real t_0, Real amplitude) {  // waveform.cpp:80-85
  return std::make_unique<Waveform>([tau, t_0](Real t) { return std::exp(-std::pow((t - t_0) / tau, 2)); },
                                    amplitude);
}
std::unique_ptr<Waveform> Waveform::cosineModulatedGaussian(Real tau, Real t_0, Real f, Real amplitude) {
  return std::make_unique<Waveform>(
      [tau, t_0, f](Real t) {
        return std::exp(-std::pow((t - t_0) / tau, 2)) * std::cos(2 * kPi * f * (t - t_0));
      },
      amplitude);
}
std::unique_ptr<Waveform> Waveform::sine(Real f, Real amplitude) {
  return std::make_unique<Waveform>([f](Real t) { return std::sin(2 * kPi * f * t); }, amplitude);
}
std::unique_ptr<Waveform> Waveform::cosine(Real f, Real amplitude) {
  return std::make_unique<Waveform>([f](Real t) { return std::cos(2 * kPi * f * t); }, amplitude);
}
std::unique_ptr<Waveform> Waveform::step(Real t_0, Real amplitude) {
  return std::make_unique<Waveform>([t_0](Real t) { return static_cast<Real>(t >= t_0); }, amplitude);
}

// ------------------------------------------------------------- objects ---
Object::Object(std::string n, std::unique_ptr<Shape> s, std::unique_ptr<Material> m)
    : _name(std::move(n)), _shape(std::move(s)), _material(std::move(m)) {}
LumpedElement::LumpedElement(std::string n, std::unique_ptr<Cube> c, Axis::XYZ xyz, std::unique_ptr<Material> m)
    : Object(std::move(n), std::move(c), std::move(m)), _xyz(xyz) {}
Resistor::Resistor(std::string n, std::unique_ptr<Cube> c, Axis::XYZ xyz, Real r, std::unique_ptr<Material> m)
    : LumpedElement(std::move(n), std::move(c), xyz, std::move(m)), _resistance(r == 0 ? 1e-10 : r) {}
Capacitor::Capacitor(std::string n, std::unique_ptr<Cube> c, Axis::XYZ xyz, Real cap, std::unique_ptr<Material> m)
    : LumpedElement(std::move(n), std::move(c), xyz, std::move(m)), _capacitance(cap) {}
Inductor::Inductor(std::string n, std::unique_ptr<Cube> c, Axis::XYZ xyz, Real l, std::unique_ptr<Material> m)
    : LumpedElement(std::move(n), std::move(c), xyz, std::move(m)), _inductance(l == 0 ? 1e-10 : l) {}
PecPlane::PecPlane(std::string n, std::unique_ptr<Cube> c) : Object(std::move(n), std::move(c), Material::createPec()) {}
static Axis::XYZ dirAxis(Axis::Direction d) { return static_cast<Axis::XYZ>(static_cast<int>(d) / 2); }
static bool dirNeg(Axis::Direction d) { return static_cast<int>(d) % 2 == 0; }
VoltageSource::VoltageSource(std::string n, std::unique_ptr<Cube> c, Axis::Direction d, Real r,
                             std::unique_ptr<Waveform> w, std::unique_ptr<Material> m)
    : LumpedElement(std::move(n), std::move(c), dirAxis(d), std::move(m)),
      _direction(d),
      _resistance(r == 0 ? 1e-20 : r),
      _waveform(std::move(w)) {}
CurrentSource::CurrentSource(std::string n, std::unique_ptr<Cube> c, Axis::Direction d, Real r,
                             std::unique_ptr<Waveform> w, std::unique_ptr<Material> m)
    : LumpedElement(std::move(n), std::move(c), dirAxis(d), std::move(m)),
      _direction(d),
      _resistance(r == 0 ? 1e-20 : r),
      _waveform(std::move(w)) {}
PML::PML(int t, Axis::Direction d, int o, Real sr, Real amin, Real amax, Real kmax)
    : _thickness(t), _direction(d), _order(o), _sigma_ratio(sr), _alpha_min(amin), _alpha_max(amax),
      _kappa_max(kmax) {}
TFSF3D::TFSF3D(std::size_t x, std::size_t y, std::size_t z, Real th, Real ph, Real ps, std::unique_ptr<Waveform> w)
    : TFSF(x, y, z, th, ph, ps, std::move(w), Kind::THREE) {}
TFSF2D::TFSF2D(std::size_t x, std::size_t y, Real ph, std::unique_ptr<Waveform> w)  // tfsf_2d.cpp:10-13
    : TFSF(x, y, 0, constant::PI * 0.5, ph, 0, std::move(w), Kind::TWO) {}
TFSF1D::TFSF1D(Index z, bool forward, std::unique_ptr<Waveform> w)  // tfsf_1d.cpp:14-23
    : TFSF(0, 0, z, forward ? 0 : constant::PI, 0, 0, std::move(w), Kind::ONE), _forward(forward), _start(z) {}

Array3D<Real>& EMF::field(Field f) {
  switch (f) {
    case Field::EX: return ex;
    case Field::EY: return ey;
    case Field::EZ: return ez;
    case Field::HX: return hx;
    case Field::HY: return hy;
    case Field::HZ: return hz;
    default: throw XFDTDException("EMF::field: unsupported field");
  }
}

FieldMonitor::FieldMonitor(std::unique_ptr<Shape> s, EMF::Field f, std::string n, std::string d)
    : Monitor(std::move(n), std::move(d)), _shape(std::move(s)), _field(f) {}
MovieMonitor::MovieMonitor(std::unique_ptr<Monitor> frame, std::size_t interval, std::string n, std::string d)
    : Monitor(std::move(n), std::move(d)), _interval(interval) {
  auto* fm = dynamic_cast<FieldMonitor*>(frame.get());
  if (!fm) throw XFDTDException("MovieMonitor frame must be a FieldMonitor");
  frame.release();
  _frame.reset(fm);
  if (_interval == 0) throw XFDTDException("MovieMonitor interval must be positive");
}
FieldTimeMonitor::FieldTimeMonitor(std::string n, std::unique_ptr<Cube> c, std::string d, EMF::Field f)
    : Monitor(std::move(n), std::move(d)), _cube(std::move(c)), _field(f) {}
VoltageMonitor::VoltageMonitor(std::string n, std::unique_ptr<Cube> c, Axis::Direction dir, std::string d)
    : Monitor(std::move(n), std::move(d)), _cube(std::move(c)), _direction(dir) {}
CurrentMonitor::CurrentMonitor(std::string n, std::unique_ptr<Cube> c, Axis::Direction dir, std::string d)
    : Monitor(std::move(n), std::move(d)), _cube(std::move(c)), _direction(dir) {}

namespace {
// Monitor::output (monitor.cpp:42-54): <dir>/<name>.npy, the directory
// created when missing; an empty directory name is the working directory.
std::string outputPath(const Monitor& m) {
  if (m.outputDir().empty()) return m.name() + ".npy";
  std::filesystem::create_directories(m.outputDir());
  return (std::filesystem::path(m.outputDir()) / (m.name() + ".npy")).string();
}

// TimeMonitor::output (time_monitor.cpp:11-21): xt::stack(time, data), a
// (2, nt) array.
void writeTimeSeries(const Monitor& m, const std::vector<Real>& time, const std::vector<Real>& data) {
  if (time.size() != data.size()) throw XFDTDException(m.name() + ": no recording to output (run first)");
  std::vector<Real> st(time);
  st.insert(st.end(), data.begin(), data.end());
  npy::write(outputPath(m), {2, time.size()}, st.data());
}
}  // namespace

void FieldMonitor::output() {
  npy::write(outputPath(*this), {_data.shape[0], _data.shape[1], _data.shape[2]}, _data.data.data());
}
void FieldTimeMonitor::output() { writeTimeSeries(*this, _time, _data); }
void VoltageMonitor::output() { writeTimeSeries(*this, _time, _data); }
void CurrentMonitor::output() { writeTimeSeries(*this, _time, _data); }

// time_param.cpp:10-21
Real TimeParam::calculateDt(Real cfl, Real dx, Real dy, Real dz) {
  return cfl / (constant::C_0 * std::sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz)));
}
Real TimeParam::calculateDt(Real cfl, Real dx, Real dy) {
  return cfl / (constant::C_0 * std::sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy)));
}
Real TimeParam::calculateDt(Real cfl, Real dz) { return cfl / (constant::C_0 * std::sqrt(1.0 / (dz * dz))); }

TimeParam::TimeParam(Real dt, Index size, Index start_time_step)
    : _dt(dt), _start(start_time_step), _size(size), _current(start_time_step) {
  // time_param.cpp:63-71: linspace((start + 1) dt, (end) dt) and the half steps
  const Real a = static_cast<Real>(start_time_step), b = static_cast<Real>(start_time_step + size);
  _e_time.data = linspaceAssign((a + 1) * dt, b * dt, size);
  _h_time.data = linspaceAssign((a + 0.5) * dt, (b + 0.5) * dt, size);
}

namespace {
// The reference's default visitor (simulation.cpp:148-268): start / end
// lines with elapsed times and a progress bar on stderr.
class DefaultSimulationFlagVisitor : public SimulationFlagVisitor {
 public:
  auto initStep(SimulationInitFlag flag) -> void override {
    using clock = std::chrono::high_resolution_clock;
    auto elapsed = [](clock::time_point t0) {
      return std::to_string(std::chrono::duration_cast<std::chrono::milliseconds>(clock::now() - t0).count()) +
             " ms";
    };
    switch (flag) {
      case SimulationInitFlag::SimulationStart:
        _sim = clock::now();
        std::fprintf(stderr, "Simulation start\n");
        break;
      case SimulationInitFlag::SimulationEnd:
        std::fprintf(stderr, "Simulation end\nElapsed time: %s\n", elapsed(_sim).c_str());
        break;
      case SimulationInitFlag::InitStart:
        _init = clock::now();
        std::fprintf(stderr, "Simulation init start\n");
        break;
      case SimulationInitFlag::InitEnd:
        std::fprintf(stderr, "Simulation init end\nElapsed time: %s\n", elapsed(_init).c_str());
        break;
      case SimulationInitFlag::UpdateStart:
        _update = clock::now();
        std::fprintf(stderr, "Simulation update start\n");
        break;
      case SimulationInitFlag::UpdateEnd:
        std::fprintf(stderr, "Simulation update end\nElapsed time: %s\n", elapsed(_update).c_str());
        break;
    }
  }
  auto iteratorStep(SimulationIteratorFlag flag, Index cur, Index start, Index end) -> void override {
    if (flag != SimulationIteratorFlag::NextStep || end == 0) return;
    using clock = std::chrono::high_resolution_clock;
    const double p = static_cast<double>(cur) / static_cast<double>(end);
    const int width = 50, pos = static_cast<int>(width * p);
    std::string bar(width, ' ');
    for (int i = 0; i < width; ++i) bar[i] = i < pos ? '=' : (i == pos ? '>' : ' ');
    const double ms = std::chrono::duration<double, std::milli>(clock::now() - _update).count();
    const double left = cur > start ? ms * static_cast<double>(end - cur) / static_cast<double>(cur - start) : 0.0;
    std::fprintf(stderr, "\rProgress: [%s] %3d%% %zu/%zu Elapsed time: %.0f ms. Estimated remaining time: %.0f ms%s",
                 bar.c_str(), static_cast<int>(p * 100.0), cur, end, ms, left, cur == end ? "\n" : "");
    std::fflush(stderr);
  }

 private:
  std::chrono::high_resolution_clock::time_point _sim, _init, _update;
};
}  // namespace

// ------------------------------------------------------------- network ---
namespace {
// dft (util/dft.h:9-26): X(f) = dt sum_j x_j exp(-2 i pi f ((j + 1) dt + shift)),
// each term in the reference's evaluation order.
std::vector<std::complex<Real>> seriesDft(const std::vector<Real>& x, Real dt, const std::vector<Real>& freqs,
                                          Real shift) {
  const std::complex<Real> II(0.0, 1.0);
  std::vector<std::complex<Real>> res(freqs.size());
  for (std::size_t i = 0; i < freqs.size(); ++i) {
    std::complex<Real> sum{0.0, 0.0};
    for (std::size_t j = 0; j < x.size(); ++j) {
      const Real t = static_cast<Real>(j + 1) * dt + shift;
      sum += x[j] * std::exp(-static_cast<Real>(2.0) * II * kPi * freqs[i] * t);
    }
    res[i] = sum * dt;
  }
  return res;
}
}  // namespace

Port::Port(std::size_t index, bool is_source, std::complex<Real> impedance,
           std::shared_ptr<CurrentMonitor> current_monitor, std::shared_ptr<VoltageMonitor> voltage_monitor)
    : _index(index),
      _is_source(is_source),
      _impedance(impedance),
      _current(std::move(current_monitor)),
      _voltage(std::move(voltage_monitor)) {}

// Port::calculateSParameters (port.cpp:44-130), one process.
void Port::calculateSParameters(const std::vector<Real>& frequencies) {
  if (!_current || !_voltage) throw XFDTDException("Port: monitor missing");
  const auto z = _impedance;
  const Real k = std::sqrt(std::real(z));
  const auto i = seriesDft(_current->data(), _dt, frequencies, -1.5 * _dt);
  const auto v = seriesDft(_voltage->data(), _dt, frequencies, 0);
  _a.resize(frequencies.size());
  _b.resize(frequencies.size());
  for (std::size_t f = 0; f < frequencies.size(); ++f) {
    _a[f] = Real{0.5} * (v[f] + z * i[f]) / k;
    _b[f] = Real{0.5} * (v[f] - std::conj(z) * i[f]) / k;
  }
}

void Network::init(Real dt) {
  std::vector<std::size_t> seen;
  for (auto& p : _ports) {
    if (!p) throw XFDTDException("Network: null port");
    p->init(dt);
    if (std::find(seen.begin(), seen.end(), p->index()) != seen.end())
      throw XFDTDException("Port index must be unique");
    seen.push_back(p->index());
  }
}

// Network::output (network.cpp:47-84): S(j, i) = b_j / a_i for every source
// port i, keyed 10 j + i.
void Network::output() {
  for (auto& p : _ports) p->calculateSParameters(_frequencies);
  _s.clear();
  for (std::size_t i = 0; i < _ports.size(); ++i) {
    if (!_ports[i]->isSource()) continue;
    for (std::size_t j = 0; j < _ports.size(); ++j) {
      const int key = static_cast<int>(_ports[j]->index() * 10 + _ports[i]->index());
      std::vector<std::complex<Real>> sv(_frequencies.size());
      for (std::size_t f = 0; f < sv.size(); ++f) sv[f] = _ports[j]->b()[f] / _ports[i]->a()[f];
      auto it = std::find_if(_s.begin(), _s.end(), [&](const auto& e) { return e.first == key; });
      if (it != _s.end())
        it->second = std::move(sv);
      else
        _s.emplace_back(key, std::move(sv));
    }
  }
  if (_output_dir.empty()) return;
  std::filesystem::create_directories(_output_dir);
  for (const auto& e : _s)
    npy::writeComplex(_output_dir + "/s" + std::to_string(e.first) + ".npy", {e.second.size()},
                      reinterpret_cast<const double*>(e.second.data()));
  npy::write(_output_dir + "/frequencies.npy", {_frequencies.size()}, _frequencies.data());
}

// ---------------------------------------------------------- simulation ---
namespace {

// One staggered array in the reference extent, k fastest.
// Large host array whose pages are first touched by the threads that fill
// it (an uninitialised allocation followed by a parallel sweep), so page
// faults are taken in parallel rather than by one zero-filling thread.
struct Buffer {
  std::unique_ptr<Real[]> p;
  std::size_t n = 0;
  Real& operator[](std::size_t i) { return p[i]; }
  Real operator[](std::size_t i) const { return p[i]; }
  Real* data() { return p.get(); }
  const Real* data() const { return p.get(); }
  std::size_t size() const { return n; }
  void allocate(std::size_t count) {
    p.reset(count ? new Real[count] : nullptr);  // uninitialised
    n = count;
  }
  void release() {
    p.reset();
    n = 0;
  }
};

struct Field3 {
  std::size_t n[3] = {0, 0, 0};
  Buffer v;
  // Allocates without initialising; the caller writes every element.
  void allocate(const std::size_t e[3]) {
    for (int a = 0; a < 3; ++a) n[a] = e[a];
    v.allocate(e[0] * e[1] * e[2]);
  }
  void resize(const std::size_t e[3], Real fill) {
    allocate(e);
    const std::size_t plane = n[1] * n[2];
    parallelFor(n[0], [&](std::size_t lo, std::size_t hi) {
      std::fill(v.data() + lo * plane, v.data() + hi * plane, fill);
    });
  }
  Real& at(std::size_t i, std::size_t j, std::size_t k) { return v[(i * n[1] + j) * n[2] + k]; }
  Real at(std::size_t i, std::size_t j, std::size_t k) const { return v[(i * n[1] + j) * n[2] + k]; }
};

// Extent of component c of E (h = false) or H.
void extent(bool h, int c, const std::size_t N[3], std::size_t out[3]) {
  for (int a = 0; a < 3; ++a) out[a] = h ? (a == c ? N[a] + 1 : N[a]) : (a == c ? N[a] : N[a] + 1);
}

const char* kCoefNames[18] = {"cexe", "cexhy", "cexhz", "ceye", "ceyhz", "ceyhx", "ceze", "cezhx", "cezhy",
                              "chxh", "chxey", "chxez", "chyh", "chyez", "chyex", "chzh", "chzex", "chzey"};

// The device batch shared by a SimulationBatch and the Simulations whose
// contexts are its cases; destroyed with the last of them.
struct BatchCore {
  xfd_batch* b = nullptr;
  ~BatchCore() {
    if (b) xfd_batch_destroy(b);
  }
};

}  // namespace

struct Simulation::Impl {
  Real d[3], cfl;
  Precision precision = Precision::FP64;
  std::vector<std::shared_ptr<Object>> objects;
  std::vector<std::shared_ptr<WaveformSource>> sources;
  std::vector<std::shared_ptr<Boundary>> boundaries;
  std::vector<std::shared_ptr<Monitor>> monitors;
  std::vector<std::shared_ptr<NFFFTFrequencyDomain>> nf2ffs;
  std::vector<std::shared_ptr<Network>> networks;
  std::vector<xfd_nf2ff_desc> nf2ff_d;
  std::vector<std::shared_ptr<NFFFTTimeDomain>> nf2ff_tds;
  std::vector<xfd_tdnf_desc> tdnf_d;

  // grid (grid_space_generator.cpp, grid_space.cpp:526-544)
  std::size_t N[3] = {0, 0, 0};
  std::vector<Real> e_node[3], h_node[3], e_size[3], h_size[3];
  Real min_d[3];
  // GridSpace::Dimension: 3, 2 (z infinite, one cell along z) or 1 (x and
  // y infinite, one cell along each); the based cell sizes of the grid
  int dim = 3;
  Real based[3] = {0, 0, 0};
  Real dt = 0;
  Index steps = 0;
  bool initialised = false;

  // material (eps, mu, sigma_e, sigma_m per component) and coefficients
  Field3 eps[3], mu[3], se[3], sm[3];
  Field3 coef[18];
  // Per-cell index (into `objects`) of the object whose shape took the cell
  // in the staircase; kept only when some object is dispersive.
  std::vector<int32_t> owner;
  // ADE: scheme, pole count, per-object coefficient rows, per-cell ids
  xfd_ade_desc ade_d{};
  bool has_ade = false;
  std::vector<Real> ade_coeff;
  std::vector<int16_t> ade_cell;
  mutable std::map<std::string, Array3D<Real>> coef_view;

  // device hand-off
  struct SlabData {
    xfd_cpml_desc d{};
    std::vector<Real> ae, be, ah, bh;
    std::vector<Real> cpsi[4];
  };
  std::vector<SlabData> slabs;
  struct LumpedData {
    xfd_lumped_desc d{};
    std::vector<Real> coff, coff_j;
    std::vector<Real> wave;
  };
  std::vector<LumpedData> lumped;
  struct TfsfData {
    xfd_tfsf_desc d{};
    std::vector<Real> proj[6];
    std::vector<Real> e_inc, h_inc;
  };
  std::vector<TfsfData> tfsf;
  struct GatherPlan {
    xfd_gather_desc d{};
    Monitor* monitor;
    bool time_series;
  };
  std::vector<GatherPlan> gathers;
  // voltage / current monitors as weighted sums (xfd_line_desc)
  struct LinePlan {
    Monitor* monitor;
    std::vector<int32_t> terms;  // field, i, j, k, group
    std::vector<Real> weights;
    Real scale = 1;
  };
  std::vector<LinePlan> lines;
  std::vector<xfd_line_desc> ld;

  // Adjoint: design cells with their relative permittivity, the DFT
  // objective (a box of points of one E component at one frequency) and the
  // results of the last forward().
  bool has_design = false;
  std::size_t design_o[3] = {0, 0, 0}, design_n[3] = {0, 0, 0};
  std::vector<Real> design_eps;
  bool has_objective = false;
  int obj_field = 2;
  std::size_t obj_o[3] = {0, 0, 0}, obj_n[3] = {0, 0, 0};
  Real obj_freq = 0;
  int obj_gather = -1;  // index in gathers
  std::vector<std::complex<Real>> obj_dft;
  Real obj_value = 0;
  bool forward_done = false;

  // flag visitors (simulation.h:64-66), the run's time axis and the
  // FieldMonitors added on their own with their node box
  std::vector<std::shared_ptr<SimulationFlagVisitor>> visitors;
  std::shared_ptr<CalculationParam> calc;
  struct FieldBox {
    FieldMonitor* monitor;
    int32_t o[3], e[3];
  };
  std::vector<FieldBox> field_boxes;

  xfd_ctx* ctx = nullptr;
  // Set when ctx is a case of a SimulationBatch: the batch core owns it.
  std::shared_ptr<BatchCore> batch_core;
  std::shared_ptr<EMF> emf = std::make_shared<EMF>();
  double loop_seconds = 0;
  bool profiling = false;
  uint32_t profiling_classes = ~0u;
  std::vector<KernelStat> stats;
  long long dev_bytes = 0;
  int coef_layout = 0;  // 0 dense, 1 16-bit class ids, 2 32-bit class ids
  int cpml_fused = 0;

  ~Impl() { releaseDevice(); }

  void releaseDevice() {
    for (auto& n : nf2ffs) n->data().device = nullptr;
    if (ctx && !batch_core) xfd_destroy(ctx);
    ctx = nullptr;
    batch_core.reset();
  }

  // ---------------------------------------------------------------- grid --
  Real tolerance() const {  // grid_space.cpp:459-468, per dimension
    if (dim == 1) return min_d[2] / 1000;
    if (dim == 2) return std::min(min_d[0], min_d[1]) / 1000;
    return std::min({min_d[0], min_d[1], min_d[2]}) / 1000;
  }

  std::size_t nearest(int a, Real x) const {  // xt::argmin(abs(e_node - x)), first minimum
    // the one-cell axes of a 1D/2D grid: GridSpace2D::handleTransformZ
    // (grid_space_2d.cpp:116-122), GridSpace1D::handleTransformX/Y
    // (grid_space_1d.cpp:117-135): 0 below +inf, else 1
    if ((dim == 2 && a == 2) || (dim == 1 && a < 2)) return x < std::numeric_limits<Real>::infinity() ? 0 : 1;
    const auto& e = e_node[a];
    std::size_t best = 0;
    Real bv = std::abs(e[0] - x);
    for (std::size_t i = 1; i < e.size(); ++i) {
      const Real v = std::abs(e[i] - x);
      if (v < bv) {
        bv = v;
        best = i;
      }
    }
    return best;
  }

  void gridBox(const Shape& s, std::size_t o[3], std::size_t e[3]) const {  // grid_space.cpp:282-293
    auto c = s.wrappedCube();
    for (int a = 0; a < 3; ++a) {
      o[a] = nearest(a, c->origin()[a]);
      e[a] = nearest(a, c->end()[a]);
    }
  }

  void buildGrid() {
    // Shape::makeWrappedCube (shape.cpp:8-38) -> extendDomain
    // (grid_space_generator.cpp:36-99) -> generateUniformGridSpace (:16-34):
    // 3D (:183-217), 2D when only z is infinite (:152-181), 1D when x and y
    // are (:126-150; decideDimension :101-124)
    if (objects.empty()) throw XFDTDSimulationException("No object is added");
    Real lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::numeric_limits<Real>::max();
      hi[a] = std::numeric_limits<Real>::lowest();
    }
    for (auto& o : objects) {
      auto c = o->shape().wrappedCube();
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], c->origin()[a]);
        hi[a] = std::max(hi[a], c->end()[a]);
      }
    }
    Cube dom(Vector(lo[0], lo[1], lo[2]), Vector(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]));
    auto dimension = [](const Cube& c) {
      const bool inf[3] = {std::isinf(c.size()[0]), std::isinf(c.size()[1]), std::isinf(c.size()[2])};
      if (inf[0] && inf[1] && !inf[2]) return 1;
      if (!inf[0] && !inf[1] && inf[2]) return 2;
      if (!inf[0] && !inf[1] && !inf[2]) return 3;
      return 0;
    };
    bool any_pml = false;
    Real plo[3] = {dom.origin()[0], dom.origin()[1], dom.origin()[2]};
    Real phi[3] = {dom.end()[0], dom.end()[1], dom.end()[2]};
    const int dim0 = dimension(dom);
    for (auto& b : boundaries) {
      auto* p = dynamic_cast<PML*>(b.get());
      if (!p) continue;
      any_pml = true;
      const int a = ax(dirAxis(p->direction()));
      // grid_space_generator.cpp:52-65: a 1D grid takes PML along z only, a
      // 2D grid along x and y only
      if (dim0 == 1 && a != 2) throw XFDTDGridSpaceException("PML has to be in Z direction");
      if (dim0 == 2 && a == 2) throw XFDTDGridSpaceException("PML has to be in X or Y direction");
      if (p->thickness() < 0) continue;
      if (dirNeg(p->direction()))
        plo[a] -= p->thickness() * d[a];
      else
        phi[a] += p->thickness() * d[a];
    }
    if (any_pml)
      dom = Cube(Vector(plo[0], plo[1], plo[2]), Vector(phi[0] - plo[0], phi[1] - plo[1], phi[2] - plo[2]));
    dim = dimension(dom);
    if (dim == 0) throw XFDTDGridSpaceException("GridSpace dimension is undefined");
    // based sizes: (dx, dy, dz) in 3D; GridSpace2D passes based_dx for both
    // in-plane axes and 1 for z (grid_space_2d.cpp:14-23); GridSpace1D
    // (1, 1, dz) (grid_space_1d.cpp:12-19)
    const bool cells[3] = {dim != 1, dim != 1, dim != 2};
    based[0] = dim == 1 ? 1 : d[0];
    based[1] = dim == 1 ? 1 : (dim == 2 ? d[0] : d[1]);
    based[2] = dim == 2 ? 1 : d[2];
    for (int a = 0; a < 3; ++a) {
      if (!cells[a]) continue;
      N[a] = static_cast<std::size_t>(std::round((dom.end()[a] - dom.origin()[a]) / d[a]));
    }
    for (int a = 0; a < 3; ++a)
      if (cells[a] && N[a] <= 1) throw XFDTDGridSpaceException("GridSpace is too small");
    const Real inf = std::numeric_limits<Real>::infinity();
    for (int a = 0; a < 3; ++a) {
      if (!cells[a]) {
        // the one-cell axis (grid_space_2d.cpp:96-106, grid_space_1d.cpp:
        // 64-75): nodes {-inf, inf}, one cell of size 1 centred at 0
        N[a] = 1;
        e_node[a] = {-inf, inf};
        h_node[a] = {0.0};
        e_size[a] = {1.0};
        h_size[a] = {1.0};
        min_d[a] = 1;
        continue;
      }
      // region end: min + (max - min) with max = min + d n in 3D
      // (generateGridSpace3D), min + n d in 1D/2D (generateGridSpace1D/2D)
      const Real mn = dom.origin()[a];
      const Real end = dim == 3 ? mn + ((mn + d[a] * N[a]) - mn) : mn + static_cast<Real>(N[a]) * d[a];
      e_node[a] = linspaceAssign(mn, end, N[a] + 1);
      // GridSpace::correctGridSpaceForOne (grid_space.cpp:517-544) with the
      // axis' based size for the half cells beyond the end nodes
      const auto& e = e_node[a];
      const std::size_t m = N[a];
      h_node[a].resize(m);
      e_size[a].resize(m);
      h_size[a].resize(m + 1);
      for (std::size_t i = 0; i < m; ++i) {
        h_node[a][i] = (e[i] + e[i + 1]) / 2;
        e_size[a][i] = e[i + 1] - e[i];
      }
      Real prev = e[0] - based[a] / 2;
      for (std::size_t i = 0; i < m; ++i) {
        h_size[a][i] = h_node[a][i] - prev;
        prev = h_node[a][i];
      }
      h_size[a][m] = (e[m] + based[a] / 2) - prev;
      min_d[a] = *std::min_element(e_size[a].begin(), e_size[a].end());
    }
    // TimeParam::calculateDt (time_param.cpp:10-21) for the grid's dimension
    // (Simulation::makeTimeParam, simulation.cpp:654-676)
    if (dim == 1)
      dt = cfl / (kC0 * std::sqrt(1.0 / (min_d[2] * min_d[2])));
    else if (dim == 2)
      dt = cfl / (kC0 * std::sqrt(1.0 / (min_d[0] * min_d[0]) + 1.0 / (min_d[1] * min_d[1])));
    else
      dt = cfl / (kC0 * std::sqrt(1.0 / (min_d[0] * min_d[0]) + 1.0 / (min_d[1] * min_d[1]) +
                                  1.0 / (min_d[2] * min_d[2])));
  }

  // ------------------------------------------------------------ material --
  void buildMaterial() {
    for (int c = 0; c < 3; ++c) {  // material_param.cpp:365-393
      std::size_t ee[3], hh[3];
      extent(false, c, N, ee);
      extent(true, c, N, hh);
      eps[c].resize(ee, kEps0);
      se[c].resize(ee, kSigmaZero);
      mu[c].resize(hh, kMu0);
      sm[c].resize(hh, kSigmaZero);
    }
    // Staircase (object.cpp:114-198): a cell takes the object's properties
    // when its centre is inside the shape; objects in add order, the last
    // wins.  Only cells whose centre lies inside the shape's bounding box
    // widened by the tolerance can pass the test, so only those are visited.
    const Real tol = tolerance();
    bool any_dispersive = false;
    for (auto& o : objects) any_dispersive = any_dispersive || o->material().dispersion();
    owner.clear();
    if (any_dispersive) owner.assign(N[0] * N[1] * N[2], -1);
    for (std::size_t oi = 0; oi < objects.size(); ++oi) {
      auto& o = objects[oi];
      if (dynamic_cast<PecPlane*>(o.get())) continue;  // simulation.cpp:611-613
      const auto& p = o->material().emProperty();
      auto box = o->shape().wrappedCube();
      std::size_t lo[3], hi[3];
      for (int a = 0; a < 3; ++a) {
        const auto& h = h_node[a];
        const Real l = box->origin()[a] - tol * 2, u = box->end()[a] + tol * 2;
        lo[a] = static_cast<std::size_t>(std::lower_bound(h.begin(), h.end(), l) - h.begin());
        hi[a] = static_cast<std::size_t>(std::upper_bound(h.begin(), h.end(), u) - h.begin());
        if (std::isinf(box->origin()[a])) lo[a] = 0;
        if (std::isinf(box->end()[a])) hi[a] = N[a];
      }
      if (lo[0] >= hi[0] || lo[1] >= hi[1] || lo[2] >= hi[2]) continue;
      const Shape& shape = o->shape();
      parallelFor(hi[0] - lo[0], [&](std::size_t b, std::size_t e) {
        for (std::size_t i = lo[0] + b; i < lo[0] + e; ++i)
          for (std::size_t j = lo[1]; j < hi[1]; ++j)
            for (std::size_t k = lo[2]; k < hi[2]; ++k) {
              if (!shape.isInside(Vector(h_node[0][i], h_node[1][j], h_node[2][k]), tol)) continue;
              if (!owner.empty()) owner[(i * N[1] + j) * N[2] + k] = static_cast<int32_t>(oi);
              for (int c = 0; c < 3; ++c) {
                eps[c].at(i, j, k) = p.epsilon();
                mu[c].at(i, j, k) = p.mu();
                se[c].at(i, j, k) = p.sigmaE();
                sm[c].at(i, j, k) = p.sigmaM();
              }
            }
      });
    }
    if (has_design)  // the design cells' permittivity replaces the objects'
      for (std::size_t i = 0; i < design_n[0]; ++i)
        for (std::size_t j = 0; j < design_n[1]; ++j)
          for (std::size_t k = 0; k < design_n[2]; ++k) {
            const Real e = design_eps[(i * design_n[1] + j) * design_n[2] + k] * kEps0;
            for (int c = 0; c < 3; ++c) eps[c].at(design_o[0] + i, design_o[1] + j, design_o[2] + k) = e;
          }
    // Edge averaging (calculation_param.cpp:43-168): eps/sigma_e the mean of
    // the four cells around an edge, mu/sigma_m the harmonic mean of the two
    // cells across a face; evaluated from the pre-averaging values.
    for (int c = 0; c < 3; ++c) {
      averageEdges(eps[c], c);
      averageEdges(se[c], c);
      averageFaces(mu[c], c);
      averageFaces(sm[c], c);
    }
    for (auto& o : objects)  // simulation.cpp:629-635
      if (auto* pp = dynamic_cast<PecPlane*>(o.get())) applyPecPlane(*pp);
  }

  // PecPlane::correctMaterialSpace (pec_plane.cpp:12-67): the plane's two
  // tangential E components take the material's sigma_e on the plane.
  void applyPecPlane(const PecPlane& pp) {
    std::size_t o[3], e[3];
    gridBox(pp.shape(), o, e);
    const Real sg = pp.material().emProperty().sigmaE();
    auto fill = [&](Field3& f, std::size_t i0, std::size_t i1, std::size_t j0, std::size_t j1, std::size_t k0,
                    std::size_t k1) {
      for (std::size_t i = i0; i < i1; ++i)
        for (std::size_t j = j0; j < j1; ++j)
          for (std::size_t k = k0; k < k1; ++k) f.at(i, j, k) = sg;
    };
    if (o[0] == e[0]) {
      fill(se[1], o[0], o[0] + 1, o[1], e[1], o[2], e[2] + 1);
      fill(se[2], o[0], o[0] + 1, o[1], e[1] + 1, o[2], e[2]);
    } else if (o[1] == e[1]) {
      fill(se[2], o[0], e[0] + 1, o[1], o[1] + 1, o[2], e[2]);
      fill(se[0], o[0], e[0], o[1], o[1] + 1, o[2], e[2] + 1);
    } else if (o[2] == e[2]) {
      fill(se[0], o[0], e[0], o[1], e[1] + 1, o[2], o[2] + 1);
      fill(se[1], o[0], e[0] + 1, o[1], e[1], o[2], o[2] + 1);
    } else {
      throw XFDTDSimulationException(pp.name() + " PecPlane is not a plane!");
    }
  }

  // In-place averaging from pre-averaging values without a full copy: the
  // new plane i depends only on the old planes i and i-1, so each thread
  // takes a run of planes, saves the old plane just below its run before
  // anyone writes, and then walks its planes downwards keeping a copy of the
  // old plane being overwritten.  `fn(cur, below, q)` returns the new value
  // at in-plane offset q from the old planes.
  template <class Fn>
  void planeUpdate(Field3& f, std::size_t i_start, std::size_t j_start, std::size_t k_start, Fn&& fn) {
    const std::size_t n0 = N[0], plane = f.n[1] * f.n[2];
    if (i_start >= n0) return;
    const std::size_t count = n0 - i_start;
    const std::size_t parts = std::max<std::size_t>(1, std::min<std::size_t>(
                                                            {std::max(1u, std::thread::hardware_concurrency()),
                                                             count, 32}));
    const std::size_t step = (count + parts - 1) / parts;
    std::vector<std::vector<Real>> below(parts);
    for (std::size_t p = 0; p < parts; ++p) {  // save old planes below each run
      const std::size_t lo = i_start + p * step;
      if (lo >= n0 || lo == 0) continue;
      below[p].assign(f.v.data() + (lo - 1) * plane, f.v.data() + lo * plane);
    }
    std::vector<std::thread> th;
    for (std::size_t p = 0; p < parts; ++p) {
      const std::size_t lo = i_start + p * step, hi = std::min(n0, lo + step);
      if (lo >= hi) break;
      th.emplace_back([&, p, lo, hi] {
        std::vector<Real> cur(plane);
        for (std::size_t i = hi; i-- > lo;) {
          Real* out = f.v.data() + i * plane;
          std::copy(out, out + plane, cur.begin());
          const Real* prev = i > lo ? f.v.data() + (i - 1) * plane : (i > 0 ? below[p].data() : nullptr);
          for (std::size_t j = j_start; j < N[1]; ++j)
            for (std::size_t k = k_start; k < N[2]; ++k) {
              const std::size_t q = j * f.n[2] + k;
              out[q] = fn(cur.data(), prev, q);
            }
        }
      });
    }
    for (auto& t : th) t.join();
  }

  // calculation_param.cpp:55-83 (eps) and :105-134 (sigma_e): the mean of
  // the four cells around an edge, in the reference's operand order.
  void averageEdges(Field3& f, int c) {
    const std::size_t r = f.n[2];  // in-plane row stride (j)
    if (c == 0)
      planeUpdate(f, 0, 1, 1, [r](const Real* o, const Real*, std::size_t q) {
        return 0.25 * (o[q] + o[q - r] + o[q - 1] + o[q - r - 1]);
      });
    else if (c == 1)
      planeUpdate(f, 1, 0, 1, [](const Real* o, const Real* b, std::size_t q) {
        return 0.25 * (o[q] + b[q] + o[q - 1] + b[q - 1]);
      });
    else
      planeUpdate(f, 1, 1, 0, [r](const Real* o, const Real* b, std::size_t q) {
        return 0.25 * (o[q] + b[q] + o[q - r] + b[q - r]);
      });
  }

  // calculation_param.cpp:84-104 (mu) and :135-167 (sigma_m): the harmonic
  // mean 2ab/(a+b) of the two cells across a face.
  void averageFaces(Field3& f, int c) {
    const std::size_t r = f.n[2];
    auto h = [](Real a, Real q) { return 2 * a * q / (a + q); };
    if (c == 0)
      planeUpdate(f, 1, 0, 0, [h](const Real* o, const Real* b, std::size_t q) { return h(o[q], b[q]); });
    else if (c == 1)
      planeUpdate(f, 0, 1, 0, [h, r](const Real* o, const Real*, std::size_t q) { return h(o[q], o[q - r]); });
    else
      planeUpdate(f, 0, 0, 1, [h](const Real* o, const Real*, std::size_t q) { return h(o[q], o[q - 1]); });
  }

  // -------------------------------------------------------- coefficients --
  void buildCoefficients() {  // calculation_param.cpp:170-264
    for (int h = 0; h < 2; ++h)
      for (int c = 0; c < 3; ++c) {
        std::size_t ext[3];
        extent(h, c, N, ext);
        Field3& cs = coef[h * 9 + c * 3];
        Field3& ca = coef[h * 9 + c * 3 + 1];
        Field3& cb = coef[h * 9 + c * 3 + 2];
        cs.allocate(ext);  // every element is written below
        ca.allocate(ext);
        cb.allocate(ext);
        // E_c: e_size on c, h_size elsewhere; H_c the opposite
        const std::vector<Real>* sz[3];
        for (int a = 0; a < 3; ++a) sz[a] = (h != 0) == (a == c) ? &h_size[a] : &e_size[a];
        const int A = (c + 1) % 3, B = (c + 2) % 3;
        const Field3& m = h ? mu[c] : eps[c];
        const Field3& s = h ? sm[c] : se[c];
        const Real dtv = dt;
        parallelFor(ext[0], [&](std::size_t lo, std::size_t hi) {
          for (std::size_t i = lo; i < hi; ++i)
            for (std::size_t j = 0; j < ext[1]; ++j)
              for (std::size_t k = 0; k < ext[2]; ++k) {
                const std::size_t id[3] = {i, j, k};
                // xt::meshgrid broadcasts a one-element size array (the
                // one-cell axis of a 1D/2D grid) over a component extent of 2
                const Real da = (*sz[A])[sz[A]->size() == 1 ? 0 : id[A]];
                const Real db = (*sz[B])[sz[B]->size() == 1 ? 0 : id[B]];
                const std::size_t p = (i * ext[1] + j) * ext[2] + k;
                const Real e = m.v[p], sg = s.v[p];
                cs.v[p] = (2 * e - dtv * sg) / (2 * e + dtv * sg);
                if (!h) {
                  ca.v[p] = -(2 * dtv / db) / (2 * e + sg * dtv);
                  cb.v[p] = (2 * dtv / da) / (2 * e + sg * dtv);
                } else {
                  ca.v[p] = (2 * dtv / db) / (2 * e + sg * dtv);
                  cb.v[p] = -(2 * dtv / da) / (2 * e + sg * dtv);
                }
              }
        });
      }
  }

  // ------------------------------------------------------- lumped elements --
  // LumpedElement ranges (lumped_element.h:239-264) and the meshgrid of cell
  // sizes over them; the patch of resistor.cpp:85-127,
  // voltage_source.cpp:109-161 and current_source.cpp:111-161.
  void buildLumped(LumpedElement& le) {
    std::size_t o[3], e[3];
    gridBox(le.shape(), o, e);
    const int m = ax(le.xyz()), A = (m + 1) % 3, B = (m + 2) % 3;
    std::size_t lo[3], cnt[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = o[a];
      cnt[a] = a == m ? e[a] - o[a] : e[a] - o[a] + 1;
    }
    const Real gmain = static_cast<Real>(e[m] - o[m]);
    const std::size_t ga = e[A] - o[A] + 1, gb = e[B] - o[B] + 1;
    const std::size_t gc = e[m] - o[m];
    (void)gmain;
    auto size_of = [&](int a, std::size_t idx) { return a == m ? e_size[a][idx] : h_size[a][idx]; };

    if (auto* cap = dynamic_cast<Capacitor*>(&le)) {
      // capacitor.cpp:24-25, 58, 73-81: C' = C nc / (nc nb), beta =
      // 2 C' dc / (da db); the E coefficients with +beta in numerator and
      // denominator
      const Real cf = cap->capacitance() * gc / (gc * gb);
      Field3& cece = coef[m * 3];
      Field3& cecha = coef[m * 3 + 1];
      Field3& cechb = coef[m * 3 + 2];
      for (std::size_t i = 0; i < cnt[0]; ++i)
        for (std::size_t j = 0; j < cnt[1]; ++j)
          for (std::size_t k = 0; k < cnt[2]; ++k) {
            const std::size_t g[3] = {lo[0] + i, lo[1] + j, lo[2] + k};
            const Real da = size_of(A, g[A]), db = size_of(B, g[B]), dc = size_of(m, g[m]);
            const Real beta = 2 * cf * dc / (da * db);
            const Real e0 = eps[m].at(g[0], g[1], g[2]), s0 = se[m].at(g[0], g[1], g[2]);
            cece.at(g[0], g[1], g[2]) = (2 * e0 - dt * s0 + beta) / (2 * e0 + dt * s0 + beta);
            cecha.at(g[0], g[1], g[2]) = -2 * dt / ((2 * e0 + dt * s0 + beta) * db);
            cechb.at(g[0], g[1], g[2]) = 2 * dt / ((2 * e0 + dt * s0 + beta) * da);
          }
      return;
    }
    if (auto* ind = dynamic_cast<Inductor*>(&le)) {
      // inductor.cpp:29-31, 69, 84: L' = L na nb / nc, cjcec = dt dc /
      // (L' da db), cecjc = -2 dt / (2 eps + dt sigma); no coefficient patch
      const Real lf = ind->inductance() * ga * gb / gc;
      LumpedData ld;
      const std::size_t vol = cnt[0] * cnt[1] * cnt[2];
      ld.coff.resize(vol);
      ld.coff_j.resize(vol);
      for (std::size_t i = 0; i < cnt[0]; ++i)
        for (std::size_t j = 0; j < cnt[1]; ++j)
          for (std::size_t k = 0; k < cnt[2]; ++k) {
            const std::size_t g[3] = {lo[0] + i, lo[1] + j, lo[2] + k};
            const Real da = size_of(A, g[A]), db = size_of(B, g[B]), dc = size_of(m, g[m]);
            const std::size_t q = (i * cnt[1] + j) * cnt[2] + k;
            ld.coff_j[q] = dt * dc / (lf * da * db);
            ld.coff[q] = -(2 * dt) / (2 * eps[m].at(g[0], g[1], g[2]) + dt * se[m].at(g[0], g[1], g[2]));
          }
      ld.d.kind = XFD_LUMPED_INDUCTOR;
      ld.d.component = m;
      for (int a = 0; a < 3; ++a) {
        ld.d.origin[a] = static_cast<int32_t>(lo[a]);
        ld.d.size[a] = static_cast<int32_t>(cnt[a]);
      }
      lumped.push_back(std::move(ld));
      return;
    }
    auto* res = dynamic_cast<Resistor*>(&le);
    auto* vs = dynamic_cast<VoltageSource*>(&le);
    auto* cs = dynamic_cast<CurrentSource*>(&le);
    if (!res && !vs && !cs) throw XFDTDSimulationException("unsupported lumped element " + le.name());
    const Real R = res ? res->resistance() : vs ? vs->resistance() : cs->resistance();
    const Real rf = R * ga * gb / gc;
    Real factor = 0;
    if (vs) factor = vs->waveform().amplitude() / gc;
    if (cs) factor = cs->waveform().amplitude() / (ga * gb);
    if ((vs && dirNeg(vs->direction())) || (cs && dirNeg(cs->direction()))) factor *= -1;

    Field3& cece = coef[m * 3];
    Field3& cecha = coef[m * 3 + 1];
    Field3& cechb = coef[m * 3 + 2];
    const Field3& ep = eps[m];
    const Field3& sg = se[m];
    LumpedData ld;
    ld.coff.resize(cnt[0] * cnt[1] * cnt[2]);
    for (std::size_t i = 0; i < cnt[0]; ++i)
      for (std::size_t j = 0; j < cnt[1]; ++j)
        for (std::size_t k = 0; k < cnt[2]; ++k) {
          const std::size_t g[3] = {lo[0] + i, lo[1] + j, lo[2] + k};
          const Real da = size_of(A, g[A]), db = size_of(B, g[B]), dc = size_of(m, g[m]);
          Real beta, alpha = 0;
          if (vs) {
            alpha = da * db * rf;
            beta = dt * dc / alpha;
          } else {
            beta = dt * dc / (rf * da * db);
          }
          const Real e0 = ep.at(g[0], g[1], g[2]), s0 = sg.at(g[0], g[1], g[2]);
          cece.at(g[0], g[1], g[2]) = (2 * e0 - dt * s0 - beta) / (2 * e0 + dt * s0 + beta);
          cecha.at(g[0], g[1], g[2]) = -2 * dt / ((2 * e0 + dt * s0 + beta) * db);
          cechb.at(g[0], g[1], g[2]) = 2 * dt / ((2 * e0 + dt * s0 + beta) * da);
          const std::size_t q = (i * cnt[1] + j) * cnt[2] + k;
          if (vs) ld.coff[q] = -2 * dt / ((2 * e0 + dt * s0 + beta) * alpha) * factor;
          if (cs) ld.coff[q] = -2 * dt / ((2 * e0 + dt * s0 + beta) * (db * da)) * factor;
        }
    if (res) return;
    ld.d.component = m;
    for (int a = 0; a < 3; ++a) {
      ld.d.origin[a] = static_cast<int32_t>(lo[a]);
      ld.d.size[a] = static_cast<int32_t>(cnt[a]);
    }
    Waveform& w = vs ? vs->waveform() : cs->waveform();
    w.init(hTime());  // voltage_source.cpp:163-167
    ld.wave = w.value();
    lumped.push_back(std::move(ld));
  }

  std::vector<Real> eTime() const { return linspaceAssign(1 * dt, steps * dt, steps); }
  std::vector<Real> hTime() const { return linspaceAssign(0.5 * dt, (steps + 0.5) * dt, steps); }

  // ----------------------------------------------------------------- NF2FF --
  // NFFFT::init + FDPlaneData::initDFT (nffft.cpp:73-223,
  // nffft_fd_data.cpp:65-76): surfaces and phase tables of each NF2FF.
  void buildNf2ff() {
    nf2ff_d.clear();
    for (auto& n : nf2ffs) {
      const std::size_t dist[3] = {n->distance(0), n->distance(1), n->distance(2)};
      auto& D = n->data();
      D.setup(dist, N, e_node, h_node, e_size, n->frequencies(), dt, steps);
      xfd_nf2ff_desc d{};
      for (int s6 = 0; s6 < 6; ++s6)
        for (int a = 0; a < 3; ++a) {
          // unused surfaces are passed as an empty box at a legal position
          const std::size_t lo = D.used[s6] ? D.box_lo[s6][a] : 1;
          const std::size_t hi = D.used[s6] ? D.box_hi[s6][a] : 1;
          d.lo[s6][a] = static_cast<int32_t>(lo);
          d.hi[s6][a] = static_cast<int32_t>(hi);
        }
      d.nf = static_cast<int32_t>(n->freqCount());
      d.repeat = n->referenceRepeat() ? d.nf : 1;
      d.nt = static_cast<int64_t>(D.nt);
      nf2ff_d.push_back(d);
    }
    tdnf_d.clear();
    for (auto& n : nf2ff_tds) {  // time-domain NF2FFs
      const std::size_t dist[3] = {n->distance(0), n->distance(1), n->distance(2)};
      auto& D = n->data();
      D.setup(dist, N, e_node, h_node, e_size, n->theta(), n->phi(), dt, steps);
      xfd_tdnf_desc d{};
      for (int s6 = 0; s6 < 6; ++s6)
        for (int a = 0; a < 3; ++a) {
          d.lo[s6][a] = static_cast<int32_t>(D.used[s6] ? D.box_lo[s6][a] : 1);
          d.hi[s6][a] = static_cast<int32_t>(D.used[s6] ? D.box_hi[s6][a] : 1);
        }
      d.num = static_cast<int64_t>(D.num);
      tdnf_d.push_back(d);
    }
  }

  // ------------------------------------------------------------------- ADE --
  // Simulation::buildDispersiveSpace (simulation.cpp:745-829) and
  // Object::handleDispersion (object.cpp:52-98): the scheme follows the kinds
  // of dispersive material present; each dispersive object's cells get the
  // ADE E coefficients of its material (debye_ade_method.cpp:39-122,
  // drude_ade_method.cpp:48-133, m_lor_ade_method.cpp:66-159).
  void buildDispersive() {
    has_ade = false;
    ade_coeff.clear();
    ade_cell.clear();
    std::size_t poles = 0;
    bool is_drude = false, is_debye = false, is_lor = false;
    for (auto& o : objects) {
      if (dynamic_cast<PecPlane*>(o.get())) continue;
      auto* m = dynamic_cast<const LinearDispersiveMaterial*>(&o->material());
      if (!m) continue;
      poles = std::max(poles, m->numPoles());
      if (m->kind() == LinearDispersiveMaterial::Kind::Drude) is_drude = true;
      else if (m->kind() == LinearDispersiveMaterial::Kind::Debye) is_debye = true;
      else is_lor = true;
    }
    if (poles == 0) return;
    int scheme = XFD_ADE_MLORENTZ;
    if (is_drude && !is_debye && !is_lor) scheme = XFD_ADE_DRUDE;
    else if (!is_drude && is_debye && !is_lor) scheme = XFD_ADE_DEBYE;
    const std::size_t stride = 2 + 7 * poles;
    ade_cell.assign(N[0] * N[1] * N[2], int16_t(-1));
    int16_t nmat = 0;
    for (std::size_t oi = 0; oi < objects.size(); ++oi) {
      auto* m = dynamic_cast<const LinearDispersiveMaterial*>(&objects[oi]->material());
      if (!m || dynamic_cast<PecPlane*>(objects[oi].get())) continue;
      if (nmat == 32767) throw XFDTDSimulationException("too many dispersive objects");
      std::vector<Real> row(stride, 0.0);
      Real cece = 0, num_a = 0, num_b = 0, den = 1;  // cecha = num_a / (den db), cechb = num_b / (den da)
      const Real sigma = m->emProperty().sigmaE(), ei = m->epsilonInf();
      const std::size_t np = m->numPoles();
      if (scheme != XFD_ADE_MLORENTZ) {
        if (np != poles)
          throw XFDTDSimulationException("Debye/Drude storage needs every material to have the storage's pole count");
        std::vector<Real> kk(np), beta(np);
        for (std::size_t p = 0; p < np; ++p) {
          if (scheme == XFD_ADE_DEBYE) {
            const Real tau = m->b1()[p], de = m->a0()[p];
            kk[p] = (2 * tau - dt) / (2 * tau + dt);
            beta[p] = 2 * kEps0 * de * dt / (2 * tau + dt);
          } else {
            const Real g = m->b1()[p], wp = std::sqrt(m->a0()[p]);
            kk[p] = (1 - g * dt / 2) / (1 + g * dt / 2);
            beta[p] = (kEps0 * wp * wp * dt / 2) / (1 + g * dt / 2);
          }
        }
        Real sum_beta = 0;
        for (std::size_t p = 0; p < np; ++p) sum_beta += beta[p];
        for (std::size_t p = 0; p < np; ++p) {
          Real* r = row.data() + 2 + 7 * p;
          r[1] = scheme == XFD_ADE_DEBYE ? beta[p] / dt : beta[p];
          r[3] = kk[p];
          r[5] = 0.5 * (1 + kk[p]);
        }
        Real a, b;
        if (scheme == XFD_ADE_DEBYE) {
          a = (2 * kEps0 * ei + sum_beta - dt * sigma) / (2 * kEps0 * ei + sum_beta + dt * sigma);
          b = (2 * dt) / (2 * kEps0 * ei + sum_beta + dt * sigma);
        } else {
          a = (2 * kEps0 * ei - dt * sum_beta - dt * sigma) / (2 * kEps0 * ei - dt * sum_beta + dt * sigma);
          b = (2 * dt) / (2 * kEps0 * ei - dt * sum_beta + dt * sigma);
        }
        row[0] = -b;
        cece = a;
        num_a = -b;
        num_b = b;
      } else {
        std::vector<Real> A(np), B(np), C(np), D(np), E(np), F(np);
        for (std::size_t p = 0; p < np; ++p) {
          const Real a0 = m->a0()[p], a1 = m->a1()[p], b0 = m->b0()[p], b1 = m->b1()[p], b2 = m->b2()[p];
          A[p] = (a1 * kEps0 / (dt * dt)) + (0.5 * a0 * kEps0 / dt);
          B[p] = (-2 * a1 * kEps0) / (dt * dt);
          C[p] = (a1 * kEps0 / (dt * dt)) - (0.5 * a0 * kEps0 / dt);
          D[p] = (2 * b2 / (dt * dt)) - b0;
          E[p] = (-b2 / (dt * dt)) + (b1 / (2 * dt));
          F[p] = (b2 / (dt * dt)) + (b1 / (2 * dt));
        }
        for (std::size_t p = 0; p < np; ++p) {
          Real* r = row.data() + 2 + 7 * p;
          r[0] = A[p] / F[p];
          r[1] = B[p] / F[p];
          r[2] = C[p] / F[p];
          r[3] = D[p] / F[p];
          r[4] = E[p] / F[p];
          r[5] = 0.5 * (1 + D[p] / F[p]);
          r[6] = 0.5 * (E[p] / F[p]);
        }
        Real sum_a_f = 0, sum_b_f = 0, sum_c_f = 0;
        for (std::size_t p = 0; p < np; ++p) sum_a_f += A[p] / F[p];
        for (std::size_t p = 0; p < np; ++p) sum_b_f += B[p] / F[p];
        for (std::size_t p = 0; p < np; ++p) sum_c_f += C[p] / F[p];
        const Real temp = (ei * kEps0) / dt + 0.5 * sum_a_f + 0.5 * sigma;
        row[0] = -1 / temp;
        row[1] = -0.5 * sum_c_f / temp;
        cece = ((ei * kEps0) / dt - 0.5 * sum_b_f - 0.5 * sigma) / temp;
        num_a = -1;
        num_b = 1;
        den = temp;
      }
      const bool lor = scheme == XFD_ADE_MLORENTZ;
      const int16_t id = nmat++;
      ade_coeff.insert(ade_coeff.end(), row.begin(), row.end());
      const int32_t me = static_cast<int32_t>(oi);
      parallelFor(N[0], [&](std::size_t lo, std::size_t hi) {
        for (std::size_t i = lo; i < hi; ++i)
          for (std::size_t j = 0; j < N[1]; ++j)
            for (std::size_t k = 0; k < N[2]; ++k) {
              const std::size_t cell = (i * N[1] + j) * N[2] + k;
              if (owner[cell] != me) continue;
              ade_cell[cell] = id;
              const Real d[3] = {h_size[0][i], h_size[1][j], h_size[2][k]};
              for (int c = 0; c < 3; ++c) {
                const Real da = d[(c + 1) % 3], db = d[(c + 2) % 3];
                coef[3 * c].at(i, j, k) = cece;
                coef[3 * c + 1].at(i, j, k) = lor ? num_a / (den * db) : num_a / db;
                coef[3 * c + 2].at(i, j, k) = lor ? num_b / (den * da) : num_b / da;
              }
            }
      });
    }
    ade_d.scheme = scheme;
    ade_d.poles = static_cast<int32_t>(poles);
    ade_d.n_materials = nmat;
    has_ade = true;
  }

  // ------------------------------------------------------------------ CPML --
  void buildPml(const PML& p) {
    // PML::init (pml.cpp:98-196): E planes start one inside on the negative
    // side; calRecursiveConvolutionCoeff (:526-633); correctCoefficient*
    // (:286-524): c_psi = coefficient * size, then coefficient /= kappa.
    const int c = ax(dirAxis(p.direction())), A = (c + 1) % 3, B = (c + 2) % 3;
    const std::size_t n = static_cast<std::size_t>(std::abs(p.thickness()));
    if (n == 0 || n > N[c]) throw XFDTDException("PML thickness does not fit the grid");
    const bool neg = dirNeg(p.direction());
    const std::size_t e0 = neg ? 1 : N[c] - n, h0 = neg ? 0 : N[c] - n;
    if (e0 + n > N[c]) throw XFDTDException("PML slab leaves the grid");
    SlabData s;
    s.d.axis = c;
    s.d.n = static_cast<int32_t>(n);
    s.d.e_start = static_cast<int32_t>(e0);
    s.d.h_start = static_cast<int32_t>(h0);
    std::vector<Real> hs(h_size[c].begin() + e0, h_size[c].begin() + e0 + n);
    std::vector<Real> es(e_size[c].begin() + h0, e_size[c].begin() + h0 + n);
    const Real ord = static_cast<Real>(p.order());
    auto depth = [&](const std::vector<Real>& sz, Real frac) {
      std::vector<Real> r(n);
      Real acc = 0;
      for (std::size_t i = 0; i < n; ++i) {
        const Real inner = sz[i] - frac * sz[i];
        r[i] = i == 0 ? inner : acc + inner;
        acc += sz[i];
      }
      if (neg) std::reverse(r.begin(), r.end());
      for (auto& x : r) x = x / acc;
      return r;
    };
    auto smax = [&](Real dl) { return p.sigmaRatio() * (p.order() + 1) / (150 * kPi * dl); };
    const Real amin = p.alphaMin(), amax = p.alphaMax(), kmax = p.kappaMax();
    std::vector<Real> ke(n), kh(n);
    s.ae.resize(n);
    s.be.resize(n);
    s.ah.resize(n);
    s.bh.resize(n);
    {
      const Real sm_e = smax(*std::min_element(hs.begin(), hs.end()));
      const auto rho = depth(hs, 0.75);
      for (std::size_t i = 0; i < n; ++i) {
        const Real pw = std::pow(rho[i], ord);
        const Real sg = sm_e * pw, al = amin + (amax - amin) * (1 - rho[i]), ka = 1 + (kmax - 1) * pw;
        ke[i] = ka;
        s.be[i] = std::exp((-dt / kEps0) * (sg / ka + al));
        s.ae[i] = (s.be[i] - 1) * sg / ((sg + ka * al) * ka * hs[i]);
      }
    }
    {
      const Real e2m = kMu0 / kEps0;
      const Real sm_m = smax(*std::min_element(es.begin(), es.end())) * e2m;
      const auto rho = depth(es, 0.25);
      for (std::size_t i = 0; i < n; ++i) {
        const Real pw = std::pow(rho[i], ord);
        const Real sg = sm_m * pw, al = (amin + (amax - amin) * (1 - rho[i])) * e2m,
                   ka = 1 + (kmax - 1) * pw;
        kh[i] = ka;
        s.bh[i] = std::exp((-dt / kMu0) * (sg / ka + al));
        s.ah[i] = (s.bh[i] - 1) * sg / ((sg + ka * al) * ka * es[i]);
      }
    }
    // Patched coefficients and psi extents: E_a via its dual-B coefficient
    // (extended along B), E_b via dual-A (extended along A); H_a extended
    // along A, H_b along B.  Slot of the coefficient of component q for its
    // tangential-A / tangential-B dual is 3q+1 / 3q+2.
    struct P {
      int slot;
      int ext_axis;
      std::size_t start;
      const std::vector<Real>* size;
      const std::vector<Real>* kappa;
    };
    // E_a's derivative along c (its tangential-B axis) acts on its
    // tangential-A dual, whose coefficient is slot 3A+1 (x slab: ceyhz);
    // E_b's acts on its tangential-B dual, slot 3B+2 (x slab: cezhy).
    const P parts[4] = {{3 * A + 1, B, e0, &hs, &ke},
                        {3 * B + 2, A, e0, &hs, &ke},
                        {9 + 3 * A + 1, A, h0, &es, &kh},
                        {9 + 3 * B + 2, B, h0, &es, &kh}};
    for (int q = 0; q < 4; ++q) {
      const P& pp = parts[q];
      std::size_t ex[3];
      ex[c] = n;
      ex[A] = N[A];
      ex[B] = N[B];
      ex[pp.ext_axis] += 1;
      auto& cpsi = s.cpsi[q];
      cpsi.resize(ex[0] * ex[1] * ex[2]);
      Field3& cf = coef[pp.slot];
      parallelFor(ex[0], [&](std::size_t lo, std::size_t hi) {
        for (std::size_t x = lo; x < hi; ++x)
          for (std::size_t y = 0; y < ex[1]; ++y)
            for (std::size_t z = 0; z < ex[2]; ++z) {
              std::size_t g[3] = {x, y, z};
              const std::size_t layer = g[c];
              g[c] += pp.start;
              Real& v = cf.at(g[0], g[1], g[2]);
              cpsi[(x * ex[1] + y) * ex[2] + z] = v * (*pp.size)[layer];
              v = v / (*pp.kappa)[layer];
            }
      });
    }
    slabs.push_back(std::move(s));
  }

  // ------------------------------------------------------------------ TFSF --
  void buildTfsf(TFSF& t) {
    // tfsf.cpp:14-30 (angles), :231-254 (transforms), :118-175 (box, aux
    // line, 1D coefficients), :256-339 (projections), :36-78 (incident
    // tables from the 1D auxiliary FDTD on eTime - dt).
    TfsfData td;
    const Real st = std::sin(t.theta()), ct = std::cos(t.theta());
    const Real sp = std::sin(t.phi()), cp = std::cos(t.phi());
    const Real ss = std::sin(t.psi()), cs = std::cos(t.psi());
    const Real k[3] = {st * cp, st * sp, ct};
    const Real va[3] = {-sp, ct * cp, st * cp}, vb[3] = {cp, ct * sp, st * sp}, vc[3] = {0, -st, ct};
    const Real ke[3] = {ss, cs, 0};
    auto dot = [](const Real* a, const Real* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    const Real te[3] = {dot(va, ke), dot(vb, ke), dot(vc, ke)};
    const Real th[3] = {k[1] * te[2] - k[2] * te[1], k[2] * te[0] - k[0] * te[2], k[0] * te[1] - k[1] * te[0]};
    const std::size_t off[3] = {t.x(), t.y(), t.z()};
    std::size_t org[3], sz[3];
    for (int a = 0; a < 3; ++a) {
      if (2 * off[a] >= N[a]) throw XFDTDException("TFSF box is empty");
      org[a] = off[a];
      sz[a] = N[a] - 2 * off[a];
    }
    const Real rd = 1 / std::sqrt(std::pow(st, 4) * (std::pow(cp, 4) + std::pow(sp, 4)) + std::pow(ct, 4));
    const std::size_t aux = static_cast<std::size_t>(
        std::ceil(rd * (std::sqrt(std::pow(static_cast<Real>(sz[0]), 2) + std::pow(static_cast<Real>(sz[1]), 2) +
                                  std::pow(static_cast<Real>(sz[2]), 2)))) +
        4 + 1);
    for (int a = 0; a < 3; ++a) {
      const std::size_t inj = k[a] >= 0 ? org[a] : org[a] + sz[a];
      const Real inc = static_cast<Real>(inj) - (2 * k[a] / rd);
      auto& pi = td.proj[a];
      auto& ph = td.proj[3 + a];
      pi.resize(sz[a] + 1);
      ph.resize(sz[a] + 2);
      for (std::size_t i = 0; i < pi.size(); ++i) pi[i] = (i + org[a] - inc) * k[a] * rd;
      for (std::size_t i = 0; i < ph.size(); ++i) ph[i] = (i + org[a] - inc - 0.5) * k[a] * rd;
    }
    // tfsf.cpp:153-161: the auxiliary line's cell is the based dz (1D and
    // 3D) or based dx (2D) over ratio_delta
    const Real sdl = (dim == 2 ? based[0] : based[2]) / rd;
    const Real ceie = 1, chih = 1;
    const Real ceih = -(dt / (kEps0 * sdl)), chie = -(dt / (kMu0 * sdl));
    const Real abc0 = (kC0 * dt - sdl) / (kC0 * dt + sdl), abc1 = 2 * sdl / (kC0 * dt + sdl);

    std::vector<Real> tm = eTime();
    for (auto& x : tm) x = x - dt;
    t.waveform().init(tm);
    const auto& src = t.waveform().value();
    const std::size_t nt = steps, nl = aux - 1;
    td.e_inc.assign(nt * aux, 0);
    td.h_inc.assign(nt * nl, 0);
    auto E = [&](std::size_t r, std::size_t l) -> Real& { return td.e_inc[r * aux + l]; };
    auto H = [&](std::size_t r, std::size_t l) -> Real& { return td.h_inc[r * nl + l]; };
    if (nt > 0) {
      E(0, 0) = src[0];
      for (std::size_t l = 1; l < nl; ++l) H(0, l) = chih * H(0, l) + chie * (E(0, l + 1) - E(0, l));
      for (std::size_t r = 1; r < nt; ++r) {
        if (nl > 1) E(r, 0) = src[r];
        for (std::size_t l = 1; l < nl; ++l) E(r, l) = ceie * E(r - 1, l) + ceih * (H(r - 1, l) - H(r - 1, l - 1));
        const Real a1 = r < 2 ? 0.0 : E(r - 2, nl - 1);
        const Real a2 = r < 2 ? 0.0 : E(r - 2, nl);
        E(r, nl) = -a1 + abc0 * (E(r, nl - 1) + a2) + abc1 * (E(r - 1, nl) + E(r - 1, nl - 1));
        for (std::size_t l = 0; l < nl; ++l) H(r, l) = chih * H(r - 1, l) + chie * (E(r, l + 1) - E(r, l));
      }
    }
    for (int a = 0; a < 3; ++a) {
      td.d.origin[a] = static_cast<int32_t>(org[a]);
      td.d.size[a] = static_cast<int32_t>(sz[a]);
      td.d.transform_e[a] = te[a];
      td.d.transform_h[a] = th[a];
      td.d.ca[a] = dt / (kEps0 * based[a]);  // tfsf.cpp:177-205
      td.d.cb[a] = dt / (kMu0 * based[a]);
    }
    td.d.nt = static_cast<int64_t>(nt);
    td.d.aux = static_cast<int64_t>(aux);
    td.d.kind = static_cast<int32_t>(t.kind());
    auto* t1 = dynamic_cast<TFSF1D*>(&t);
    td.d.forward = t1 && t1->forward() ? 1 : 0;
    tfsf.push_back(std::move(td));
  }

  // -------------------------------------------------------------- monitors --
  void monitorBox(const Shape& s, int32_t o[3], int32_t e[3]) const {
    // Monitor::defaultInit (monitor.cpp:103-127) and the one-cell shift of a
    // box starting at index 0 (field_monitor.cpp:24-67): on x and y unless
    // the grid is 1D, on z only in 3D (field_monitor.cpp:29-38,
    // field_time_monitor.cpp:43-52)
    std::size_t so[3], se_[3];
    gridBox(s, so, se_);
    for (int a = 0; a < 3; ++a) {
      const bool shift = a < 2 ? dim != 1 : dim == 3;
      o[a] = static_cast<int32_t>(so[a] == 0 && shift ? 1 : so[a]);
      e[a] = static_cast<int32_t>(se_[a]);
    }
  }

  // VoltageMonitor::init / update (voltage_monitor.cpp:21-227) for one
  // domain: E along the direction's axis c over the cube's range, the a and
  // b ranges extended by one and kept off the grid faces, each term weighted
  // by -dc / (na nb) (negated for a negative direction), summed in i, j, k
  // order.
  LinePlan voltagePlan(VoltageMonitor& vm) {
    LinePlan p;
    p.monitor = &vm;
    std::size_t o[3], e[3];
    gridBox(vm.cube(), o, e);
    const int dir = static_cast<int>(vm.direction()), c = dir / 2, a = (c + 1) % 3, b = (c + 2) % 3;
    std::size_t r[3][2];
    for (int q : {a, b}) {
      r[q][0] = o[q];
      r[q][1] = e[q] + 1;
      if (r[q][0] <= 0) r[q][0] = 1;
      if (N[q] <= r[q][1]) r[q][1] = N[q] - 1;
    }
    r[c][0] = o[c];
    r[c][1] = e[c];
    for (int q = 0; q < 3; ++q)
      if (r[q][0] >= r[q][1]) throw XFDTDException(vm.name() + " globalTask is not valid");
    const std::size_t cnt = (r[a][1] - r[a][0]) * (r[b][1] - r[b][0]);
    for (std::size_t i = r[0][0]; i < r[0][1]; ++i)
      for (std::size_t j = r[1][0]; j < r[1][1]; ++j)
        for (std::size_t k = r[2][0]; k < r[2][1]; ++k) {
          const std::size_t pc = c == 0 ? i : (c == 1 ? j : k);
          Real w = -h_size[c][pc] / static_cast<Real>(cnt);
          if (dir % 2 == 0) w *= -1.0;
          p.terms.insert(p.terms.end(), {c, static_cast<int32_t>(i), static_cast<int32_t>(j), static_cast<int32_t>(k), 0});
          p.weights.push_back(w);
        }
    p.scale = 1.0;
    vm.time() = eTime();
    vm.data().assign(steps, 0);
    return p;
  }

  // CurrentMonitor::init / update (current_monitor.cpp:18-385) for one
  // domain: the H samples around the cube's last plane along the direction,
  // weighted by +-e_size; the two loop sides are the two groups.
  LinePlan currentPlan(CurrentMonitor& cm) {
    LinePlan p;
    p.monitor = &cm;
    std::size_t o[3], e[3];
    gridBox(cm.cube(), o, e);
    const int dir = static_cast<int>(cm.direction()), c = dir / 2, a = (c + 1) % 3, b = (c + 2) % 3;
    if (e[c] == 0) throw XFDTDException("CurrentMonitor::init(): global_c_e is 0");
    if (o[a] == 0 || e[a] == N[a]) throw XFDTDException("CurrentMonitor::init(): global x range is invalid");
    if (o[b] == 0 || e[b] == N[b]) throw XFDTDException("CurrentMonitor::init(): global y range is invalid");
    // make_ranges (:138-169): the side along one axis exists when the other
    // axis' face lies strictly inside the grid
    auto ranges = [](std::size_t b_start, std::size_t b_end, std::size_t b_upper, std::size_t a_upper,
                     std::size_t st, std::size_t en, std::size_t rn[2], std::size_t rp[2]) {
      if (st <= 0) st = 1;
      en = en + 1;
      if (a_upper <= en) en = a_upper - 1;
      rn[0] = rn[1] = rp[0] = rp[1] = 0;
      if (b_start != 0 && b_start != b_upper - 1 && b_start != b_upper) {
        rn[0] = st;
        rn[1] = en;
      }
      if (b_end != b_upper && b_end != b_upper - 1 && b_end != 0) {
        rp[0] = st;
        rp[1] = en;
      }
    };
    std::size_t ha_n[2], ha_p[2], hb_n[2], hb_p[2];
    ranges(o[b], e[b], N[b], N[a], o[a], e[a], ha_n, ha_p);
    ranges(o[a], e[a], N[a], N[b], o[b], e[b], hb_n, hb_p);
    auto valid = [](const std::size_t* q) { return q[0] < q[1]; };
    // da / db: e_size along a / b from the first valid side (the
    // reference's views)
    const std::size_t* ra = valid(ha_n) ? ha_n : ha_p;
    const std::size_t* rb = valid(hb_n) ? hb_n : hb_p;
    auto da = [&](std::size_t x) { return e_size[a][ra[0] + x]; };
    auto db = [&](std::size_t x) { return e_size[b][rb[0] + x]; };
    // node ranges: a, b the cube's; c its last plane
    std::size_t lo[3], hi[3];
    lo[a] = o[a];
    hi[a] = e[a];
    lo[b] = o[b];
    hi[b] = e[b];
    lo[c] = e[c] - 1;
    hi[c] = e[c];
    auto add = [&](int field, std::size_t i, std::size_t j, std::size_t k, Real w, int group) {
      p.terms.insert(p.terms.end(), {field, static_cast<int32_t>(i), static_cast<int32_t>(j), static_cast<int32_t>(k), group});
      p.weights.push_back(w);
    };
    const std::size_t is = lo[0], ie = hi[0], js = lo[1], je = hi[1], ks = lo[2], ke = hi[2];
    if (c == 0) {  // integral_y (group 0) then integral_z (group 1)
      if (valid(ha_n)) for (std::size_t j = ha_n[0]; j < ha_n[1]; ++j) add(4, ie - 1, j, ks - 1, da(j - ha_n[0]), 0);
      if (valid(ha_p)) for (std::size_t j = ha_p[0]; j < ha_p[1]; ++j) add(4, ie - 1, j, ke, -da(j - ha_p[0]), 0);
      if (valid(hb_n)) for (std::size_t k = hb_n[0]; k < hb_n[1]; ++k) add(5, ie - 1, js - 1, k, -db(k - hb_n[0]), 1);
      if (valid(hb_p)) for (std::size_t k = hb_p[0]; k < hb_p[1]; ++k) add(5, ie - 1, je, k, db(k - hb_p[0]), 1);
    } else if (c == 1) {  // integral_z (group 0) then integral_x (group 1)
      if (valid(ha_n)) for (std::size_t k = ha_n[0]; k < ha_n[1]; ++k) add(5, is - 1, je - 1, k, da(k - ha_n[0]), 0);
      if (valid(ha_p)) for (std::size_t k = ha_p[0]; k < ha_p[1]; ++k) add(5, ie, je - 1, k, -da(k - ha_p[0]), 0);
      if (valid(hb_n)) for (std::size_t i = hb_n[0]; i < hb_n[1]; ++i) add(3, i, je - 1, ks - 1, -db(i - hb_n[0]), 1);
      if (valid(hb_p)) for (std::size_t i = hb_p[0]; i < hb_p[1]; ++i) add(3, i, je - 1, ke, db(i - hb_p[0]), 1);
    } else {  // integral_x (group 0) then integral_y (group 1)
      if (valid(ha_n)) for (std::size_t i = ha_n[0]; i < ha_n[1]; ++i) add(3, i, js - 1, ke - 1, da(i - ha_n[0]), 0);
      if (valid(ha_p)) for (std::size_t i = ha_p[0]; i < ha_p[1]; ++i) add(3, i, je, ke - 1, -da(i - ha_p[0]), 0);
      if (valid(hb_n)) for (std::size_t j = hb_n[0]; j < hb_n[1]; ++j) add(4, is - 1, j, ke - 1, -db(j - hb_n[0]), 1);
      if (valid(hb_p)) for (std::size_t j = hb_p[0]; j < hb_p[1]; ++j) add(4, ie, j, ke - 1, db(j - hb_p[0]), 1);
    }
    p.scale = dir % 2 == 1 ? 1.0 : -1.0;
    cm.time() = hTime();
    cm.data().assign(steps, 0);
    return p;
  }

  void buildMonitors() {
    lines.clear();
    field_boxes.clear();
    for (auto& m : monitors) {
      GatherPlan gp;
      gp.monitor = m.get();
      if (auto* mv = dynamic_cast<MovieMonitor*>(m.get())) {
        FieldMonitor& f = mv->frame();
        gp.d.field = fieldId(f.field());
        monitorBox(f.shape(), gp.d.origin, gp.d.end);
        gp.d.first = 0;
        gp.d.interval = static_cast<int64_t>(mv->frameInterval());
        gp.d.count = steps == 0 ? 0 : static_cast<int64_t>((steps - 1) / mv->frameInterval() + 1);
        gp.time_series = false;
        mv->frames().clear();
      } else if (auto* ft = dynamic_cast<FieldTimeMonitor*>(m.get())) {
        gp.d.field = fieldId(ft->field());
        int32_t o[3], e[3];
        monitorBox(ft->cube(), o, e);
        // field_time_monitor.cpp:81-95 keeps the last cell of the box
        for (int a = 0; a < 3; ++a) {
          gp.d.origin[a] = e[a] > o[a] ? e[a] - 1 : o[a];
          gp.d.end[a] = e[a] > o[a] ? e[a] : o[a];
        }
        gp.d.first = 0;
        gp.d.interval = 1;
        gp.d.count = static_cast<int64_t>(steps);
        gp.time_series = true;
        ft->time() = ft->field() <= EMF::Field::EM ? eTime() : hTime();
        ft->data().assign(steps, 0);
      } else if (auto* vm = dynamic_cast<VoltageMonitor*>(m.get())) {
        lines.push_back(voltagePlan(*vm));
        continue;
      } else if (auto* cm = dynamic_cast<CurrentMonitor*>(m.get())) {
        lines.push_back(currentPlan(*cm));
        continue;
      } else if (auto* fm = dynamic_cast<FieldMonitor*>(m.get())) {
        // on its own a FieldMonitor records nothing while the loop runs
        // (field_monitor.cpp:80): its box of the final fields is taken after
        // the run
        FieldBox fb{fm, {0, 0, 0}, {0, 0, 0}};
        monitorBox(fm->shape(), fb.o, fb.e);
        field_boxes.push_back(fb);
        continue;
      } else {
        throw XFDTDException("unsupported monitor type");
      }
      gathers.push_back(gp);
    }
    obj_gather = -1;
    if (has_objective) {  // every step, the objective's box
      GatherPlan gp;
      gp.monitor = nullptr;
      gp.time_series = false;
      gp.d.field = obj_field;
      for (int a = 0; a < 3; ++a) {
        gp.d.origin[a] = static_cast<int32_t>(obj_o[a]);
        gp.d.end[a] = static_cast<int32_t>(obj_o[a] + obj_n[a]);
      }
      gp.d.first = 0;
      gp.d.interval = 1;
      gp.d.count = static_cast<int64_t>(steps);
      obj_gather = static_cast<int>(gathers.size());
      gathers.push_back(gp);
    }
  }

  static int fieldId(EMF::Field f) {
    switch (f) {
      case EMF::Field::EX: return 0;
      case EMF::Field::EY: return 1;
      case EMF::Field::EZ: return 2;
      case EMF::Field::HX: return 3;
      case EMF::Field::HY: return 4;
      case EMF::Field::HZ: return 5;
      default: throw XFDTDException("monitor field must be one component");
    }
  }

  // ---------------------------------------------------------------- init --
  // The host precompute does not depend on the device precision, so a second
  // init() with the same run length (after setPrecision) reuses it and only
  // rebuilds the device context.
  bool precomputed = false;
  Index precomputed_steps = 0;

  void init(Index n) {
    if (!precomputed || precomputed_steps != n) precompute(n);
    releaseDevice();
    createDevice();
    initialised = true;
  }

  void precompute(Index n) {
    initialised = false;
    precomputed = false;
    releaseDevice();
    slabs.clear();
    lumped.clear();
    tfsf.clear();
    gathers.clear();
    steps = n;
    const bool verbose = std::getenv("XFD_VERBOSE") != nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto stage = [&](const char* name) {
      if (!verbose) return;
      auto t1 = std::chrono::steady_clock::now();
      std::fprintf(stderr, "[xfdtd] precompute %-12s %8.3f s\n", name,
                   std::chrono::duration<double>(t1 - t0).count());
      t0 = t1;
    };
    buildGrid();
    stage("grid");
    buildMaterial();
    stage("material");
    buildCoefficients();
    stage("coefficients");
    // Simulation::correctUpdateCoefficient (simulation.cpp:638-652):
    // objects (lumped patches) in add order, then boundaries.
    for (auto& o : objects)
      if (auto* le = dynamic_cast<LumpedElement*>(o.get())) buildLumped(*le);
    stage("lumped");
    for (auto& b : boundaries)
      if (auto* p = dynamic_cast<PML*>(b.get())) buildPml(*p);
    stage("cpml");
    for (auto& s : sources) {
      auto* t = dynamic_cast<TFSF*>(s.get());
      if (!t) throw XFDTDException("unsupported waveform source");
      buildTfsf(*t);
    }
    stage("tfsf");
    buildDispersive();  // simulation.cpp:651, after every other coefficient patch
    stage("dispersive");
    // Simulation::makeUpdator (simulation.cpp:688-743): the ADE updators
    // are 3D only, any other dispersive case throws
    if (has_ade && dim != 3) throw XFDTDSimulationException("don't support this type of updator yet.");
    // the NF2FF surfaces (nffft.cpp:97-223) close a 3D box
    if (dim != 3 && (!nf2ffs.empty() || !nf2ff_tds.empty()))
      throw XFDTDSimulationException("NF2FF needs a 3D grid space");
    buildMonitors();
    buildNf2ff();
    coef_view.clear();
    stage("monitors");
    for (auto& nw : networks) nw->init(dt);  // Simulation::init (simulation.cpp:353-355)
    calc = std::make_shared<CalculationParam>(std::make_shared<TimeParam>(dt, steps, 0));
    precomputed = true;
    precomputed_steps = n;
  }

  // Device descriptor of the precomputed scenario.  The arrays it points to
  // are held here until the next call.
  std::vector<xfd_cpml_desc> cp;
  std::vector<xfd_lumped_desc> lp;
  std::vector<xfd_tfsf_desc> tp;
  std::vector<xfd_gather_desc> gp;

  xfd_desc makeDesc() {
    xfd_desc dd{};
    dd.nx = static_cast<int32_t>(N[0]);
    dd.ny = static_cast<int32_t>(N[1]);
    dd.nz = static_cast<int32_t>(N[2]);
    dd.dtype = precision == Precision::FP64 ? XFD_F64 : XFD_F32;
    dd.dim = dim;
    for (int s = 0; s < 18; ++s) dd.coeff[s] = coef[s].v.data();
    cp.clear();
    lp.clear();
    tp.clear();
    gp.clear();
    for (auto& s : slabs) {
      s.d.a_e = s.ae.data();
      s.d.b_e = s.be.data();
      s.d.a_h = s.ah.data();
      s.d.b_h = s.bh.data();
      s.d.c_ea_psi_hb = s.cpsi[0].data();
      s.d.c_eb_psi_ha = s.cpsi[1].data();
      s.d.c_ha_psi_eb = s.cpsi[2].data();
      s.d.c_hb_psi_ea = s.cpsi[3].data();
      cp.push_back(s.d);
    }
    for (auto& l : lumped) {
      l.d.coff = l.coff.data();
      l.d.coff_j = l.coff_j.empty() ? nullptr : l.coff_j.data();
      l.d.nt = static_cast<int64_t>(l.wave.size());
      l.d.waveform = l.wave.empty() ? nullptr : l.wave.data();
      lp.push_back(l.d);
    }
    for (auto& t : tfsf) {
      t.d.px_int = t.proj[0].data();
      t.d.py_int = t.proj[1].data();
      t.d.pz_int = t.proj[2].data();
      t.d.px_half = t.proj[3].data();
      t.d.py_half = t.proj[4].data();
      t.d.pz_half = t.proj[5].data();
      t.d.e_inc = t.e_inc.data();
      t.d.h_inc = t.h_inc.data();
      tp.push_back(t.d);
    }
    for (auto& g : gathers) gp.push_back(g.d);
    dd.n_cpml = static_cast<int32_t>(cp.size());
    dd.cpml = cp.data();
    dd.n_lumped = static_cast<int32_t>(lp.size());
    dd.lumped = lp.data();
    dd.n_tfsf = static_cast<int32_t>(tp.size());
    dd.tfsf = tp.data();
    dd.n_gather = static_cast<int32_t>(gp.size());
    dd.gather = gp.data();
    if (has_ade) {
      ade_d.coeff = ade_coeff.data();
      ade_d.cell_material = ade_cell.data();
      dd.ade = &ade_d;
    }
    for (std::size_t q = 0; q < nf2ffs.size(); ++q) {
      auto& D = nf2ffs[q]->data();
      nf2ff_d[q].te = reinterpret_cast<const double*>(D.te.data());
      nf2ff_d[q].th = reinterpret_cast<const double*>(D.th.data());
    }
    dd.n_nf2ff = static_cast<int32_t>(nf2ff_d.size());
    dd.nf2ff = nf2ff_d.empty() ? nullptr : nf2ff_d.data();
    for (std::size_t q = 0; q < nf2ff_tds.size(); ++q) {
      auto& D = nf2ff_tds[q]->data();
      tdnf_d[q].s_e = D.s_e.data();
      tdnf_d[q].s_h = D.s_h.data();
      tdnf_d[q].ds = D.ds.data();
    }
    dd.n_tdnf = static_cast<int32_t>(tdnf_d.size());
    dd.tdnf = tdnf_d.empty() ? nullptr : tdnf_d.data();
    ld.clear();
    for (auto& l : lines) {
      xfd_line_desc x{};
      x.n_terms = static_cast<int32_t>(l.weights.size());
      x.terms = l.terms.data();
      x.weights = l.weights.data();
      x.scale = l.scale;
      x.count = static_cast<int64_t>(steps);
      ld.push_back(x);
    }
    dd.n_line = static_cast<int32_t>(ld.size());
    dd.line = ld.empty() ? nullptr : ld.data();
    return dd;
  }

  void createDevice() {
    const xfd_desc dd = makeDesc();
    xfd_ctx* c = nullptr;
    if (xfd_create(&dd, &c) != XFD_OK) throw XFDTDSimulationException(std::string("xfd_create: ") + xfd_last_error());
    attach(c, nullptr);
  }

  // Takes a device context: owned, or a case of the batch `core`.
  void attach(xfd_ctx* c, std::shared_ptr<BatchCore> core) {
    ctx = c;
    batch_core = std::move(core);
    initialised = true;
    for (std::size_t q = 0; q < nf2ffs.size(); ++q) {  // far-field sums on the device
      auto& D = nf2ffs[q]->data();
      const NFFFTFrequencyDomain::Data* Dp = &D;
      D.device = [this, q, Dp](std::size_t f, Real k, const std::vector<Real>& t, const std::vector<Real>& p,
                               const Vector& o, std::vector<std::complex<Real>>& out) {
        const std::size_t na = t.size();
        out.assign(4 * na, {0, 0});
        std::vector<double> buf(8 * na);
        const double org[3] = {o[0], o[1], o[2]};
        const double* en[3] = {Dp->e_node[0].data(), Dp->e_node[1].data(), Dp->e_node[2].data()};
        const double* hn[3] = {Dp->h_node[0].data(), Dp->h_node[1].data(), Dp->h_node[2].data()};
        const double* es[3] = {Dp->e_size[0].data(), Dp->e_size[1].data(), Dp->e_size[2].data()};
        if (xfd_nf2ff_far_field(ctx, static_cast<int32_t>(q), static_cast<int32_t>(f), static_cast<int32_t>(na),
                                t.data(), p.data(), org, k, en, hn, es, buf.data()) != XFD_OK)
          throw XFDTDSimulationException(std::string("xfd_nf2ff_far_field: ") + xfd_last_error());
        for (std::size_t n = 0; n < na; ++n)
          for (int w = 0; w < 4; ++w) out[w * na + n] = {buf[n * 8 + 2 * w], buf[n * 8 + 2 * w + 1]};
      };
    }
    int64_t b = 0;
    xfd_device_bytes(ctx, &b);
    dev_bytes = b;
    int32_t cl = 0;
    xfd_coef_layout(ctx, &cl, nullptr);
    coef_layout = cl;
    int32_t fu = 0;
    xfd_cpml_fused(ctx, &fu);
    cpml_fused = fu;
  }

  void sendFlag(SimulationInitFlag f) {
    for (auto& v : visitors) v->initStep(f);
  }

  // The time loop (Domain::run, domain.cpp:80-137): with visitors it is
  // submitted in about 100 pieces, each followed by NextStep.
  void runLoop() {
    sendFlag(SimulationInitFlag::UpdateStart);
    const Index piece = visitors.empty() ? std::max<Index>(steps, 1) : std::max<Index>(1, (steps + 99) / 100);
    double total = 0;
    for (Index t = 0; t < steps; t += piece) {
      const Index n = std::min(piece, steps - t);
      advance(t, n);
      total += loop_seconds;
      for (auto& v : visitors) v->iteratorStep(SimulationIteratorFlag::NextStep, t + n, 0, steps);
    }
    loop_seconds = total;
    finish();
    sendFlag(SimulationInitFlag::UpdateEnd);
  }

  // Steps t0 .. t0+n-1 on the device; returns when they are done.
  void advance(Index t0, Index n) {
    if (!initialised || !ctx) throw XFDTDSimulationException("Simulation is not initialised");
    if (batch_core) throw XFDTDSimulationException("this Simulation is stepped by its SimulationBatch");
    if (t0 + n > steps) throw XFDTDSimulationException("steps beyond the run length given to init()");
    xfd_set_profiling_classes(ctx, profiling ? profiling_classes : 0u);
    const auto start = std::chrono::steady_clock::now();
    if (xfd_step(ctx, static_cast<int64_t>(t0), static_cast<int64_t>(n)) != XFD_OK)
      throw XFDTDSimulationException(std::string("xfd_step: ") + xfd_last_error());
    loop_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    if (calc) calc->timeParam()->setCurrentTimeStep(t0 + n);
    stats.clear();
    for (int k = 0;; ++k) {
      const char* name = nullptr;
      int64_t l = 0;
      double ms = 0;
      if (xfd_kernel_stats(ctx, k, &name, &l, &ms) != XFD_OK) break;
      stats.push_back({name, static_cast<long long>(l), ms});
    }
  }

  void finish() {
    if (!ctx) throw XFDTDSimulationException("Simulation is not initialised");
    collectOutputs();
  }

  // ------------------------------------------------------------- adjoint --
  void checkAdjointSetup() const {
    if (!has_design) throw XFDTDException("no design region (setDesignRegion)");
    if (!has_objective) throw XFDTDException("no objective (setDftObjective)");
  }

  // DFT weight of the value recorded after step n: dt exp(-j 2 pi f (n+1) dt).
  std::complex<Real> dftWeight(Index n) const {
    const Real ph = -2 * kPi * obj_freq * static_cast<Real>(n + 1) * dt;
    return std::complex<Real>(dt * std::cos(ph), dt * std::sin(ph));
  }

  Real forward(Index n, Index interval) {
    checkAdjointSetup();
    if (!precomputed || precomputed_steps != n) precompute(n);
    releaseDevice();
    createDevice();
    forward_done = false;
    if (xfd_forward_checkpointed(ctx, static_cast<int64_t>(n), static_cast<int64_t>(interval)) != XFD_OK)
      throw XFDTDSimulationException(std::string("xfd_forward_checkpointed: ") + xfd_last_error());
    collectOutputs();
    const std::size_t np = obj_n[0] * obj_n[1] * obj_n[2];
    std::vector<Real> series(np * n);
    if (!series.empty() && xfd_read_gather(ctx, obj_gather, series.data()) != XFD_OK)
      throw XFDTDSimulationException(std::string("xfd_read_gather: ") + xfd_last_error());
    obj_dft.assign(np, {0, 0});
    for (Index t = 0; t < n; ++t) {
      const auto w = dftWeight(t);
      for (std::size_t p = 0; p < np; ++p) obj_dft[p] += w * series[t * np + p];
    }
    obj_value = 0;
    for (auto& d : obj_dft) obj_value += std::norm(d);
    forward_done = true;
    return obj_value;
  }

  std::vector<Real> backward() {
    if (!forward_done || !ctx) throw XFDTDSimulationException("backward() needs a forward() first");
    const Index n = steps;
    const std::size_t np = obj_n[0] * obj_n[1] * obj_n[2];
    // dJ/dF_p after step t = 2 Re(conj(D_p) w_t)
    std::vector<Real> src(np * n);
    for (Index t = 0; t < n; ++t) {
      const auto w = dftWeight(t);
      for (std::size_t p = 0; p < np; ++p) src[p * n + t] = 2 * std::real(std::conj(obj_dft[p]) * w);
    }
    std::vector<int32_t> pts;
    for (std::size_t i = 0; i < obj_n[0]; ++i)
      for (std::size_t j = 0; j < obj_n[1]; ++j)
        for (std::size_t k = 0; k < obj_n[2]; ++k) {
          pts.push_back(obj_field);
          pts.push_back(static_cast<int32_t>(obj_o[0] + i));
          pts.push_back(static_cast<int32_t>(obj_o[1] + j));
          pts.push_back(static_cast<int32_t>(obj_o[2] + k));
        }
    xfd_adjoint_desc d{};
    int32_t eo[3][3], en[3][3];
    for (int a = 0; a < 3; ++a) {
      d.origin[a] = static_cast<int32_t>(design_o[a]);
      d.size[a] = static_cast<int32_t>(design_n[a]);
    }
    // the design edges and their coefficient derivatives
    std::vector<Real> beta, gamma;
    for (int q = 0; q < 3; ++q) {
      for (int a = 0; a < 3; ++a) {
        eo[q][a] = d.origin[a];
        en[q][a] = a == q ? d.size[a] : d.size[a] + 1;
      }
      for (int i = 0; i < en[q][0]; ++i)
        for (int j = 0; j < en[q][1]; ++j)
          for (int k = 0; k < en[q][2]; ++k) {
            const std::size_t ii = eo[q][0] + i, jj = eo[q][1] + j, kk = eo[q][2] + k;
            const Real e = eps[q].at(ii, jj, kk), sg = se[q].at(ii, jj, kk);
            const Real den = 2 * e + sg * dt;
            const Real b = -2 / den, sv = (2 * e - dt * sg) / den;
            beta.push_back(b);
            gamma.push_back(4 * sg * dt / (den * den) - b * sv);
          }
    }
    d.beta = beta.data();
    d.gamma = gamma.data();
    d.n_points = static_cast<int32_t>(np);
    d.points = pts.data();
    d.nt = static_cast<int64_t>(n);
    d.source = src.data();
    std::vector<Real> ge(beta.size());
    if (xfd_backward(ctx, &d, ge.data()) != XFD_OK)
      throw XFDTDSimulationException(std::string("xfd_backward: ") + xfd_last_error());
    // eps of an edge is the mean of its four cells (averageEdges), so each
    // design cell collects eps0/4 of the gradient of its twelve edges
    std::vector<Real> g(design_n[0] * design_n[1] * design_n[2], 0);
    std::size_t off = 0;
    for (int q = 0; q < 3; ++q) {
      const int A = (q + 1) % 3, B = (q + 2) % 3;
      for (int i = 0; i < en[q][0]; ++i)
        for (int j = 0; j < en[q][1]; ++j)
          for (int k = 0; k < en[q][2]; ++k, ++off) {
            const int e[3] = {i, j, k};
            for (int da = -1; da <= 0; ++da)
              for (int db = -1; db <= 0; ++db) {
                int c[3] = {e[0], e[1], e[2]};
                c[A] += da;
                c[B] += db;
                if (c[A] < 0 || c[B] < 0 || c[A] >= d.size[A] || c[B] >= d.size[B]) continue;
                g[(static_cast<std::size_t>(c[0]) * design_n[1] + c[1]) * design_n[2] + c[2]] +=
                    0.25 * kEps0 * ge[off];
              }
          }
    }
    return g;
  }

  void collectOutputs() {
    for (std::size_t gi = 0; gi < gathers.size(); ++gi) {
      auto& g = gathers[gi];
      if (!g.monitor) continue;
      const std::size_t n0 = g.d.end[0] - g.d.origin[0], n1 = g.d.end[1] - g.d.origin[1],
                        n2 = g.d.end[2] - g.d.origin[2];
      const std::size_t fe = n0 * n1 * n2;
      std::vector<Real> buf(fe * static_cast<std::size_t>(g.d.count));
      if (!buf.empty() && xfd_read_gather(ctx, static_cast<int32_t>(gi), buf.data()) != XFD_OK)
        throw XFDTDSimulationException(std::string("xfd_read_gather: ") + xfd_last_error());
      if (g.time_series) {  // its file is written by output() (time_monitor.cpp:11-21)
        auto* ft = static_cast<FieldTimeMonitor*>(g.monitor);
        if (fe == 1) ft->data() = buf;
      } else {
        auto* mv = static_cast<MovieMonitor*>(g.monitor);
        auto& frames = mv->frames();
        frames.clear();
        for (int64_t f = 0; f < g.d.count; ++f) {
          // FieldMonitor::output (field_monitor.cpp:82-96) leaves the data
          // unassigned (an empty array) when the box is empty
          Array3D<Real> a = fe > 0 ? Array3D<Real>(n0, n1, n2) : Array3D<Real>(0, 0, 0);
          std::memcpy(a.data.data(), buf.data() + f * fe, fe * sizeof(Real));
          char name[32];
          std::snprintf(name, sizeof(name), "%05lld", static_cast<long long>(f * g.d.interval));
          if (!mv->outputDir().empty() && fe > 0) {  // movie_monitor.cpp:26-34, monitor.cpp:42-54
            std::filesystem::create_directories(mv->outputDir());
            npy::write(mv->outputDir() + "/" + name + ".npy", {n0, n1, n2}, a.data.data());
          }
          if (f + 1 == g.d.count) mv->frame().data() = a;
          frames.emplace_back(name, std::move(a));
        }
      }
    }
    for (std::size_t m = 0; m < lines.size(); ++m) {  // voltage / current monitors
      std::vector<Real> buf(steps);
      if (!buf.empty() && xfd_read_line(ctx, static_cast<int32_t>(m), buf.data()) != XFD_OK)
        throw XFDTDSimulationException(std::string("xfd_read_line: ") + xfd_last_error());
      // files are written by output() (time_monitor.cpp:11-21)
      if (auto* vm = dynamic_cast<VoltageMonitor*>(lines[m].monitor))
        vm->data() = buf;
      else if (auto* cm = dynamic_cast<CurrentMonitor*>(lines[m].monitor))
        cm->data() = buf;
    }
    // final fields (Simulation::emf)
    Array3D<Real>* dst[6] = {&emf->ex, &emf->ey, &emf->ez, &emf->hx, &emf->hy, &emf->hz};
    for (int f = 0; f < 6; ++f) {
      std::size_t ext[3];
      extent(f >= 3, f % 3, N, ext);
      *dst[f] = Array3D<Real>(ext[0], ext[1], ext[2]);
      if (xfd_read_field(ctx, f, dst[f]->data.data()) != XFD_OK)
        throw XFDTDSimulationException(std::string("xfd_read_field: ") + xfd_last_error());
    }
    for (auto& fb : field_boxes) {  // FieldMonitor::output's view (field_monitor.cpp:82-96)
      const Array3D<Real>& src = emf->field(fb.monitor->field());
      const std::size_t n0 = fb.e[0] > fb.o[0] ? fb.e[0] - fb.o[0] : 0, n1 = fb.e[1] > fb.o[1] ? fb.e[1] - fb.o[1] : 0,
                        n2 = fb.e[2] > fb.o[2] ? fb.e[2] - fb.o[2] : 0;
      Array3D<Real> a(n0, n1, n2);
      for (std::size_t i = 0; i < n0; ++i)
        for (std::size_t j = 0; j < n1; ++j)
          for (std::size_t k = 0; k < n2; ++k) a(i, j, k) = src(fb.o[0] + i, fb.o[1] + j, fb.o[2] + k);
      fb.monitor->data() = std::move(a);
    }
    for (std::size_t q = 0; q < nf2ffs.size(); ++q) {  // the surface currents
      auto& D = nf2ffs[q]->data();
      for (std::size_t f = 0; f < D.freqs.size(); ++f)
        for (int d = 0; d < 6; ++d)
          for (int w = 0; w < 4; ++w) {
            auto& v = D.cur[(f * 6 + d) * 4 + w];
            if (v.empty()) continue;
            if (xfd_read_nf2ff(ctx, static_cast<int32_t>(q), static_cast<int32_t>(f), d, w,
                               reinterpret_cast<double*>(v.data())) != XFD_OK)
              throw XFDTDSimulationException(std::string("xfd_read_nf2ff: ") + xfd_last_error());
          }
    }
    for (std::size_t q = 0; q < nf2ff_tds.size(); ++q) {  // time-domain potentials
      auto& D = nf2ff_tds[q]->data();
      for (int d = 0; d < 6; ++d)
        for (int w = 0; w < 4; ++w) {
          D.pot[d][w].assign(D.num, 0.0);
          if (xfd_read_td_nf2ff(ctx, static_cast<int32_t>(q), d, w, D.pot[d][w].data()) != XFD_OK)
            throw XFDTDSimulationException(std::string("xfd_read_td_nf2ff: ") + xfd_last_error());
        }
    }
  }
};

Simulation::Simulation(Real dx, Real dy, Real dz, Real cfl, ThreadConfig) : _impl(std::make_unique<Impl>()) {
  _impl->d[0] = dx;
  _impl->d[1] = dy;
  _impl->d[2] = dz;
  _impl->cfl = cfl;
}
Simulation::~Simulation() = default;
void Simulation::addObject(std::shared_ptr<Object> o) {
  _impl->objects.push_back(std::move(o));
  _impl->precomputed = false;
}
void Simulation::addWaveformSource(std::shared_ptr<WaveformSource> s) {
  _impl->sources.push_back(std::move(s));
  _impl->precomputed = false;
}
void Simulation::addBoundary(std::shared_ptr<Boundary> b) {
  _impl->boundaries.push_back(std::move(b));
  _impl->precomputed = false;
}
void Simulation::addMonitor(std::shared_ptr<Monitor> m) {
  _impl->monitors.push_back(std::move(m));
  _impl->precomputed = false;
}
void Simulation::addNF2FF(std::shared_ptr<NFFFTFrequencyDomain> n) {
  _impl->nf2ffs.push_back(std::move(n));
  _impl->precomputed = false;
}
void Simulation::addNF2FF(std::shared_ptr<NFFFTTimeDomain> n) {
  _impl->nf2ff_tds.push_back(std::move(n));
  _impl->precomputed = false;
}
void Simulation::addNetwork(std::shared_ptr<Network> n) {
  _impl->networks.push_back(std::move(n));
  _impl->precomputed = false;
}
const std::vector<std::shared_ptr<Network>>& Simulation::networks() const { return _impl->networks; }
void Simulation::setPrecision(Precision p) { _impl->precision = p; }
void Simulation::init(Index n) { _impl->init(n); }
void Simulation::precompute(Index n) { _impl->precompute(n); }
void Simulation::run() { _impl->runLoop(); }
void Simulation::advance(Index t0, Index n) { _impl->advance(t0, n); }
void Simulation::finish() { _impl->finish(); }
void Simulation::run(Index n) {  // simulation.cpp:284-289
  _impl->sendFlag(SimulationInitFlag::SimulationStart);
  init(n);
  run();
  _impl->sendFlag(SimulationInitFlag::SimulationEnd);
}
auto Simulation::addVisitor(std::shared_ptr<SimulationFlagVisitor> v) -> void {
  _impl->visitors.push_back(std::move(v));
}
auto Simulation::addDefaultVisitor() -> void {
  _impl->visitors.push_back(std::make_shared<DefaultSimulationFlagVisitor>());
}
auto Simulation::visitors() -> std::vector<std::shared_ptr<SimulationFlagVisitor>> { return _impl->visitors; }
const std::shared_ptr<CalculationParam>& Simulation::calculationParam() const { return _impl->calc; }
Real Simulation::dt() const { return _impl->dt; }
std::array<Index, 3> Simulation::gridSize() const { return {_impl->N[0], _impl->N[1], _impl->N[2]}; }
const std::shared_ptr<EMF>& Simulation::emf() const { return _impl->emf; }
const Array3D<Real>& Simulation::coefficient(const std::string& name) const {
  for (int s = 0; s < 18; ++s) {
    if (name != kCoefNames[s]) continue;
    auto it = _impl->coef_view.find(name);
    if (it == _impl->coef_view.end()) {
      const auto& f = _impl->coef[s];
      Array3D<Real> a(f.n[0], f.n[1], f.n[2]);
      a.data.assign(f.v.data(), f.v.data() + f.v.size());
      it = _impl->coef_view.emplace(name, std::move(a)).first;
    }
    return it->second;
  }
  throw XFDTDException("unknown coefficient " + name);
}
const std::vector<Real>& Simulation::eNode(int a) const { return _impl->e_node[a]; }
const std::vector<Real>& Simulation::hSize(int a) const { return _impl->h_size[a]; }
double Simulation::loopSeconds() const { return _impl->loop_seconds; }
void Simulation::setProfiling(bool on) {
  _impl->profiling = on;
  _impl->profiling_classes = ~0u;
}
void Simulation::setProfilingClasses(unsigned mask) {
  _impl->profiling = mask != 0;
  _impl->profiling_classes = mask;
}
std::vector<Simulation::KernelStat> Simulation::kernelStats() const { return _impl->stats; }
long long Simulation::deviceBytes() const { return _impl->dev_bytes; }
void* Simulation::deviceContext() const { return _impl->ctx; }
bool Simulation::classedCoefficients() const { return _impl->coef_layout != 0; }
int Simulation::coefficientLayout() const { return _impl->coef_layout; }
int Simulation::fusedCpml() const { return _impl->cpml_fused; }

void Simulation::setDesignRegion(const std::array<Index, 3>& origin, const std::array<Index, 3>& size,
                                 std::vector<Real> eps_r) {
  auto& I = *_impl;
  if (eps_r.size() != size[0] * size[1] * size[2] || eps_r.empty())
    throw XFDTDException("design permittivity must have one value per design cell");
  for (Real e : eps_r)
    if (!(e > 0)) throw XFDTDException("design permittivity must be positive");
  for (int a = 0; a < 3; ++a) {
    I.design_o[a] = origin[a];
    I.design_n[a] = size[a];
  }
  I.design_eps = std::move(eps_r);
  I.has_design = true;
  I.precomputed = false;
}

void Simulation::setDftObjective(EMF::Field field, const std::array<Index, 3>& origin,
                                 const std::array<Index, 3>& size, Real frequency) {
  auto& I = *_impl;
  const int f = Impl::fieldId(field);
  if (f >= 3) throw XFDTDException("the objective records an E component");
  if (size[0] * size[1] * size[2] == 0) throw XFDTDException("empty objective box");
  I.obj_field = f;
  for (int a = 0; a < 3; ++a) {
    I.obj_o[a] = origin[a];
    I.obj_n[a] = size[a];
  }
  I.obj_freq = frequency;
  I.has_objective = true;
  I.precomputed = false;
}

Real Simulation::forward(Index time_step, Index checkpoint_interval) {
  return _impl->forward(time_step, checkpoint_interval);
}
std::vector<Real> Simulation::backward() { return _impl->backward(); }
std::vector<std::complex<Real>> Simulation::objectiveDft() const { return _impl->obj_dft; }

// ------------------------------------------------------ SimulationBatch --
struct SimulationBatch::Impl {
  std::vector<Simulation*> sims;
  std::shared_ptr<BatchCore> core;
  Index steps = 0;
  bool graph = false;
  double loop_seconds = 0, device_ms = 0;

  xfd_batch* batch() const {
    if (!core) throw XFDTDSimulationException("SimulationBatch is not initialised");
    return core->b;
  }
  static void check(int rc, const char* what) {
    if (rc != XFD_OK) throw XFDTDSimulationException(std::string(what) + ": " + xfd_last_error());
  }
};

SimulationBatch::SimulationBatch() : _impl(std::make_unique<Impl>()) {}
SimulationBatch::~SimulationBatch() = default;

void SimulationBatch::add(Simulation& sim) {
  _impl->sims.push_back(&sim);
  _impl->core.reset();
}

std::size_t SimulationBatch::size() const { return _impl->sims.size(); }

void SimulationBatch::init(Index n) {
  auto& I = *_impl;
  if (I.sims.empty()) throw XFDTDSimulationException("empty SimulationBatch");
  I.core.reset();
  std::vector<xfd_desc> descs;
  for (auto* s : I.sims) {
    auto& S = *s->_impl;
    if (!S.precomputed || S.precomputed_steps != n) S.precompute(n);
    S.releaseDevice();
    descs.push_back(S.makeDesc());
  }
  auto core = std::make_shared<BatchCore>();
  Impl::check(xfd_batch_create(static_cast<int32_t>(descs.size()), descs.data(), &core->b), "xfd_batch_create");
  for (std::size_t i = 0; i < I.sims.size(); ++i) {
    xfd_ctx* c = nullptr;
    Impl::check(xfd_batch_case(core->b, static_cast<int32_t>(i), &c), "xfd_batch_case");
    I.sims[i]->_impl->attach(c, core);
  }
  Impl::check(xfd_batch_set_graph(core->b, I.graph ? 1 : 0), "xfd_batch_set_graph");
  I.core = std::move(core);
  I.steps = n;
}

void SimulationBatch::setGraph(bool on) {
  _impl->graph = on;
  if (_impl->core) Impl::check(xfd_batch_set_graph(_impl->batch(), on ? 1 : 0), "xfd_batch_set_graph");
}

void SimulationBatch::tune(int reps) { Impl::check(xfd_batch_tune(_impl->batch(), reps), "xfd_batch_tune"); }

void SimulationBatch::advance(Index t0, Index n) {
  auto& I = *_impl;
  xfd_batch* b = I.batch();
  if (t0 + n > I.steps) throw XFDTDSimulationException("steps beyond the run length given to init()");
  const auto start = std::chrono::steady_clock::now();
  Impl::check(xfd_batch_step(b, static_cast<int64_t>(t0), static_cast<int64_t>(n)), "xfd_batch_step");
  I.loop_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
  Impl::check(xfd_batch_last_ms(b, &I.device_ms), "xfd_batch_last_ms");
}

void SimulationBatch::finish() {
  _impl->batch();
  for (auto* s : _impl->sims) s->_impl->finish();
}

void SimulationBatch::run() {
  advance(0, _impl->steps);
  finish();
}

void SimulationBatch::run(Index n) {
  init(n);
  run();
}

double SimulationBatch::loopSeconds() const { return _impl->loop_seconds; }
double SimulationBatch::deviceMilliseconds() const { return _impl->device_ms; }

SimulationBatch::Layout SimulationBatch::layout() const {
  int32_t shared = 0, ichunk = 0, grid = 0, lk = 0, rj = 0;
  Impl::check(xfd_batch_info(_impl->batch(), &shared, &ichunk, &grid), "xfd_batch_info");
  Impl::check(xfd_batch_tile(_impl->batch(), &lk, &rj), "xfd_batch_tile");
  return {shared != 0, ichunk, grid, lk, rj};
}

}  // namespace xfdtd
