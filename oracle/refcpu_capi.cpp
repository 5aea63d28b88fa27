// C-ABI of the refcpu oracle (TEST INFRASTRUCTURE ONLY; see refcpu.hpp).
//
// Scenarios arrive as the line-oriented spec text that
// xfdtd_core_amd/spec.py writes (one directive per line, '#' comments).  The
// grammar is documented there; this parser is an independent reading of it.
#include <cstring>
#include <sstream>

#include "refcpu.hpp"

using namespace refcpu;

namespace {
thread_local std::string g_err;

struct Handle {
  Sim sim;
};

Dir parseDir(const std::string& s) {
  static const char* n[6] = {"XN", "XP", "YN", "YP", "ZN", "ZP"};
  for (int i = 0; i < 6; ++i)
    if (s == n[i]) return static_cast<Dir>(i);
  throw Error("bad direction " + s);
}
Axis parseAxis(const std::string& s) {
  if (s == "X") return X;
  if (s == "Y") return Y;
  if (s == "Z") return Z;
  throw Error("bad axis " + s);
}
int parseField(const std::string& s) {
  static const char* n[6] = {"EX", "EY", "EZ", "HX", "HY", "HZ"};
  for (int i = 0; i < 6; ++i)
    if (s == n[i]) return i;
  throw Error("bad field " + s);
}

Real num(std::istringstream& is) {
  std::string t;
  if (!(is >> t)) throw Error("missing number");
  return std::strtod(t.c_str(), nullptr);
}
std::string word(std::istringstream& is) {
  std::string t;
  if (!(is >> t)) throw Error("missing token");
  return t;
}

// waveform.cpp:80-101 plus the user-functor form used for a derivative of
// Gaussian (waveform.h:13).
Waveform parseWaveform(std::istringstream& is) {
  std::string kind = word(is);
  Waveform w;
  if (kind == "gaussian") {
    Real tau = num(is), t0 = num(is);
    w.amplitude = num(is);
    w.f = [tau, t0](Real t) { return std::exp(-std::pow((t - t0) / tau, 2)); };
  } else if (kind == "cosine_modulated_gaussian") {
    Real tau = num(is), t0 = num(is), f = num(is);
    w.amplitude = num(is);
    w.f = [tau, t0, f](Real t) {
      return std::exp(-std::pow((t - t0) / tau, 2)) * std::cos(2 * PI * f * (t - t0));
    };
  } else if (kind == "sine") {
    Real f = num(is);
    w.amplitude = num(is);
    w.f = [f](Real t) { return std::sin(2 * PI * f * t); };
  } else if (kind == "cosine") {
    Real f = num(is);
    w.amplitude = num(is);
    w.f = [f](Real t) { return std::cos(2 * PI * f * t); };
  } else if (kind == "step") {
    Real t0 = num(is);
    w.amplitude = num(is);
    w.f = [t0](Real t) { return static_cast<Real>(t >= t0); };
  } else if (kind == "gaussian_derivative") {
    Real tau = num(is), t0 = num(is);
    w.amplitude = num(is);
    w.f = [tau, t0](Real t) {
      Real u = (t - t0) / tau;
      return -u * std::exp(-(u * u));
    };
  } else {
    throw Error("unknown waveform " + kind);
  }
  return w;
}

std::unique_ptr<Shape> parseShape(std::istringstream& is) {
  std::string kind = word(is);
  if (kind == "cube") {
    Real ox = num(is), oy = num(is), oz = num(is), sx = num(is), sy = num(is), sz = num(is);
    return std::make_unique<Cube>(Vec3{ox, oy, oz}, Vec3{sx, sy, sz});
  }
  if (kind == "sphere") {
    Real cx = num(is), cy = num(is), cz = num(is), r = num(is);
    return std::make_unique<Sphere>(Vec3{cx, cy, cz}, r);
  }
  if (kind == "cylinder") {
    Real cx = num(is), cy = num(is), cz = num(is), r = num(is), h = num(is);
    Axis a = parseAxis(word(is));
    return std::make_unique<Cylinder>(Vec3{cx, cy, cz}, r, h, a);
  }
  throw Error("unknown shape " + kind);
}

Material parseMaterial(std::istringstream& is) {
  std::string kind = word(is);
  if (kind == "plain") {
    Real er = num(is), mr = num(is), se = num(is), sm = num(is);
    return Material::make(er, mr, se, sm);
  }
  auto list = [&](Index n) {
    std::vector<Real> v(n);
    for (auto& x : v) x = num(is);
    return v;
  };
  if (kind == "lorentz" || kind == "drude" || kind == "debye") {
    const Real eps_inf = num(is);
    const Index n = static_cast<Index>(num(is));
    if (kind == "drude") {
      auto wp = list(n);
      auto g = list(n);
      return Material::makeDrude(eps_inf, wp, g);
    }
    if (kind == "debye") {
      auto es = list(n);
      auto tau = list(n);
      return Material::makeDebye(eps_inf, es, tau);
    }
    auto a0 = list(n), a1 = list(n), b0 = list(n), b1 = list(n), b2 = list(n);
    return Material::makeLorentz(eps_inf, a0, a1, b0, b1, b2);
  }
  throw Error("unknown material kind " + kind);
}

std::unique_ptr<Cube> parseCube(std::istringstream& is) {
  Real ox = num(is), oy = num(is), oz = num(is), sx = num(is), sy = num(is), sz = num(is);
  return std::make_unique<Cube>(Vec3{ox, oy, oz}, Vec3{sx, sy, sz});
}

void parseSpec(Sim& s, const std::string& spec) {
  std::istringstream all(spec);
  std::string line;
  bool have_sim = false;
  while (std::getline(all, line)) {
    auto h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream is(line);
    std::string d;
    if (!(is >> d)) continue;
    if (d == "sim") {
      s.dx = num(is);
      s.dy = num(is);
      s.dz = num(is);
      s.cfl = num(is);
      have_sim = true;
    } else if (d == "threads") {
      for (int a = 0; a < 3; ++a) s.tcfg[a] = static_cast<int>(num(is));
    } else if (d == "object") {
      auto o = std::make_unique<Object>();
      o->name = word(is);
      o->shape = parseShape(is);
      o->material = parseMaterial(is);
      s.objects.push_back(std::move(o));
    } else if (d == "pml") {
      auto p = std::make_unique<PML>();
      p->thickness = static_cast<int>(num(is));
      p->dir = parseDir(word(is));
      p->order = 4;
      p->sigma_ratio = 1;
      p->alpha_min = 0;
      p->alpha_max = 0.05;
      p->kappa_max = 10;
      std::string t;
      if (is >> t) {
        p->order = std::stoi(t);
        p->sigma_ratio = num(is);
        p->alpha_min = num(is);
        p->alpha_max = num(is);
        p->kappa_max = num(is);
      }
      s.pmls.push_back(std::move(p));
    } else if (d == "tfsf3d") {
      auto t = std::make_unique<TFSF3D>();
      t->x = static_cast<Index>(num(is));
      t->y = static_cast<Index>(num(is));
      t->z = static_cast<Index>(num(is));
      t->theta = num(is);
      t->phi = num(is);
      t->psi = num(is);
      t->waveform = parseWaveform(is);
      s.tfsfs.push_back(std::move(t));
    } else if (d == "tfsf2d") {  // TFSF2D (tfsf_2d.cpp:10-13)
      auto t = std::make_unique<TFSF3D>();
      t->kind = 2;
      t->x = static_cast<Index>(num(is));
      t->y = static_cast<Index>(num(is));
      t->z = 0;
      t->theta = PI * 0.5;
      t->phi = num(is);
      t->psi = 0;
      t->waveform = parseWaveform(is);
      s.tfsfs.push_back(std::move(t));
    } else if (d == "tfsf1d") {  // TFSF1D (tfsf_1d.cpp:14-23)
      auto t = std::make_unique<TFSF3D>();
      t->kind = 1;
      t->x = 0;
      t->y = 0;
      t->z = static_cast<Index>(num(is));
      t->forward = num(is) != 0;
      t->theta = t->forward ? 0 : PI;
      t->phi = 0;
      t->psi = 0;
      t->waveform = parseWaveform(is);
      s.tfsfs.push_back(std::move(t));
    } else if (d == "current_source" || d == "voltage_source") {
      auto e = std::make_unique<SourceElement>();
      e->voltage = d == "voltage_source";
      e->name = word(is);
      e->shape = parseCube(is);
      e->direction = parseDir(word(is));
      e->xyz = dirAxis(e->direction);
      e->resistance = num(is);
      if (e->resistance == 0) e->resistance = 1e-20;
      e->waveform = parseWaveform(is);
      e->material = Material::make(1, 1, 0, 0);
      s.objects.push_back(std::move(e));
    } else if (d == "resistor") {
      auto e = std::make_unique<Resistor>();
      e->name = word(is);
      e->shape = parseCube(is);
      e->xyz = parseAxis(word(is));
      e->resistance = num(is);
      if (e->resistance == 0) e->resistance = 1e-10;
      e->material = Material::make(1, 1, 0, 0);
      s.objects.push_back(std::move(e));
    } else if (d == "capacitor") {
      auto e = std::make_unique<Capacitor>();
      e->name = word(is);
      e->shape = parseCube(is);
      e->xyz = parseAxis(word(is));
      e->capacitance = num(is);
      e->material = Material::make(1, 1, 0, 0);
      s.objects.push_back(std::move(e));
    } else if (d == "inductor") {
      auto e = std::make_unique<Inductor>();
      e->name = word(is);
      e->shape = parseCube(is);
      e->xyz = parseAxis(word(is));
      e->inductance = num(is);
      if (e->inductance == 0) e->inductance = 1e-10;  // inductor.cpp:15-17
      e->material = Material::make(1, 1, 0, 0);
      s.objects.push_back(std::move(e));
    } else if (d == "pec_plane") {
      auto e = std::make_unique<PecPlane>();
      e->name = word(is);
      e->shape = parseCube(is);
      e->material = Material::make(1, 1, 1e10, 0);  // ElectroMagneticProperty::pec
      s.objects.push_back(std::move(e));
    } else if (d == "nf2ff") {
      auto n = std::make_unique<NF2FF>();
      for (int a = 0; a < 3; ++a) n->dist[a] = static_cast<Index>(num(is));
      const Index nf = static_cast<Index>(num(is));
      for (Index f = 0; f < nf; ++f) n->freqs.push_back(num(is));
      n->out_dir = word(is);
      if (n->out_dir == "-") n->out_dir.clear();
      std::string opt;
      if (is >> opt && opt == "plain_dft") n->reference_repeat = false;
      s.nf2ffs.push_back(std::move(n));
    } else if (d == "design_region") {  // ox oy oz nx ny nz eps_r... (cells, k fastest)
      for (int a = 0; a < 3; ++a) s.design_o[a] = static_cast<Index>(num(is));
      for (int a = 0; a < 3; ++a) s.design_n[a] = static_cast<Index>(num(is));
      s.design_eps.resize(s.design_n[0] * s.design_n[1] * s.design_n[2]);
      for (auto& e : s.design_eps) e = num(is);
      s.has_design = true;
    } else if (d == "dft_objective") {  // FIELD ox oy oz nx ny nz FREQ
      auto ob = std::make_unique<DftObjective>();
      ob->field = parseField(word(is));
      for (int a = 0; a < 3; ++a) ob->o[a] = static_cast<Index>(num(is));
      for (int a = 0; a < 3; ++a) ob->n[a] = static_cast<Index>(num(is));
      ob->freq = num(is);
      s.objective = std::move(ob);
    } else if (d == "movie") {
      auto m = std::make_unique<MovieMonitor>();
      auto f = std::make_unique<FieldMonitor>();
      f->field = parseField(word(is));
      f->shape = parseCube(is);
      m->interval = static_cast<Index>(num(is));
      m->name = word(is);
      m->out_dir = word(is);
      if (m->out_dir == "-") m->out_dir.clear();
      m->frame = std::move(f);
      s.monitors.push_back(std::move(m));
    } else if (d == "field_monitor") {  // FIELD CUBE NAME OUT_DIR: a FieldMonitor on its own
      auto f = std::make_unique<FieldMonitor>();
      f->field = parseField(word(is));
      f->shape = parseCube(is);
      f->name = word(is);
      f->out_dir = word(is);
      if (f->out_dir == "-") f->out_dir.clear();
      s.monitors.push_back(std::move(f));
    } else if (d == "voltage_monitor" || d == "current_monitor") {  // NAME CUBE DIR OUT_DIR
      std::unique_ptr<Monitor> mon;
      std::string name = word(is);
      auto cube = parseCube(is);
      const int dir = static_cast<int>(parseDir(word(is)));
      if (d == "voltage_monitor") {
        auto m = std::make_unique<VoltageMonitor>();
        m->shape = std::move(cube);
        m->dir = dir;
        mon = std::move(m);
      } else {
        auto m = std::make_unique<CurrentMonitor>();
        m->shape = std::move(cube);
        m->dir = dir;
        mon = std::move(m);
      }
      mon->name = name;
      mon->out_dir = word(is);
      if (mon->out_dir == "-") mon->out_dir.clear();
      s.monitors.push_back(std::move(mon));
    } else if (d == "nf2ff_td") {  // DX DY DZ THETA PHI OUT_DIR
      auto n = std::make_unique<NF2FFTD>();
      for (auto& v : n->dist) v = static_cast<Index>(num(is));
      n->theta = num(is);
      n->phi = num(is);
      n->out_dir = word(is);
      if (n->out_dir == "-") n->out_dir.clear();
      s.nf2ff_tds.push_back(std::move(n));
    } else if (d == "network") {  // NAME OUT_DIR NF F.. NP (INDEX SOURCE ZRE ZIM CURRENT VOLTAGE)..
      word(is);
      auto nw = std::make_unique<Network>();
      nw->out_dir = word(is);
      if (nw->out_dir == "-") nw->out_dir.clear();
      const Index nf = static_cast<Index>(num(is));
      for (Index q = 0; q < nf; ++q) nw->freqs.push_back(num(is));
      const Index np = static_cast<Index>(num(is));
      auto byName = [&](const std::string& n) -> Monitor* {
        for (auto& m : s.monitors)
          if (m->name == n) return m.get();
        throw Error("network: no monitor " + n);
      };
      for (Index q = 0; q < np; ++q) {
        Port p;
        p.index = static_cast<Index>(num(is));
        p.is_source = num(is) != 0;
        const Real zr = num(is), zi = num(is);
        p.z = {zr, zi};
        p.current = dynamic_cast<CurrentMonitor*>(byName(word(is)));
        p.voltage = dynamic_cast<VoltageMonitor*>(byName(word(is)));
        if (!p.current || !p.voltage) throw Error("network: port needs a current and a voltage monitor");
        nw->ports.push_back(p);
      }
      s.networks.push_back(std::move(nw));
    } else if (d == "field_time") {
      auto m = std::make_unique<FieldTimeMonitor>();
      m->name = word(is);
      m->field = parseField(word(is));
      m->shape = parseCube(is);
      m->out_dir = word(is);
      if (m->out_dir == "-") m->out_dir.clear();
      s.monitors.push_back(std::move(m));
    } else {
      throw Error("unknown directive " + d);
    }
  }
  if (!have_sim) throw Error("spec has no 'sim' line");
}

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void* ref_create(const char* spec) {
  auto h = std::make_unique<Handle>();
  if (guard([&] { parseSpec(h->sim, spec); }) != 0) return nullptr;
  return h.release();
}

void ref_destroy(void* p) { delete static_cast<Handle*>(p); }

int ref_set_write_files(void* p, int on) {
  static_cast<Handle*>(p)->sim.write_files = on != 0;
  return 0;
}

int ref_set_threads(void* p, int tx, int ty, int tz) {
  auto& s = static_cast<Handle*>(p)->sim;
  s.tcfg[0] = tx;
  s.tcfg[1] = ty;
  s.tcfg[2] = tz;
  return 0;
}

// Simulation::run(steps): init + loop + outputs.
int ref_run(void* p, int64_t steps) {
  return guard([&] { static_cast<Handle*>(p)->sim.run(static_cast<Index>(steps)); });
}

// The time loop after ref_init (fields as left by init or ref_write_field).
int ref_loop(void* p) {
  return guard([&] { static_cast<Handle*>(p)->sim.loop(); });
}

// Overwrites a field (reference extent, doubles) between init and loop.
int ref_write_field(void* p, int field, const double* src) {
  return guard([&] {
    if (field < 0 || field > 5) throw Error("bad field id");
    Arr3& a = static_cast<Handle*>(p)->sim.emf.byId(field);
    std::memcpy(a.v.data(), src, a.size() * sizeof(double));
  });
}

// Simulation::init(steps) alone, for inspecting the precompute.
int ref_init(void* p, int64_t steps) {
  return guard([&] { static_cast<Handle*>(p)->sim.initSteps(static_cast<Index>(steps)); });
}

double ref_dt(void* p) { return static_cast<Handle*>(p)->sim.dt; }
// The DFT objective after run(): J into *value, the point DFTs (interleaved
// complex) into dft when it is not null; returns the point count.
int ref_objective(void* p, double* value, double* dft) {
  auto& s = static_cast<Handle*>(p)->sim;
  if (!s.objective) return -1;
  if (value) *value = s.objective->value();
  if (dft) std::memcpy(dft, s.objective->dft.data(), s.objective->dft.size() * sizeof(Cplx));
  return static_cast<int>(s.objective->dft.size());
}

int ref_nf2ff_count(void* p) { return static_cast<int>(static_cast<Handle*>(p)->sim.nf2ffs.size()); }

// Time-domain NF2FF n: potential length, the raw potential `which` (ua ub
// wa wb) of surface d, and the far-field arrays (w_theta w_phi u_theta
// u_phi) of processFarField.
int64_t ref_nf2ff_td_num(void* p, int n) {
  int64_t r = -1;
  guard([&] { r = static_cast<int64_t>(static_cast<Handle*>(p)->sim.nf2ff_tds.at(n)->num); });
  return r;
}
int ref_nf2ff_td_potential(void* p, int n, int d, int which, double* dst) {
  return guard([&] {
    const auto& v = static_cast<Handle*>(p)->sim.nf2ff_tds.at(n)->pot[d][which];
    std::memcpy(dst, v.data(), v.size() * sizeof(double));
  });
}
int ref_nf2ff_td_far_field(void* p, int n, double* dst) {
  return guard([&] {
    auto* h = static_cast<Handle*>(p);
    std::vector<Real> out[4];
    h->sim.nf2ff_tds.at(n)->farField(h->sim, out);
    for (int w = 0; w < 4; ++w) std::memcpy(dst + w * out[w].size(), out[w].data(), out[w].size() * sizeof(double));
  });
}

// Network n: Network::output on the last run's recordings, then S(j, i)
// by key 10 j + i as nf interleaved (re, im) doubles.
int ref_network_output(void* p, int n) {
  return guard([&] { static_cast<Handle*>(p)->sim.networks.at(n)->output(); });
}
int ref_network_s(void* p, int n, int key, double* dst) {
  return guard([&] {
    const auto& s = static_cast<Handle*>(p)->sim.networks.at(n)->s;
    auto it = s.find(key);
    if (it == s.end()) throw Error("no S-parameter " + std::to_string(key));
    std::memcpy(dst, it->second.data(), it->second.size() * sizeof(std::complex<Real>));
  });
}

// The restated dft and power waves on given series, for the closed-form
// checks: out = nf complex values (interleaved); a / b likewise.
int ref_series_dft(const double* x, int64_t n, double dt, const double* freqs, int64_t nf, double shift,
                   double* out) {
  return guard([&] {
    auto r = seriesDft(std::vector<Real>(x, x + n), dt, std::vector<Real>(freqs, freqs + nf), shift);
    std::memcpy(out, r.data(), r.size() * sizeof(std::complex<Real>));
  });
}
int ref_port_waves(const double* current, const double* voltage, int64_t n, double dt, const double* freqs,
                   int64_t nf, double z_re, double z_im, double* a, double* b) {
  return guard([&] {
    CurrentMonitor cm;
    VoltageMonitor vm;
    cm.data.v.assign(current, current + n);
    vm.data.v.assign(voltage, voltage + n);
    Port port;
    port.z = {z_re, z_im};
    port.current = &cm;
    port.voltage = &vm;
    port.dt = dt;
    port.calculateSParameters(std::vector<Real>(freqs, freqs + nf));
    std::memcpy(a, port.a.data(), port.a.size() * sizeof(std::complex<Real>));
    std::memcpy(b, port.b.data(), port.b.size() * sizeof(std::complex<Real>));
  });
}

// Surface task of NF2FF `n`, surface d (XN XP YN YP ZN ZP): lo[3], hi[3]
// (hi == lo when the surface is not used).
int ref_nf2ff_surface(void* p, int n, int d, int64_t* lo, int64_t* hi) {
  return guard([&] {
    auto& s = static_cast<Handle*>(p)->sim;
    const Task& t = s.nf2ffs.at(n)->surf[d];
    for (int a = 0; a < 3; ++a) {
      lo[a] = static_cast<int64_t>(t.r(a).s);
      hi[a] = static_cast<int64_t>(t.valid() ? t.r(a).e : t.r(a).s);
    }
  });
}

// Accumulated surface currents of frequency f, surface d, which = 0 ja,
// 1 jb, 2 ma, 3 mb; interleaved (re, im) doubles.
int ref_nf2ff_current(void* p, int n, int f, int d, int which, double* out) {
  return guard([&] {
    auto& pl = static_cast<Handle*>(p)->sim.nf2ffs.at(n)->planes.at(f);
    const auto& v = which == 0 ? pl.ja[d] : which == 1 ? pl.jb[d] : which == 2 ? pl.ma[d] : pl.mb[d];
    for (size_t q = 0; q < v.size(); ++q) {
      out[2 * q] = v[q].real();
      out[2 * q + 1] = v[q].imag();
    }
  });
}

// aTheta / aPhi (a_pot = 1) or fTheta / fPhi (a_pot = 0) over theta x phi
// (one of them of size 1); out: interleaved complex.
int ref_nf2ff_potential(void* p, int n, int f, int a_pot, int theta_comp, int nt, const double* theta, int np,
                        const double* phi, const double* origin, double* out) {
  return guard([&] {
    auto& s = static_cast<Handle*>(p)->sim;
    std::vector<Real> th(theta, theta + nt), ph(phi, phi + np);
    auto v = s.nf2ffs.at(n)->potential(s, f, a_pot != 0, theta_comp != 0, th, ph, {origin[0], origin[1], origin[2]});
    for (size_t q = 0; q < v.size(); ++q) {
      out[2 * q] = v[q].real();
      out[2 * q + 1] = v[q].imag();
    }
  });
}

int ref_nf2ff_power(void* p, int n, int f, double* out) {
  return guard([&] {
    auto& s = static_cast<Handle*>(p)->sim;
    *out = s.nf2ffs.at(n)->power(s, f);
  });
}

double ref_loop_seconds(void* p) { return static_cast<Handle*>(p)->sim.loop_seconds; }
// Copies up to `cap` per-step end times (seconds since the loop start) of
// the last loop; returns how many there are.
int64_t ref_step_end_seconds(void* p, double* out, int64_t cap) {
  const auto& v = static_cast<Handle*>(p)->sim.step_end_seconds;
  for (int64_t i = 0; i < cap && i < static_cast<int64_t>(v.size()); ++i) out[i] = v[i];
  return static_cast<int64_t>(v.size());
}

int ref_grid(void* p, int64_t* n3) {
  auto& g = static_cast<Handle*>(p)->sim.g;
  n3[0] = static_cast<int64_t>(g.nx());
  n3[1] = static_cast<int64_t>(g.ny());
  n3[2] = static_cast<int64_t>(g.nz());
  return 0;
}

// Node arrays: which 0=e_node 1=h_node 2=e_size 3=h_size; returns length.
int64_t ref_grid_array(void* p, int which, int axis, double* dst) {
  auto& g = static_cast<Handle*>(p)->sim.g;
  const Arr1* a = which == 0 ? &g.e_node[axis] : which == 1 ? &g.h_node[axis]
                : which == 2 ? &g.e_size[axis] : &g.h_size[axis];
  if (dst) std::memcpy(dst, a->v.data(), a->size() * sizeof(double));
  return static_cast<int64_t>(a->size());
}

static int copy3(const Arr3& a, int64_t* shape, double* dst) {
  shape[0] = static_cast<int64_t>(a.n0);
  shape[1] = static_cast<int64_t>(a.n1);
  shape[2] = static_cast<int64_t>(a.n2);
  if (dst) std::memcpy(dst, a.v.data(), a.size() * sizeof(double));
  return 0;
}

// field: 0..5 = Ex Ey Ez Hx Hy Hz.  dst may be null to query the shape.
int ref_field(void* p, int field, int64_t* shape, double* dst) {
  if (field < 0 || field > 5) {
    g_err = "bad field id";
    return -1;
  }
  return copy3(static_cast<Handle*>(p)->sim.emf.byId(field), shape, dst);
}

int ref_coeff(void* p, const char* name, int64_t* shape, double* dst) {
  Arr3* a = static_cast<Handle*>(p)->sim.c.byName(name);
  if (!a) {
    g_err = std::string("unknown coefficient ") + name;
    return -1;
  }
  return copy3(*a, shape, dst);
}

// Material parameter arrays after averaging: kind 0=eps 1=mu 2=sigma_e 3=sigma_m.
int ref_material(void* p, int kind, int comp, int64_t* shape, double* dst) {
  auto& mp = static_cast<Handle*>(p)->sim.mp;
  const Arr3& a = kind == 0 ? mp.eps[comp] : kind == 1 ? mp.mu[comp] : kind == 2 ? mp.se[comp] : mp.sm[comp];
  return copy3(a, shape, dst);
}

// PML per-layer coefficients of boundary `b`: which 0=a_e 1=b_e 2=a_h 3=b_h
// 4=kappa_e 5=kappa_h.
int64_t ref_pml_profile(void* p, int b, int which, double* dst) {
  auto& s = static_cast<Handle*>(p)->sim;
  if (b < 0 || b >= static_cast<int>(s.pmls.size())) return -1;
  auto& pm = *s.pmls[b];
  const Arr1* a[6] = {&pm.a_e, &pm.b_e, &pm.a_h, &pm.b_h, &pm.kappa_e, &pm.kappa_h};
  if (dst) std::memcpy(dst, a[which]->v.data(), a[which]->size() * sizeof(double));
  return static_cast<int64_t>(a[which]->size());
}

// TFSF incident tables of source `i`: which 0 = e_inc, 1 = h_inc.
int ref_tfsf_table(void* p, int i, int which, int64_t* shape, double* dst) {
  auto& s = static_cast<Handle*>(p)->sim;
  if (i < 0 || i >= static_cast<int>(s.tfsfs.size())) return -1;
  const Arr2& a = which == 0 ? s.tfsfs[i]->e_inc : s.tfsfs[i]->h_inc;
  shape[0] = static_cast<int64_t>(a.n0);
  shape[1] = static_cast<int64_t>(a.n1);
  if (dst) std::memcpy(dst, a.v.data(), a.v.size() * sizeof(double));
  return 0;
}

int ref_monitor_count(void* p) {
  return static_cast<int>(static_cast<Handle*>(p)->sim.monitors.size());
}

// Records of monitor m: a movie has one record per frame; a field-time
// monitor has one record (time row and value row, shape (2, nt)).
int ref_monitor_records(void* p, int m) {
  auto& s = static_cast<Handle*>(p)->sim;
  Monitor* mo = s.monitors.at(m).get();
  if (auto mv = dynamic_cast<MovieMonitor*>(mo)) return static_cast<int>(mv->frames.size());
  if (dynamic_cast<FieldTimeMonitor*>(mo) || dynamic_cast<VoltageMonitor*>(mo) ||
      dynamic_cast<CurrentMonitor*>(mo) || dynamic_cast<FieldMonitor*>(mo))
    return 1;
  return 0;
}

int ref_monitor_record(void* p, int m, int r, int64_t* shape, double* dst) {
  auto& s = static_cast<Handle*>(p)->sim;
  Monitor* mo = s.monitors.at(m).get();
  if (auto mv = dynamic_cast<MovieMonitor*>(mo)) {
    const Arr3& a = mv->frames.at(r).second;
    return copy3(a, shape, dst);
  }
  if (auto fm = dynamic_cast<FieldMonitor*>(mo)) {  // output() after the run: the box of the fields
    fm->output(s);
    return copy3(fm->data, shape, dst);
  }
  if (auto ft = dynamic_cast<FieldTimeMonitor*>(mo)) {
    Index nt = ft->time.size();
    shape[0] = 2;
    shape[1] = static_cast<int64_t>(nt);
    shape[2] = 1;
    if (dst) {
      std::memcpy(dst, ft->time.v.data(), nt * sizeof(double));
      std::memcpy(dst + nt, ft->data.v.data(), nt * sizeof(double));
    }
    return 0;
  }
  const Arr1* time = nullptr;
  const Arr1* data = nullptr;
  if (auto vm = dynamic_cast<VoltageMonitor*>(mo)) {
    time = &vm->time;
    data = &vm->data;
  } else if (auto cm = dynamic_cast<CurrentMonitor*>(mo)) {
    time = &cm->time;
    data = &cm->data;
  }
  if (time) {
    const Index nt = time->size();
    shape[0] = 2;
    shape[1] = static_cast<int64_t>(nt);
    shape[2] = 1;
    if (dst) {
      std::memcpy(dst, time->v.data(), nt * sizeof(double));
      std::memcpy(dst + nt, data->v.data(), nt * sizeof(double));
    }
    return 0;
  }
  g_err = "monitor has no records";
  return -1;
}

}  // extern "C"
