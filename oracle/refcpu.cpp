// refcpu — FP64 CPU restatement of the XFDTD_CORE 3D time-stepping path.
// TEST INFRASTRUCTURE ONLY (oracle + "port" CPU baseline); see refcpu.hpp.
// Citations are file:line in the reference repository.
#include "refcpu.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <limits>
#include <map>
#include <sstream>
#include <thread>

namespace refcpu {

// ------------------------------------------------------------- helpers ---
Arr1 linspace(Real start, Real stop, Index n) {
  // xbuilder.hpp linspace: step = (stop - start) / max(1, n - 1); the
  // arange generator assigns by running sum and pins the last sample.
  Arr1 out(n);
  Real step = (stop - start) / std::fmax(Real(1), Real(n - 1));
  Real v = start;
  for (Index i = 0; i < n; ++i) {
    out(i) = v;
    v += step;
  }
  if (n > 1) out(n - 1) = stop;
  return out;
}

bool intersected(const Range& d, const Range& r) {  // index_task.h:84-87
  return !(r.e <= d.s || d.e <= r.s);
}
bool intersected(const Task& a, const Task& b) {
  return intersected(a.x, b.x) && intersected(a.y, b.y) && intersected(a.z, b.z);
}
std::optional<Task> intersection(const Task& a, const Task& b) {  // :97-115
  if (!intersected(a, b)) return std::nullopt;
  auto f = [](Range p, Range q) {
    return Range{std::max(p.s, q.s), std::min(p.e, q.e)};
  };
  return Task{f(a.x, b.x), f(a.y, b.y), f(a.z, b.z)};
}

static Range decomposeSize(int rank, int np, Index size) {  // decompose_task.h:9-33
  if (np <= 0) throw Error("Number of processes is less than or equal to zero");
  if (size < static_cast<Index>(np))
    throw Error("Problem size is less than number of processes");
  Index q = size / np, r = size % np;
  if (static_cast<Index>(rank) < r) {
    Index s = rank * (q + 1);
    return {s, s + q + 1};
  }
  Index s = rank * q + r;
  return {s, s + q};
}

std::vector<Task> decomposeTask(const Task& p, int nx, int ny, int nz) {
  // decompose_task.h:69-93 (row major)
  if (nx <= 0 || ny <= 0 || nz <= 0) return {p};
  auto dr = [](Range r, int np) {
    std::vector<Range> out;
    if (r.e <= r.s) throw Error("Invalid range");
    for (int i = 0; i < np; ++i) {
      auto t = decomposeSize(i, np, r.e - r.s);
      out.push_back({t.s + r.s, t.e + r.s});
    }
    return out;
  };
  auto xs = dr(p.x, nx), ys = dr(p.y, ny), zs = dr(p.z, nz);
  std::vector<Task> res;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) res.push_back({xs[i], ys[j], zs[k]});
  return res;
}

// ---------------------------------------------------------------- shapes --
Cube::Cube(Vec3 o, Vec3 s) : origin(o), size(s), end(o + s) {
  // cube.cpp:128-139: a NaN end (inf - inf) becomes +inf
  if (std::isnan(end.x)) end.x = std::numeric_limits<Real>::infinity();
  if (std::isnan(end.y)) end.y = std::numeric_limits<Real>::infinity();
  if (std::isnan(end.z)) end.z = std::numeric_limits<Real>::infinity();
}

bool Cube::isInside(const Vec3& p, Real eps) const {
  bool xi = lessEqual(origin.x, p.x, eps) && lessEqual(p.x, end.x, eps);
  bool yi = lessEqual(origin.y, p.y, eps) && lessEqual(p.y, end.y, eps);
  bool zi = lessEqual(origin.z, p.z, eps) && lessEqual(p.z, end.z, eps);
  return xi && yi && zi;
}

bool Sphere::isInside(const Vec3& p, Real eps) const {
  return lessEqual((p - center).norm(), radius, eps);
}

Cube Sphere::wrapped() const {
  return Cube(center - Vec3{radius, radius, radius},
              Vec3{radius * 2, radius * 2, radius * 2});
}

Cube Cylinder::wrapped() const {  // cylinder.cpp:73-98
  Real s = radius * 2;
  if (axis == X)
    return Cube(Vec3{center.x - height / 2, center.y - radius, center.z - radius},
                Vec3{height, s, s});
  if (axis == Y)
    return Cube(Vec3{center.x - radius, center.y - height / 2, center.z - radius},
                Vec3{s, height, s});
  return Cube(Vec3{center.x - radius, center.y - radius, center.z - height / 2},
              Vec3{s, s, height});
}

bool Cylinder::isInside(const Vec3& p, Real eps) const {  // cylinder.cpp:41-66
  if (!wrapped().isInside(p, eps)) return false;
  Real d;
  if (axis == X)
    d = std::sqrt(std::pow(p.y - center.y, 2) + std::pow(p.z - center.z, 2));
  else if (axis == Y)
    d = std::sqrt(std::pow(p.x - center.x, 2) + std::pow(p.z - center.z, 2));
  else
    d = std::sqrt(std::pow(p.x - center.x, 2) + std::pow(p.y - center.y, 2));
  return lessEqual(d, radius, 1e-8);  // default epsilon of floatCompare
}

// ------------------------------------------------------------- waveform ---
void Waveform::init(const Arr1& time) {
  value = Arr1(time.size());
  for (Index i = 0; i < time.size(); ++i) value(i) = amplitude * f(time(i));
}

Material Material::make(Real er, Real mr, Real se, Real sm) {
  Material m;
  m.eps_r = er;
  m.mu_r = mr;
  m.sigma_e = se == 0 ? SIGMA_ZERO : se;
  m.sigma_m = sm == 0 ? SIGMA_ZERO : sm;
  m.eps = er * EPSILON_0;
  m.mu = mr * MU_0;
  return m;
}

// ----------------------------------------------------------------- grid ---
Real GridSpace::eps() const {  // grid_space.cpp:459-468
  if (dim == 1) return min_d[2] / 1000;
  if (dim == 2) return std::min(min_d[0], min_d[1]) / 1000;
  return std::min({min_d[0], min_d[1], min_d[2]}) / 1000;
}

Index GridSpace::nearest(int axis, Real x) const {
  // The one-cell axes of 1D/2D grids: handleTransformZ of GridSpace2D
  // (grid_space_2d.cpp:116-122) and handleTransformX/Y of GridSpace1D
  // (grid_space_1d.cpp:117-135) give 0 below +inf, else 1.
  if ((dim == 2 && axis == 2) || (dim == 1 && axis < 2))
    return x < std::numeric_limits<Real>::infinity() ? 0 : 1;
  // xt::argmin(xt::abs(e_node - x)): first index of the minimum
  const Arr1& e = e_node[axis];
  Index best = 0;
  Real bv = std::abs(e(0) - x);
  for (Index i = 1; i < e.size(); ++i) {
    Real v = std::abs(e(i) - x);
    if (v < bv) {
      bv = v;
      best = i;
    }
  }
  return best;
}

std::array<Index, 3> GridSpace::box(const Shape& s, Index* end) const {
  Cube c = s.wrapped();
  std::array<Index, 3> o{nearest(0, c.origin.x), nearest(1, c.origin.y),
                         nearest(2, c.origin.z)};
  end[0] = nearest(0, c.end.x);
  end[1] = nearest(1, c.end.y);
  end[2] = nearest(2, c.end.z);
  return o;
}

Material Material::makeLorentz(Real eps_inf, std::vector<Real> a0, std::vector<Real> a1, std::vector<Real> b0,
                               std::vector<Real> b1, std::vector<Real> b2) {
  // MLorentzEqDecision: all five parameter arrays have the pole count
  if (a1.size() != a0.size() || b0.size() != a0.size() || b1.size() != a0.size() || b2.size() != a0.size())
    throw Error("Modified Lorentz Model paramater doesn't have same size.");
  Material m = make(1, 1, 0, 0);  // ElectroMagneticProperty::air()
  m.kind = Lorentz;
  m.eps_inf = eps_inf;
  m.a0 = std::move(a0);
  m.a1 = std::move(a1);
  m.b0 = std::move(b0);
  m.b1 = std::move(b1);
  m.b2 = std::move(b2);
  return m;
}

Material Material::makeDrude(Real eps_inf, std::vector<Real> omega_p, std::vector<Real> gamma) {
  // DrudeEqDecision: a0 = wp*wp, a1 = 0, b0 = 0, b1 = gamma, b2 = 1
  std::vector<Real> a0(omega_p.size()), zero(omega_p.size(), 0.0), one(omega_p.size(), 1.0);
  for (Index p = 0; p < omega_p.size(); ++p) a0[p] = omega_p[p] * omega_p[p];
  Material m = makeLorentz(eps_inf, a0, zero, zero, gamma, one);
  m.kind = Drude;
  m.omega_p = std::move(omega_p);
  m.gamma = std::move(gamma);
  return m;
}

Material Material::makeDebye(Real eps_inf, std::vector<Real> eps_static, std::vector<Real> tau) {
  // DebyeEqDecision: a0 = eps_s - eps_inf, a1 = 0, b0 = 1, b1 = tau, b2 = 0
  std::vector<Real> a0(eps_static.size()), zero(eps_static.size(), 0.0), one(eps_static.size(), 1.0);
  for (Index p = 0; p < eps_static.size(); ++p) a0[p] = eps_static[p] - eps_inf;
  Material m = makeLorentz(eps_inf, a0, zero, one, tau, zero);
  m.kind = Debye;
  m.eps_static = std::move(eps_static);
  m.tau = std::move(tau);
  return m;
}

// ---------------------------------------------------------------- ADE --
ADEStorage::ADEStorage(Kind k, Index p, Index x, Index y, Index z) : kind(k), P(p), nx(x), ny(y), nz(z) {
  // the arrays each storage allocates (ade_method constructors)
  const Index n4 = x * y * z * p;
  cjj.assign(n4, 0);
  cje.assign(n4, 0);
  cjsj.assign(n4, 0);
  cej_sum = Arr3(x, y, z);
  for (int c = 0; c < 3; ++c) j[c].assign(n4, 0);
  if (k == MLorentz) {
    cjjp.assign(n4, 0);
    cjen.assign(n4, 0);
    cjep.assign(n4, 0);
    cjsjp.assign(n4, 0);
    cee_p = Arr3(x, y, z);
    for (int c = 0; c < 3; ++c) {
      e_prev[c] = Arr3(x, y, z);
      j_prev[c].assign(n4, 0);
    }
  }
}

void ADEStorage::correctCoeff(Sim& s, Index i, Index jj, Index k, const Material& m) {
  if (P < m.numPoles()) throw Error("The number of poles is not enough.");
  const Real dt = s.dt;
  const Real sigma = m.sigma_e;
  const Index np = m.numPoles();
  Real cece, b_over;  // cece and the numerator of cecha / cechb
  Real cecha_num, cechb_num;
  bool lor_form = false;
  if (kind == Debye || kind == Drude) {
    // Debye- or Drude-only storage loops over the storage's pole count
    // (debye_ade_method.cpp:68-72, drude_ade_method.cpp:79-83); a material
    // with fewer poles would be read past its arrays, so require equality.
    if (np != P) throw Error("Debye/Drude storage needs every material to have the storage's pole count");
    std::vector<Real> kk(P), beta(P);
    for (Index p = 0; p < P; ++p) {
      if (kind == Debye) {  // DebyeADEMethodCoefficient (debye_ade_method.cpp:9-31)
        const Real tau = m.b1[p], de = m.a0[p];
        kk[p] = (2 * tau - dt) / (2 * tau + dt);
        beta[p] = 2 * EPSILON_0 * de * dt / (2 * tau + dt);
      } else {  // DrudeADEMethodCoefficient (drude_ade_method.cpp:9-38)
        const Real g = m.b1[p], wp = std::sqrt(m.a0[p]);  // DrudeEqDecision::omegaP
        kk[p] = (1 - g * dt / 2) / (1 + g * dt / 2);
        beta[p] = (EPSILON_0 * wp * wp * dt / 2) / (1 + g * dt / 2);
      }
    }
    Real sum_beta = 0;
    for (Index p = 0; p < P; ++p) sum_beta += beta[p];
    for (Index p = 0; p < P; ++p) {
      cjj[at(i, jj, k, p)] = kk[p];
      cje[at(i, jj, k, p)] = kind == Debye ? beta[p] / dt : beta[p];
      cjsj[at(i, jj, k, p)] = 0.5 * (1 + cjj[at(i, jj, k, p)]);
    }
    const Real ei = m.eps_inf;
    Real a, b;
    if (kind == Debye) {
      a = (2 * EPSILON_0 * ei + sum_beta - dt * sigma) / (2 * EPSILON_0 * ei + sum_beta + dt * sigma);
      b = (2 * dt) / (2 * EPSILON_0 * ei + sum_beta + dt * sigma);
    } else {  // the reference's Drude a/b subtract dt*sum_beta in both places
      a = (2 * EPSILON_0 * ei - dt * sum_beta - dt * sigma) / (2 * EPSILON_0 * ei - dt * sum_beta + dt * sigma);
      b = (2 * dt) / (2 * EPSILON_0 * ei - dt * sum_beta + dt * sigma);
    }
    cej_sum(i, jj, k) = -b;
    cece = a;
    b_over = b;
    cecha_num = -b;
    cechb_num = b;
  } else {  // MLorentzADEMethodCoefficient (m_lor_ade_method.cpp:8-39, 66-159)
    lor_form = true;
    std::vector<Real> A(np), B(np), C(np), D(np), E(np), F(np);
    for (Index p = 0; p < np; ++p) {
      const Real a0 = m.a0[p], a1 = m.a1[p], b0 = m.b0[p], b1 = m.b1[p], b2 = m.b2[p];
      A[p] = (a1 * EPSILON_0 / (dt * dt)) + (0.5 * a0 * EPSILON_0 / dt);
      B[p] = (-2 * a1 * EPSILON_0) / (dt * dt);
      C[p] = (a1 * EPSILON_0 / (dt * dt)) - (0.5 * a0 * EPSILON_0 / dt);
      D[p] = (2 * b2 / (dt * dt)) - b0;
      E[p] = (-b2 / (dt * dt)) + (b1 / (2 * dt));
      F[p] = (b2 / (dt * dt)) + (b1 / (2 * dt));
    }
    for (Index p = 0; p < np; ++p) {
      cjen[at(i, jj, k, p)] = A[p] / F[p];
      cje[at(i, jj, k, p)] = B[p] / F[p];
      cjep[at(i, jj, k, p)] = C[p] / F[p];
      cjj[at(i, jj, k, p)] = D[p] / F[p];
      cjjp[at(i, jj, k, p)] = E[p] / F[p];
      cjsj[at(i, jj, k, p)] = 0.5 * (1 + D[p] / F[p]);
      cjsjp[at(i, jj, k, p)] = 0.5 * (E[p] / F[p]);
    }
    Real sum_a_f = 0, sum_b_f = 0, sum_c_f = 0;
    for (Index p = 0; p < np; ++p) sum_a_f += A[p] / F[p];
    for (Index p = 0; p < np; ++p) sum_b_f += B[p] / F[p];
    for (Index p = 0; p < np; ++p) sum_c_f += C[p] / F[p];
    const Real ei = m.eps_inf;
    const Real temp = (ei * EPSILON_0) / dt + 0.5 * sum_a_f + 0.5 * sigma;
    cej_sum(i, jj, k) = -1 / temp;
    cee_p(i, jj, k) = -0.5 * sum_c_f / temp;
    cece = ((ei * EPSILON_0) / dt - 0.5 * sum_b_f - 0.5 * sigma) / temp;
    b_over = temp;
    cecha_num = -1;
    cechb_num = 1;
  }
  // cece / cecha / cechb of all three components at (i, j, k), with the
  // h-node sizes of the cell; (da, db) = (dy, dz), (dz, dx), (dx, dy)
  const Real d[3] = {s.g.h_size[0](i), s.g.h_size[1](jj), s.g.h_size[2](k)};
  for (int c = 0; c < 3; ++c) {
    const Real da = d[(c + 1) % 3], db = d[(c + 2) % 3];
    s.c.self(0, c)(i, jj, k) = cece;
    if (lor_form) {
      s.c.dualA(0, c)(i, jj, k) = cecha_num / (b_over * db);
      s.c.dualB(0, c)(i, jj, k) = cechb_num / (b_over * da);
    } else {
      s.c.dualA(0, c)(i, jj, k) = cecha_num / db;
      s.c.dualB(0, c)(i, jj, k) = cechb_num / da;
    }
  }
}

// ---------------------------------------------------------- coefficients --
Arr3& Coeff::self(int h, int c) {
  if (!h) return c == 0 ? cexe : c == 1 ? ceye : ceze;
  return c == 0 ? chxh : c == 1 ? chyh : chzh;
}
// coeff<attr, xyz, dual, tangentialA(xyz)> (fdtd_update_coefficient.h)
Arr3& Coeff::dualA(int h, int c) {
  if (!h) return c == 0 ? cexhy : c == 1 ? ceyhz : cezhx;
  return c == 0 ? chxey : c == 1 ? chyez : chzex;
}
Arr3& Coeff::dualB(int h, int c) {
  if (!h) return c == 0 ? cexhz : c == 1 ? ceyhx : cezhy;
  return c == 0 ? chxez : c == 1 ? chyex : chzey;
}
Arr3* Coeff::byName(const std::string& n) {
  static const char* names[18] = {"cexe", "cexhy", "cexhz", "ceye", "ceyhz", "ceyhx",
                                  "ceze", "cezhx", "cezhy", "chxh", "chxey", "chxez",
                                  "chyh", "chyez", "chyex", "chzh", "chzex", "chzey"};
  Arr3* arrs[18] = {&cexe, &cexhy, &cexhz, &ceye, &ceyhz, &ceyhx, &ceze, &cezhx, &cezhy,
                    &chxh, &chxey, &chxez, &chyh, &chyez, &chyex, &chzh, &chzex, &chzey};
  for (int i = 0; i < 18; ++i)
    if (n == names[i]) return arrs[i];
  return nullptr;
}

// -------------------------------------------------------------- objects ---
void Object::init(Sim& s) {
  gb_origin = s.g.box(*shape, gb_end.data());
}

void Object::correctMaterialSpace(Sim& s, Index index) {
  // object.cpp:114-198: every cell whose centre (h-node) is inside the shape
  // takes the object's properties; later objects overwrite earlier ones.
  // Each cell is tested and written independently, so the i planes are
  // split across host threads (the reference's XFDTD_CORE_PSTL_ENABLE
  // branch is the same per-cell lambda under std::execution::par_unseq).
  material_index = index;
  const Real eps = s.g.eps();
  const Index nx = s.g.nx(), ny = s.g.ny(), nz = s.g.nz();
  auto& mp = s.mp;
  auto planes = [&](Index i0, Index i1) {
    for (Index i = i0; i < i1; ++i)
      for (Index j = 0; j < ny; ++j)
        for (Index k = 0; k < nz; ++k) {
          Vec3 p{s.g.h_node[0](i), s.g.h_node[1](j), s.g.h_node[2](k)};
          if (!shape->isInside(p, eps)) continue;
          s.material_grid[(i * ny + j) * nz + k] = index;
          for (int a = 0; a < 3; ++a) {
            mp.eps[a](i, j, k) = material.eps;
            mp.mu[a](i, j, k) = material.mu;
            mp.se[a](i, j, k) = material.sigma_e;
            mp.sm[a](i, j, k) = material.sigma_m;
          }
        }
  };
  const Index parts = std::min<Index>(nx, std::max(1u, std::min(32u, std::thread::hardware_concurrency())));
  if (parts <= 1 || nx * ny * nz < (Index(1) << 18)) {
    planes(0, nx);
    return;
  }
  std::vector<std::thread> th;
  const Index step = (nx + parts - 1) / parts;
  for (Index p = 0; p < parts; ++p) {
    const Index i0 = p * step, i1 = std::min(nx, i0 + step);
    if (i0 < i1) th.emplace_back(planes, i0, i1);
  }
  for (auto& t : th) t.join();
}

Index LumpedElement::nodeCount(int a) const {
  Index s = a == 0 ? is : a == 1 ? js : ks;
  Index e = a == 0 ? ie : a == 1 ? je : ke;
  return a == static_cast<int>(xyz) ? e - s : e - s + 1;
}
Index LumpedElement::globalCountMain() const {
  return xyz == X ? ie - is : xyz == Y ? je - js : ke - ks;
}
Index LumpedElement::globalCountSubA() const {
  return xyz == X ? je - js + 1 : xyz == Y ? ke - ks + 1 : ie - is + 1;
}
Index LumpedElement::globalCountSubB() const {
  return xyz == X ? ke - ks + 1 : xyz == Y ? ie - is + 1 : je - js + 1;
}
Range LumpedElement::range(int a) const {
  Index s = a == 0 ? is : a == 1 ? js : ks;
  Index e = a == 0 ? ie : a == 1 ? je : ke;
  return a == static_cast<int>(xyz) ? Range{s, e} : Range{s, e + 1};
}
Task LumpedElement::indexTask() const {  // lumped_element.cpp:154-158
  return {{is, is + nodeCount(0)}, {js, js + nodeCount(1)}, {ks, ks + nodeCount(2)}};
}
void LumpedElement::init(Sim& s) {
  Object::init(s);
  is = gb_origin[0];
  js = gb_origin[1];
  ks = gb_origin[2];
  ie = gb_end[0];
  je = gb_end[1];
  ke = gb_end[2];
}

void LumpedElement::sizes(Sim& s, Arr3& da, Arr3& db, Arr3& dc) const {
  // xt::meshgrid(view(x), view(y), view(z)) over rangeX/Y/Z, with the main
  // axis taking e_size and the others h_size (current_source.cpp:55-105).
  Range r[3] = {range(0), range(1), range(2)};
  const Arr1* sz[3];
  for (int a = 0; a < 3; ++a) sz[a] = a == static_cast<int>(xyz) ? &s.g.e_size[a] : &s.g.h_size[a];
  Arr3 m[3];
  for (int a = 0; a < 3; ++a) m[a] = Arr3(r[0].size(), r[1].size(), r[2].size());
  for (Index i = 0; i < r[0].size(); ++i)
    for (Index j = 0; j < r[1].size(); ++j)
      for (Index k = 0; k < r[2].size(); ++k) {
        m[0](i, j, k) = (*sz[0])(r[0].s + i);
        m[1](i, j, k) = (*sz[1])(r[1].s + j);
        m[2](i, j, k) = (*sz[2])(r[2].s + k);
      }
  // (da, db, dc) = sizes along (tanA, tanB, main)
  da = m[tanA(xyz)];
  db = m[tanB(xyz)];
  dc = m[xyz];
}

// Applies the lumped-element coefficient patch (resistor.cpp:86-98,
// voltage_source.cpp:109-132, current_source.cpp:111-131) over the element's
// view, and fills `coff` with -2dt/(den*scale)*factor when requested.
static void lumpedPatch(Sim& s, const LumpedElement& le, const Arr3& beta, const Arr3& da,
                        const Arr3& db, const Arr3* coff_scale, bool coff_is_alpha,
                        Real factor, Arr3* coff) {
  const Real dt = s.dt;
  int c = le.xyz;
  Arr3& cece = s.c.self(0, c);
  Arr3& cecha = s.c.dualA(0, c);
  Arr3& cechb = s.c.dualB(0, c);
  Arr3& eps = s.mp.eps[c];
  Arr3& sig = s.mp.se[c];
  Range r[3] = {le.range(0), le.range(1), le.range(2)};
  if (coff) *coff = Arr3(r[0].size(), r[1].size(), r[2].size());
  for (Index i = 0; i < r[0].size(); ++i)
    for (Index j = 0; j < r[1].size(); ++j)
      for (Index k = 0; k < r[2].size(); ++k) {
        Index gi = r[0].s + i, gj = r[1].s + j, gk = r[2].s + k;
        Real e = eps(gi, gj, gk), sg = sig(gi, gj, gk), b = beta(i, j, k);
        cece(gi, gj, gk) = (2 * e - dt * sg - b) / (2 * e + dt * sg + b);
        cecha(gi, gj, gk) = -2 * dt / ((2 * e + dt * sg + b) * db(i, j, k));
        cechb(gi, gj, gk) = 2 * dt / ((2 * e + dt * sg + b) * da(i, j, k));
        if (coff) {
          Real scale = coff_is_alpha ? (*coff_scale)(i, j, k)
                                     : (db(i, j, k) * da(i, j, k));
          (*coff)(i, j, k) = -2 * dt / ((2 * e + dt * sg + b) * scale) * factor;
        }
      }
}

void Resistor::init(Sim& s) {  // resistor.cpp:31-83
  LumpedElement::init(s);
  rf = resistance * globalCountSubA() * globalCountSubB() / globalCountMain();
  sizes(s, da, db, dc);
}

void Resistor::correctUpdateCoefficient(Sim& s) {  // resistor.cpp:85-127
  beta = Arr3(dc.n0, dc.n1, dc.n2);
  for (Index n = 0; n < beta.size(); ++n)
    beta.v[n] = s.dt * dc.v[n] / (rf * da.v[n] * db.v[n]);
  lumpedPatch(s, *this, beta, da, db, nullptr, false, 0, nullptr);
}

void Capacitor::init(Sim& s) {  // capacitor.cpp:16-59
  LumpedElement::init(s);
  // cf(c, na, nb, nc) = c * nc / (nc * nb), the count product in integers
  const Index na = globalCountSubA(), nb = globalCountSubB(), nc = globalCountMain();
  (void)na;
  cf = capacitance * nc / (nc * nb);
  sizes(s, da, db, dc);
  beta = Arr3(dc.n0, dc.n1, dc.n2);
  for (Index n = 0; n < beta.size(); ++n) beta.v[n] = 2 * cf * dc.v[n] / (da.v[n] * db.v[n]);
}

void Capacitor::correctUpdateCoefficient(Sim& s) {  // capacitor.cpp:61-112
  const Real dt = s.dt;
  const int c = xyz;
  Arr3& cece = s.c.self(0, c);
  Arr3& cecha = s.c.dualA(0, c);
  Arr3& cechb = s.c.dualB(0, c);
  const Arr3& eps = s.mp.eps[c];
  const Arr3& sig = s.mp.se[c];
  Range r[3] = {range(0), range(1), range(2)};
  for (Index i = 0; i < r[0].size(); ++i)
    for (Index j = 0; j < r[1].size(); ++j)
      for (Index k = 0; k < r[2].size(); ++k) {
        const Index gi = r[0].s + i, gj = r[1].s + j, gk = r[2].s + k;
        const Real e = eps(gi, gj, gk), sg = sig(gi, gj, gk), b = beta(i, j, k);
        cece(gi, gj, gk) = (2 * e - dt * sg + b) / (2 * e + dt * sg + b);
        cecha(gi, gj, gk) = -2 * dt / ((2 * e + dt * sg + b) * db(i, j, k));
        cechb(gi, gj, gk) = 2 * dt / ((2 * e + dt * sg + b) * da(i, j, k));
      }
}

void Inductor::init(Sim& s) {  // inductor.cpp:23-74
  LumpedElement::init(s);
  lf = inductance * globalCountSubA() * globalCountSubB() / globalCountMain();
  sizes(s, da, db, dc);
  cjcec = Arr3(dc.n0, dc.n1, dc.n2);
  for (Index n = 0; n < cjcec.size(); ++n) cjcec.v[n] = s.dt * dc.v[n] / (lf * da.v[n] * db.v[n]);
  j = Arr3(nodeCount(0), nodeCount(1), nodeCount(2));
}

void Inductor::correctUpdateCoefficient(Sim& s) {  // inductor.cpp:76-105
  const Real dt = s.dt;
  const int c = xyz;
  const Arr3& eps = s.mp.eps[c];
  const Arr3& sig = s.mp.se[c];
  Range r[3] = {range(0), range(1), range(2)};
  cecjc = Arr3(r[0].size(), r[1].size(), r[2].size());
  for (Index i = 0; i < r[0].size(); ++i)
    for (Index jj = 0; jj < r[1].size(); ++jj)
      for (Index k = 0; k < r[2].size(); ++k) {
        const Index gi = r[0].s + i, gj = r[1].s + jj, gk = r[2].s + k;
        cecjc(i, jj, k) = -(2 * dt) / (2 * eps(gi, gj, gk) + dt * sig(gi, gj, gk));
      }
}

namespace {
// InductorCorrector::correctE (lumped_element_corrrector.cpp:40-99): per
// edge in (i, j, k) order, E += cecjc * J, then J += cjcec * E.
struct InductorCorrector : Corrector {
  Sim* s;
  Inductor* el;
  Task task, local;
  void correctE(Index) override {
    Arr3& e = s->emf.f(0, el->xyz);
    for (Index i = 0; i < task.x.size(); ++i)
      for (Index j = 0; j < task.y.size(); ++j)
        for (Index k = 0; k < task.z.size(); ++k) {
          const Index li = local.x.s + i, lj = local.y.s + j, lk = local.z.s + k;
          Real& ev = e(task.x.s + i, task.y.s + j, task.z.s + k);
          ev += el->cecjc(li, lj, lk) * el->j(li, lj, lk);
          el->j(li, lj, lk) += el->cjcec(li, lj, lk) * ev;
        }
  }
  void correctH(Index) override {}
};
}  // namespace

std::unique_ptr<Corrector> Inductor::generateCorrector(Sim& s, const Task& t) {  // inductor.cpp:111-135
  Task dom = indexTask();
  if (!intersected(t, dom)) return nullptr;
  auto in = intersection(t, dom);
  if (!in) return nullptr;
  auto c = std::make_unique<InductorCorrector>();
  c->s = &s;
  c->el = this;
  c->task = *in;
  c->local = {{in->x.s - dom.x.s, in->x.e - dom.x.s},
              {in->y.s - dom.y.s, in->y.e - dom.y.s},
              {in->z.s - dom.z.s, in->z.e - dom.z.s}};
  return c;
}

void PecPlane::correctMaterialSpace(Sim& s, Index) {  // pec_plane.cpp:12-67
  const Real sg = material.sigma_e;
  const Index is = gb_origin[0], js = gb_origin[1], ks = gb_origin[2];
  const Index ie = gb_end[0], je = gb_end[1], ke = gb_end[2];
  auto fill = [&](Arr3& a, Index i0, Index i1, Index j0, Index j1, Index k0, Index k1) {
    for (Index i = i0; i < i1; ++i)
      for (Index j = j0; j < j1; ++j)
        for (Index k = k0; k < k1; ++k) a(i, j, k) = sg;
  };
  if (is == ie) {
    fill(s.mp.se[1], is, is + 1, js, je, ks, ke + 1);
    fill(s.mp.se[2], is, is + 1, js, je + 1, ks, ke);
    return;
  }
  if (js == je) {
    fill(s.mp.se[2], is, ie + 1, js, js + 1, ks, ke);
    fill(s.mp.se[0], is, ie, js, js + 1, ks, ke + 1);
    return;
  }
  if (ks == ke) {
    fill(s.mp.se[0], is, ie, js, je + 1, ks, ks + 1);
    fill(s.mp.se[1], is, ie + 1, js, je, ks, ks + 1);
    return;
  }
  throw Error(name + " PecPlane is not a plane!");
}

void SourceElement::init(Sim& s) {
  // voltage_source.cpp:47-107, current_source.cpp:41-109
  LumpedElement::init(s);
  rf = resistance * globalCountSubA() * globalCountSubB() / globalCountMain();
  if (voltage)
    amp_factor = waveform.amplitude / globalCountMain();
  else
    amp_factor = waveform.amplitude / (globalCountSubA() * globalCountSubB());
  if (dirNegative(direction)) amp_factor *= -1;
  sizes(s, da, db, dc);
  if (!voltage) {  // current_source.cpp:107-108: beta fixed at init
    beta = Arr3(dc.n0, dc.n1, dc.n2);
    for (Index n = 0; n < beta.size(); ++n)
      beta.v[n] = (s.dt * dc.v[n]) / (rf * da.v[n] * db.v[n]);
  }
}

void SourceElement::correctUpdateCoefficient(Sim& s) {
  if (voltage) {  // voltage_source.cpp:109-114
    alpha = Arr3(dc.n0, dc.n1, dc.n2);
    beta = Arr3(dc.n0, dc.n1, dc.n2);
    for (Index n = 0; n < alpha.size(); ++n) {
      alpha.v[n] = da.v[n] * db.v[n] * rf;
      beta.v[n] = s.dt * dc.v[n] / alpha.v[n];
    }
    lumpedPatch(s, *this, beta, da, db, &alpha, true, amp_factor, &coff);
  } else {
    lumpedPatch(s, *this, beta, da, db, nullptr, false, amp_factor, &coff);
  }
}

void SourceElement::initTimeDependentVariable(Sim& s) { waveform.init(s.hTime()); }

namespace {
// VoltageSourceCorrector / CurrentSourceCorrector::correctE
// (lumped_element_corrrector.cpp:8-38): E += coff * w[t] over the element.
struct SourceCorrector : Corrector {
  Sim* s;
  SourceElement* el;
  Task task, local;
  void correctE(Index t) override {
    Arr3& e = s->emf.f(0, el->xyz);
    Real w = el->waveform.value(t);
    for (Index i = 0; i < task.x.size(); ++i)
      for (Index j = 0; j < task.y.size(); ++j)
        for (Index k = 0; k < task.z.size(); ++k)
          e(task.x.s + i, task.y.s + j, task.z.s + k) +=
              el->coff(local.x.s + i, local.y.s + j, local.z.s + k) * w;
  }
  void correctH(Index) override {}
};
}  // namespace

std::unique_ptr<Corrector> SourceElement::generateCorrector(Sim& s, const Task& t) {
  Task dom = indexTask();
  if (!intersected(t, dom)) return nullptr;
  auto in = intersection(t, dom);
  auto c = std::make_unique<SourceCorrector>();
  c->s = &s;
  c->el = this;
  c->task = *in;
  c->local = {{in->x.s - dom.x.s, in->x.e - dom.x.s},
              {in->y.s - dom.y.s, in->y.e - dom.y.s},
              {in->z.s - dom.z.s, in->z.e - dom.z.s}};
  return c;
}

// ------------------------------------------------------------------- PML ---
void PML::init(Sim& s) {  // pml.cpp:98-196 (single process: node == global)
  n = static_cast<Index>(std::abs(thickness));
  main = dirAxis(dir);
  Index N = s.g.n(main);
  switch (dir) {
    case XN: case YN: case ZN:
      g_e_start = 1;
      g_h_start = 0;
      break;
    default:
      g_e_start = N - n;
      g_h_start = N - n;
  }
  Index na = s.g.e_size[tanA(main)].size();
  Index nb = s.g.e_size[tanB(main)].size();
  auto x = fromABC(main, 0, 0, g_h_start);
  auto xe = fromABC(main, na, nb, g_h_start + n);
  Task slab{{x[0], xe[0]}, {x[1], xe[1]}, {x[2], xe[2]}};
  Task whole{{0, s.g.nx()}, {0, s.g.ny()}, {0, s.g.nz()}};
  auto t = intersection(whole, slab);
  if (!t) {
    node_task = Task{};
    return;
  }
  node_task = *t;
  Index cstart = node_task.r(main).s;
  node_e_start = dirNegative(dir) ? cstart + 1 : cstart;
  node_h_start = cstart;
}

void PML::recursiveConvolution(Sim& s) {  // pml.cpp:526-633
  auto rho = [&](const Arr1& size, Real frac) {
    Arr1 interval(n), d(n);
    for (Index i = 0; i < n; ++i) interval(i) = size(i) - frac * size(i);
    d(0) = interval(0);
    Real sum = 0;
    for (Index i = 1; i < n; ++i) {
      sum += size(i - 1);
      d(i) = sum + interval(i);
    }
    sum += size(n - 1);
    if (dirNegative(dir)) std::reverse(d.v.begin(), d.v.end());
    for (Index i = 0; i < n; ++i) d(i) = d(i) / sum;
    return d;
  };
  auto sigma_max = [&](Real dl) {
    return sigma_ratio * (order + 1) / (150 * PI * dl);
  };
  const Real dt = s.dt;
  const Real ord = static_cast<Real>(order);

  Real min_h = *std::min_element(g_h_size.v.begin(), g_h_size.v.end());
  Real smax_e = sigma_max(min_h);
  Arr1 rho_e = rho(g_h_size, 0.75);
  kappa_e = Arr1(n);
  a_e = Arr1(n);
  b_e = Arr1(n);
  for (Index i = 0; i < n; ++i) {
    Real sg = smax_e * std::pow(rho_e(i), ord);
    Real al = alpha_min + (alpha_max - alpha_min) * (1 - rho_e(i));
    Real ka = 1 + (kappa_max - 1) * std::pow(rho_e(i), ord);
    kappa_e(i) = ka;
    b_e(i) = std::exp((-dt / EPSILON_0) * (sg / ka + al));
    a_e(i) = (b_e(i) - 1) * sg / ((sg + ka * al) * ka * g_h_size(i));
  }

  Real e2m = MU_0 / EPSILON_0;
  Real min_e = *std::min_element(g_e_size.v.begin(), g_e_size.v.end());
  Real smax_m = sigma_max(min_e) * e2m;
  Arr1 rho_m = rho(g_e_size, 0.25);
  kappa_h = Arr1(n);
  a_h = Arr1(n);
  b_h = Arr1(n);
  for (Index i = 0; i < n; ++i) {
    Real sg = smax_m * std::pow(rho_m(i), ord);
    Real al = (alpha_min + (alpha_max - alpha_min) * (1 - rho_m(i))) * e2m;
    Real ka = 1 + (kappa_max - 1) * std::pow(rho_m(i), ord);
    kappa_h(i) = ka;
    b_h(i) = std::exp((-dt / MU_0) * (sg / ka + al));
    a_h(i) = (b_h(i) - 1) * sg / ((sg + ka * al) * ka * g_e_size(i));
  }
}

void PML::correctUpdateCoefficient(Sim& s) {  // pml.cpp:214-244
  if (!node_task.valid()) return;
  g_h_size = Arr1(n);
  g_e_size = Arr1(n);
  for (Index i = 0; i < n; ++i) {
    g_h_size(i) = s.g.h_size[main](g_e_start + i);
    g_e_size(i) = s.g.e_size[main](g_h_start + i);
  }
  recursiveConvolution(s);

  // pml.cpp:286-524, written once for the three axes through the (a,b,c)
  // rotation.  For slab axis c: E_a is corrected through the c-derivative of
  // H_b (coefficient dualB of E_a... see below) and E_b through H_a.
  const Axis c = main, a = tanA(c), b = tanB(c);
  const Index nn[3] = {node_task.x.size(), node_task.y.size(), node_task.z.size()};
  // psi arrays: E_a / H_b are extended (+1) along b; E_b / H_a along a.
  auto ext = [&](Axis along) {
    Index m[3] = {nn[0], nn[1], nn[2]};
    m[along] += 1;
    return Arr3(m[0], m[1], m[2]);
  };
  c_ea_psi_hb = ext(b);
  ea_psi_hb = ext(b);
  c_eb_psi_ha = ext(a);
  eb_psi_ha = ext(a);
  c_ha_psi_eb = ext(a);
  ha_psi_eb = ext(a);
  c_hb_psi_ea = ext(b);
  hb_psi_ea = ext(b);

  // The E_a coefficient multiplying (H_b(c) - H_b(c-1)) is the E_a
  // coefficient for its tangential-B dual when E_a's own tangential axes are
  // (b', c') ... resolved concretely: X slab -> ceyhz, cezhy; Y slab ->
  // cezhx, cexhz; Z slab -> cexhy, ceyhx (pml.cpp:300-305, 377-381, 457-461).
  Arr3* ceahb;
  Arr3* cebha;
  Arr3* chaeb;
  Arr3* chbea;
  if (c == X) {
    ceahb = &s.c.ceyhz; cebha = &s.c.cezhy; chaeb = &s.c.chyez; chbea = &s.c.chzey;
  } else if (c == Y) {
    ceahb = &s.c.cezhx; cebha = &s.c.cexhz; chaeb = &s.c.chzex; chbea = &s.c.chxez;
  } else {
    ceahb = &s.c.cexhy; cebha = &s.c.ceyhx; chaeb = &s.c.chxey; chbea = &s.c.chyex;
  }
  auto patch = [&](Arr3& cpsi, Arr3& coef, Index start, const Arr1& dl, const Arr1& kappa,
                   Index gstart) {
    // loops i, j, k over the psi extent; the main-axis index is offset by the
    // node start (pml.cpp:307-327 and the Y/Z analogues)
    for (Index i = 0; i < cpsi.n0; ++i)
      for (Index j = 0; j < cpsi.n1; ++j)
        for (Index k = 0; k < cpsi.n2; ++k) {
          Index idx[3] = {i, j, k};
          Index layer = idx[c];
          idx[c] = layer + start;
          Index g = idx[c];  // global == node (single process)
          Real& cf = coef(idx[0], idx[1], idx[2]);
          cpsi(i, j, k) = cf * dl(g - gstart);
          cf = cf / kappa(g - gstart);
        }
  };
  patch(c_ea_psi_hb, *ceahb, node_e_start, g_h_size, kappa_e, g_e_start);
  patch(c_eb_psi_ha, *cebha, node_e_start, g_h_size, kappa_e, g_e_start);
  patch(c_ha_psi_eb, *chaeb, node_h_start, g_e_size, kappa_h, g_h_start);
  patch(c_hb_psi_ea, *chbea, node_h_start, g_e_size, kappa_h, g_h_start);
}

namespace {
// correctPML (pml_corrector.h:14-52)
template <bool IsE>
void correctPML(Arr3& field, Arr3& psi, const Arr1& ca, const Arr1& cb, const Arr3& dual,
                const Arr3& cpsi, Axis xyz, Index is, Index ie, Index js, Index je, Index ks,
                Index ke, Index gstart, Index nstart) {
  for (Index i = is; i < ie; ++i)
    for (Index j = js; j < je; ++j)
      for (Index k = ks; k < ke; ++k) {
        auto abc = toABC(xyz, i, j, k);
        Index gc = abc[2];
        Real av = ca(gc - gstart), bv = cb(gc - gstart);
        Index layer = abc[2] - nstart;
        auto l = fromABC(xyz, abc[0], abc[1], layer);
        Real cp = cpsi(l[0], l[1], l[2]);
        Real& pv = psi(l[0], l[1], l[2]);
        Real dp = dual(i, j, k);
        auto q = fromABC(xyz, abc[0], abc[1], IsE ? abc[2] - 1 : abc[2] + 1);
        Real dq = dual(q[0], q[1], q[2]);
        if (IsE)
          pv = bv * pv + av * (dp - dq);
        else
          pv = bv * pv + av * (dq - dp);
        field(i, j, k) += cp * pv;
      }
}

struct PMLCorrector : Corrector {  // pml_corrector.h:54-382
  Sim* s;
  PML* p;
  Task task;
  void correctE(Index) override {
    if (!task.valid()) return;
    const Axis c = p->main, a = tanA(c), b = tanB(c);
    Index st[3] = {task.x.s, task.y.s, task.z.s};
    Index off = 0;
    {
      auto abc = toABC(c, st[0], st[1], st[2]);
      if (abc[2] == 0) {
        auto q = fromABC(c, abc[0], abc[1], abc[2] + 1);
        off = 1;
        st[0] = q[0];
        st[1] = q[1];
        st[2] = q[2];
      }
    }
    auto nt = toABC(c, p->node_task.x.e, p->node_task.y.e, p->node_task.z.e);
    auto te = toABC(c, task.x.e, task.y.e, task.z.e);
    {  // E_a with dual H_b: [a_s,a_e) x [b_s, b_e(+1)) x [c_s, c_e + off)
      auto e = (te[1] == nt[1]) ? fromABC(c, te[0], te[1] + 1, te[2] + off)
                                : fromABC(c, te[0], te[1], te[2] + off);
      correctPML<true>(s->emf.f(0, a), p->ea_psi_hb, p->a_e, p->b_e, s->emf.f(1, b),
                       p->c_ea_psi_hb, c, st[0], e[0], st[1], e[1], st[2], e[2], p->g_e_start,
                       p->node_e_start);
    }
    {  // E_b with dual H_a: [a_s, a_e(+1)) x [b_s, b_e) x [c_s, c_e + off)
      auto e = (te[0] == nt[0]) ? fromABC(c, te[0] + 1, te[1], te[2] + off)
                                : fromABC(c, te[0], te[1], te[2] + off);
      correctPML<true>(s->emf.f(0, b), p->eb_psi_ha, p->a_e, p->b_e, s->emf.f(1, a),
                       p->c_eb_psi_ha, c, st[0], e[0], st[1], e[1], st[2], e[2], p->g_e_start,
                       p->node_e_start);
    }
  }
  void correctH(Index) override {
    if (!task.valid()) return;
    const Axis c = p->main, a = tanA(c), b = tanB(c);
    auto nt = toABC(c, p->node_task.x.e, p->node_task.y.e, p->node_task.z.e);
    auto te = toABC(c, task.x.e, task.y.e, task.z.e);
    {  // H_a with dual E_b: [a_s, a_e(+1)) x [b_s, b_e) x [c_s, c_e)
      auto e = (te[0] == nt[0]) ? fromABC(c, te[0] + 1, te[1], te[2])
                                : fromABC(c, te[0], te[1], te[2]);
      correctPML<false>(s->emf.f(1, a), p->ha_psi_eb, p->a_h, p->b_h, s->emf.f(0, b),
                        p->c_ha_psi_eb, c, task.x.s, e[0], task.y.s, e[1], task.z.s, e[2],
                        p->g_h_start, p->node_h_start);
    }
    {  // H_b with dual E_a: [a_s, a_e) x [b_s, b_e(+1)) x [c_s, c_e)
      auto e = (te[1] == nt[1]) ? fromABC(c, te[0], te[1] + 1, te[2])
                                : fromABC(c, te[0], te[1], te[2]);
      correctPML<false>(s->emf.f(1, b), p->hb_psi_ea, p->a_h, p->b_h, s->emf.f(0, a),
                        p->c_hb_psi_ea, c, task.x.s, e[0], task.y.s, e[1], task.z.s, e[2],
                        p->g_h_start, p->node_h_start);
    }
  }
};
}  // namespace

std::unique_ptr<Corrector> PML::generateCorrector(Sim& s, const Task& t) {
  // pml.cpp:246-284
  if (!node_task.valid() || !intersected(t, node_task)) return nullptr;
  auto in = intersection(t, node_task);
  auto c = std::make_unique<PMLCorrector>();
  c->s = &s;
  c->p = this;
  c->task = *in;
  return c;
}

// ------------------------------------------------------------------ TFSF ---
Task TFSF3D::globalTask() const {
  return {{box_origin[0], box_origin[0] + box_size[0]},
          {box_origin[1], box_origin[1] + box_size[1]},
          {box_origin[2], box_origin[2] + box_size[2]}};
}

void TFSF3D::init(Sim& s) {
  // constructor (tfsf.cpp:14-30) and initTransform (tfsf.cpp:231-254)
  st = std::sin(theta);
  ct = std::cos(theta);
  sp = std::sin(phi);
  cp = std::cos(phi);
  ss = std::sin(psi);
  cs = std::cos(psi);
  k = Vec3{st * cp, st * sp, ct};
  Vec3 va{-sp, ct * cp, st * cp};
  Vec3 vb{cp, ct * sp, st * sp};
  Vec3 vc{0, -st, ct};
  Vec3 ke{ss, cs, 0};
  transform_e = Vec3{va.dot(ke), vb.dot(ke), vc.dot(ke)};
  transform_h = k.cross(transform_e);

  // defaultInit (tfsf.cpp:118-175)
  Index sz[3] = {s.g.nx(), s.g.ny(), s.g.nz()};
  Index off[3] = {x, y, z};
  for (int a = 0; a < 3; ++a) {
    box_origin[a] = 0 + off[a];
    box_size[a] = sz[a] - 2 * off[a];
  }
  ratio_delta = 1 / std::sqrt(std::pow(st, 4) * (std::pow(cp, 4) + std::pow(sp, 4)) +
                              std::pow(ct, 4));
  aux = static_cast<Index>(
      std::ceil(ratio_delta * (std::sqrt(std::pow(static_cast<Real>(box_size[0]), 2) +
                                         std::pow(static_cast<Real>(box_size[1]), 2) +
                                         std::pow(static_cast<Real>(box_size[2]), 2)))) +
      4 + 1);

  // calculateProjection (tfsf.cpp:256-302); injection corner (:304-339)
  std::array<Index, 3> inj;
  for (int a = 0; a < 3; ++a)
    inj[a] = (k[a] >= 0) ? box_origin[a] : box_origin[a] + box_size[a];
  Real inc[3];
  for (int a = 0; a < 3; ++a) inc[a] = static_cast<Real>(inj[a]) - (2 * k[a] / ratio_delta);
  auto get_int = [&](Index nn, Real inc_p, Index org, Real kk) {
    Arr1 arr(nn);
    for (Index i = 0; i < nn; ++i) arr(i) = (i + org - inc_p) * kk * ratio_delta;
    return arr;
  };
  auto get_half = [&](Index nn, Index org, Real inc_p, Real kk) {
    Arr1 arr(nn);
    for (Index i = 0; i < nn; ++i) arr(i) = (i + org - inc_p - 0.5) * kk * ratio_delta;
    return arr;
  };
  px_int = get_int(box_size[0] + 1, inc[0], box_origin[0], k.x);
  py_int = get_int(box_size[1] + 1, inc[1], box_origin[1], k.y);
  pz_int = get_int(box_size[2] + 1, inc[2], box_origin[2], k.z);
  px_half = get_half(box_size[0] + 2, box_origin[0], inc[0], k.x);
  py_half = get_half(box_size[1] + 2, box_origin[1], inc[1], k.y);
  pz_half = get_half(box_size[2] + 2, box_origin[2], inc[2], k.z);

  const Real dt = s.dt;
  // tfsf.cpp:153-161: based dz in 1D and 3D, based dx in 2D
  scaled_dl = (s.g.dim == 2 ? s.g.based[0] : s.g.based[2]) / ratio_delta;
  ceie = 1;
  chih = 1;
  ceihi = -(dt / (EPSILON_0 * scaled_dl));
  chiei = -(dt / (MU_0 * scaled_dl));
  abc0 = (C_0 * dt - scaled_dl) / (C_0 * dt + scaled_dl);
  abc1 = 2 * scaled_dl / (C_0 * dt + scaled_dl);
  for (int a = 0; a < 3; ++a) {  // tfsf.cpp:177-205
    ca[a] = dt / (EPSILON_0 * s.g.based[a]);
    cb[a] = dt / (MU_0 * s.g.based[a]);
  }
}

void TFSF3D::initTimeDependentVariable(Sim& s) {  // tfsf.cpp:36-78
  Arr1 t = s.eTime();
  for (Index i = 0; i < t.size(); ++i) t(i) = t(i) - s.dt;
  waveform.init(t);
  const Index nt = s.size;
  e_inc = Arr2(nt, aux);
  h_inc = Arr2(nt, aux - 1);
  auto& e = e_inc;
  auto& h = h_inc;
  const Index nl = aux - 1;
  const Arr1& src = waveform.value;
  if (nt == 0) return;
  e(0, 0) = src(0);
  for (Index l = 1; l < nl; ++l) h(0, l) = chih * h(0, l) + chiei * (e(0, l + 1) - e(0, l));
  for (Index tt = 1; tt < nt; ++tt) {
    for (Index l = 1; l < nl; ++l) {
      e(tt, 0) = src(tt);
      e(tt, l) = ceie * e(tt - 1, l) + ceihi * (h(tt - 1, l) - h(tt - 1, l - 1));
    }
    Real e_l_prev_t_prev_prev = (tt < 2) ? 0.0 : e(tt - 2, nl - 1);
    Real e_l_t_prev_prev = (tt < 2) ? 0.0 : e(tt - 2, nl);
    e(tt, nl) = -e_l_prev_t_prev_prev + abc0 * (e(tt, nl - 1) + e_l_t_prev_prev) +
                abc1 * (e(tt - 1, nl) + e(tt - 1, nl - 1));
    for (Index l = 0; l < nl; ++l) h(tt, l) = chih * h(tt - 1, l) + chiei * (e(tt, l + 1) - e(tt, l));
  }
}

namespace {
struct TFSFCorrector : Corrector {  // tfsf_corrector.h:12-351, tfsf_corrector.cpp
  Sim* s;
  TFSF3D* f;
  Task task, node, global;
  bool flag[6];

  // incident field component `comp` (0..2) of attribute (H if h) at global
  // (i, j, k), time row t (tfsf_corrector.cpp:16-122)
  Real inc(bool h, int comp, Index t, Index i, Index j, Index k) const {
    Index ii = i - global.x.s, jj = j - global.y.s, kk = k - global.z.s;
    Real p;
    if (!h) {
      if (comp == 0) p = f->px_half(ii + 1) + f->py_int(jj) + f->pz_int(kk);
      else if (comp == 1) p = f->px_int(ii) + f->py_half(jj + 1) + f->pz_int(kk);
      else p = f->px_int(ii) + f->py_int(jj) + f->pz_half(kk + 1);
    } else {
      if (comp == 0) p = f->px_int(ii) + f->py_half(jj + 1) + f->pz_half(kk + 1) - 0.5;
      else if (comp == 1) p = f->px_half(ii + 1) + f->py_int(jj) + f->pz_half(kk + 1) - 0.5;
      else p = f->px_half(ii + 1) + f->py_half(jj + 1) + f->pz_int(kk) - 0.5;
    }
    Index idx = static_cast<Index>(p);
    Real w = p - idx;
    const Arr2& tab = h ? f->h_inc : f->e_inc;
    Real tr = h ? f->transform_h[comp] : f->transform_e[comp];
    return tr * ((1 - w) * tab(t, idx) + w * tab(t, idx + 1));
  }

  Index nodeGlobalEnd(int axis) const { return node.r(axis).e; }

  // correctTFSF<direction, attribute, field_xyz> (tfsf_corrector.h:89-229)
  void correct(Dir d, bool h, Axis fxyz, Index t) {
    if (!flag[d]) return;
    const Axis xyz = dirAxis(d);
    const Axis da = tanA(xyz), db = tanB(xyz);
    const Axis dual_field = (fxyz == da) ? db : da;
    auto s0 = toABC(xyz, task.x.s, task.y.s, task.z.s);
    auto e0 = toABC(xyz, task.x.e, task.y.e, task.z.e);
    Index as = s0[0], bs = s0[1], cs = s0[2], ae = e0[0], be = e0[1], ce = e0[2];
    const int offset_a = ((!h && fxyz == db) || (h && fxyz == da)) ? 1 : 0;
    const int offset_b = offset_a == 1 ? 0 : 1;
    if (offset_a == 1 && ae == nodeGlobalEnd(da)) ae += 1;
    if (offset_b == 1 && be == nodeGlobalEnd(db)) be += 1;
    const bool neg = dirNegative(d);
    cs = neg ? cs : ce;
    ce = cs + 1;
    const int flagc = 1 * (h ? -1 : 1) * ((fxyz == da) ? -1 : 1) * (neg ? -1 : 1);
    auto lo = fromABC(xyz, as, bs, cs);
    auto hi = fromABC(xyz, ae, be, ce);
    const Real coef_base = h ? f->cb[xyz] : f->ca[xyz];
    Arr3& field = s->emf.f(h ? 1 : 0, fxyz);
    for (Index i = lo[0]; i < hi[0]; ++i)
      for (Index j = lo[1]; j < hi[1]; ++j)
        for (Index k = lo[2]; k < hi[2]; ++k) {
          auto abc = toABC(xyz, i, j, k);
          const Real coefficient = coef_base * flagc;
          if (neg) {
            if (!h) {
              auto q = fromABC(xyz, abc[0], abc[1], abc[2] - 1);
              Real v = inc(!h, dual_field, t, q[0], q[1], q[2]);
              field(i, j, k) += v * coefficient;
            } else {
              Real v = inc(!h, dual_field, t, i, j, k);
              auto q = fromABC(xyz, abc[0], abc[1], abc[2] - 1);
              field(q[0], q[1], q[2]) += v * coefficient;
            }
            continue;
          }
          Real v = inc(!h, dual_field, t, i, j, k);
          field(i, j, k) += v * coefficient;
        }
  }

  void correctE(Index t) override {  // tfsf_corrector.cpp:177-203
    if (f->kind == 1) {  // TFSF1DCorrector::correctE (tfsf_corrector.cpp:124-136)
      correct(f->forward ? ZN : ZP, false, X, t);
      return;
    }
    if (f->kind == 2) {  // TFSF2DCorrector::correctE (tfsf_corrector.cpp:153-163)
      correct(XN, false, Z, t);
      correct(XP, false, Z, t);
      correct(YN, false, Z, t);
      correct(YP, false, Z, t);
      return;
    }
    correct(XN, false, Z, t);
    correct(XP, false, Z, t);
    correct(YN, false, Z, t);
    correct(YP, false, Z, t);
    correct(ZN, false, Y, t);
    correct(ZP, false, Y, t);
    correct(XN, false, Y, t);
    correct(XP, false, Y, t);
    correct(ZN, false, X, t);
    correct(ZP, false, X, t);
    correct(YN, false, X, t);
    correct(YP, false, X, t);
  }
  void correctH(Index t) override {  // tfsf_corrector.cpp:205-232
    if (f->kind == 1) {  // TFSF1DCorrector::correctH (tfsf_corrector.cpp:138-151)
      correct(f->forward ? ZN : ZP, true, Y, t);
      return;
    }
    if (f->kind == 2) {  // TFSF2DCorrector::correctH (tfsf_corrector.cpp:165-175)
      correct(XN, true, Y, t);
      correct(XP, true, Y, t);
      correct(YN, true, X, t);
      correct(YP, true, X, t);
      return;
    }
    correct(XN, true, Y, t);
    correct(XP, true, Y, t);
    correct(YN, true, X, t);
    correct(YP, true, X, t);
    correct(ZN, true, X, t);
    correct(ZP, true, X, t);
    correct(XN, true, Z, t);
    correct(XP, true, Z, t);
    correct(YN, true, Z, t);
    correct(YP, true, Z, t);
    correct(ZN, true, Y, t);
    correct(ZP, true, Y, t);
  }
};
}  // namespace

std::unique_ptr<Corrector> TFSF3D::generateCorrector(Sim& s, const Task& t) {
  // tfsf_3d.cpp:21-45 (node offset 0 in a single process)
  Task g = globalTask();
  auto in = intersection(t, g);
  if (!in) return nullptr;
  auto c = std::make_unique<TFSFCorrector>();
  c->s = &s;
  c->f = this;
  c->task = *in;
  c->global = g;
  Task whole{{0, s.g.nx()}, {0, s.g.ny()}, {0, s.g.nz()}};
  auto nt = intersection(whole, g);
  c->node = nt ? *nt : Task{};
  c->flag[XN] = in->x.s == g.x.s;
  c->flag[XP] = in->x.e == g.x.e;
  c->flag[YN] = in->y.s == g.y.s;
  c->flag[YP] = in->y.e == g.y.e;
  c->flag[ZN] = in->z.s == g.z.s;
  c->flag[ZP] = in->z.e == g.z.e;
  return c;
}

// -------------------------------------------------------------- monitors --
static void monitorBox(Sim& s, const Shape& sh, Index o[3], Index e[3]) {
  // Monitor::defaultInit (monitor.cpp:103-127) then the +1 shift of a box
  // that starts at index 0 (field_monitor.cpp:24-67).
  Index end[3];
  auto org = s.g.box(sh, end);
  for (int a = 0; a < 3; ++a) {
    o[a] = org[a];
    e[a] = end[a];
    // origin +1, size -1 (end unchanged) on x and y unless the grid is 1D,
    // on z only in 3D (field_monitor.cpp:29-38, field_time_monitor.cpp:43-52)
    const bool shift = a < 2 ? s.g.dim != 1 : s.g.dim == 3;
    if (o[a] == 0 && shift) o[a] = 1;
  }
}

void FieldMonitor::init(Sim& s) { monitorBox(s, *shape, o, e); }

void FieldMonitor::output(Sim& s) {
  bool valid = o[0] < e[0] && o[1] < e[1] && o[2] < e[2];  // monitor.cpp:155
  if (!valid) return;
  Arr3& f = s.emf.byId(field);
  data = Arr3(e[0] - o[0], e[1] - o[1], e[2] - o[2]);
  for (Index i = 0; i < data.n0; ++i)
    for (Index j = 0; j < data.n1; ++j)
      for (Index k = 0; k < data.n2; ++k) data(i, j, k) = f(o[0] + i, o[1] + j, o[2] + k);
  if (s.write_files && !out_dir.empty()) {
    std::filesystem::create_directories(out_dir);
    writeNpy(out_dir + "/" + name + ".npy", {data.n0, data.n1, data.n2}, data.v.data());
  }
}

void MovieMonitor::init(Sim& s) {
  frame->out_dir = out_dir;
  frame->init(s);
}

void MovieMonitor::update(Sim& s, Index) {
  if (count % interval == 0) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%05zu", count);
    frame->name = buf;
    frame->output(s);
    frames.emplace_back(buf, frame->data);
  }
  ++count;
}

void FieldTimeMonitor::init(Sim& s) { monitorBox(s, *shape, o, e); }

void FieldTimeMonitor::initTimeDependentVariable(Sim& s) {
  time = (field < 3) ? s.eTime() : s.hTime();
  data = Arr1(time.size());
}

void FieldTimeMonitor::update(Sim& s, Index t) {
  Arr3& f = s.emf.byId(field);
  for (Index i = o[0]; i < e[0]; ++i)
    for (Index j = o[1]; j < e[1]; ++j)
      for (Index k = o[2]; k < e[2]; ++k) data(t) = f(i, j, k);
}

void FieldTimeMonitor::output(Sim& s) {
  bool valid = o[0] < e[0] && o[1] < e[1] && o[2] < e[2];
  if (!valid || !s.write_files || out_dir.empty()) return;
  std::vector<Real> st(2 * time.size());
  for (Index i = 0; i < time.size(); ++i) {
    st[i] = time(i);
    st[time.size() + i] = data(i);
  }
  std::filesystem::create_directories(out_dir);
  writeNpy(out_dir + "/" + name + ".npy", {2, time.size()}, st.data());
}

// ------------------------------------------------ voltage / current ----
namespace {
// xfdtd::transform::xYZToABC for a direction's axis: (a, b, c) =
// X: (y, z, x), Y: (z, x, y), Z: (x, y, z).
void abcAxes(int axis, int& a, int& b, int& c) {
  c = axis;
  a = (axis + 1) % 3;
  b = (axis + 2) % 3;
}

void writeTimeSeries(Sim& s, const std::string& dir, const std::string& name, const Arr1& time, const Arr1& data) {
  if (!s.write_files || dir.empty()) return;
  std::vector<Real> st(2 * time.size());
  for (Index i = 0; i < time.size(); ++i) {
    st[i] = time(i);
    st[time.size() + i] = data(i);
  }
  std::filesystem::create_directories(dir);
  writeNpy(dir + "/" + name + ".npy", {2, time.size()}, st.data());
}
}  // namespace

void VoltageMonitor::init(Sim& s) {  // voltage_monitor.cpp:21-189, one process
  Index end[3];
  const auto org = s.g.box(*shape, end);
  int a, b, c;
  abcAxes(dir / 2, a, b, c);
  const Index n[3] = {s.g.nx(), s.g.ny(), s.g.nz()};
  auto clip = [](Index lo, Index up, Index& st, Index& en) {  // new_range
    if (st <= lo) st = lo + 1;
    if (up <= en) en = up - 1;
  };
  // a and b: the cube's range extended by one, kept off the grid faces
  for (int q : {a, b}) {
    r[q][0] = org[q];
    r[q][1] = end[q] + 1;
    clip(0, n[q], r[q][0], r[q][1]);
  }
  r[c][0] = org[c];  // c: the cube's range (the node clipping is undone at the global ends)
  r[c][1] = end[c];
  for (int q = 0; q < 3; ++q)
    if (r[q][0] >= r[q][1]) throw Error(name + " globalTask is not valid");
  const Index cnt = (r[a][1] - r[a][0]) * (r[b][1] - r[b][0]);
  coff = Arr1(r[c][1] - r[c][0]);
  for (Index q = 0; q < coff.size(); ++q) {
    coff(q) = -s.g.h_size[c](r[c][0] + q) / static_cast<Real>(cnt);
    if (dir % 2 == 0) coff(q) *= -1.0;  // negative direction
  }
}

void VoltageMonitor::initTimeDependentVariable(Sim& s) {  // voltage_monitor.cpp:229-233
  time = s.eTime();
  data = Arr1(time.size());
}

void VoltageMonitor::update(Sim& s, Index t) {  // voltage_monitor.cpp:191-227
  const int c = dir / 2;
  Arr3& f = s.emf.byId(c);
  for (Index i = r[0][0]; i < r[0][1]; ++i)
    for (Index j = r[1][0]; j < r[1][1]; ++j)
      for (Index k = r[2][0]; k < r[2][1]; ++k) {
        const Index p[3] = {i, j, k};
        data(t) += coff(p[c] - r[c][0]) * f(i, j, k);
      }
}

void VoltageMonitor::output(Sim& s) { writeTimeSeries(s, out_dir, name, time, data); }

void CurrentMonitor::init(Sim& s) {  // current_monitor.cpp:18-270, one process
  Index end[3];
  const auto org = s.g.box(*shape, end);
  int a, b, c;
  abcAxes(dir / 2, a, b, c);
  const Index n[3] = {s.g.nx(), s.g.ny(), s.g.nz()};
  if (end[c] == 0) throw Error("CurrentMonitor::init(): global_c_e is 0");
  if (org[a] == 0 || end[a] == n[a]) throw Error("CurrentMonitor::init(): global x range is invalid");
  if (org[b] == 0 || end[b] == n[b]) throw Error("CurrentMonitor::init(): global y range is invalid");
  r[a][0] = org[a];
  r[a][1] = end[a];
  r[b][0] = org[b];
  r[b][1] = end[b];
  r[c][0] = end[c] - 1;  // the last plane along the direction
  r[c][1] = end[c];
  // make_ranges (:138-169): the loop side along one axis, present when the
  // other axis' face lies strictly inside the grid
  auto ranges = [](Index b_start, Index b_end, Index b_upper, Index a_upper, Index st, Index en, Index* rn,
                   Index* rp) {
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
  ranges(org[b], end[b], n[b], n[a], org[a], end[a], ha_n, ha_p);
  ranges(org[a], end[a], n[a], n[b], org[b], end[b], hb_n, hb_p);
  auto valid = [](const Index* q) { return q[0] < q[1]; };
  auto sizes = [&](int axis, const Index* q) {
    Arr1 v(q[1] - q[0]);
    for (Index i = 0; i < v.size(); ++i) v(i) = s.g.e_size[axis](q[0] + i);
    return v;
  };
  // da along a, db along b (eSize of those axes)
  if (valid(ha_n))
    da = sizes(a, ha_n);
  else if (valid(ha_p))
    da = sizes(a, ha_p);
  if (valid(hb_n))
    db = sizes(b, hb_n);
  else if (valid(hb_p))
    db = sizes(b, hb_p);
  positive = dir % 2 == 1 ? 1.0 : -1.0;
}

void CurrentMonitor::initTimeDependentVariable(Sim& s) {  // current_monitor.cpp:387-391
  time = s.hTime();
  data = Arr1(time.size());
}

void CurrentMonitor::update(Sim& s, Index t) {  // current_monitor.cpp:272-385
  auto valid = [](const Index* q) { return q[0] < q[1]; };
  if (!(valid(ha_n) || valid(ha_p) || valid(hb_n) || valid(hb_p))) return;
  const Index is = r[0][0], ie = r[0][1], js = r[1][0], je = r[1][1], ks = r[2][0], ke = r[2][1];
  auto& e = s.emf;
  Real ix = 0, iy = 0, iz = 0;
  const int c = dir / 2;
  if (c == 0) {
    if (valid(ha_n))
      for (Index j = ha_n[0]; j < ha_n[1]; ++j) iy += e.hy(ie - 1, j, ks - 1) * da(j - ha_n[0]);
    if (valid(ha_p))
      for (Index j = ha_p[0]; j < ha_p[1]; ++j) iy -= e.hy(ie - 1, j, ke) * da(j - ha_p[0]);
    if (valid(hb_n))
      for (Index k = hb_n[0]; k < hb_n[1]; ++k) iz -= e.hz(ie - 1, js - 1, k) * db(k - hb_n[0]);
    if (valid(hb_p))
      for (Index k = hb_p[0]; k < hb_p[1]; ++k) iz += e.hz(ie - 1, je, k) * db(k - hb_p[0]);
    data(t) = positive * (iz + iy);
  } else if (c == 1) {
    if (valid(ha_n))
      for (Index k = ha_n[0]; k < ha_n[1]; ++k) iz += e.hz(is - 1, je - 1, k) * da(k - ha_n[0]);
    if (valid(ha_p))
      for (Index k = ha_p[0]; k < ha_p[1]; ++k) iz -= e.hz(ie, je - 1, k) * da(k - ha_p[0]);
    if (valid(hb_n))
      for (Index i = hb_n[0]; i < hb_n[1]; ++i) ix -= e.hx(i, je - 1, ks - 1) * db(i - hb_n[0]);
    if (valid(hb_p))
      for (Index i = hb_p[0]; i < hb_p[1]; ++i) ix += e.hx(i, je - 1, ke) * db(i - hb_p[0]);
    data(t) = positive * (ix + iz);
  } else {
    if (valid(ha_n))
      for (Index i = ha_n[0]; i < ha_n[1]; ++i) ix += e.hx(i, js - 1, ke - 1) * da(i - ha_n[0]);
    if (valid(ha_p))
      for (Index i = ha_p[0]; i < ha_p[1]; ++i) ix -= e.hx(i, je, ke - 1) * da(i - ha_p[0]);
    if (valid(hb_n))
      for (Index j = hb_n[0]; j < hb_n[1]; ++j) iy -= e.hy(is - 1, j, ke - 1) * db(j - hb_n[0]);
    if (valid(hb_p))
      for (Index j = hb_p[0]; j < hb_p[1]; ++j) iy += e.hy(ie, j, ke - 1) * db(j - hb_p[0]);
    data(t) = positive * (ix + iy);
  }
}

void CurrentMonitor::output(Sim& s) { writeTimeSeries(s, out_dir, name, time, data); }

// ----------------------------------------------------------------- NF2FF --
namespace {
// GCC's complex product for finite operands: (ac - bd, ad + bc)
Cplx cmul(const Cplx& a, const Cplx& b) {
  return {a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real()};
}
// interpolate::interpolateSurfaceCenter (interpolate_scheme.h:18-138): the
// tangential field component `f` (0..5) at the centre of the face of the
// surface normal to axis c at (i, j, k)
Real surfaceCenter(Sim& s, int c, int f, Index i, Index j, Index k) {
  const Arr3& a = s.emf.byId(f);
  const Real h = 0.5, q = 0.25;
  if (c == 0) {
    if (f == 5) return q * (a(i, j, k + 1) + a(i, j, k) + a(i - 1, j, k + 1) + a(i - 1, j, k));
    if (f == 4) return q * (a(i, j + 1, k) + a(i, j, k) + a(i - 1, j + 1, k) + a(i - 1, j, k));
    if (f == 2) return h * (a(i, j + 1, k) + a(i, j, k));
    if (f == 1) return h * (a(i, j, k + 1) + a(i, j, k));
  } else if (c == 1) {
    if (f == 5) return q * (a(i, j, k + 1) + a(i, j, k) + a(i, j - 1, k + 1) + a(i, j - 1, k));
    if (f == 3) return q * (a(i + 1, j, k) + a(i, j, k) + a(i + 1, j - 1, k) + a(i, j - 1, k));
    if (f == 2) return h * (a(i + 1, j, k) + a(i, j, k));
    if (f == 0) return h * (a(i, j, k + 1) + a(i, j, k));
  } else {
    if (f == 3) return q * (a(i + 1, j, k) + a(i, j, k) + a(i, j, k - 1) + a(i + 1, j, k - 1));
    if (f == 4) return q * (a(i, j, k) + a(i, j + 1, k) + a(i, j, k - 1) + a(i, j + 1, k - 1));
    if (f == 0) return h * (a(i, j + 1, k) + a(i, j, k));
    if (f == 1) return h * (a(i + 1, j, k) + a(i, j, k));
  }
  throw Error("Invalid field");
}
}  // namespace

namespace {
// NFFFT::initGlobal / makeNodeAxisTask (nffft.cpp:88-223) with node ==
// global: the global surfaces are the box faces; a node surface is empty
// when it lies on plane 0 or 1 (negative side) or on the last plane.
void nffftSurfaces(Sim& s, const Index dist[3], Task node[6], Task global[6]) {
  const Index n[3] = {s.g.nx(), s.g.ny(), s.g.nz()};
  Index lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    if (dist[a] == 0) throw Error("distance == 0");
    lo[a] = dist[a];
    hi[a] = n[a] - dist[a];
    if (hi[a] <= lo[a]) throw Error("Global box is not legal");
  }
  for (int d = 0; d < 6; ++d) {
    const int c = d / 2;
    const bool neg = d % 2 == 0;
    Range r[3];
    for (int a = 0; a < 3; ++a) r[a] = {lo[a], hi[a]};
    const Index cs = neg ? lo[c] : hi[c];
    r[c] = {cs, cs + 1};
    global[d] = {r[0], r[1], r[2]};
    node[d] = ((neg && cs <= 1) || (!neg && cs >= n[c])) ? Task{} : global[d];
  }
  bool any = false;
  for (int d = 0; d < 6; ++d) any = any || node[d].valid();
  if (!any) throw Error("Invalid task");
}
}  // namespace

void NF2FF::init(Sim& s) {  // nffft.cpp:73-223 with node == global
  Task global[6];
  nffftSurfaces(s, dist, surf, global);
  planes.clear();
  for (Real f : freqs) {
    Plane p;
    p.freq = f;
    for (int d = 0; d < 6; ++d) {
      const Index cells = surf[d].valid() ? surf[d].x.size() * surf[d].y.size() * surf[d].z.size() : 0;
      p.ja[d].assign(cells, Cplx{});
      p.jb[d].assign(cells, Cplx{});
      p.ma[d].assign(cells, Cplx{});
      p.mb[d].assign(cells, Cplx{});
    }
    planes.push_back(std::move(p));
  }
}

void NF2FF::initTimeDependentVariable(Sim& s) {  // nffft_fd_data.cpp:65-76
  const Real dt = s.dt;
  const Cplx II{0, 1};
  for (auto& p : planes) {
    p.te.assign(s.size, Cplx{});
    p.th.assign(s.size, Cplx{});
    for (Index t = 0; t < s.size; ++t) {
      p.te[t] = dt * std::exp(-II * static_cast<Real>(2.0) * PI * (p.freq * (t + 1) * dt));
      p.th[t] = dt * std::exp(-II * static_cast<Real>(2.0) * PI * (p.freq * (t + static_cast<Real>(0.5)) * dt));
    }
  }
}

void NF2FF::update(Sim& s, Index t) {
  // every FDPlaneData::update runs once per frequency of the NFFFT
  // (nffft_frequency_domain.cpp:45-53); calculateJ then calculateM per
  // surface (nffft_fd_data_update.cpp:7-122)
  const Index repeats = reference_repeat ? freqs.size() : 1;
  for (Index rep = 0; rep < repeats; ++rep)
    for (auto& p : planes)
      for (int pass = 0; pass < 2; ++pass)
        for (int d = 0; d < 6; ++d) {
          const Task& tk = surf[d];
          if (!tk.valid()) continue;
          const int c = d / 2, A = (c + 1) % 3, B = (c + 2) % 3;
          const bool pos = d % 2 == 1;
          const bool is_j = pass == 0;
          const Real ca = is_j ? (pos ? -1.0 : 1.0) : (pos ? 1.0 : -1.0);
          const Real cb = is_j ? (pos ? 1.0 : -1.0) : (pos ? -1.0 : 1.0);
          const int fa = (is_j ? 3 : 0) + A, fb = (is_j ? 3 : 0) + B;
          const Cplx tr = is_j ? p.th[t] : p.te[t];
          auto& va = is_j ? p.ja[d] : p.ma[d];
          auto& vb = is_j ? p.jb[d] : p.mb[d];
          Index q = 0;
          for (Index i = tk.x.s; i < tk.x.e; ++i)
            for (Index j = tk.y.s; j < tk.y.e; ++j)
              for (Index k = tk.z.s; k < tk.z.e; ++k, ++q) {
                va[q] += ca * surfaceCenter(s, c, fb, i, j, k) * tr;
                vb[q] += cb * surfaceCenter(s, c, fa, i, j, k) * tr;
              }
        }
}

std::vector<Cplx> NF2FF::potential(Sim& s, Index fi, bool a_pot, bool theta_comp, const std::vector<Real>& theta,
                                   const std::vector<Real>& phi, const Vec3& origin) const {
  if (theta.empty() || phi.empty()) throw Error("Invalid theta or phi size");
  if (theta.size() != 1 && phi.size() != 1) throw Error("There must be only one theta or phi");
  const Index na = theta.size() * phi.size();
  auto th = [&](Index n) { return theta[theta.size() == 1 ? 0 : n]; };
  auto ph = [&](Index n) { return phi[phi.size() == 1 ? 0 : n]; };
  // cCSToSCSTransformMatrix (transform.h:24-58) of axis x/y/z
  auto tmat = [&](int axis, Index n) -> Real {
    const Real t = th(n), p = ph(n);
    if (theta_comp) {
      if (axis == 0) return std::cos(t) * std::cos(p);
      if (axis == 1) return std::cos(t) * std::sin(p);
      return -1 * std::sin(t);
    }
    if (axis == 0) return -1 * std::sin(p);
    if (axis == 1) return std::cos(p);
    return 0.0;
  };
  const Plane& P = planes[fi];
  const Real k = 2.0 * PI * P.freq / C_0;
  const Cplx ik = Cplx{0, 1} * k;
  std::vector<Cplx> total(na, Cplx{});
  for (int d = 0; d < 6; ++d) {
    std::vector<Cplx> data(na, Cplx{});
    const Task& tk = surf[d];
    if (tk.valid()) {
      const int c = d / 2, A = (c + 1) % 3, B = (c + 2) % 3;
      const auto& va = a_pot ? P.ja[d] : P.ma[d];
      const auto& vb = a_pot ? P.jb[d] : P.mb[d];
      std::vector<Real> st(na), ct(na), sp(na), cp(na), stsp(na), stcp(na);
      std::vector<Cplx> ta(na), tb(na);
      for (Index n = 0; n < na; ++n) {
        ct[n] = std::cos(th(n));
        st[n] = std::sin(th(n));
        cp[n] = std::cos(ph(n));
        sp[n] = std::sin(ph(n));
        stsp[n] = st[n] * sp[n];
        stcp[n] = st[n] * cp[n];
        ta[n] = tmat(A, n);
        tb[n] = tmat(B, n);
      }
      Index q = 0;
      for (Index i = tk.x.s; i < tk.x.e; ++i)
        for (Index j = tk.y.s; j < tk.y.e; ++j)
          for (Index kk = tk.z.s; kk < tk.z.e; ++kk, ++q) {
            // rVector / ds (nffft_fd_data_process.cpp:196-225)
            Vec3 rv;
            Real ds;
            if (c == 0) {
              rv = {s.g.e_node[0](i), s.g.h_node[1](j), s.g.h_node[2](kk)};
              ds = s.g.e_size[1](j) * s.g.e_size[2](kk);
            } else if (c == 1) {
              rv = {s.g.h_node[0](i), s.g.e_node[1](j), s.g.h_node[2](kk)};
              ds = s.g.e_size[2](kk) * s.g.e_size[0](i);
            } else {
              rv = {s.g.h_node[0](i), s.g.h_node[1](j), s.g.e_node[2](kk)};
              ds = s.g.e_size[0](i) * s.g.e_size[1](j);
            }
            const Vec3 r = rv - origin;
            for (Index n = 0; n < na; ++n) {
              const Real v = r.x * stcp[n] + r.y * stsp[n] + r.z * ct[n];
              const Cplx phase = std::exp(ik * v);
              const Cplx sum = cmul(va[q], ta[n]) + cmul(vb[q], tb[n]);
              data[n] += cmul(sum, phase) * ds;
            }
          }
    }
    for (Index n = 0; n < na; ++n) total[n] += data[n];
  }
  return total;
}

Real NF2FF::power(Sim& s, Index fi) const {
  const Plane& P = planes[fi];
  Cplx acc{0.0};
  for (int d = 0; d < 6; ++d) {
    const Task& tk = surf[d];
    Cplx pw{0.0};
    if (tk.valid()) {
      const int c = d / 2;
      Index q = 0;
      for (Index i = tk.x.s; i < tk.x.e; ++i)
        for (Index j = tk.y.s; j < tk.y.e; ++j)
          for (Index kk = tk.z.s; kk < tk.z.e; ++kk, ++q) {
            Real ds;
            if (c == 0) ds = s.g.e_size[1](j) * s.g.e_size[2](kk);
            else if (c == 1) ds = s.g.e_size[2](kk) * s.g.e_size[0](i);
            else ds = s.g.e_size[0](i) * s.g.e_size[1](j);
            pw += ds * (cmul(std::conj(P.ja[d][q]), P.mb[d][q]) - cmul(std::conj(P.jb[d][q]), P.ma[d][q]));
          }
      const Real negative = d % 2 == 0 ? -1.0 : 1.0;
      pw = negative * pw;
    }
    acc = acc + pw;
  }
  return 0.5 * std::real(acc);
}

// ------------------------------------------------------------ simulation --
Arr1 Sim::eTime() const {
  return linspace((start_step + 1) * dt, (start_step + size) * dt, size);
}
Arr1 Sim::hTime() const {
  return linspace((start_step + 0.5) * dt, (start_step + size + 0.5) * dt, size);
}

void Sim::generateGridSpace() {
  // Shape::makeWrappedCube (shape.cpp:8-38)
  Real mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = std::numeric_limits<Real>::max();
    mx[a] = std::numeric_limits<Real>::lowest();
  }
  if (objects.empty()) throw Error("No object is added");
  for (auto& o : objects) {
    Cube c = o->shape->wrapped();
    for (int a = 0; a < 3; ++a) {
      mn[a] = std::min(mn[a], c.origin[a]);
      mx[a] = std::max(mx[a], c.end[a]);
    }
  }
  Cube dom(Vec3{mn[0], mn[1], mn[2]}, Vec3{mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]});
  // GridSpaceGenerator::decideDimension (grid_space_generator.cpp:101-124):
  // x and y infinite -> 1D, only z infinite -> 2D, none -> 3D
  auto decide = [](const Cube& c) {
    const bool ix = std::isinf(c.size.x), iy = std::isinf(c.size.y), iz = std::isinf(c.size.z);
    if (ix && iy && !iz) return 1;
    if (!ix && !iy && iz) return 2;
    if (!ix && !iy && !iz) return 3;
    return 0;
  };
  // GridSpaceGenerator::extendDomain (grid_space_generator.cpp:36-99)
  Real d[3] = {dx, dy, dz};
  if (!pmls.empty()) {
    const int dim0 = decide(dom);
    Real lo[3] = {dom.origin.x, dom.origin.y, dom.origin.z};
    Real hi[3] = {dom.end.x, dom.end.y, dom.end.z};
    for (auto& p : pmls) {
      int a = dirAxis(p->dir);
      if (dim0 == 1 && a != Z) throw Error("PML has to be in Z direction");
      if (dim0 == 2 && a == Z) throw Error("PML has to be in X or Y direction");
      if (p->thickness < 0) continue;
      if (dirNegative(p->dir))
        lo[a] -= p->thickness * d[a];
      else
        hi[a] += p->thickness * d[a];
    }
    dom = Cube(Vec3{lo[0], lo[1], lo[2]}, Vec3{hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
  }
  g.dim = decide(dom);  // generateUniformGridSpace (:16-34)
  if (g.dim == 0) throw Error("GridSpace dimension is undefined");
  const Real inf = std::numeric_limits<Real>::infinity();
  // per-axis node arrays (h_node, e_size, h_size) of an axis with cells:
  // GridSpace::correctGridSpaceForOne (grid_space.cpp:517-544)
  auto correctOne = [&](int a, Real dl) {
    const Arr1& e = g.e_node[a];
    Index m = e.size() - 1;
    g.h_node[a] = Arr1(m);
    g.e_size[a] = Arr1(m);
    for (Index i = 0; i < m; ++i) {
      g.h_node[a](i) = (e(i) + e(i + 1)) / 2;
      g.e_size[a](i) = e(i + 1) - e(i);
    }
    std::vector<Real> tmp(m + 2);
    tmp[0] = e(0) - dl / 2;
    for (Index i = 0; i < m; ++i) tmp[i + 1] = g.h_node[a](i);
    tmp[m + 1] = e(m) + dl / 2;
    g.h_size[a] = Arr1(m + 1);
    for (Index i = 0; i < m + 1; ++i) g.h_size[a](i) = tmp[i + 1] - tmp[i];
    Real md = std::numeric_limits<Real>::max();
    for (Index i = 0; i < m; ++i) md = std::min(md, g.e_size[a](i));
    g.min_d[a] = md;
  };
  // the one-cell axis of a 1D/2D grid (grid_space_2d.cpp:96-106,
  // grid_space_1d.cpp:64-75): nodes {-inf, inf}, one cell of size 1 centred
  // at 0
  auto oneCell = [&](int a) {
    g.e_node[a] = Arr1(2);
    g.e_node[a](0) = -inf;
    g.e_node[a](1) = inf;
    g.h_node[a] = Arr1(1, 0.0);
    g.e_size[a] = Arr1(1, 1.0);
    g.h_size[a] = Arr1(1, 1.0);
    g.min_d[a] = 1;
  };
  if (g.dim == 1) {
    // generateGridSpace1D (grid_space_generator.cpp:126-150): z only; the
    // GridSpace1D base sizes are (1, 1, dz) (grid_space_1d.cpp:12-19)
    const Real min_z = dom.origin.z, max_z0 = dom.end.z;
    const Index nz = static_cast<Index>(std::round((max_z0 - min_z) / dz));
    if (nz <= 1) throw Error("GridSpace is too small");
    Cube region(Vec3{-inf, -inf, min_z}, Vec3{inf, inf, nz * dz});
    g.based[0] = 1;
    g.based[1] = 1;
    g.based[2] = dz;
    g.e_node[2] = linspace(region.origin.z, region.end.z, nz + 1);
    correctOne(2, g.based[2]);
    oneCell(0);
    oneCell(1);
    return;
  }
  if (g.dim == 2) {
    // generateGridSpace2D (grid_space_generator.cpp:152-181); the GridSpace2D
    // constructor passes based_dx for both in-plane sizes and 1 for z
    // (grid_space_2d.cpp:14-23)
    Real lo[2] = {dom.origin.x, dom.origin.y};
    Index nn[2];
    for (int a = 0; a < 2; ++a) {
      nn[a] = static_cast<Index>(std::round(((a == 0 ? dom.end.x : dom.end.y) - lo[a]) / d[a]));
    }
    if (nn[0] <= 1 || nn[1] <= 1) throw Error("GridSpace is too small");
    Cube region(Vec3{lo[0], lo[1], -inf}, Vec3{nn[0] * dx, nn[1] * dy, inf});
    g.based[0] = dx;
    g.based[1] = dx;
    g.based[2] = 1;
    g.e_node[0] = linspace(region.origin.x, region.end.x, nn[0] + 1);
    g.e_node[1] = linspace(region.origin.y, region.end.y, nn[1] + 1);
    correctOne(0, g.based[0]);
    correctOne(1, g.based[1]);
    oneCell(2);
    return;
  }
  // generateGridSpace3D (grid_space_generator.cpp:183-217)
  Real lo[3] = {dom.origin.x, dom.origin.y, dom.origin.z};
  Real hi[3] = {dom.end.x, dom.end.y, dom.end.z};
  Index nn[3];
  for (int a = 0; a < 3; ++a) {
    nn[a] = static_cast<Index>(std::round((hi[a] - lo[a]) / d[a]));
    if (nn[a] <= 1) throw Error("GridSpace is too small");
  }
  for (int a = 0; a < 3; ++a) hi[a] = lo[a] + d[a] * nn[a];
  Cube region(Vec3{lo[0], lo[1], lo[2]}, Vec3{hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
  for (int a = 0; a < 3; ++a) {
    g.based[a] = d[a];
    g.e_node[a] = linspace(region.origin[a], region.end[a], nn[a] + 1);
  }
  // GridSpace3D::correctGridSpace (grid_space_3d.cpp:111-137, grid_space.cpp:526-544)
  for (int a = 0; a < 3; ++a) correctOne(a, d[a]);
}

void Sim::generateEMF() {
  const Index nx = g.nx(), ny = g.ny(), nz = g.nz();
  emf.ex = Arr3(nx, ny + 1, nz + 1);
  emf.ey = Arr3(nx + 1, ny, nz + 1);
  emf.ez = Arr3(nx + 1, ny + 1, nz);
  emf.hx = Arr3(nx + 1, ny, nz);
  emf.hy = Arr3(nx, ny + 1, nz);
  emf.hz = Arr3(nx, ny, nz + 1);
}

static void extentOf(int h, int c, Index nx, Index ny, Index nz, Index out[3]) {
  // electric component c is nx+1 on the other two axes; magnetic on its own
  Index n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) out[a] = h ? (a == c ? n[a] + 1 : n[a]) : (a == c ? n[a] : n[a] + 1);
}

void Sim::buildDispersiveSpace() {  // simulation.cpp:745-829
  ade.reset();
  Index num_pole = 0;
  bool is_drude = false, is_debye = false, is_lor = false;
  for (const auto& m : material_array) {
    if (!m.dispersive()) continue;
    num_pole = std::max(num_pole, m.numPoles());
    if (m.kind == Material::Drude) is_drude = true;
    else if (m.kind == Material::Debye) is_debye = true;
    else is_lor = true;
  }
  if (num_pole == 0) return;
  ADEStorage::Kind k = ADEStorage::MLorentz;
  if (is_drude && !is_debye && !is_lor) k = ADEStorage::Drude;
  else if (!is_drude && is_debye && !is_lor) k = ADEStorage::Debye;
  ade = std::make_unique<ADEStorage>(k, num_pole, g.nx(), g.ny(), g.nz());
  // Object::handleDispersion (object.cpp:52-98): every cell whose material
  // index is the object's gets the object's ADE coefficients
  const Index nx = g.nx(), ny = g.ny(), nz = g.nz();
  for (auto& o : objects) {
    if (!o->material.dispersive()) continue;
    for (Index i = 0; i < nx; ++i)
      for (Index j = 0; j < ny; ++j)
        for (Index kk = 0; kk < nz; ++kk)
          if (material_grid[(i * ny + j) * nz + kk] == o->material_index)
            ade->correctCoeff(*this, i, j, kk, o->material);
  }
}

void Sim::generateMaterialSpace() {  // simulation.cpp:593-636
  const Index nx = g.nx(), ny = g.ny(), nz = g.nz();
  material_grid.assign(nx * ny * nz, std::numeric_limits<Index>::max());
  for (int c = 0; c < 3; ++c) {  // material_param.cpp:365-393
    Index ee[3], hh[3];
    extentOf(0, c, nx, ny, nz, ee);
    extentOf(1, c, nx, ny, nz, hh);
    mp.eps[c] = Arr3(ee[0], ee[1], ee[2], EPSILON_0);
    mp.se[c] = Arr3(ee[0], ee[1], ee[2], SIGMA_ZERO);
    mp.mu[c] = Arr3(hh[0], hh[1], hh[2], MU_0);
    mp.sm[c] = Arr3(hh[0], hh[1], hh[2], SIGMA_ZERO);
  }
  Index m_index = 0;
  for (auto& o : objects) {
    if (o->isPecPlane()) continue;
    o->correctMaterialSpace(*this, m_index);
    material_array.push_back(o->material);
    ++m_index;
  }
  if (has_design)
    for (Index i = 0; i < design_n[0]; ++i)
      for (Index j = 0; j < design_n[1]; ++j)
        for (Index k = 0; k < design_n[2]; ++k) {
          const Real e = design_eps[(i * design_n[1] + j) * design_n[2] + k] * EPSILON_0;
          for (int c = 0; c < 3; ++c) mp.eps[c](design_o[0] + i, design_o[1] + j, design_o[2] + k) = e;
        }
  averageMaterial();
  for (auto& o : objects)
    if (o->isPecPlane()) o->correctMaterialSpace(*this, Index(-1));
}

void Sim::averageMaterial() {  // calculation_param.cpp:43-168
  const Index nx = g.nx(), ny = g.ny(), nz = g.nz();
  auto arith = [&](Arr3& a, int c) {
    // average of the four cells sharing the edge; evaluated from a copy
    Arr3 o = a;
    Index is = c == 0 ? 0 : 1, js = c == 1 ? 0 : 1, ks = c == 2 ? 0 : 1;
    for (Index i = is; i < nx; ++i)
      for (Index j = js; j < ny; ++j)
        for (Index k = ks; k < nz; ++k) {
          Real v;
          if (c == 0)
            v = o(i, j, k) + o(i, j - 1, k) + o(i, j, k - 1) + o(i, j - 1, k - 1);
          else if (c == 1)
            v = o(i, j, k) + o(i - 1, j, k) + o(i, j, k - 1) + o(i - 1, j, k - 1);
          else
            v = o(i, j, k) + o(i - 1, j, k) + o(i, j - 1, k) + o(i - 1, j - 1, k);
          a(i, j, k) = 0.25 * v;
        }
  };
  auto harm = [&](Arr3& a, int c) {
    Arr3 o = a;
    Index is = c == 0 ? 1 : 0, js = c == 1 ? 1 : 0, ks = c == 2 ? 1 : 0;
    for (Index i = is; i < nx; ++i)
      for (Index j = js; j < ny; ++j)
        for (Index k = ks; k < nz; ++k) {
          Real p = o(i, j, k);
          Real q = c == 0 ? o(i - 1, j, k) : c == 1 ? o(i, j - 1, k) : o(i, j, k - 1);
          a(i, j, k) = 2 * p * q / (p + q);
        }
  };
  for (int c = 0; c < 3; ++c) arith(mp.eps[c], c);
  for (int c = 0; c < 3; ++c) harm(mp.mu[c], c);
  for (int c = 0; c < 3; ++c) arith(mp.se[c], c);
  for (int c = 0; c < 3; ++c) harm(mp.sm[c], c);
}

void Sim::calculateCoefficient() {  // calculation_param.cpp:170-264
  const Index nx = g.nx(), ny = g.ny(), nz = g.nz();
  for (int h = 0; h < 2; ++h)
    for (int c = 0; c < 3; ++c) {
      Index ext[3];
      extentOf(h, c, nx, ny, nz, ext);
      Arr3& cs = this->c.self(h, c);
      Arr3& ca = this->c.dualA(h, c);
      Arr3& cb = this->c.dualB(h, c);
      cs = Arr3(ext[0], ext[1], ext[2]);
      ca = Arr3(ext[0], ext[1], ext[2]);
      cb = Arr3(ext[0], ext[1], ext[2]);
      // sizes per axis: E component c uses e_size on axis c and h_size on
      // the others; H component c uses h_size on c and e_size elsewhere.
      const Arr1* sz[3];
      for (int a = 0; a < 3; ++a)
        sz[a] = h ? (a == c ? &g.h_size[a] : &g.e_size[a]) : (a == c ? &g.e_size[a] : &g.h_size[a]);
      const Axis A = tanA(static_cast<Axis>(c)), B = tanB(static_cast<Axis>(c));
      Arr3& m = h ? mp.mu[c] : mp.eps[c];
      Arr3& sg = h ? mp.sm[c] : mp.se[c];
      for (Index i = 0; i < ext[0]; ++i)
        for (Index j = 0; j < ext[1]; ++j)
          for (Index k = 0; k < ext[2]; ++k) {
            Index id[3] = {i, j, k};
            // xt::meshgrid of the size arrays broadcasts against the field
            // extent: the one-cell axis of a 1D/2D grid holds a single size
            // (1) that every index along it takes, also where the component
            // extends to 2 (Ex/Ey/Hz along z in 2D, Ey/Ez/Hx along x in 1D)
            Real da = (*sz[A])(sz[A]->size() == 1 ? 0 : id[A]);
            Real db = (*sz[B])(sz[B]->size() == 1 ? 0 : id[B]);
            Real e = m(i, j, k), s = sg(i, j, k);
            if (!h) {
              cs(i, j, k) = (2 * e - dt * s) / (2 * e + dt * s);
              ca(i, j, k) = -(2 * dt / db) / (2 * e + s * dt);
              cb(i, j, k) = (2 * dt / da) / (2 * e + s * dt);
            } else {
              cs(i, j, k) = (2 * e - dt * s) / (2 * e + dt * s);
              ca(i, j, k) = (2 * dt / db) / (2 * e + s * dt);
              cb(i, j, k) = -(2 * dt / da) / (2 * e + s * dt);
            }
          }
    }
}

void Sim::init() {  // simulation.cpp:319-371
  if (objects.empty()) throw Error("No object is added");
  generateGridSpace();
  generateEMF();
  // Simulation::makeTimeParam (simulation.cpp:654-676) and
  // TimeParam::calculateDt (time_param.cpp:10-21) for the grid's dimension
  if (g.dim == 1)
    dt = cfl / (C_0 * std::sqrt(1.0 / (g.min_d[2] * g.min_d[2])));
  else if (g.dim == 2)
    dt = cfl / (C_0 * std::sqrt(1.0 / (g.min_d[0] * g.min_d[0]) + 1.0 / (g.min_d[1] * g.min_d[1])));
  else
    dt = cfl / (C_0 * std::sqrt(1.0 / (g.min_d[0] * g.min_d[0]) + 1.0 / (g.min_d[1] * g.min_d[1]) +
                                1.0 / (g.min_d[2] * g.min_d[2])));
  for (auto& o : objects) o->init(*this);
  for (auto& p : pmls) p->init(*this);
  for (auto& t : tfsfs) t->init(*this);
  generateMaterialSpace();
  calculateCoefficient();
  for (auto& o : objects) o->correctUpdateCoefficient(*this);  // simulation.cpp:638-652
  for (auto& p : pmls) p->correctUpdateCoefficient(*this);
  buildDispersiveSpace();  // simulation.cpp:651
  // Simulation::makeUpdator (simulation.cpp:688-743): the ADE updators are
  // 3D only; any other dispersive case throws
  if (ade && g.dim != 3) throw Error("don't support this type of updator yet.");
  // the NF2FF surfaces (nffft.cpp:97-223) assume a 3D box; a 1D/2D grid has
  // no closed surface to integrate over, so this restatement refuses it
  if (g.dim != 3 && (!nf2ffs.empty() || !nf2ff_tds.empty()))
    throw Error("NF2FF needs a 3D grid space");
  for (auto& m : monitors) m->init(*this);
  for (auto& n : networks) n->init(*this);  // simulation.cpp:353-355
  for (auto& n : nf2ffs) n->init(*this);  // simulation.cpp:350-358
  for (auto& n : nf2ff_tds) n->init(*this);
}

void Sim::initSteps(Index n) {  // simulation.cpp:373-389
  init();
  start_step = 0;
  size = n;
  current = 0;
  for (auto& o : objects) o->initTimeDependentVariable(*this);
  for (auto& t : tfsfs) t->initTimeDependentVariable(*this);
  for (auto& n : nf2ffs) n->initTimeDependentVariable(*this);  // simulation.cpp:373-389
  for (auto& n : nf2ff_tds) n->initTimeDependentVariable(*this);
  for (auto& m : monitors) m->initTimeDependentVariable(*this);
  if (objective) objective->dft.assign(objective->n[0] * objective->n[1] * objective->n[2], Cplx(0, 0));
}

namespace {
// Domain::run (domain.cpp:82-135) for one thread task.
struct Domain {
  Sim* s;
  Task task;
  bool master;
  std::vector<std::unique_ptr<Corrector>> correctors;

  void updateH() {  // basic_updator.cpp:23-41, update_scheme.h:64-108
    auto& f = s->emf;
    auto& c = s->c;
    for (Index i = task.x.s; i < task.x.e; ++i)
      for (Index j = task.y.s; j < task.y.e; ++j)
        for (Index k = task.z.s; k < task.z.e; ++k)
          f.hx(i, j, k) = c.chxh(i, j, k) * f.hx(i, j, k) +
                          c.chxey(i, j, k) * (f.ey(i, j, k + 1) - f.ey(i, j, k)) +
                          c.chxez(i, j, k) * (f.ez(i, j + 1, k) - f.ez(i, j, k));
    for (Index i = task.x.s; i < task.x.e; ++i)
      for (Index j = task.y.s; j < task.y.e; ++j)
        for (Index k = task.z.s; k < task.z.e; ++k)
          f.hy(i, j, k) = c.chyh(i, j, k) * f.hy(i, j, k) +
                          c.chyez(i, j, k) * (f.ez(i + 1, j, k) - f.ez(i, j, k)) +
                          c.chyex(i, j, k) * (f.ex(i, j, k + 1) - f.ex(i, j, k));
    for (Index i = task.x.s; i < task.x.e; ++i)
      for (Index j = task.y.s; j < task.y.e; ++j)
        for (Index k = task.z.s; k < task.z.e; ++k)
          f.hz(i, j, k) = c.chzh(i, j, k) * f.hz(i, j, k) +
                          c.chzex(i, j, k) * (f.ex(i, j + 1, k) - f.ex(i, j, k)) +
                          c.chzey(i, j, k) * (f.ey(i + 1, j, k) - f.ey(i, j, k));
  }

  // TemplateADEUpdateScheme::updateE (template_ade_update_scheme.h:14-88)
  // for Drude / Debye storage and MLorentzUpdator::updateE
  // (m_lor_ade_updator.cpp:21-87) for modified-Lorentz storage, over the
  // same ranges as the plain update, for every cell.
  void updateEAde(int comp) {
    auto& f = s->emf;
    auto& c = s->c;
    ADEStorage& st = *s->ade;
    const int A = (comp + 1) % 3, B = (comp + 2) % 3;
    Arr3& e = f.f(0, comp);
    const Arr3& ha = f.f(1, A);
    const Arr3& hb = f.f(1, B);
    const Arr3& cs = c.self(0, comp);
    const Arr3& ca = c.dualA(0, comp);
    const Arr3& cb = c.dualB(0, comp);
    Index lo[3] = {task.x.s, task.y.s, task.z.s};
    const Index hi[3] = {task.x.e, task.y.e, task.z.e};
    if (lo[A] == 0) lo[A] = 1;
    if (lo[B] == 0) lo[B] = 1;
    const bool lor = st.kind == ADEStorage::MLorentz;
    std::vector<Real>& J = st.j[comp];
    for (Index i = lo[0]; i < hi[0]; ++i)
      for (Index j = lo[1]; j < hi[1]; ++j)
        for (Index k = lo[2]; k < hi[2]; ++k) {
          Index ia[3] = {i, j, k}, ib[3] = {i, j, k};
          ia[B] -= 1;  // dual A neighbour one back along B
          ib[A] -= 1;  // dual B neighbour one back along A
          Real j_sum = 0;
          for (Index p = 0; p < st.P; ++p) {
            const Index q = st.at(i, j, k, p);
            if (lor)
              j_sum += st.cjsj[q] * J[q] + st.cjsjp[q] * st.j_prev[comp][q];
            else
              j_sum += st.cjsj[q] * J[q];
          }
          const Real e_prev = lor ? st.e_prev[comp](i, j, k) : 0.0;
          const Real cee_p = lor ? st.cee_p(i, j, k) : 0.0;
          const Real e_cur = e(i, j, k);
          const Real e_next = cs(i, j, k) * e(i, j, k) + ca(i, j, k) * (ha(i, j, k) - ha(ia[0], ia[1], ia[2])) +
                              cb(i, j, k) * (hb(i, j, k) - hb(ib[0], ib[1], ib[2]));
          e(i, j, k) = cee_p * e_prev + e_next + st.cej_sum(i, j, k) * j_sum;
          const Real en = e(i, j, k);
          for (Index p = 0; p < st.P; ++p) {
            const Index q = st.at(i, j, k, p);
            if (st.kind == ADEStorage::Debye) {  // debye_ade_updator.cpp:41-49
              J[q] = st.cjj[q] * J[q] + st.cje[q] * (en - e_cur);
            } else if (st.kind == ADEStorage::Drude) {  // drude_ade_updator.cpp:31-39
              J[q] = st.cjj[q] * J[q] + st.cje[q] * (en + e_cur);
            } else {  // m_lor_ade_updator.cpp:89-112
              const Real jn = st.cjen[q] * en + st.cje[q] * e_cur + st.cjep[q] * e_prev + st.cjj[q] * J[q] +
                              st.cjjp[q] * st.j_prev[comp][q];
              st.j_prev[comp][q] = J[q];
              J[q] = jn;
            }
          }
          if (lor) st.e_prev[comp](i, j, k) = e_cur;
        }
  }

  void updateE() {  // basic_updator_3d.cpp:38-69
    auto& f = s->emf;
    auto& c = s->c;
    if (s->ade) {
      for (int comp = 0; comp < 3; ++comp) updateEAde(comp);
      return;
    }
    const Index xs = task.x.s, ys = task.y.s, zs = task.z.s;
    const Index xe = task.x.e, ye = task.y.e, ze = task.z.e;
    const Index xs1 = xs == 0 ? 1 : xs, ys1 = ys == 0 ? 1 : ys, zs1 = zs == 0 ? 1 : zs;
    if (s->g.dim == 1) {
      // BasicUpdatorTEM::updateE (basic_updator.cpp:90-109): Ex on the one
      // x/y edge, its Hz term written as zeros
      for (Index k = zs1; k < ze; ++k)
        f.ex(0, 0, k) = c.cexe(0, 0, k) * f.ex(0, 0, k) + c.cexhy(0, 0, k) * (f.hy(0, 0, k) - f.hy(0, 0, k - 1)) +
                        0.0 * (0.0 - 0.0);
      return;
    }
    // a 2D grid (nz = 1) runs BasicUpdatorTE::updateE (basic_updator_te.cpp:
    // 23-38): the Ez update below; the Ex and Ey ranges [1, nz) are empty
    for (Index i = xs; i < xe; ++i)
      for (Index j = ys1; j < ye; ++j)
        for (Index k = zs1; k < ze; ++k)
          f.ex(i, j, k) = c.cexe(i, j, k) * f.ex(i, j, k) +
                          c.cexhy(i, j, k) * (f.hy(i, j, k) - f.hy(i, j, k - 1)) +
                          c.cexhz(i, j, k) * (f.hz(i, j, k) - f.hz(i, j - 1, k));
    for (Index i = xs1; i < xe; ++i)
      for (Index j = ys; j < ye; ++j)
        for (Index k = zs1; k < ze; ++k)
          f.ey(i, j, k) = c.ceye(i, j, k) * f.ey(i, j, k) +
                          c.ceyhz(i, j, k) * (f.hz(i, j, k) - f.hz(i - 1, j, k)) +
                          c.ceyhx(i, j, k) * (f.hx(i, j, k) - f.hx(i, j, k - 1));
    for (Index i = xs1; i < xe; ++i)
      for (Index j = ys1; j < ye; ++j)
        for (Index k = zs; k < ze; ++k)
          f.ez(i, j, k) = c.ceze(i, j, k) * f.ez(i, j, k) +
                          c.cezhx(i, j, k) * (f.hx(i, j, k) - f.hx(i, j - 1, k)) +
                          c.cezhy(i, j, k) * (f.hy(i, j, k) - f.hy(i - 1, j, k));
  }
};
}  // namespace

void Sim::run(Index n) {  // simulation.cpp:284-289
  initSteps(n);
  loop();
}

void Sim::loop() {  // simulation.cpp:291-317 with the Domain::run bodies
  Task problem{{0, g.nx()}, {0, g.ny()}, {0, g.nz()}};
  auto tasks = decomposeTask(problem, tcfg[0], tcfg[1], tcfg[2]);
  if (tasks.size() > std::max(1u, std::thread::hardware_concurrency()))
    throw Error("The number of threads is too large");
  std::vector<Domain> doms(tasks.size());
  // simulation.cpp:437-489: object correctors (whole problem) on the master
  // domain, then per-task TFSF and PML correctors, in add order.
  for (Index id = 0; id < tasks.size(); ++id) {
    auto& d = doms[id];
    d.s = this;
    d.task = tasks[id];
    d.master = id == 0;
    if (d.master)
      for (auto& o : objects)
        if (auto c = o->generateCorrector(*this, problem)) d.correctors.push_back(std::move(c));
    for (auto& t : tfsfs)
      if (auto c = t->generateCorrector(*this, tasks[id])) d.correctors.push_back(std::move(c));
    for (auto& p : pmls)
      if (auto c = p->generateCorrector(*this, tasks[id])) d.correctors.push_back(std::move(c));
  }
  std::barrier bar(static_cast<std::ptrdiff_t>(doms.size()));
  std::chrono::steady_clock::time_point t0;
  auto body = [&](Domain& d) {
    for (Index t = start_step; t < start_step + size; ++t) {
      d.updateH();
      bar.arrive_and_wait();
      for (auto& c : d.correctors) c->correctH(t);
      bar.arrive_and_wait();
      bar.arrive_and_wait();
      d.updateE();
      bar.arrive_and_wait();
      for (auto& c : d.correctors) c->correctE(t);
      bar.arrive_and_wait();
      if (d.master) {
        current = t;
        for (auto& m : monitors) m->update(*this, t);
        for (auto& n : nf2ffs) n->update(*this, t);  // Domain::record (domain.cpp:176-184)
        for (auto& n : nf2ff_tds) n->update(*this, t);
        if (objective) {
          auto& ob = *objective;
          const Real ph = -2 * PI * ob.freq * static_cast<Real>(t + 1) * dt;
          const Cplx w(dt * std::cos(ph), dt * std::sin(ph));
          Arr3& f = emf.byId(ob.field);
          Index p = 0;
          for (Index i = 0; i < ob.n[0]; ++i)
            for (Index j = 0; j < ob.n[1]; ++j)
              for (Index k = 0; k < ob.n[2]; ++k, ++p) ob.dft[p] += w * f(ob.o[0] + i, ob.o[1] + j, ob.o[2] + k);
        }
      }
      bar.arrive_and_wait();
      bar.arrive_and_wait();
      if (d.master)
        step_end_seconds.push_back(
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
  };
  step_end_seconds.clear();
  t0 = std::chrono::steady_clock::now();
  {
    std::vector<std::thread> th;
    for (Index id = 1; id < doms.size(); ++id) th.emplace_back(body, std::ref(doms[id]));
    body(doms[0]);
    for (auto& x : th) x.join();
  }
  loop_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  current = start_step + size;
  for (auto& m : monitors) m->output(*this);
}

// ------------------------------------------------------------ NF2FF (TD) --
namespace {
// rVector (nffft_td_data.h:22-35): E node along the surface normal, H nodes
// across it (the same for E and H)
Vec3 tdR(Sim& s, int c, Index i, Index j, Index k) {
  if (c == 0) return {s.g.e_node[0](i), s.g.h_node[1](j), s.g.h_node[2](k)};
  if (c == 1) return {s.g.h_node[0](i), s.g.e_node[1](j), s.g.h_node[2](k)};
  return {s.g.h_node[0](i), s.g.h_node[1](j), s.g.e_node[2](k)};
}
// surfaceArea (nffft_td_data.h:37-47)
Real tdArea(Sim& s, int c, Index i, Index j, Index k) {
  if (c == 0) return s.g.e_size[1](j) * s.g.e_size[2](k);
  if (c == 1) return s.g.e_size[2](k) * s.g.e_size[0](i);
  return s.g.e_size[0](i) * s.g.e_size[1](j);
}
Vec3 tdDirection(Real theta, Real phi) {  // observationDirection (:520-522)
  return {std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
}
}  // namespace

void NF2FFTD::init(Sim& s) {
  Task global[6];
  nffftSurfaces(s, dist, surf, global);
  // distanceRange (nfffft_time_domain.cpp:410-480) over the global surfaces
  const Vec3 u = tdDirection(theta, phi);
  dmin = std::numeric_limits<Real>::max();
  dmax = std::numeric_limits<Real>::lowest();
  for (int d = 0; d < 6; ++d) {
    const Task& tk = global[d];
    for (Index k = tk.z.s; k < tk.z.e; ++k)
      for (Index j = tk.y.s; j < tk.y.e; ++j)
        for (Index i = tk.x.s; i < tk.x.e; ++i) {
          const Vec3 r = tdR(s, d / 2, i, j, k);
          const Real dis = r.x * u.x + r.y * u.y + r.z * u.z;
          dmin = std::min(dmin, dis);
          dmax = std::max(dmax, dis);
        }
  }
  for (int d = 0; d < 6; ++d) {
    const Index cells = surf[d].valid() ? surf[d].x.size() * surf[d].y.size() * surf[d].z.size() : 0;
    for (int w = 0; w < 4; ++w) prev[d][w].assign(cells, 0.0);
  }
}

void NF2FFTD::initTimeDependentVariable(Sim& s) {  // nfffft_time_domain.cpp:53-66
  const Real aux = (dmax - dmin) / (C_0 * s.dt);
  num = static_cast<Index>(std::ceil(aux)) + s.size + 4;
  for (int d = 0; d < 6; ++d)
    for (int w = 0; w < 4; ++w) pot[d][w].assign(num, 0.0);
}

void NF2FFTD::update(Sim& s, Index t) {
  const Vec3 u = tdDirection(theta, phi);
  for (int attr = 0; attr < 2; ++attr)  // E on every surface, then H
    for (int d = 0; d < 6; ++d) {
      const Task& tk = surf[d];
      if (!tk.valid()) continue;
      const int c = d / 2, A = (c + 1) % 3, B = (c + 2) % 3;
      const bool is_e = attr == 0, neg = d % 2 == 0;
      const int fa = (is_e ? 0 : 3) + A, fb = (is_e ? 0 : 3) + B;
      const Real offset = is_e ? 0.5 : 0.0;
      Real coeff_a = is_e ? 1.0 : -1.0;
      if (neg) coeff_a *= -1.0;
      Real coeff_b = is_e ? -1.0 : 1.0;
      if (neg) coeff_b *= -1.0;
      auto& pa = pot[d][is_e ? 0 : 2];
      auto& pb = pot[d][is_e ? 1 : 3];
      auto& prev_a = prev[d][is_e ? 0 : 2];
      auto& prev_b = prev[d][is_e ? 1 : 3];
      const Real dt = s.dt;
      Index q = 0;
      for (Index i = tk.x.s; i < tk.x.e; ++i)
        for (Index j = tk.y.s; j < tk.y.e; ++j)
          for (Index k = tk.z.s; k < tk.z.e; ++k, ++q) {
            const Vec3 r = tdR(s, c, i, j, k);
            const Real time_delay = dmax - (r.x * u.x + r.y * u.y + r.z * u.z);
            const Real time_step_delay = time_delay / (C_0 * dt);
            const Index nn = static_cast<Index>(std::floor(time_step_delay + offset + t));
            const Real coeff_f = time_step_delay + offset + t - nn;
            const Real ds = tdArea(s, c, i, j, k);
            const Real a_avg = surfaceCenter(s, c, fa, i, j, k);
            const Real b_avg = surfaceCenter(s, c, fb, i, j, k);
            const Real delta_a = a_avg - prev_a[q];
            const Real delta_b = b_avg - prev_b[q];
            pa.at(nn) += coeff_a * (1 - coeff_f) * ds * delta_b;
            pa.at(nn + 1) += coeff_a * coeff_f * ds * delta_b;
            pb.at(nn) += coeff_b * (1 - coeff_f) * ds * delta_a;
            pb.at(nn + 1) += coeff_b * coeff_f * ds * delta_a;
            prev_a[q] = a_avg;
            prev_b[q] = b_avg;
          }
    }
}

void NF2FFTD::farField(Sim& s, std::vector<Real> out[4]) const {
  // wa() ... ub(): the potentials / (C_0 dt 4 pi) (nffft_td_data.h:239-261)
  auto sc = [&](int d, int w) {
    std::vector<Real> v(num);
    for (Index n = 0; n < num; ++n) v[n] = pot[d][w][n] / (C_0 * s.dt * 4 * PI);
    return v;
  };
  // wx = zn.wa + zp.wa + yn.wb + yp.wb, ... (nfffft_time_domain.cpp:484-512)
  auto sum4 = [&](std::vector<Real> a, const std::vector<Real>& b, const std::vector<Real>& c,
                  const std::vector<Real>& d) {
    for (Index n = 0; n < num; ++n) a[n] = a[n] + b[n] + c[n] + d[n];
    return a;
  };
  const auto wx = sum4(sc(ZN, 2), sc(ZP, 2), sc(YN, 3), sc(YP, 3));
  const auto wy = sum4(sc(XN, 2), sc(XP, 2), sc(ZN, 3), sc(ZP, 3));
  const auto wz = sum4(sc(XN, 3), sc(XP, 3), sc(YN, 2), sc(YP, 2));
  const auto ux = sum4(sc(ZN, 0), sc(ZP, 0), sc(YN, 1), sc(YP, 1));
  const auto uy = sum4(sc(XN, 0), sc(XP, 0), sc(ZN, 1), sc(ZP, 1));
  const auto uz = sum4(sc(XN, 1), sc(XP, 1), sc(YN, 0), sc(YP, 0));
  const Real ct = std::cos(theta), st = std::sin(theta), cp = std::cos(phi), sp = std::sin(phi);
  for (int w = 0; w < 4; ++w) out[w].assign(num, 0.0);
  for (Index n = 0; n < num; ++n) {
    out[0][n] = wx[n] * ct * cp + wy[n] * ct * sp - wz[n] * st;
    out[1][n] = -wx[n] * sp + wy[n] * cp;
    out[2][n] = ux[n] * ct * cp + uy[n] * ct * sp - uz[n] * st;
    out[3][n] = -ux[n] * sp + uy[n] * cp;
  }
}

// --------------------------------------------------------------- network --
std::vector<std::complex<Real>> seriesDft(const std::vector<Real>& x, Real dt, const std::vector<Real>& freqs,
                                          Real shift) {
  const std::complex<Real> II(0.0, 1.0);  // constant.h:23
  std::vector<std::complex<Real>> res(freqs.size());
  for (std::size_t i = 0; i < freqs.size(); ++i) {
    std::complex<Real> sum{0.0, 0.0};
    for (std::size_t j = 0; j < x.size(); ++j) {
      auto t = (j + 1) * dt + shift;
      sum += x[j] * std::exp(-static_cast<Real>(2.0) * II * PI * freqs[i] * t);
    }
    res[i] = sum * dt;
  }
  return res;
}

void Port::calculateSParameters(const std::vector<Real>& freqs) {
  if (!current || !voltage) throw Error("Port: monitor missing");
  auto k = std::sqrt(std::real(z));
  auto i = seriesDft(current->data.v, dt, freqs, -1.5 * dt);  // port.cpp:122-124
  auto v = seriesDft(voltage->data.v, dt, freqs);
  a.assign(freqs.size(), {});
  b.assign(freqs.size(), {});
  for (std::size_t f = 0; f < freqs.size(); ++f) {
    a[f] = Real{0.5} * (v[f] + z * i[f]) / k;
    b[f] = Real{0.5} * (v[f] - std::conj(z) * i[f]) / k;
  }
}

void Network::init(Sim& sim) {
  std::map<Index, int> seen;
  for (auto& p : ports) {
    p.dt = sim.dt;  // Port::init (port.cpp:37-42)
    if (!seen.emplace(p.index, 0).second) throw Error("Port index must be unique");
  }
}

void Network::output() {
  for (auto& p : ports) p.calculateSParameters(freqs);
  s.clear();
  for (std::size_t i = 0; i < ports.size(); ++i) {
    if (!ports[i].is_source) continue;
    for (std::size_t j = 0; j < ports.size(); ++j) {
      std::vector<std::complex<Real>> sv(freqs.size());
      for (std::size_t f = 0; f < freqs.size(); ++f) sv[f] = ports[j].b[f] / ports[i].a[f];
      s[static_cast<int>(ports[j].index * 10 + ports[i].index)] = sv;
    }
  }
  if (out_dir.empty()) return;
  std::filesystem::create_directories(out_dir);
  for (const auto& e : s)
    writeNpyComplex(out_dir + "/s" + std::to_string(e.first) + ".npy", {e.second.size()}, e.second.data());
  writeNpy(out_dir + "/frequencies.npy", {freqs.size()}, freqs.data());
}

// ------------------------------------------------------------------- npy --
namespace {
void writeNpyRaw(const std::string& path, const std::vector<Index>& shape, const char* descr, const void* data,
                 std::size_t item_bytes) {
  std::string sh = "(";
  for (Index i = 0; i < shape.size(); ++i) {
    sh += std::to_string(shape[i]);
    if (shape.size() == 1 || i + 1 < shape.size()) sh += ", ";
  }
  sh += ")";
  std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
  Index total = 10 + hdr.size() + 1;
  Index pad = (64 - total % 64) % 64;
  hdr += std::string(pad, ' ') + "\n";
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) throw Error("cannot write " + path);
  const char magic[] = "\x93NUMPY\x01\x00";
  std::fwrite(magic, 1, 8, fp);
  uint16_t hl = static_cast<uint16_t>(hdr.size());
  std::fwrite(&hl, 2, 1, fp);
  std::fwrite(hdr.data(), 1, hdr.size(), fp);
  Index n = 1;
  for (auto v : shape) n *= v;
  std::fwrite(data, item_bytes, n, fp);
  std::fclose(fp);
}
}  // namespace

void writeNpy(const std::string& path, const std::vector<Index>& shape, const Real* data) {
  writeNpyRaw(path, shape, "<f8", data, sizeof(Real));
}

// complex128, interleaved (re, im) as xt::dump_npy writes std::complex<double>
void writeNpyComplex(const std::string& path, const std::vector<Index>& shape, const std::complex<Real>* data) {
  writeNpyRaw(path, shape, "<c16", data, sizeof(std::complex<Real>));
}

}  // namespace refcpu
