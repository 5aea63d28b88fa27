// refcpu — FP64 CPU restatement of the XFDTD_CORE 3D time-stepping path.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
// against, and the "port" CPU baseline that bench.py times.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product never does.
//
// It restates, from the reference sources read as text, the reference's own
// algorithm and its evaluation order (every function cites the file:line it
// follows; paths are relative to the reference repository root).  The
// reference itself cannot be compiled here: it hard-requires xtensor <= 0.25
// (CMakeLists.txt:35-41), which is absent, so there is no oracle/_ref build.
//
// Restatement assumptions at the xtensor boundary (parity unpinned at the
// ulp level, see the header of each helper):
//   * xt::linspace assigned into a container uses the generator's assign_to
//     path: cumulative addition of the step, last element forced to `stop`.
//   * view-assignment whose right-hand side aliases the target is evaluated
//     through a temporary (pre-assignment values), as xtensor's
//     xsemantic_base::operator= does.
//   * xt::norm_l2 of a 3-vector is sqrt((x*x + y*y) + z*z).
//   * no FMA contraction (the reference builds -std=c++20 -O3 without -march,
//     CMakeLists.txt:17-19); this file is compiled with -ffp-contract=off.
#pragma once

#include <array>
#include <barrier>
#include <cmath>
#include <cstddef>
#include <complex>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace refcpu {

using Real = double;
using Index = std::size_t;

// include/xfdtd/common/constant.h:11-18
inline constexpr Real PI = 3.141592653589793238462643383279502884;
inline constexpr Real EPSILON_0 = 8.854187817e-12;
inline constexpr Real MU_0 = 4 * PI * 1e-7;
inline const Real C_0 = 1 / std::sqrt(EPSILON_0 * MU_0);
inline constexpr Real SIGMA_ZERO = 1e-20;  // material.h:40-41

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------- arrays --
struct Arr1 {
  std::vector<Real> v;
  Arr1() = default;
  explicit Arr1(Index n, Real x = 0) : v(n, x) {}
  Real& operator()(Index i) { return v[i]; }
  Real operator()(Index i) const { return v[i]; }
  Index size() const { return v.size(); }
};

struct Arr2 {
  Index n0 = 0, n1 = 0;
  std::vector<Real> v;
  Arr2() = default;
  Arr2(Index a, Index b) : n0(a), n1(b), v(a * b, 0.0) {}
  Real& operator()(Index i, Index j) { return v[i * n1 + j]; }
  Real operator()(Index i, Index j) const { return v[i * n1 + j]; }
};

// Row-major with k fastest (include/xfdtd/common/type_define.h:16-31).
struct Arr3 {
  Index n0 = 0, n1 = 0, n2 = 0;
  std::vector<Real> v;
  Arr3() = default;
  Arr3(Index a, Index b, Index c, Real x = 0) : n0(a), n1(b), n2(c), v(a * b * c, x) {}
  Real& operator()(Index i, Index j, Index k) { return v[(i * n1 + j) * n2 + k]; }
  Real operator()(Index i, Index j, Index k) const { return v[(i * n1 + j) * n2 + k]; }
  Index size() const { return v.size(); }
};

// xt::linspace(start, stop, n) assigned to a container (xbuilder.hpp
// arange_generator::assign_to in xtensor 0.25): running sum of the step,
// final element set to stop.  Used by grid_space_generator.cpp:210-212 and
// time_param.cpp:63-71.
Arr1 linspace(Real start, Real stop, Index n);

// ------------------------------------------------------------ index task --
// include/xfdtd/common/index_task.h:10-120
struct Range {
  Index s = 0, e = 0;
  Index size() const { return e - s; }
  bool valid() const { return s < e; }
};
struct Task {
  Range x, y, z;
  bool valid() const { return x.valid() && y.valid() && z.valid(); }
  Range r(int axis) const { return axis == 0 ? x : axis == 1 ? y : z; }
};
bool intersected(const Range& a, const Range& b);
bool intersected(const Task& a, const Task& b);
std::optional<Task> intersection(const Task& a, const Task& b);
// src/include/util/decompose_task.h:9-93
std::vector<Task> decomposeTask(const Task& problem, int nx, int ny, int nz);

// ------------------------------------------------------------------ axes --
enum Axis { X = 0, Y = 1, Z = 2 };
enum Dir { XN = 0, XP, YN, YP, ZN, ZP };
inline Axis dirAxis(Dir d) { return static_cast<Axis>(static_cast<int>(d) / 2); }
inline bool dirNegative(Dir d) { return static_cast<int>(d) % 2 == 0; }
// coordinate_system.h:191-217: tangential A/B of c are (y,z), (z,x), (x,y).
inline Axis tanA(Axis c) { return static_cast<Axis>((c + 1) % 3); }
inline Axis tanB(Axis c) { return static_cast<Axis>((c + 2) % 3); }
// util/transform/abc_xyz.h:66-112: (x,y,z) -> (a,b,c) with c the main axis.
inline std::array<Index, 3> toABC(Axis c, Index x, Index y, Index z) {
  if (c == X) return {y, z, x};
  if (c == Y) return {z, x, y};
  return {x, y, z};
}
inline std::array<Index, 3> fromABC(Axis c, Index a, Index b, Index cc) {
  if (c == X) return {cc, a, b};
  if (c == Y) return {b, cc, a};
  return {a, b, cc};
}

// ---------------------------------------------------------------- vector --
struct Vec3 {
  Real x = 0, y = 0, z = 0;
  Real operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
  Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  Vec3 operator/(Real s) const { return {x / s, y / s, z / s}; }
  Real norm() const { return std::sqrt((x * x + y * y) + z * z); }
  Real dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
  Vec3 cross(const Vec3& o) const {  // vector.cpp:77-80
    return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x};
  }
};

// ---------------------------------------------------------------- shapes --
struct Cube;
struct Shape {
  virtual ~Shape() = default;
  virtual bool isInside(const Vec3& p, Real eps) const = 0;
  virtual Cube wrapped() const = 0;
};
struct Cube : Shape {
  Vec3 origin, size, end;
  Cube() = default;
  Cube(Vec3 o, Vec3 s);  // cube.cpp:60-64 (end = origin + size)
  bool isInside(const Vec3& p, Real eps) const override;  // cube.cpp:107-118
  Cube wrapped() const override { return *this; }
};
struct Sphere : Shape {
  Vec3 center;
  Real radius;
  Sphere(Vec3 c, Real r) : center(c), radius(r) {}
  bool isInside(const Vec3& p, Real eps) const override;  // sphere.cpp:29-32
  Cube wrapped() const override;                            // sphere.cpp:34-37
};
struct Cylinder : Shape {
  Vec3 center;
  Real radius, height;
  Axis axis;
  Cylinder(Vec3 c, Real r, Real h, Axis a) : center(c), radius(r), height(h), axis(a) {}
  bool isInside(const Vec3& p, Real eps) const override;
  Cube wrapped() const override;
};
// src/include/util/float_compare.h:24-42 (LessEqual)
inline bool lessEqual(Real lhs, Real rhs, Real eps) {
  return std::abs(lhs - rhs) <= eps ? true : lhs < rhs;
}

// ------------------------------------------------------------- waveform ---
struct Waveform {
  std::function<Real(Real)> f;
  Real amplitude = 1;
  Arr1 value;
  void init(const Arr1& time);  // waveform.cpp:28-34
};

// ------------------------------------------------------------- material ---
// src/material/material.cpp:11-26 and dispersive_material_equation.cpp
struct Material {
  Real eps_r = 1, mu_r = 1, sigma_e = SIGMA_ZERO, sigma_m = SIGMA_ZERO;
  Real eps = EPSILON_0, mu = MU_0;
  // Linear dispersion in the reference's modified-Lorentz form
  // (a0 + a1 jw) / (b0 + b1 jw + b2 (jw)^2) per pole.
  enum Kind { Plain = 0, Lorentz, Drude, Debye } kind = Plain;
  Real eps_inf = 1;
  std::vector<Real> a0, a1, b0, b1, b2;
  // Drude / Debye native parameters (drude_medium.cpp, debye_medium.cpp)
  std::vector<Real> omega_p, gamma, eps_static, tau;
  bool dispersive() const { return kind != Plain; }
  Index numPoles() const { return a0.size(); }
  static Material make(Real eps_r, Real mu_r, Real se, Real sm);
  // MLorentzMaterial::makeMLorentz (dipersive_material.cpp:47-56)
  static Material makeLorentz(Real eps_inf, std::vector<Real> a0, std::vector<Real> a1,
                              std::vector<Real> b0, std::vector<Real> b1, std::vector<Real> b2);
  // DrudeMedium::makeDrudeMedium + DrudeEqDecision (drude_medium.cpp:9-14,
  // dispersive_material_equation.cpp:55-59)
  static Material makeDrude(Real eps_inf, std::vector<Real> omega_p, std::vector<Real> gamma);
  // DebyeMedium::makeDebyeMedium + DebyeEqDecision (debye_medium.cpp:10-17,
  // dispersive_material_equation.cpp:48-53)
  static Material makeDebye(Real eps_inf, std::vector<Real> eps_static, std::vector<Real> tau);
};

struct Sim;

// ADEMethodStorage and its Drude / Debye / modified-Lorentz forms
// (ade_method.h, drude_ade_method.cpp, debye_ade_method.cpp,
// m_lor_ade_method.cpp): per-cell (nx, ny, nz) and per-pole (p fastest)
// coefficient and state arrays.
struct ADEStorage {
  enum Kind { Drude = 0, Debye = 1, MLorentz = 2 } kind;
  Index P = 0, nx = 0, ny = 0, nz = 0;
  std::vector<Real> cjj, cjjp, cjen, cje, cjep, cjsj, cjsjp;  // (nx, ny, nz, P)
  Arr3 cej_sum, cee_p;                                       // (nx, ny, nz)
  Arr3 e_prev[3];
  std::vector<Real> j[3], j_prev[3];                         // (nx, ny, nz, P)
  ADEStorage(Kind k, Index p, Index x, Index y, Index z);
  Index at(Index i, Index jj, Index k, Index p) const { return ((i * ny + jj) * nz + k) * P + p; }
  void correctCoeff(Sim& s, Index i, Index jj, Index k, const Material& m);
};

// ----------------------------------------------------------------- grid ---
struct GridSpace {
  // GridSpace::Dimension (grid_space.h): 3, or 2 (z infinite: one cell
  // along z, grid_space_2d.cpp) or 1 (x and y infinite: one cell along each,
  // grid_space_1d.cpp)
  int dim = 3;
  Real based[3];
  Real min_d[3];
  Arr1 e_node[3], h_node[3], e_size[3], h_size[3];
  Index n(int a) const { return h_node[a].size(); }
  Index nx() const { return n(0); }
  Index ny() const { return n(1); }
  Index nz() const { return n(2); }
  Real eps() const;                              // grid_space.cpp:459-468
  Index nearest(int axis, Real x) const;         // grid_space.cpp:383-396
  std::array<Index, 3> box(const Shape& s, Index* end) const;  // :282-287
};

// ---------------------------------------------------------------- fields --
struct EMF {
  Arr3 ex, ey, ez, hx, hy, hz;
  // field index: 0..2 = E x/y/z, 3..5 = H x/y/z
  Arr3& f(int attr_h, int c) {
    static_assert(true);
    if (!attr_h) return c == 0 ? ex : c == 1 ? ey : ez;
    return c == 0 ? hx : c == 1 ? hy : hz;
  }
  Arr3& byId(int id) { return f(id >= 3, id % 3); }
};

// 18 Yee coefficient arrays (fdtd_update_coefficient.h:102-121).
struct Coeff {
  Arr3 cexe, cexhy, cexhz, ceye, ceyhz, ceyhx, ceze, cezhx, cezhy;
  Arr3 chxh, chxey, chxez, chyh, chyez, chyex, chzh, chzex, chzey;
  Arr3& self(int attr_h, int c);
  // coefficient of the dual field along tangential A / B of axis c.
  Arr3& dualA(int attr_h, int c);
  Arr3& dualB(int attr_h, int c);
  Arr3* byName(const std::string& n);
};

struct MaterialParam {
  Arr3 eps[3], mu[3], se[3], sm[3];
};

struct Sim;

// ----------------------------------------------------------- correctors ---
struct Corrector {
  virtual ~Corrector() = default;
  virtual void correctE(Index t) = 0;
  virtual void correctH(Index t) = 0;
};

// -------------------------------------------------------------- objects ---
struct Object {
  std::string name;
  std::unique_ptr<Shape> shape;
  Material material;
  Index material_index = 0;
  // grid box of the shape's wrapped cube (object.cpp:40-43)
  std::array<Index, 3> gb_origin{}, gb_end{};
  virtual ~Object() = default;
  virtual bool isPecPlane() const { return false; }
  virtual void init(Sim& s);
  virtual void correctMaterialSpace(Sim& s, Index idx);  // object.cpp:114-198
  virtual void correctUpdateCoefficient(Sim&) {}
  virtual void initTimeDependentVariable(Sim&) {}
  virtual std::unique_ptr<Corrector> generateCorrector(Sim&, const Task&) { return nullptr; }
};

struct LumpedElement : Object {
  Axis xyz = Z;
  Index is = 0, ie = 0, js = 0, je = 0, ks = 0, ke = 0;
  Index nodeCount(int a) const;                 // lumped_element.cpp:31-56
  Index globalCountMain() const;                // :97-108
  Index globalCountSubA() const;                // :110-121
  Index globalCountSubB() const;                // :123-134
  Range range(int a) const;                     // lumped_element.h:239-264
  void init(Sim& s) override;                   // lumped_element.cpp:14-27
  // meshgrid of (x, y, z) sizes over the element's ranges; returns da, db, dc
  void sizes(Sim& s, Arr3& da, Arr3& db, Arr3& dc) const;
  Task indexTask() const;
};

struct Resistor : LumpedElement {
  Real resistance;
  Real rf = 0;
  Arr3 da, db, dc, beta;
  void init(Sim& s) override;
  void correctUpdateCoefficient(Sim& s) override;
};

// Capacitor: coefficient-only patch with beta = 2 C' dc / (da db)
// (capacitor.cpp:16-112).
struct Capacitor : LumpedElement {
  Real capacitance;
  Real cf = 0;
  Arr3 da, db, dc, beta;
  void init(Sim& s) override;
  void correctUpdateCoefficient(Sim& s) override;
};

// Inductor: a current state J per edge, E += cecjc*J; J += cjcec*E after
// the E update (inductor.cpp:23-135, lumped_element_corrrector.cpp:40-99).
struct Inductor : LumpedElement {
  Real inductance;
  Real lf = 0;
  Arr3 da, db, dc, cecjc, cjcec, j;
  void init(Sim& s) override;
  void correctUpdateCoefficient(Sim& s) override;
  std::unique_ptr<Corrector> generateCorrector(Sim& s, const Task& t) override;
};

// PecPlane: sigma_e of a PEC material on the plane's tangential edges,
// written after material averaging (pec_plane.cpp:12-67,
// simulation.cpp:629-635).
struct PecPlane : Object {
  bool isPecPlane() const override { return true; }
  void correctMaterialSpace(Sim& s, Index idx) override;
};

struct SourceElement : LumpedElement {
  Dir direction;
  Real resistance;
  Waveform waveform;
  bool voltage = true;
  Real rf = 0, amp_factor = 0;
  Arr3 da, db, dc, beta, alpha, coff;
  void init(Sim& s) override;
  void correctUpdateCoefficient(Sim& s) override;
  void initTimeDependentVariable(Sim& s) override;
  std::unique_ptr<Corrector> generateCorrector(Sim& s, const Task& t) override;
};

// ------------------------------------------------------------- boundary ---
struct PML {
  int thickness;
  Index n;
  Dir dir;
  Axis main;
  int order;
  Real sigma_ratio, alpha_min, alpha_max, kappa_max;
  Index g_e_start = 0, g_h_start = 0, node_e_start = 0, node_h_start = 0;
  Task node_task;  // slab in grid indices (pml.cpp:153-187)
  Arr1 g_h_size, g_e_size, kappa_e, kappa_h, a_e, b_e, a_h, b_h;
  Arr3 c_ea_psi_hb, c_eb_psi_ha, c_ha_psi_eb, c_hb_psi_ea;
  Arr3 ea_psi_hb, eb_psi_ha, ha_psi_eb, hb_psi_ea;
  void init(Sim& s);                      // pml.cpp:98-196
  void correctUpdateCoefficient(Sim& s);  // pml.cpp:214-244, 286-524
  void recursiveConvolution(Sim& s);      // pml.cpp:526-633
  std::unique_ptr<Corrector> generateCorrector(Sim& s, const Task& t);
};

// ----------------------------------------------------------------- TFSF ---
struct TFSF3D {
  Index x, y, z;
  Real theta, phi, psi;
  // 3: TFSF3D; 2: TFSF2D(x, y, phi) = TFSF(x, y, 0, pi/2, phi, 0)
  // (tfsf_2d.cpp:10-13); 1: TFSF1D(z, forward) = TFSF(0, 0, z, forward ? 0 :
  // pi, 0, 0) (tfsf_1d.cpp:14-23).  The kind picks the corrector calls.
  int kind = 3;
  bool forward = true;
  Waveform waveform;
  Real st, ct, sp, cp, ss, cs;
  Vec3 k, transform_e, transform_h;
  std::array<Index, 3> box_origin{}, box_size{};
  Real ratio_delta = 1, scaled_dl = 0;
  Index aux = 0;
  Real ceie = 1, chih = 1, ceihi = 0, chiei = 0, abc0 = 0, abc1 = 0;
  Arr1 px_int, py_int, pz_int, px_half, py_half, pz_half;
  Arr2 e_inc, h_inc;
  Real ca[3], cb[3];
  void init(Sim& s);                       // tfsf.cpp:118-175
  void initTimeDependentVariable(Sim& s);  // tfsf.cpp:36-78
  std::unique_ptr<Corrector> generateCorrector(Sim& s, const Task& t);
  Task globalTask() const;
};

// -------------------------------------------------------------- monitors --
struct Monitor {
  std::string name, out_dir;
  virtual ~Monitor() = default;
  virtual void init(Sim&) {}
  virtual void initTimeDependentVariable(Sim&) {}
  virtual void update(Sim&, Index t) = 0;
  virtual void output(Sim&) {}
};

struct FieldMonitor : Monitor {
  std::unique_ptr<Shape> shape;
  int field = 0;
  Index o[3]{}, e[3]{};
  Arr3 data;
  void init(Sim& s) override;  // field_monitor.cpp:17-78
  void update(Sim&, Index) override {}
  void output(Sim& s) override;  // field_monitor.cpp:82-96
};

struct MovieMonitor : Monitor {
  std::unique_ptr<FieldMonitor> frame;
  Index interval = 1, count = 0;
  std::vector<std::pair<std::string, Arr3>> frames;
  void init(Sim& s) override;
  void update(Sim& s, Index t) override;  // movie_monitor.cpp:26-34
};

struct FieldTimeMonitor : Monitor {
  std::unique_ptr<Shape> shape;
  int field = 0;
  Index o[3]{}, e[3]{};
  Arr1 time, data;
  void init(Sim& s) override;                       // field_time_monitor.cpp:15-79
  void initTimeDependentVariable(Sim& s) override;  // :97-111
  void update(Sim& s, Index t) override;            // :81-95
  void output(Sim& s) override;                     // time_monitor.cpp:11-21
};

// VoltageMonitor (voltage_monitor.cpp:15-235) and CurrentMonitor
// (current_monitor.cpp:12-385) for one process: the node task is the global
// task.  `dir` is an Axis::Direction (XN XP YN YP ZN ZP = 0..5).
struct VoltageMonitor : Monitor {
  std::unique_ptr<Shape> shape;
  int dir = 5;
  Index r[3][2]{};  // is ie, js je, ks ke
  Arr1 coff;
  Arr1 time, data;
  void init(Sim& s) override;
  void initTimeDependentVariable(Sim& s) override;
  void update(Sim& s, Index t) override;
  void output(Sim& s) override;  // time_monitor.cpp:11-21
};

struct CurrentMonitor : Monitor {
  std::unique_ptr<Shape> shape;
  int dir = 5;
  Index r[3][2]{};                 // is ie, js je, ks ke
  Index ha_n[2]{}, ha_p[2]{}, hb_n[2]{}, hb_p[2]{};  // [start, end), empty when start >= end
  Arr1 da, db;
  Real positive = 1;
  Arr1 time, data;
  void init(Sim& s) override;
  void initTimeDependentVariable(Sim& s) override;
  void update(Sim& s, Index t) override;
  void output(Sim& s) override;
};

// ------------------------------------------------------------ simulation --
// ------------------------------------------------------------ network ----
// dft (include/xfdtd/util/dft.h:9-26): X(f) = dt * sum_j x_j
// exp(-2 i pi f ((j + 1) dt + shift)), terms in the reference's order.
std::vector<std::complex<Real>> seriesDft(const std::vector<Real>& x, Real dt, const std::vector<Real>& freqs,
                                          Real shift = 0);

// Port (network/port.h, port.cpp:44-130) for one process: the DFT of the
// current record (shifted by -1.5 dt) and of the voltage record, and the
// power waves a = 0.5 (V + z I) / sqrt(Re z), b = 0.5 (V - conj(z) I) /
// sqrt(Re z).
struct Port {
  Index index = 0;
  bool is_source = false;
  std::complex<Real> z{50, 0};
  const CurrentMonitor* current = nullptr;
  const VoltageMonitor* voltage = nullptr;
  Real dt = 0;
  std::vector<std::complex<Real>> a, b;
  void calculateSParameters(const std::vector<Real>& freqs);
};

// Network (network.cpp:32-84): S(j, i) = b_j / a_i for every source port i,
// keyed 10 j + i; "s{key}.npy" (complex128) and "frequencies.npy".
struct Network {
  std::string out_dir;
  std::vector<Real> freqs;
  std::vector<Port> ports;
  std::map<int, std::vector<std::complex<Real>>> s;
  void init(Sim& sim);  // network.cpp:32-45
  void output();        // network.cpp:47-84
};

// ---------------------------------------------------------------- NF2FF ---
// NFFFTFrequencyDomain (nffft.cpp:39-236, nffft_frequency_domain.cpp) with
// one FDPlaneData per frequency (nffft_fd_data.cpp, nffft_fd_data_update.cpp,
// nffft_fd_data_process.cpp), one process: node surfaces == global ones.
using Cplx = std::complex<Real>;
struct NF2FF {
  Index dist[3] = {1, 1, 1};
  std::vector<Real> freqs;
  std::string out_dir;
  bool reference_repeat = true;  // false: one accumulation per step
  // surfaces XN XP YN YP ZN ZP (nffft.cpp:97-223); invalid ones are empty
  Task surf[6];
  struct Plane {
    Real freq = 0;
    // per surface: (ja, jb) and (ma, mb) on the surface task's shape
    std::vector<Cplx> ja[6], jb[6], ma[6], mb[6];
    std::vector<Cplx> te, th;  // DFT phase tables (nffft_fd_data.cpp:65-76)
  };
  std::vector<Plane> planes;
  void init(Sim& s);
  void initTimeDependentVariable(Sim& s);
  void update(Sim& s, Index t);  // nffft_frequency_domain.cpp:45-53
  // potentials (nffft_fd_data_process.cpp:44-160): A (J) or F (M),
  // theta or phi component, over a theta / phi cut (one of them size 1)
  std::vector<Cplx> potential(Sim& s, Index f, bool a_pot, bool theta_comp, const std::vector<Real>& theta,
                              const std::vector<Real>& phi, const Vec3& origin) const;
  Real power(Sim& s, Index f) const;  // nffft_fd_data_process.cpp:28-42, 162-194
};

// NFFFTTimeDomain (nfffft_time_domain.cpp, src/include/nffft/nffft_td_data.h)
// for one process: per step, the time derivative of each surface cell's
// face-centre tangential E and H, scaled by its area, binned by the
// retarded time along one observation direction (theta, phi) into the
// surface's potential arrays (ua, ub from E; wa, wb from H).
struct NF2FFTD {
  Index dist[3] = {1, 1, 1};
  Real theta = 0, phi = 0;
  std::string out_dir;
  Task surf[6];  // node tasks (nffft.cpp:152-223), invalid ones empty
  Real dmin = 0, dmax = 0;  // distanceRange over the global surfaces (:410-480)
  Index num = 0;            // potential length (:53-66)
  std::vector<Real> prev[6][4];  // per surface: ea, eb, ha, hb previous values, loop order
  std::vector<Real> pot[6][4];   // per surface: ua, ub, wa, wb
  void init(Sim& s);
  void initTimeDependentVariable(Sim& s);
  void update(Sim& s, Index t);  // nfffft_time_domain.cpp:68-82, nffft_td_data.h:159-231
  // processFarField (nfffft_time_domain.cpp:84-160): w_theta, w_phi, u_theta, u_phi
  void farField(Sim& s, std::vector<Real> out[4]) const;
};

// The DFT objective of the adjoint form (no reference counterpart): over the
// grid points [o, o+n) of one E component, D_p = sum_t dt exp(-j 2 pi f
// (t+1) dt) F_p^{t+1}, accumulated after each step's corrections;
// J = sum_p |D_p|^2.
struct DftObjective {
  int field = 2;
  Index o[3] = {0, 0, 0}, n[3] = {0, 0, 0};
  Real freq = 0;
  std::vector<Cplx> dft;
  Real value() const {
    Real v = 0;
    for (auto& d : dft) v += std::norm(d);
    return v;
  }
};

struct Sim {
  Real dx, dy, dz, cfl;
  // design cells of the adjoint form: relative permittivity per cell (k
  // fastest), replacing the objects' permittivity before the averaging
  bool has_design = false;
  Index design_o[3] = {0, 0, 0}, design_n[3] = {0, 0, 0};
  std::vector<Real> design_eps;
  std::unique_ptr<DftObjective> objective;
  int tcfg[3] = {1, 1, 1};
  std::vector<std::unique_ptr<Object>> objects;
  std::vector<std::unique_ptr<PML>> pmls;
  std::vector<std::unique_ptr<TFSF3D>> tfsfs;
  std::vector<std::unique_ptr<Monitor>> monitors;
  std::vector<std::unique_ptr<NF2FF>> nf2ffs;
  std::vector<std::unique_ptr<Network>> networks;
  std::vector<std::unique_ptr<NF2FFTD>> nf2ff_tds;
  bool write_files = true;

  GridSpace g;
  EMF emf;
  Coeff c;
  MaterialParam mp;
  std::vector<Material> material_array;
  std::vector<Index> material_grid;  // per-cell material index (grid_space.cpp:357-368)
  std::unique_ptr<ADEStorage> ade;   // Simulation::_ade_method_storage
  Real dt = 0;
  Index start_step = 0, size = 0, current = 0;
  double loop_seconds = 0;
  // seconds from the loop start to the end of each step (timing only)
  std::vector<double> step_end_seconds;

  void generateGridSpace();             // simulation.cpp:496-512
  void generateEMF();                   // simulation.cpp:514-528
  void generateMaterialSpace();         // simulation.cpp:593-636
  void averageMaterial();               // calculation_param.cpp:43-168
  void calculateCoefficient();          // calculation_param.cpp:170-264
  void buildDispersiveSpace();          // simulation.cpp:745-829, object.cpp:52-98
  void init();                          // simulation.cpp:319-371
  void initSteps(Index n);              // simulation.cpp:373-389
  void run(Index n);                    // simulation.cpp:284-289
  void loop();                          // simulation.cpp:291-317
  Arr1 eTime() const;                   // time_param.cpp:63-66
  Arr1 hTime() const;                   // time_param.cpp:68-71
};

void writeNpy(const std::string& path, const std::vector<Index>& shape,
              const Real* data);
void writeNpyComplex(const std::string& path, const std::vector<Index>& shape,
                     const std::complex<Real>* data);

}  // namespace refcpu
