// Host side of the frequency-domain NF2FF: surfaces, DFT phase tables, and
// the far-field potentials and radiated power formed from the surface
// currents that the device accumulated (FDPlaneData, nffft_fd_data*.cpp).
// Built like the rest of the host precompute (FP64, no FMA contraction).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstdio>
#include <filesystem>
#include <sstream>
#include <iomanip>
#include <limits>

#include "../../../include/xfdtd_core_amd.hpp"
#include "nf2ff_data.hpp"
#include "npy.hpp"

namespace xfdtd {

namespace {
using Cplx = std::complex<Real>;
constexpr Real kPiN = 3.141592653589793238462643383279502884;
const Real kC0N = 1 / std::sqrt(8.854187817e-12 * (4 * kPiN * 1e-7));

// GCC's complex product for finite operands: (ac - bd, ad + bc)
Cplx cmul(const Cplx& a, const Cplx& b) {
  return {a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real()};
}
}  // namespace

NFFFTFrequencyDomain::NFFFTFrequencyDomain(std::size_t dx, std::size_t dy, std::size_t dz,
                                           std::vector<Real> frequencies, std::string output_dir)
    : _distance{dx, dy, dz},
      _frequencies(std::move(frequencies)),
      _output_dir(std::move(output_dir)),
      _data(std::make_unique<Data>()) {
  // nffft.cpp:19-33
  if (dx == 0) throw XFDTDException("distance_x == 0");
  if (dy == 0) throw XFDTDException("distance_y == 0");
  if (dz == 0) throw XFDTDException("distance_z == 0");
}
NFFFTFrequencyDomain::~NFFFTFrequencyDomain() = default;

void NFFFTFrequencyDomain::Data::setup(const std::size_t dist[3], const std::size_t n[3],
                                       const std::vector<Real> node_e[3], const std::vector<Real> node_h[3],
                                       const std::vector<Real> size_e[3], const std::vector<Real>& freq, Real dt,
                                       std::size_t steps) {
  // NFFFT::initGlobal / makeNodeAxisTask (nffft.cpp:73-223) for one process
  std::size_t lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = dist[a];
    hi[a] = n[a] > dist[a] ? n[a] - dist[a] : 0;
    if (hi[a] <= lo[a]) throw XFDTDException("NFFFT: Global box is not legal");
  }
  bool any = false;
  for (int d = 0; d < 6; ++d) {
    const int c = d / 2;
    const bool neg = d % 2 == 0;
    const std::size_t cs = neg ? lo[c] : hi[c];
    used[d] = !((neg && cs <= 1) || (!neg && cs >= n[c]));
    for (int a = 0; a < 3; ++a) {  // an unused surface is the empty task, as IndexTask{}
      box_lo[d][a] = used[d] ? (a == c ? cs : lo[a]) : 0;
      box_hi[d][a] = used[d] ? (a == c ? cs + 1 : hi[a]) : 0;
    }
    any = any || used[d];
  }
  if (!any) throw XFDTDException("NFFFT: Invalid task");
  for (int a = 0; a < 3; ++a) {
    e_node[a] = node_e[a];
    h_node[a] = node_h[a];
    e_size[a] = size_e[a];
  }
  freqs = freq;
  // FDPlaneData::initDFT (nffft_fd_data.cpp:65-76)
  const Cplx II{0, 1};
  te.assign(freqs.size() * steps, Cplx{});
  th.assign(freqs.size() * steps, Cplx{});
  for (std::size_t f = 0; f < freqs.size(); ++f)
    for (std::size_t t = 0; t < steps; ++t) {
      te[f * steps + t] = dt * std::exp(-II * static_cast<Real>(2.0) * kPiN * (freqs[f] * (t + 1) * dt));
      th[f * steps + t] =
          dt * std::exp(-II * static_cast<Real>(2.0) * kPiN * (freqs[f] * (t + static_cast<Real>(0.5)) * dt));
    }
  nt = steps;
  cur.assign(freqs.size() * 6 * 4, {});
  for (std::size_t f = 0; f < freqs.size(); ++f)
    for (int d = 0; d < 6; ++d)
      for (int w = 0; w < 4; ++w) cur[(f * 6 + d) * 4 + w].assign(cells(d), Cplx{});
}

std::size_t NFFFTFrequencyDomain::Data::cells(int d) const {
  if (!used[d]) return 0;
  return (box_hi[d][0] - box_lo[d][0]) * (box_hi[d][1] - box_lo[d][1]) * (box_hi[d][2] - box_lo[d][2]);
}

std::array<std::size_t, 6> NFFFTFrequencyDomain::surface(int d) const {
  const auto& D = *_data;
  return {D.box_lo[d][0], D.box_lo[d][1], D.box_lo[d][2], D.box_hi[d][0], D.box_hi[d][1], D.box_hi[d][2]};
}

const std::vector<std::complex<Real>>& NFFFTFrequencyDomain::current(std::size_t f, int d, int which) const {
  if (f >= _frequencies.size() || d < 0 || d > 5 || which < 0 || which > 3)
    throw XFDTDException("NFFFT: bad current index");
  return _data->cur.at((f * 6 + d) * 4 + which);
}

bool NFFFTFrequencyDomain::deviceFarField() const {
  const char* env = std::getenv("XFD_FARFIELD");
  return _data->device && !(env && std::string(env) == "host");
}

std::vector<std::complex<Real>> NFFFTFrequencyDomain::potential(std::size_t fi, Potential pot, bool theta_comp,
                                                                 const std::vector<Real>& theta,
                                                                 const std::vector<Real>& phi,
                                                                 const Vector& origin) const {
  // FDPlaneData::getPotential / calculatePotential (nffft_fd_data_process.cpp:44-160)
  if (theta.empty() || phi.empty()) throw XFDTDException("Invalid theta or phi size");
  if (theta.size() != 1 && phi.size() != 1) throw XFDTDException("There must be only one theta or phi");
  const auto& D = *_data;
  const std::size_t na = theta.size() * phi.size();
  auto th = [&](std::size_t n) { return theta[theta.size() == 1 ? 0 : n]; };
  auto ph = [&](std::size_t n) { return phi[phi.size() == 1 ? 0 : n]; };
  // cCSToSCSTransformMatrix (transform.h:24-58)
  auto tmat = [&](int axis, std::size_t n) -> Real {
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
  const Real freq = _frequencies.at(fi);
  const Real k = 2.0 * kPiN * freq / kC0N;
  if (deviceFarField()) {  // the same sums on the GPU (xfd_nf2ff_far_field)
    std::vector<Real> t(na), p(na);
    for (std::size_t n = 0; n < na; ++n) {
      t[n] = th(n);
      p[n] = ph(n);
    }
    std::vector<Cplx> all;
    D.device(fi, k, t, p, origin, all);
    const std::size_t w = (pot == Potential::A ? 0 : 2) + (theta_comp ? 0 : 1);
    return std::vector<Cplx>(all.begin() + w * na, all.begin() + (w + 1) * na);
  }
  const Cplx ik = Cplx{0, 1} * k;
  std::vector<Cplx> total(na, Cplx{});
  std::vector<Real> st(na), ct(na), stsp(na), stcp(na);
  for (std::size_t n = 0; n < na; ++n) {
    ct[n] = std::cos(th(n));
    st[n] = std::sin(th(n));
    stsp[n] = st[n] * std::sin(ph(n));
    stcp[n] = st[n] * std::cos(ph(n));
  }
  for (int d = 0; d < 6; ++d) {
    std::vector<Cplx> data(na, Cplx{});
    if (D.used[d]) {
      const int c = d / 2, A = (c + 1) % 3, B = (c + 2) % 3;
      const int wa = pot == Potential::A ? 0 : 2;
      const auto& va = D.cur[(fi * 6 + d) * 4 + wa];
      const auto& vb = D.cur[(fi * 6 + d) * 4 + wa + 1];
      std::vector<Cplx> ta(na), tb(na);
      for (std::size_t n = 0; n < na; ++n) {
        ta[n] = tmat(A, n);
        tb[n] = tmat(B, n);
      }
      std::size_t q = 0;
      for (std::size_t i = D.box_lo[d][0]; i < D.box_hi[d][0]; ++i)
        for (std::size_t j = D.box_lo[d][1]; j < D.box_hi[d][1]; ++j)
          for (std::size_t kk = D.box_lo[d][2]; kk < D.box_hi[d][2]; ++kk, ++q) {
            // rVector / ds (nffft_fd_data_process.cpp:196-225)
            Real r0, r1, r2, ds;
            if (c == 0) {
              r0 = D.e_node[0][i];
              r1 = D.h_node[1][j];
              r2 = D.h_node[2][kk];
              ds = D.e_size[1][j] * D.e_size[2][kk];
            } else if (c == 1) {
              r0 = D.h_node[0][i];
              r1 = D.e_node[1][j];
              r2 = D.h_node[2][kk];
              ds = D.e_size[2][kk] * D.e_size[0][i];
            } else {
              r0 = D.h_node[0][i];
              r1 = D.h_node[1][j];
              r2 = D.e_node[2][kk];
              ds = D.e_size[0][i] * D.e_size[1][j];
            }
            const Real rx = r0 - origin[0], ry = r1 - origin[1], rz = r2 - origin[2];
            for (std::size_t n = 0; n < na; ++n) {
              const Real v = rx * stcp[n] + ry * stsp[n] + rz * ct[n];
              const Cplx phase = std::exp(ik * v);
              const Cplx sum = cmul(va[q], ta[n]) + cmul(vb[q], tb[n]);
              data[n] += cmul(sum, phase) * ds;
            }
          }
    }
    for (std::size_t n = 0; n < na; ++n) total[n] += data[n];
  }
  return total;
}

std::array<std::vector<std::complex<Real>>, 4> NFFFTFrequencyDomain::potentials(std::size_t f,
                                                                                 const std::vector<Real>& theta,
                                                                                 const std::vector<Real>& phi,
                                                                                 const Vector& origin) const {
  std::array<std::vector<Cplx>, 4> r;
  if (deviceFarField()) {
    const std::size_t na = theta.size() * phi.size();
    if (theta.empty() || phi.empty()) throw XFDTDException("Invalid theta or phi size");
    if (theta.size() != 1 && phi.size() != 1) throw XFDTDException("There must be only one theta or phi");
    std::vector<Real> t(na), p(na);
    for (std::size_t n = 0; n < na; ++n) {
      t[n] = theta[theta.size() == 1 ? 0 : n];
      p[n] = phi[phi.size() == 1 ? 0 : n];
    }
    std::vector<Cplx> all;
    _data->device(f, 2.0 * kPiN * _frequencies.at(f) / kC0N, t, p, origin, all);
    for (int w = 0; w < 4; ++w) r[w].assign(all.begin() + w * na, all.begin() + (w + 1) * na);
    return r;
  }
  r[0] = potential(f, Potential::A, true, theta, phi, origin);
  r[1] = potential(f, Potential::A, false, theta, phi, origin);
  r[2] = potential(f, Potential::F, true, theta, phi, origin);
  r[3] = potential(f, Potential::F, false, theta, phi, origin);
  return r;
}

Real NFFFTFrequencyDomain::power(std::size_t fi) const {
  // FDPlaneData::power / calculatePower (nffft_fd_data_process.cpp:28-42, 162-194)
  const auto& D = *_data;
  Cplx acc{0.0};
  for (int d = 0; d < 6; ++d) {
    Cplx pw{0.0};
    if (D.used[d]) {
      const int c = d / 2;
      const auto& ja = D.cur[(fi * 6 + d) * 4 + 0];
      const auto& jb = D.cur[(fi * 6 + d) * 4 + 1];
      const auto& ma = D.cur[(fi * 6 + d) * 4 + 2];
      const auto& mb = D.cur[(fi * 6 + d) * 4 + 3];
      std::size_t q = 0;
      for (std::size_t i = D.box_lo[d][0]; i < D.box_hi[d][0]; ++i)
        for (std::size_t j = D.box_lo[d][1]; j < D.box_hi[d][1]; ++j)
          for (std::size_t kk = D.box_lo[d][2]; kk < D.box_hi[d][2]; ++kk, ++q) {
            Real ds;
            if (c == 0) ds = D.e_size[1][j] * D.e_size[2][kk];
            else if (c == 1) ds = D.e_size[2][kk] * D.e_size[0][i];
            else ds = D.e_size[0][i] * D.e_size[1][j];
            pw += ds * (cmul(std::conj(ja[q]), mb[q]) - cmul(std::conj(jb[q]), ma[q]));
          }
      const Real negative = d % 2 == 0 ? -1.0 : 1.0;
      pw = negative * pw;
    }
    acc = acc + pw;
  }
  return 0.5 * std::real(acc);
}

void NFFFTFrequencyDomain::processFarField(const std::vector<Real>& theta, Real phi, const std::string& sub_dir,
                                           const Vector& origin) const {
  const std::vector<Real> ph{phi};
  const auto dir = std::filesystem::path(_output_dir) / sub_dir;
  for (std::size_t f = 0; f < _frequencies.size(); ++f) {
    std::stringstream ss;
    ss << std::fixed << std::setprecision(2) << _frequencies[f] / 1e9 << "GHz";
    std::filesystem::create_directories(dir);
    const char* names[4] = {"_a_theta.npy", "_f_phi.npy", "_a_phi.npy", "_f_theta.npy"};
    const int order[4] = {0, 3, 1, 2};  // a_theta, f_phi, a_phi, f_theta in potentials() order
    const auto all = potentials(f, theta, ph, origin);
    for (int q = 0; q < 4; ++q) {
      const auto& v = all[order[q]];
      npy::writeComplex((dir / (ss.str() + names[q])).string(), {v.size()}, reinterpret_cast<const double*>(v.data()));
    }
  }
}

void NFFFTFrequencyDomain::processFarField(Real theta, const std::vector<Real>& phi, const std::string& sub_dir,
                                           const Vector& origin) const {
  const std::vector<Real> th{theta};
  const auto dir = std::filesystem::path(_output_dir) / sub_dir;
  for (std::size_t f = 0; f < _frequencies.size(); ++f) {
    std::stringstream ss;
    ss << std::fixed << std::setprecision(2) << _frequencies[f] / 1e9 << "GHz";
    std::filesystem::create_directories(dir);
    const char* names[4] = {"_a_theta.npy", "_f_phi.npy", "_a_phi.npy", "_f_theta.npy"};
    const int order[4] = {0, 3, 1, 2};  // a_theta, f_phi, a_phi, f_theta in potentials() order
    const auto all = potentials(f, th, phi, origin);
    for (int q = 0; q < 4; ++q) {
      const auto& v = all[order[q]];
      npy::writeComplex((dir / (ss.str() + names[q])).string(), {v.size()}, reinterpret_cast<const double*>(v.data()));
    }
  }
}

void NFFFTFrequencyDomain::outputRadiationPower() const {
  // nffft_frequency_domain.cpp:65-92: stack(frequencies, power) -> power.npy
  const std::size_t nf = _frequencies.size();
  std::vector<Real> out(2 * nf);
  for (std::size_t f = 0; f < nf; ++f) {
    out[f] = _frequencies[f];
    out[nf + f] = power(f);
  }
  std::filesystem::create_directories(_output_dir.empty() ? "." : _output_dir);
  npy::write((std::filesystem::path(_output_dir) / "power.npy").string(), {2, nf}, out.data());
}

// ------------------------------------------------------ time domain ------
NFFFTTimeDomain::NFFFTTimeDomain(std::size_t dx, std::size_t dy, std::size_t dz, Real theta, Real phi)
    : NFFFTTimeDomain(dx, dy, dz, "", theta, phi) {}

NFFFTTimeDomain::NFFFTTimeDomain(std::size_t dx, std::size_t dy, std::size_t dz, std::string output_dir, Real theta,
                                 Real phi)
    : _distance{dx, dy, dz}, _output_dir(std::move(output_dir)), _theta(theta), _phi(phi),
      _data(std::make_unique<Data>()) {
  if (dx == 0) throw XFDTDException("distance_x == 0");  // nffft.cpp:19-33
  if (dy == 0) throw XFDTDException("distance_y == 0");
  if (dz == 0) throw XFDTDException("distance_z == 0");
}
NFFFTTimeDomain::~NFFFTTimeDomain() = default;

// NFFFTTimeDomain::init / initTimeDependentVariable (nfffft_time_domain.cpp:
// 35-66): the node surfaces (as the frequency-domain NF2FF), the distance
// range of the six global surfaces along the observation direction
// (:410-480), and per node-surface cell the time-step delay
// (dmax - r . u) / (c dt) plus the E (0.5) or H (0) offset and the face
// area (nffft_td_data.h:22-47, 178-199), in the reference's expressions.
void NFFFTTimeDomain::Data::setup(const std::size_t dist[3], const std::size_t n[3],
                                  const std::vector<Real> node_e[3], const std::vector<Real> node_h[3],
                                  const std::vector<Real> size_e[3], Real theta, Real phi, Real dt_,
                                  std::size_t steps) {
  std::size_t lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = dist[a];
    hi[a] = n[a] > dist[a] ? n[a] - dist[a] : 0;
    if (hi[a] <= lo[a]) throw XFDTDException("NFFFT: Global box is not legal");
  }
  std::size_t glo[6][3], ghi[6][3];
  bool any = false;
  for (int d = 0; d < 6; ++d) {
    const int c = d / 2;
    const bool neg = d % 2 == 0;
    const std::size_t cs = neg ? lo[c] : hi[c];
    used[d] = !((neg && cs <= 1) || (!neg && cs >= n[c]));
    for (int a = 0; a < 3; ++a) {
      glo[d][a] = a == c ? cs : lo[a];
      ghi[d][a] = a == c ? cs + 1 : hi[a];
      box_lo[d][a] = used[d] ? glo[d][a] : 0;
      box_hi[d][a] = used[d] ? ghi[d][a] : 0;
    }
    any = any || used[d];
  }
  if (!any) throw XFDTDException("NFFFT: Invalid task");
  // observationDirection (:520-522)
  const Real ux = std::sin(theta) * std::cos(phi), uy = std::sin(theta) * std::sin(phi), uz = std::cos(theta);
  auto r = [&](int c, std::size_t i, std::size_t j, std::size_t k, Real& x, Real& y, Real& z) {
    x = c == 0 ? node_e[0][i] : node_h[0][i];
    y = c == 1 ? node_e[1][j] : node_h[1][j];
    z = c == 2 ? node_e[2][k] : node_h[2][k];
  };
  Real dmin = std::numeric_limits<Real>::max(), dmax = std::numeric_limits<Real>::lowest();
  for (int d = 0; d < 6; ++d)
    for (std::size_t k = glo[d][2]; k < ghi[d][2]; ++k)
      for (std::size_t j = glo[d][1]; j < ghi[d][1]; ++j)
        for (std::size_t i = glo[d][0]; i < ghi[d][0]; ++i) {
          Real x, y, z;
          r(d / 2, i, j, k, x, y, z);
          const Real dis = x * ux + y * uy + z * uz;
          dmin = std::min(dmin, dis);
          dmax = std::max(dmax, dis);
        }
  dt = dt_;
  const Real aux = (dmax - dmin) / (kC0N * dt);
  num = static_cast<std::size_t>(std::ceil(aux)) + steps + 4;
  s_e.clear();
  s_h.clear();
  ds.clear();
  for (int d = 0; d < 6; ++d) {
    if (!used[d]) continue;
    const int c = d / 2;
    for (std::size_t i = box_lo[d][0]; i < box_hi[d][0]; ++i)
      for (std::size_t j = box_lo[d][1]; j < box_hi[d][1]; ++j)
        for (std::size_t k = box_lo[d][2]; k < box_hi[d][2]; ++k) {
          Real x, y, z;
          r(c, i, j, k, x, y, z);
          const Real time_delay = dmax - (x * ux + y * uy + z * uz);
          const Real time_step_delay = time_delay / (kC0N * dt);
          s_e.push_back(time_step_delay + 0.5);
          s_h.push_back(time_step_delay + 0.0);
          ds.push_back(c == 0 ? size_e[1][j] * size_e[2][k]
                              : (c == 1 ? size_e[2][k] * size_e[0][i] : size_e[0][i] * size_e[1][j]));
        }
  }
  for (auto& p : pot)
    for (auto& v : p) v.clear();
}

std::size_t NFFFTTimeDomain::length() const { return _data->num; }

const std::vector<Real>& NFFFTTimeDomain::potential(int d, int which) const {
  if (d < 0 || d >= 6 || which < 0 || which >= 4) throw XFDTDException("NFFFTTimeDomain: bad potential");
  return _data->pot[d][which];
}

// processFarField (nfffft_time_domain.cpp:84-160) with wa() .. ub() =
// potential / (c dt 4 pi) (nffft_td_data.h:239-261) and wx .. uz the sums of
// the surfaces' potentials along each axis (nfffft_time_domain.cpp:484-512).
std::array<std::vector<Real>, 4> NFFFTTimeDomain::farField() const {
  const Data& D = *_data;
  const std::size_t num = D.num;
  for (auto& p : D.pot)
    for (auto& v : p)
      if (v.size() != num) throw XFDTDException("NFFFTTimeDomain: no potentials (run the simulation first)");
  auto sc = [&](int d, int w) {
    std::vector<Real> v(num);
    for (std::size_t q = 0; q < num; ++q) v[q] = D.pot[d][w][q] / (kC0N * D.dt * 4 * kPiN);
    return v;
  };
  auto sum4 = [&](std::vector<Real> a, const std::vector<Real>& b, const std::vector<Real>& c,
                  const std::vector<Real>& d) {
    for (std::size_t q = 0; q < num; ++q) a[q] = a[q] + b[q] + c[q] + d[q];
    return a;
  };
  enum { XN, XP, YN, YP, ZN, ZP };
  const auto wx = sum4(sc(ZN, 2), sc(ZP, 2), sc(YN, 3), sc(YP, 3));
  const auto wy = sum4(sc(XN, 2), sc(XP, 2), sc(ZN, 3), sc(ZP, 3));
  const auto wz = sum4(sc(XN, 3), sc(XP, 3), sc(YN, 2), sc(YP, 2));
  const auto ux = sum4(sc(ZN, 0), sc(ZP, 0), sc(YN, 1), sc(YP, 1));
  const auto uy = sum4(sc(XN, 0), sc(XP, 0), sc(ZN, 1), sc(ZP, 1));
  const auto uz = sum4(sc(XN, 1), sc(XP, 1), sc(YN, 0), sc(YP, 0));
  const Real ct = std::cos(_theta), st = std::sin(_theta), cp = std::cos(_phi), sp = std::sin(_phi);
  std::array<std::vector<Real>, 4> out;
  for (auto& o : out) o.assign(num, 0.0);
  for (std::size_t q = 0; q < num; ++q) {
    out[0][q] = wx[q] * ct * cp + wy[q] * ct * sp - wz[q] * st;
    out[1][q] = -wx[q] * sp + wy[q] * cp;
    out[2][q] = ux[q] * ct * cp + uy[q] * ct * sp - uz[q] * st;
    out[3][q] = -ux[q] * sp + uy[q] * cp;
  }
  return out;
}

void NFFFTTimeDomain::processFarField() const {
  const auto out = farField();
  const char* names[4] = {"w_theta.npy", "w_phi.npy", "u_theta.npy", "u_phi.npy"};
  std::filesystem::create_directories(_output_dir.empty() ? "." : _output_dir);
  for (int w = 0; w < 4; ++w)
    npy::write((std::filesystem::path(_output_dir) / names[w]).string(), {out[w].size()}, out[w].data());
}

}  // namespace xfdtd
