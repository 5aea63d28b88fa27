// State of one NFFFTFrequencyDomain on the host: the six surfaces, the DFT
// phase tables handed to the device, a copy of the grid it needs for the far
// field, and the surface currents copied back after the run.
#pragma once

#include <complex>
#include <cstddef>
#include <functional>
#include <vector>

#include "../../../include/xfdtd_core_amd.hpp"

namespace xfdtd {

struct NFFFTFrequencyDomain::Data {
  bool used[6] = {false, false, false, false, false, false};
  std::size_t box_lo[6][3] = {}, box_hi[6][3] = {};
  std::vector<Real> e_node[3], h_node[3], e_size[3];
  std::vector<Real> freqs;
  std::vector<std::complex<Real>> te, th;  // nf x nt
  std::size_t nt = 0;
  std::vector<std::vector<std::complex<Real>>> cur;  // index (f * 6 + d) * 4 + which
  // Set while a device context holds this NF2FF's currents: computes, for
  // frequency f and the directions (theta[q], phi[q]), the potentials
  // A_theta, A_phi, F_theta, F_phi (out: 4 x na, in that order) on the GPU.
  std::function<void(std::size_t f, Real k, const std::vector<Real>& theta, const std::vector<Real>& phi,
                     const Vector& origin, std::vector<std::complex<Real>>& out)>
      device;
  void setup(const std::size_t dist[3], const std::size_t n[3], const std::vector<Real> node_e[3],
             const std::vector<Real> node_h[3], const std::vector<Real> size_e[3], const std::vector<Real>& freq,
             Real dt, std::size_t steps);
  std::size_t cells(int d) const;
};

// State of one NFFFTTimeDomain on the host: the surfaces, per cell of the
// used surfaces (loop order) the time-step delays plus offsets and the face
// areas handed to the device, and the potentials copied back after the run.
struct NFFFTTimeDomain::Data {
  bool used[6] = {false, false, false, false, false, false};
  std::size_t box_lo[6][3] = {}, box_hi[6][3] = {};
  std::vector<Real> s_e, s_h, ds;
  std::size_t num = 0;
  Real dt = 0;
  std::vector<Real> pot[6][4];  // ua ub wa wb per surface
  void setup(const std::size_t dist[3], const std::size_t n[3], const std::vector<Real> node_e[3],
             const std::vector<Real> node_h[3], const std::vector<Real> size_e[3], Real theta, Real phi, Real dt,
             std::size_t steps);
};

}  // namespace xfdtd
