// CPML slabs applied by the sweeps in the same pass as the Yee update: each
// slab's parameters travel as kernel arguments (scalar loads), and a point's
// membership is the AND of per-axis masks, uniform over a plane (i), a row
// (j) or fixed over a lane's march (k), so points outside every slab cost
// one test.  Used by the two sweeps (every slab, in boundary order) and by
// the single-pass step.
#pragma once

#include "layout.hpp"

namespace xfd {

__device__ __forceinline__ int pick3(int c, int x, int y, int z) { return c == 0 ? x : (c == 1 ? y : z); }

constexpr int kMaxStepSlabs = 8;
constexpr int kMaxStepChunk = 1024;

// One CPML slab as the sweeps see it: its box (grid indices), the
// components' extents along (a, b), the per-layer recursive-convolution
// coefficients and c_psi, and psi to read (rd) and to write (wr): the same
// array, except for the single-pass step's H side (its ping-pong).
template <class T>
struct StepSlab {
  int axis, start;
  int lo[3], ext[3];
  int extA[2], extB[2];
  const T* cpsiA;
  const T* cpsiB;
  const T* rdA;
  const T* rdB;
  T* wrA;
  T* wrB;
  const T* ca;
  const T* cb;
};

template <class T>
struct StepSlabs {
  int n = 0;
  StepSlab<T> s[kMaxStepSlabs];
};

// Bit n set: coordinate p lies in slab n's box along axis `ax`.  A point is
// in slab n when its three coordinates are, so the membership mask of a
// point is the AND of three per-axis masks; the i mask is uniform over a
// plane, the j mask over a row, and the k mask fixed over a lane's march, so
// each is taken once and points outside every slab cost one test.
template <class T>
__device__ __forceinline__ uint32_t axisMask(const StepSlabs<T>& L, int ax, int p) {
  uint32_t m = 0;
  for (int n = 0; n < L.n; ++n)
    if (static_cast<unsigned>(p - L.s[n].lo[ax]) < static_cast<unsigned>(L.s[n].ext[ax])) m |= 1u << n;
  return m;
}

// The union of a per-lane mask over the wavefront.
__device__ __forceinline__ uint32_t waveUnion(uint32_t m) {
  uint32_t u = 0;
  for (int n = 0; n < kMaxStepSlabs; ++n)
    if (__any((m >> n) & 1u)) u |= 1u << n;
  return u;
}

// The CPML corrections of point (i, j, k) for every slab of L containing
// it (bits of m), in boundary order, as cpmlFused: component A = (c+1)%3
// takes the dual difference dA[c], component B = (c+2)%3 takes dB[c];
//   psi = b psi + a d,  v += c_psi psi.
// Write = false (recomputed neighbour points) leaves psi as it is.
// mu is m's union over the wavefront (the slab loop stays uniform, so the
// slab's parameters are scalar loads); lanes outside slab n skip it.
template <class T, bool Write>
__device__ __forceinline__ void stepSlabs(const StepSlabs<T>& L, uint32_t mu, uint32_t m, int i, int j, int k,
                                          T& v0, T& v1, T& v2, const T dA0, const T dA1, const T dA2, const T dB0,
                                          const T dB1, const T dB2) {
  mu = __builtin_amdgcn_readfirstlane(mu);
  while (mu) {
    const int n = __builtin_ctz(mu);
    mu &= mu - 1;
    if (!((m >> n) & 1u)) continue;
    const StepSlab<T>& S = L.s[n];
    const int u = i - S.lo[0], v = j - S.lo[1], w = k - S.lo[2];
    const int c = S.axis, a = c == 2 ? 0 : c + 1, b = c == 0 ? 2 : c - 1;
    const int64_t t = (static_cast<int64_t>(u) * S.ext[1] + v) * S.ext[2] + w;
    const int layer = pick3(c, i, j, k) - S.start;
    const T ca = S.ca[layer], cb = S.cb[layer];
    const int pa = pick3(a, i, j, k), pb = pick3(b, i, j, k);
    if (pa < S.extA[0] && pb < S.extA[1]) {
      const T d = c == 0 ? dA0 : (c == 1 ? dA1 : dA2);
      const T psi = cb * S.rdA[t] + ca * d;
      if (Write) S.wrA[t] = psi;
      const T add = S.cpsiA[t] * psi;
      const T n0 = a == 0 ? v0 + add : v0, n1 = a == 1 ? v1 + add : v1, n2 = a == 2 ? v2 + add : v2;
      v0 = n0;
      v1 = n1;
      v2 = n2;
    }
    if (pa < S.extB[0] && pb < S.extB[1]) {
      const T d = c == 0 ? dB0 : (c == 1 ? dB1 : dB2);
      const T psi = cb * S.rdB[t] + ca * d;
      if (Write) S.wrB[t] = psi;
      const T add = S.cpsiB[t] * psi;
      const T n0 = b == 0 ? v0 + add : v0, n1 = b == 1 ? v1 + add : v1, n2 = b == 2 ? v2 + add : v2;
      v0 = n0;
      v1 = n1;
      v2 = n2;
    }
  }
}

}  // namespace xfd
