// One pass per time step: the H half-step, its CPML corrections, the E
// half-step and its CPML corrections in a single launch.
//
// The reference runs updateH, the H correctors, updateE and the E correctors
// as separate sweeps over the grid (src/domain/domain.cpp:82-135), so every
// field is read twice and written once per step.  Here each lane marches
// along i over a chunk of planes and computes E(i) right after H(i), with
// the H values of its j-1 / k-1 neighbours recomputed or shuffled in and
// H(i-1) carried in registers (k_step).
// The recomputations read the neighbours' old state, which
// their owners overwrite in the same launch; so E, H and the H-side CPML psi
// are double-buffered: the step reads set A and writes set B, and the
// buffers swap roles every step.  The E-side psi is read and written by its
// owner only and stays in place.
//
// The arithmetic of every point is the separate sweeps' and k_cpml's, in
// the same order (Yee terms left to right, then every slab containing the
// point in boundary order), so FP64 results are bitwise those of the
// reference's order.  Corrections that are not point-local -- TFSF H
// records, which would have to reach the recomputed copies -- keep
// the separate sweeps; E records (lumped sources, TFSF E) follow the pass as
// before, which is the reference's order wherever no record shares a point
// with a slab.
//
// Points on the far grid planes i = nx, j = ny, k = nz are not visited: no
// Yee range reaches them and the only corrections that do (the CPML slabs'
// transverse extension) couple them to each other alone, starting from and
// keeping zero, so both buffer sets hold the same values there throughout.
#pragma once

#include "kernels.hpp"

namespace xfd {

// H at (i, j, k) (BasicUpdator::updateH, src/updator/basic_updator.cpp:
// 23-41, then the H-side slabs of mask m) from set A; ex0, ey_i, ez_i are E
// at the point, ey_n / ez_n return E at i + 1 (the next plane's ey_i,
// ez_i).  NeedY = false skips Hy and the loads only it uses (vy returns 0).
template <class T, class Coef, bool Write, bool NeedY = true>
__device__ __forceinline__ void stepH(const Grid& g, const FieldPtrs<T>& A, const Coef& C, const StepSlabs<T>& S,
                                      uint32_t mu, uint32_t m, int i, int j, int k, int64_t idx, T ex0, T ey_i,
                                      T ez_i, T& ey_n, T& ez_n, T& vx, T& vy, T& vz) {
  const int64_t pz = g.pz, pl = g.plane;
  const T ex_j = A.f[0][idx + pz];
  const T ex_k = NeedY ? A.f[0][idx + 1] : T(0);
  const T ey_k = A.f[1][idx + 1];
  const T ez_j = A.f[2][idx + pz];
  ey_n = A.f[1][idx + pl];
  ez_n = NeedY ? A.f[2][idx + pl] : T(0);
  T s, a, b;
  C.get(3, idx, s, a, b);
  vx = s * A.f[3][idx] + a * (ey_k - ey_i) + b * (ez_j - ez_i);
  vy = T(0);
  if (NeedY) {
    C.get(4, idx, s, a, b);
    vy = s * A.f[4][idx] + a * (ez_n - ez_i) + b * (ex_k - ex0);
  }
  C.get(5, idx, s, a, b);
  vz = s * A.f[5][idx] + a * (ex_j - ex0) + b * (ey_n - ey_i);
  // A: Hy(x) <- Ez, Hz(y) <- Ex, Hx(z) <- Ey;  B: Hz(x) <- Ey, Hx(y) <- Ez, Hy(z) <- Ex
  if (mu)
    stepSlabs<T, Write>(S, mu, m, i, j, k, vx, vy, vz, ez_n - ez_i, ex_j - ex0, ey_k - ey_i, ey_n - ey_i,
                        ez_j - ez_i, ex_k - ex0);
}

// The step pass over the sweep tiles (64 k-lanes by kRowsJ rows, marching a
// chunk of i-planes; work-group order as the Yee sweeps').  Each lane
// computes, per plane, H at its point and the H values its E update needs
// from the neighbours that the lanes and rows around it own:
//   * H(i-1): the previous plane of the march (registers); the chunk's first
//     plane recomputes H(i0-1) without storing it;
//   * Hz, Hx at j-1: recomputed by every lane from set A (the row is another
//     wavefront's, its loads hit the caches);
//   * Hy, Hx at k-1: the neighbouring lane's values (a lane shuffle), lane 0
//     recomputing them at k0-1.
// No LDS and no barrier: every wavefront runs on its own.
template <class T, class Coef>
__global__ __launch_bounds__(kTileThreads) void k_step(Grid g, FieldPtrs<T> A, FieldPtrs<T> B, Coef C,
                                                       StepSlabs<T> SH, StepSlabs<T> SE, SweepMap M, int ichunk) {
  int bx, by, bz;
  M.coords(bx, by, bz);
  const int lane = threadIdx.x;
  const int k = bx * kLanesK + lane;
  const int j = by * kRowsJ + threadIdx.y;
  const int i0 = bz * ichunk;
  if (j >= g.ny || i0 >= g.nx) return;  // whole rows leave together: the shuffles below stay uniform
  const bool act = k < g.nz;
  const int i1 = min(i0 + ichunk, g.nx);
  const int64_t pz = g.pz, pl = g.plane;
  int64_t idx = (static_cast<int64_t>(i0) * g.NY + j) * pz + k;
  // slab membership per axis (H and E slabs), for the point and for its
  // recomputed j-1 and k-1 neighbours
  // (i and j are uniform over a wavefront, k is not: its masks also get
  // their wavefront union)
  const uint32_t hj = axisMask(SH, 1, j), hjm = axisMask(SH, 1, j - 1), ej = axisMask(SE, 1, j);
  const uint32_t hk = axisMask(SH, 2, k), hkm = axisMask(SH, 2, k - 1), ek = axisMask(SE, 2, k);
  const uint32_t hku = waveUnion(act ? hk : 0u), hkmu = waveUnion(act ? hkm : 0u), eku = waveUnion(act ? ek : 0u);
  T hz_p = T(0), hy_p = T(0), ey_i = T(0), ez_i = T(0);
  if (act) {
    ey_i = A.f[1][idx];
    ez_i = A.f[2][idx];
    if (i0 > 0) {  // H(i0 - 1), recomputed, not stored
      const int64_t p = idx - pl;
      T en, zn, hx, hy, hz;
      const uint32_t hp = axisMask(SH, 0, i0 - 1) & hj;
      stepH<T, Coef, false>(g, A, C, SH, hp & hku, hp & hk, i0 - 1, j, k, p, A.f[0][p], A.f[1][p], A.f[2][p], en, zn,
                            hx, hy, hz);
      hz_p = hz;
      hy_p = hy;
    }
  }
  for (int i = i0; i < i1; ++i, idx += pl) {
    const uint32_t hi = axisMask(SH, 0, i);
    T hx = T(0), hy = T(0), hz = T(0), ex0 = T(0), ey_n = T(0), ez_n = T(0);
    if (act) {
      ex0 = A.f[0][idx];
      stepH<T, Coef, true>(g, A, C, SH, hi & hj & hku, hi & hj & hk, i, j, k, idx, ex0, ey_i, ez_i, ey_n, ez_n, hx, hy,
                           hz);
    }
    T hy_km = __shfl_up(hy, 1), hx_km = __shfl_up(hx, 1);
    if (!act) continue;
    if (lane == 0 && k >= 1) {  // H at k0 - 1
      const int64_t p = idx - 1;
      T en, zn, vx, vy, vz;
      stepH<T, Coef, false>(g, A, C, SH, hi & hj & hkmu, hi & hj & hkm, i, j, k - 1, p, A.f[0][p], A.f[1][p],
                            A.f[2][p], en, zn, vx, vy, vz);
      hy_km = vy;
      hx_km = vx;
    }
    // Hz and Hx at j - 1, computed unconditionally (at the point itself on
    // the j = 0 face, where the result is not used) so that their loads are
    // not held behind a branch
    T hz_jm = T(0), hx_jm = T(0);
    {
      const int64_t p = j >= 1 ? idx - pz : idx;
      const uint32_t mj = j >= 1 ? hi & hjm : 0u;
      T en, zn, vy;
      stepH<T, Coef, false, false>(g, A, C, SH, mj & hku, mj & hk, i, j - 1, k, p, A.f[0][p], A.f[1][p], A.f[2][p],
                                   en, zn, hx_jm, vy, hz_jm);
    }
    stOwn(B.f[3] + idx, hx);
    stOwn(B.f[4] + idx, hy);
    stOwn(B.f[5] + idx, hz);
    // BasicUpdator3D::updateE (src/updator/basic_updator_3d.cpp:38-69) and the
    // E-side slabs; every lane stores all three components (unchanged where
    // the Yee ranges leave a face out)
    const T dzHy = k >= 1 ? hy - hy_km : T(0);
    const T dzHx = k >= 1 ? hx - hx_km : T(0);
    const T dyHz = j >= 1 ? hz - hz_jm : T(0);
    const T dyHx = j >= 1 ? hx - hx_jm : T(0);
    const T dxHz = hz - hz_p, dxHy = hy - hy_p;
    T vx = ex0, vy = ey_i, vz = ez_i;
    {  // the Yee ranges select (the coefficient loads unconditional)
      T s0, a0, b0, s1, a1, b1, s2, a2, b2;
      C.getOwn(0, idx, s0, a0, b0);
      C.getOwn(1, idx, s1, a1, b1);
      C.getOwn(2, idx, s2, a2, b2);
      const T ux = s0 * vx + a0 * dzHy + b0 * dyHz;
      const T uy = s1 * vy + a1 * dxHz + b1 * dzHx;
      const T uz = s2 * vz + a2 * dyHx + b2 * dxHy;
      vx = (j >= 1 && k >= 1) ? ux : vx;
      vy = (i >= 1 && k >= 1) ? uy : vy;
      vz = (i >= 1 && j >= 1) ? uz : vz;
    }
    // A: Ey(x) <- Hz, Ez(y) <- Hx, Ex(z) <- Hy;  B: Ez(x) <- Hy, Ex(y) <- Hz, Ey(z) <- Hx
    const uint32_t ei = axisMask(SE, 0, i) & ej;
    if (ei & eku) stepSlabs<T, true>(SE, ei & eku, ei & ek, i, j, k, vx, vy, vz, dxHz, dyHx, dzHy, dxHy, dyHz, dzHx);
    stOwn(B.f[0] + idx, vx);
    stOwn(B.f[1] + idx, vy);
    stOwn(B.f[2] + idx, vz);
    hz_p = hz;
    hy_p = hy;
    ey_i = ey_n;
    ez_i = ez_n;
  }
}

template <class T>
struct StepLaunch {
  static void step(hipStream_t s, const Grid& g, const FieldPtrs<T>& A, const FieldPtrs<T>& B,
                   const CoefPtrs<T>& C, const StepSlabs<T>& SH, const StepSlabs<T>& SE, int ichunk);
  static void step(hipStream_t s, const Grid& g, const FieldPtrs<T>& A, const FieldPtrs<T>& B,
                   const ClassCoef<T>& C, const StepSlabs<T>& SH, const StepSlabs<T>& SE, int ichunk);
};

}  // namespace xfd
