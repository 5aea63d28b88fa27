// Reverse-time discrete adjoint of the time step (no reference counterpart).
//
// One forward step n is the composition of linear maps on the state
// (E, H, CPML psi) plus source terms that do not depend on the state:
//   Yee H over R_H -> TFSF/lumped H records -> CPML H -> Yee E over R_E ->
//   E records -> CPML E -> recording.
// The adjoint runs the transposes in reverse order.  With lambda the adjoint
// fields and mu the adjoint of each psi array, reverse step n is
//   inject    lambdaE += dJ/dE^{n+1}                       (k_adj_source)
//   gradient  g_e += lambdaE_e (beta_e E^{n+1}_e + gamma_e E^n_e)  (k_adj_grad)
//   muE       mu = b mu + c_psi lambdaE          (CPML E, per slab)
//   adj H     lambdaH = s_H lambdaH + C_E^T (a/b_E lambdaE)      (k_adj_h)
//             + D^T (a mu_E)                     (CPML E gather, per slab)
//   muH       mu = b mu + c_psi lambdaH          (CPML H, per slab)
//   adj E     lambdaE = s_E lambdaE + C_H^T (a/b_H lambdaH)      (k_adj_e)
//             + D^T (a mu_H)                     (CPML H gather, per slab)
// where the H self-coefficient s_H of step n+1 is applied lazily at the
// start of reverse step n, which is exact because nothing else touches
// lambdaH in between.  Points outside the Yee ranges (grid faces) are not
// updated by the Yee sweeps, so their self factor is 1 and they pass on no
// curl term, but they can still be written by CPML corrections and are
// therefore carried.
#pragma once

#include "kernels.hpp"

namespace xfd {

// Ranges of the Yee updates (BasicUpdator::updateH / BasicUpdator3D::updateE).
__device__ __forceinline__ bool inRH(const Grid& g, int i, int j, int k) {
  return i >= 0 && j >= 0 && k >= 0 && i < g.nx && j < g.ny && k < g.nz;
}
__device__ __forceinline__ bool inRE(const Grid& g, int q, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0) return false;
  if (q == 0) return i < g.nx && j >= 1 && j < g.ny && k >= 1 && k < g.nz;
  if (q == 1) return i >= 1 && i < g.nx && j < g.ny && k >= 1 && k < g.nz;
  return i >= 1 && i < g.nx && j >= 1 && j < g.ny && k < g.nz;
}
// Component extents: E_q is short along q, H_q long along q.
__device__ __forceinline__ bool inExt(const Grid& g, bool h, int q, int i, int j, int k) {
  const int n[3] = {g.nx, g.ny, g.nz};
  const int p[3] = {i, j, k};
  for (int a = 0; a < 3; ++a) {
    const int e = (h == (a == q)) ? n[a] + 1 : n[a];
    if (p[a] < 0 || p[a] >= e) return false;
  }
  return true;
}

// Coefficients of field q at padded index idx: self, dual-A and dual-B.
template <class T, class Coef>
__device__ __forceinline__ T coefA(const Coef& C, int q, int64_t idx) {
  T s, a, b;
  C.get(q, idx, s, a, b);
  return a;
}
template <class T, class Coef>
__device__ __forceinline__ T coefB(const Coef& C, int q, int64_t idx) {
  T s, a, b;
  C.get(q, idx, s, a, b);
  return b;
}

// Adjoint of the Yee E update for the H side, with the lazy H self factor:
// for every H point m
//   lambdaH_m = (m in R_H ? s_H(m) : 1) lambdaH_m
//             + sum over E points e in R_E that read H_m of coef(e, m) lambdaE_e.
// Forward E reads (sweepE): Ex(e) = .. + a_ex (Hy(e) - Hy(e-k)) + b_ex (Hz(e) - Hz(e-j)),
// Ey(e) = .. + a_ey (Hz(e) - Hz(e-i)) + b_ey (Hx(e) - Hx(e-k)),
// Ez(e) = .. + a_ez (Hx(e) - Hx(e-j)) + b_ez (Hy(e) - Hy(e-i)).
template <class T, class Coef>
__global__ __launch_bounds__(kTileThreads) void k_adj_h_point(Grid g, FieldPtrs<T> L, Coef C) {
  const int k = blockIdx.x * kLanesK + threadIdx.x;
  const int j = blockIdx.y * kRowsJ + threadIdx.y;
  const int i = blockIdx.z;
  if (k > g.nz || j > g.ny || i > g.nx) return;
  const int64_t pz = g.pz, pl = g.plane;
  const int64_t idx = (static_cast<int64_t>(i) * g.NY + j) * pz + k;
  const T* ex = L.f[0];
  const T* ey = L.f[1];
  const T* ez = L.f[2];
  const bool rh = inRH(g, i, j, k);
  // Hx
  if (inExt(g, true, 0, i, j, k)) {
    T v = L.f[3][idx];
    if (rh) {
      T s, a, b;
      C.get(3, idx, s, a, b);
      v = s * v;
    }
    if (inRE(g, 1, i, j, k)) v += coefB<T>(C, 1, idx) * ey[idx];
    if (inRE(g, 1, i, j, k + 1)) v -= coefB<T>(C, 1, idx + 1) * ey[idx + 1];
    if (inRE(g, 2, i, j, k)) v += coefA<T>(C, 2, idx) * ez[idx];
    if (inRE(g, 2, i, j + 1, k)) v -= coefA<T>(C, 2, idx + pz) * ez[idx + pz];
    L.f[3][idx] = v;
  }
  // Hy
  if (inExt(g, true, 1, i, j, k)) {
    T v = L.f[4][idx];
    if (rh) {
      T s, a, b;
      C.get(4, idx, s, a, b);
      v = s * v;
    }
    if (inRE(g, 0, i, j, k)) v += coefA<T>(C, 0, idx) * ex[idx];
    if (inRE(g, 0, i, j, k + 1)) v -= coefA<T>(C, 0, idx + 1) * ex[idx + 1];
    if (inRE(g, 2, i, j, k)) v += coefB<T>(C, 2, idx) * ez[idx];
    if (inRE(g, 2, i + 1, j, k)) v -= coefB<T>(C, 2, idx + pl) * ez[idx + pl];
    L.f[4][idx] = v;
  }
  // Hz
  if (inExt(g, true, 2, i, j, k)) {
    T v = L.f[5][idx];
    if (rh) {
      T s, a, b;
      C.get(5, idx, s, a, b);
      v = s * v;
    }
    if (inRE(g, 0, i, j, k)) v += coefB<T>(C, 0, idx) * ex[idx];
    if (inRE(g, 0, i, j + 1, k)) v -= coefB<T>(C, 0, idx + pz) * ex[idx + pz];
    if (inRE(g, 1, i, j, k)) v += coefA<T>(C, 1, idx) * ey[idx];
    if (inRE(g, 1, i + 1, j, k)) v -= coefA<T>(C, 1, idx + pl) * ey[idx + pl];
    L.f[5][idx] = v;
  }
}

// Adjoint of the Yee H update for the E side and the E self factor:
// for every E point e
//   lambdaE_e = (e in R_E ? s_E(e) : 1) lambdaE_e
//             + sum over H points m in R_H that read E_e of coef(m, e) lambdaH_m.
// Forward H reads (sweepH): Hx(m) = .. + a_hx (Ey(m+k) - Ey(m)) + b_hx (Ez(m+j) - Ez(m)),
// Hy(m) = .. + a_hy (Ez(m+i) - Ez(m)) + b_hy (Ex(m+k) - Ex(m)),
// Hz(m) = .. + a_hz (Ex(m+j) - Ex(m)) + b_hz (Ey(m+i) - Ey(m)).
template <class T, class Coef>
__global__ __launch_bounds__(kTileThreads) void k_adj_e_point(Grid g, FieldPtrs<T> L, Coef C) {
  const int k = blockIdx.x * kLanesK + threadIdx.x;
  const int j = blockIdx.y * kRowsJ + threadIdx.y;
  const int i = blockIdx.z;
  if (k > g.nz || j > g.ny || i > g.nx) return;
  const int64_t pz = g.pz, pl = g.plane;
  const int64_t idx = (static_cast<int64_t>(i) * g.NY + j) * pz + k;
  const T* hx = L.f[3];
  const T* hy = L.f[4];
  const T* hz = L.f[5];
  // Ex
  if (inExt(g, false, 0, i, j, k)) {
    T v = L.f[0][idx];
    if (inRE(g, 0, i, j, k)) {
      T s, a, b;
      C.get(0, idx, s, a, b);
      v = s * v;
    }
    if (inRH(g, i, j, k)) v -= coefB<T>(C, 4, idx) * hy[idx] + coefA<T>(C, 5, idx) * hz[idx];
    if (inRH(g, i, j, k - 1)) v += coefB<T>(C, 4, idx - 1) * hy[idx - 1];
    if (inRH(g, i, j - 1, k)) v += coefA<T>(C, 5, idx - pz) * hz[idx - pz];
    L.f[0][idx] = v;
  }
  // Ey
  if (inExt(g, false, 1, i, j, k)) {
    T v = L.f[1][idx];
    if (inRE(g, 1, i, j, k)) {
      T s, a, b;
      C.get(1, idx, s, a, b);
      v = s * v;
    }
    if (inRH(g, i, j, k)) v -= coefA<T>(C, 3, idx) * hx[idx] + coefB<T>(C, 5, idx) * hz[idx];
    if (inRH(g, i, j, k - 1)) v += coefA<T>(C, 3, idx - 1) * hx[idx - 1];
    if (inRH(g, i - 1, j, k)) v += coefB<T>(C, 5, idx - pl) * hz[idx - pl];
    L.f[1][idx] = v;
  }
  // Ez
  if (inExt(g, false, 2, i, j, k)) {
    T v = L.f[2][idx];
    if (inRE(g, 2, i, j, k)) {
      T s, a, b;
      C.get(2, idx, s, a, b);
      v = s * v;
    }
    if (inRH(g, i, j, k)) v -= coefB<T>(C, 3, idx) * hx[idx] + coefA<T>(C, 4, idx) * hy[idx];
    if (inRH(g, i, j - 1, k)) v += coefB<T>(C, 3, idx - pz) * hx[idx - pz];
    if (inRH(g, i - 1, j, k)) v += coefA<T>(C, 4, idx - pl) * hy[idx - pl];
    L.f[2][idx] = v;
  }
}

// The same two transposes as sweeps over the forward sweeps' tiles (64
// k-lanes by 4 j-rows, each lane marching a chunk of i-planes), written over
// products: with P the masked products of one field's coefficients and its
// adjoint value (zero outside the field's update range), every term above is
// a difference of products at the point and at one neighbour, like a
// forward update.  The i neighbour's products are carried from the previous
// plane of the march; the k and j neighbours' are computed from their loads
// (the k neighbour's lines are the lane's own, the j neighbour's another
// wavefront's row).  Every load of a plane is issued before any store of it
// and all indices stay inside the padded arrays, so the range tests select
// instead of branching and a plane costs one round trip to memory.

// a_q lambdaE_q and b_q lambdaE_q of the E point at idx, q = 0, 1, 2 (zero
// outside R_E_q).
template <class T>
struct EProd {
  T a0, b0, a1, b1, a2, b2;
};
template <class T, class Coef>
__device__ __forceinline__ EProd<T> eProd(const Grid& g, const T* __restrict__ ex, const T* __restrict__ ey,
                                          const T* __restrict__ ez, const Coef& C, int i, int j, int k,
                                          int64_t idx) {
  T s0, a0, b0, s1, a1, b1, s2, a2, b2;
  C.get(0, idx, s0, a0, b0);
  C.get(1, idx, s1, a1, b1);
  C.get(2, idx, s2, a2, b2);
  const T l0 = ex[idx], l1 = ey[idx], l2 = ez[idx];
  const bool r0 = inRE(g, 0, i, j, k), r1 = inRE(g, 1, i, j, k), r2 = inRE(g, 2, i, j, k);
  EProd<T> p;
  p.a0 = r0 ? a0 * l0 : T(0);
  p.b0 = r0 ? b0 * l0 : T(0);
  p.a1 = r1 ? a1 * l1 : T(0);
  p.b1 = r1 ? b1 * l1 : T(0);
  p.a2 = r2 ? a2 * l2 : T(0);
  p.b2 = r2 ? b2 * l2 : T(0);
  return p;
}

// a_q lambdaH_q and b_q lambdaH_q of the H point at idx, q = 3, 4, 5 (zero
// outside R_H).
template <class T>
struct HProd {
  T a3, b3, a4, b4, a5, b5;
};
template <class T, class Coef>
__device__ __forceinline__ HProd<T> hProd(const Grid& g, const T* __restrict__ hx, const T* __restrict__ hy,
                                          const T* __restrict__ hz, const Coef& C, int i, int j, int k,
                                          int64_t idx) {
  T s3, a3, b3, s4, a4, b4, s5, a5, b5;
  C.get(3, idx, s3, a3, b3);
  C.get(4, idx, s4, a4, b4);
  C.get(5, idx, s5, a5, b5);
  const T l3 = hx[idx], l4 = hy[idx], l5 = hz[idx];
  const bool r = inRH(g, i, j, k);
  HProd<T> p;
  p.a3 = r ? a3 * l3 : T(0);
  p.b3 = r ? b3 * l3 : T(0);
  p.a4 = r ? a4 * l4 : T(0);
  p.b4 = r ? b4 * l4 : T(0);
  p.a5 = r ? a5 * l5 : T(0);
  p.b5 = r ? b5 * l5 : T(0);
  return p;
}

// lambdaH over R_H = [0,nx) x [0,ny) x [0,nz) (the H points outside it are
// far faces, which no E update reads):
//   lambdaHx = s_Hx lambdaHx + Pb1 - Pb1(k+1) + Pa2 - Pa2(j+1)
//   lambdaHy = s_Hy lambdaHy + Pa0 - Pa0(k+1) + Pb2 - Pb2(i+1)
//   lambdaHz = s_Hz lambdaHz + Pb0 - Pb0(j+1) + Pa1 - Pa1(i+1)
// Plane i+1's products become the next plane's own.  The neighbours read
// here (i+1 <= nx, j+1 <= ny, k+1 <= nz) lie inside the padded arrays.
template <class T, class Coef>
__global__ __launch_bounds__(kTileThreads) void k_adj_h(Grid g, FieldPtrs<T> L, Coef C, SweepMap M, int ichunk) {
  int bx, by, bz;
  M.coords(bx, by, bz);
  const int k = bx * kLanesK + threadIdx.x;
  const int j = by * kRowsJ + threadIdx.y;
  const int i0 = bz * ichunk;
  if (k >= g.nz || j >= g.ny || i0 >= g.nx) return;
  const int i1 = min(i0 + ichunk, g.nx);
  const int64_t pz = g.pz, pl = g.plane;
  int64_t idx = (static_cast<int64_t>(i0) * g.NY + j) * pz + k;
  const T* __restrict__ ex = L.f[0];
  const T* __restrict__ ey = L.f[1];
  const T* __restrict__ ez = L.f[2];
  T* __restrict__ hx = L.f[3];
  T* __restrict__ hy = L.f[4];
  T* __restrict__ hz = L.f[5];
  EProd<T> own = eProd(g, ex, ey, ez, C, i0, j, k, idx);
  for (int i = i0; i < i1; ++i, idx += pl) {
    const EProd<T> nxt = eProd(g, ex, ey, ez, C, i + 1, j, k, idx + pl);
    T s, a, b;
    C.get(1, idx + 1, s, a, b);
    const T b1kc = b, l1k = ey[idx + 1];
    C.get(0, idx + 1, s, a, b);
    const T a0kc = a, l0k = ex[idx + 1];
    C.get(2, idx + pz, s, a, b);
    const T a2jc = a, l2j = ez[idx + pz];
    C.get(0, idx + pz, s, a, b);
    const T b0jc = b, l0j = ex[idx + pz];
    T s3, s4, s5;  // the point's own values: streamed (non-temporal)
    C.getOwn(3, idx, s3, a, b);
    C.getOwn(4, idx, s4, a, b);
    C.getOwn(5, idx, s5, a, b);
    const T vx0 = ldOwn(hx + idx), vy0 = ldOwn(hy + idx), vz0 = ldOwn(hz + idx);
    const T b1k = inRE(g, 1, i, j, k + 1) ? b1kc * l1k : T(0);
    const T a0k = inRE(g, 0, i, j, k + 1) ? a0kc * l0k : T(0);
    const T a2j = inRE(g, 2, i, j + 1, k) ? a2jc * l2j : T(0);
    const T b0j = inRE(g, 0, i, j + 1, k) ? b0jc * l0j : T(0);
    T v = s3 * vx0;
    v += own.b1;
    v -= b1k;
    v += own.a2;
    v -= a2j;
    stOwn(hx + idx, v);
    v = s4 * vy0;
    v += own.a0;
    v -= a0k;
    v += own.b2;
    v -= nxt.b2;
    stOwn(hy + idx, v);
    v = s5 * vz0;
    v += own.b0;
    v -= b0j;
    v += own.a1;
    v -= nxt.a1;
    stOwn(hz + idx, v);
    own = nxt;
  }
}

// lambdaE over [0,nx] x [0,ny] x [0,nz] (each component within its extent;
// the far-face points outside R_E keep self factor 1 but still gather the
// terms of the H points in R_H that read them):
//   lambdaEx = (s_Ex) lambdaEx - (Pb4 + Pa5) + Pb4(k-1) + Pa5(j-1)
//   lambdaEy = (s_Ey) lambdaEy - (Pa3 + Pb5) + Pa3(k-1) + Pb5(i-1)
//   lambdaEz = (s_Ez) lambdaEz - (Pb3 + Pa4) + Pb3(j-1) + Pa4(i-1)
// Plane i-1's products are carried from the previous plane.  The k-1 / j-1
// neighbours are read at the point itself where they would fall below the
// grid (their products are zero there).
template <class T, class Coef>
__global__ __launch_bounds__(kTileThreads) void k_adj_e(Grid g, FieldPtrs<T> L, Coef C, SweepMap M, int ichunk) {
  int bx, by, bz;
  M.coords(bx, by, bz);
  const int k = bx * kLanesK + threadIdx.x;
  const int j = by * kRowsJ + threadIdx.y;
  const int i0 = bz * ichunk;
  if (k > g.nz || j > g.ny || i0 > g.nx) return;
  const int i1 = min(i0 + ichunk, g.nx + 1);
  const int64_t pz = g.pz, pl = g.plane;
  int64_t idx = (static_cast<int64_t>(i0) * g.NY + j) * pz + k;
  const int64_t dk = k >= 1 ? 1 : 0, dj = j >= 1 ? pz : 0;
  T* __restrict__ ex = L.f[0];
  T* __restrict__ ey = L.f[1];
  T* __restrict__ ez = L.f[2];
  const T* __restrict__ hx = L.f[3];
  const T* __restrict__ hy = L.f[4];
  const T* __restrict__ hz = L.f[5];
  HProd<T> prev{T(0), T(0), T(0), T(0), T(0), T(0)};
  if (i0 > 0) prev = hProd(g, hx, hy, hz, C, i0 - 1, j, k, idx - pl);
  for (int i = i0; i < i1; ++i, idx += pl) {
    const HProd<T> own = hProd(g, hx, hy, hz, C, i, j, k, idx);
    T s, a, b;
    C.get(3, idx - dk, s, a, b);
    const T a3kc = a, l3k = hx[idx - dk];
    C.get(4, idx - dk, s, a, b);
    const T b4kc = b, l4k = hy[idx - dk];
    C.get(5, idx - dj, s, a, b);
    const T a5jc = a, l5j = hz[idx - dj];
    C.get(3, idx - dj, s, a, b);
    const T b3jc = b, l3j = hx[idx - dj];
    T s0, s1, s2;  // the point's own values: streamed (non-temporal)
    C.getOwn(0, idx, s0, a, b);
    C.getOwn(1, idx, s1, a, b);
    C.getOwn(2, idx, s2, a, b);
    const T vx0 = ldOwn(ex + idx), vy0 = ldOwn(ey + idx), vz0 = ldOwn(ez + idx);
    const bool rk = inRH(g, i, j, k - 1), rj = inRH(g, i, j - 1, k);
    const T a3k = rk ? a3kc * l3k : T(0);
    const T b4k = rk ? b4kc * l4k : T(0);
    const T a5j = rj ? a5jc * l5j : T(0);
    const T b3j = rj ? b3jc * l3j : T(0);
    // every point stores all three components, a component outside its
    // extent (Ex at i = nx, Ey at j = ny, Ez at k = nz: padding) its old value,
    // so no load waits behind a branch
    T v = inRE(g, 0, i, j, k) ? s0 * vx0 : vx0;
    v -= own.b4 + own.a5;
    v += b4k;
    v += a5j;
    stOwn(ex + idx, i < g.nx ? v : vx0);
    v = inRE(g, 1, i, j, k) ? s1 * vy0 : vy0;
    v -= own.a3 + own.b5;
    v += a3k;
    v += prev.b5;
    stOwn(ey + idx, j < g.ny ? v : vy0);
    v = inRE(g, 2, i, j, k) ? s2 * vz0 : vz0;
    v -= own.b3 + own.a4;
    v += b3j;
    v += prev.a4;
    stOwn(ez + idx, k < g.nz ? v : vz0);
    prev = own;
  }
}

// CPML adjoint, first half: for every psi point of one slab and half-step,
// mu = b mu + c_psi lambda_field (the slab's psi arrays in `s` hold mu, and
// the field pointers are the adjoint fields).
template <class T>
__device__ __forceinline__ void adjMuCell(const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& s, int64_t t) {
  const int64_t vol = static_cast<int64_t>(s.ext[0]) * s.ext[1] * s.ext[2];
  if (t >= vol) return;
  const int w = static_cast<int>(t % s.ext[2]);
  const int v = static_cast<int>((t / s.ext[2]) % s.ext[1]);
  const int u = static_cast<int>(t / (static_cast<int64_t>(s.ext[2]) * s.ext[1]));
  const int p[3] = {s.lo[0] + u, s.lo[1] + v, s.lo[2] + w};
  const int c = s.axis, a = (c + 1) % 3, b = (c + 2) % 3;
  const T cb = s.cb[p[c] - s.start];
  const int64_t idx = (static_cast<int64_t>(p[0]) * g.NY + p[1]) * g.pz + p[2];
  if (p[a] < s.extA[0] && p[b] < s.extA[1]) s.psiA[t] = cb * s.psiA[t] + s.cpsiA[t] * L.f[s.fieldA][idx];
  if (p[a] < s.extB[0] && p[b] < s.extB[1]) s.psiB[t] = cb * s.psiB[t] + s.cpsiB[t] * L.f[s.fieldB][idx];
}

// CPML adjoint, second half: the transpose of the psi difference.  E slab:
// psi(P) reads dual(P) - dual(P - c), so dual point P gathers
// + a(P) mu(P) - a(P + c) mu(P + c); H slab: psi(P) reads dual(P + c) -
// dual(P), so P gathers + a(P - c) mu(P - c) - a(P) mu(P).  One thread per
// dual point of the slab box widened by one plane (below for E, above for H).
template <class T, bool IsE>
__device__ __forceinline__ void adjGatherCell(const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& s,
                                              int64_t t) {
  const int c = s.axis, a = (c + 1) % 3, b = (c + 2) % 3;
  int wext[3] = {s.ext[0], s.ext[1], s.ext[2]};
  wext[c] += 1;
  const int64_t vol = static_cast<int64_t>(wext[0]) * wext[1] * wext[2];
  if (t >= vol) return;
  const int w = static_cast<int>(t % wext[2]);
  const int v = static_cast<int>((t / wext[2]) % wext[1]);
  const int u = static_cast<int>(t / (static_cast<int64_t>(wext[2]) * wext[1]));
  int p[3] = {s.lo[0] + u, s.lo[1] + v, s.lo[2] + w};
  if (IsE) p[c] -= 1;  // dual planes lo-1 .. lo+ext-1
  if (p[c] < 0) return;
  const int64_t idx = (static_cast<int64_t>(p[0]) * g.NY + p[1]) * g.pz + p[2];
  // psi points that read this dual point: P (sign +1 for E, -1 for H) and
  // P + c (E, sign -1) or P - c (H, sign +1)
  const int pin = IsE ? p[c] : p[c] - 1;      // the psi plane with the + sign
  const int pout = IsE ? p[c] + 1 : p[c];     // the psi plane with the - sign
  auto local = [&](int pc) -> int64_t {       // psi index in the slab box, or -1
    if (pc < s.lo[c] || pc >= s.lo[c] + s.ext[c]) return -1;
    int q[3] = {p[0] - s.lo[0], p[1] - s.lo[1], p[2] - s.lo[2]};
    q[c] = pc - s.lo[c];
    return (static_cast<int64_t>(q[0]) * s.ext[1] + q[1]) * s.ext[2] + q[2];
  };
  const int64_t tp = local(pin), tm = local(pout);
  // component A of the slab reads dual B; component B reads dual A
  if (p[a] < s.extA[0] && p[b] < s.extA[1]) {
    T acc = T(0);
    if (tp >= 0) acc += s.ca[pin - s.start] * s.psiA[tp];
    if (tm >= 0) acc -= s.ca[pout - s.start] * s.psiA[tm];
    if (tp >= 0 || tm >= 0) L.f[s.dualB][idx] += acc;
  }
  if (p[a] < s.extB[0] && p[b] < s.extB[1]) {
    T acc = T(0);
    if (tp >= 0) acc += s.ca[pin - s.start] * s.psiB[tp];
    if (tm >= 0) acc -= s.ca[pout - s.start] * s.psiB[tm];
    if (tp >= 0 || tm >= 0) L.f[s.dualA][idx] += acc;
  }
}

// Transposed inductor corrections.  The forward record of an inductor edge
// (InductorCorrector::correctE) is E' = E + c J followed by J' = J + w E';
// its transpose, applied in reverse record order per corrected point, is
// lambdaE += w lambdaJ, then lambdaJ += c lambdaE (lambdaJ: one adjoint
// current per inductor edge).  The other record kinds add terms that do not
// depend on the state and have no transpose.
template <class T>
__global__ __launch_bounds__(256) void k_adj_records(FieldPtrs<T> L, const Record* __restrict__ rec,
                                                     const int64_t* __restrict__ seg, int64_t nseg, T* lamJ) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const int64_t r0 = seg[s], r1 = seg[s + 1];
  T* f = L.f[rec[r0].field];
  const int64_t target = rec[r0].target;
  double le = static_cast<double>(f[target]);
  bool any = false;
  for (int64_t r = r1 - 1; r >= r0; --r) {
    const Record R = rec[r];
    if (R.kind != 2) continue;
    any = true;
    const double lj = static_cast<double>(lamJ[R.pool]);
    le = le + R.w * lj;
    lamJ[R.pool] = static_cast<T>(lj + R.coef * le);
  }
  if (any) f[target] = static_cast<T>(le);
}

// Transposed dispersive E update.  Per cell and component the forward update
// (adeUpdate, after the plain curl update en = s E + curl terms) is
//   enew = ceep Eprev + en + cej sum_p (c5_p J_p + c6_p Jp_p)
//   J_p' = c3_p J_p + c1_p (enew - E)                       Debye
//   J_p' = c3_p J_p + c1_p (enew + E)                       Drude
//   J_p' = c0_p enew + c1_p E + c2_p Eprev + c3_p J_p + c4_p Jp_p,
//   Jp_p' = J_p, Eprev' = E                                 modified Lorentz
//   E' = enew
// (c_k: the pole's coefficients in the material table, ceep and c6 only in
// the modified-Lorentz scheme).  Its transpose splits around the Yee
// transposes: before them (k_adj_ade_pre) lambdaE becomes the adjoint of
// enew,
//   lambda_enew = lambdaE' + sum_p d_p lambdaJ_p'   (d_p = c1_p, or c0_p),
// which the curl transpose (k_adj_h) and the self factor (k_adj_e, s times
// lambda_enew) then read like any lambdaE; the pole and previous-E adjoints
// take their old-state values, and the remaining part of lambdaE,
//   X = sum_p e_p lambdaJ_p' (+ lambdaEprev')       (e_p = -c1_p, +c1_p, c1_p),
// is kept in x and added after the Yee transposes (k_adj_ade_post).
// LA holds the adjoint state in the forward state's layout (j: lambdaJ,
// jp: lambdaJp, eprev: lambdaEprev).
template <class T, int S>
__device__ __forceinline__ void adeAdjPoint(const AdeArgs<T>& A, const AdeArgs<T>& LA, int m, int q, int64_t idx,
                                            T* lam, T* x) {
  const T* c = A.tab + static_cast<int64_t>(m) * A.stride;
  const T lE = lam[idx];
  T le = lE;
  T xv = S == kAdeMLorentz ? LA.eprev[q][idx] : T(0);
  T lep = T(0);
  for (int p = 0; p < A.poles; ++p) {
    const T* cp = c + 2 + 7 * p;
    const T lj = LA.j[q][p * A.total + idx];
    le += (S == kAdeMLorentz ? cp[0] : cp[1]) * lj;
    xv += (S == kAdeDebye ? -cp[1] : cp[1]) * lj;
    if (S == kAdeMLorentz) lep += cp[2] * lj;
  }
  for (int p = 0; p < A.poles; ++p) {
    const T* cp = c + 2 + 7 * p;
    const int64_t o = p * A.total + idx;
    const T lj = LA.j[q][o];
    if (S == kAdeMLorentz) {
      const T ljp = LA.jp[q][o];
      LA.j[q][o] = c[0] * cp[5] * le + cp[3] * lj + ljp;
      LA.jp[q][o] = c[0] * cp[6] * le + cp[4] * lj;
    } else {
      LA.j[q][o] = c[0] * cp[5] * le + cp[3] * lj;
    }
  }
  if (S == kAdeMLorentz) LA.eprev[q][idx] = c[1] * le + lep;
  lam[idx] = le;
  x[idx] = xv;
}

// Over the E sweep's points [0,nx) x [0,ny) x [0,nz), the components the
// forward ADE update touches (material id set, inside the Yee range).
template <class T, int S>
__global__ __launch_bounds__(256) void k_adj_ade_pre(Grid g, FieldPtrs<T> L, AdeArgs<T> A, AdeArgs<T> LA,
                                                     FieldPtrs<T> X) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(g.nx) * g.ny * g.nz) return;
  const int k = static_cast<int>(t % g.nz);
  const int j = static_cast<int>((t / g.nz) % g.ny);
  const int i = static_cast<int>(t / (static_cast<int64_t>(g.nz) * g.ny));
  const int64_t idx = (static_cast<int64_t>(i) * g.NY + j) * g.pz + k;
  const int m = A.mat[idx];
  if (m < 0) return;
  if (j >= 1 && k >= 1) adeAdjPoint<T, S>(A, LA, m, 0, idx, L.f[0], X.f[0]);
  if (i >= 1 && k >= 1) adeAdjPoint<T, S>(A, LA, m, 1, idx, L.f[1], X.f[1]);
  if (i >= 1 && j >= 1) adeAdjPoint<T, S>(A, LA, m, 2, idx, L.f[2], X.f[2]);
}

template <class T>
__global__ __launch_bounds__(256) void k_adj_ade_post(Grid g, FieldPtrs<T> L, const int16_t* __restrict__ mat,
                                                      FieldPtrs<T> X) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(g.nx) * g.ny * g.nz) return;
  const int k = static_cast<int>(t % g.nz);
  const int j = static_cast<int>((t / g.nz) % g.ny);
  const int i = static_cast<int>(t / (static_cast<int64_t>(g.nz) * g.ny));
  const int64_t idx = (static_cast<int64_t>(i) * g.NY + j) * g.pz + k;
  if (mat[idx] < 0) return;
  if (j >= 1 && k >= 1) L.f[0][idx] += X.f[0][idx];
  if (i >= 1 && k >= 1) L.f[1][idx] += X.f[1][idx];
  if (i >= 1 && j >= 1) L.f[2][idx] += X.f[2][idx];
}

template <class T>
__global__ __launch_bounds__(256) void k_adj_cpml_mu(Grid g, FieldPtrs<T> L, SlabArgs<T> s) {
  adjMuCell(g, L, s, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
}
template <class T, bool IsE>
__global__ __launch_bounds__(256) void k_adj_cpml_gather(Grid g, FieldPtrs<T> L, SlabArgs<T> s) {
  adjGatherCell<T, IsE>(g, L, s, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
}
// Two slabs in one launch, work-groups [0, nb0) on s0 and the rest on s1:
// any two for the mu update (each writes its own arrays and reads lambda
// only); for the gather, two whose widened boxes share no dual point.
template <class T>
__global__ __launch_bounds__(256) void k_adj_cpml_mu2(Grid g, FieldPtrs<T> L, SlabArgs<T> s0, SlabArgs<T> s1,
                                                      unsigned nb0) {
  if (blockIdx.x < nb0)
    adjMuCell(g, L, s0, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
  else
    adjMuCell(g, L, s1, static_cast<int64_t>(blockIdx.x - nb0) * blockDim.x + threadIdx.x);
}
template <class T, bool IsE>
__global__ __launch_bounds__(256) void k_adj_cpml_gather2(Grid g, FieldPtrs<T> L, SlabArgs<T> s0, SlabArgs<T> s1,
                                                          unsigned nb0) {
  if (blockIdx.x < nb0)
    adjGatherCell<T, IsE>(g, L, s0, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
  else
    adjGatherCell<T, IsE>(g, L, s1, static_cast<int64_t>(blockIdx.x - nb0) * blockDim.x + threadIdx.x);
}

// Design edges: the E edges of the design cells, as three boxes (Ex, Ey,
// Ez), k fastest; off[q] is the first edge of component q, off[3] the total.
struct DesignArgs {
  int o[3][3];  // box origin per component
  int e[3][3];  // box extent per component
  int64_t off[4];
};

__device__ __forceinline__ int64_t designIndex(const Grid& g, const DesignArgs& d, int64_t t, int& q) {
  q = t >= d.off[2] ? 2 : (t >= d.off[1] ? 1 : 0);
  const int64_t l = t - d.off[q];
  const int w = static_cast<int>(l % d.e[q][2]);
  const int v = static_cast<int>((l / d.e[q][2]) % d.e[q][1]);
  const int u = static_cast<int>(l / (static_cast<int64_t>(d.e[q][2]) * d.e[q][1]));
  return (static_cast<int64_t>(d.o[q][0] + u) * g.NY + (d.o[q][1] + v)) * g.pz + (d.o[q][2] + w);
}

// Copies the design edges' E values into history slot `slot`.
template <class T>
__global__ __launch_bounds__(kTileThreads) void k_adj_hist(Grid g, FieldPtrs<T> F, DesignArgs d, T* hist) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= d.off[3]) return;
  int q;
  const int64_t idx = designIndex(g, d, t, q);
  hist[t] = F.f[q][idx];
}

// g_e += lambdaE_e (beta_e E^{n+1}_e + gamma_e E^n_e), accumulated in FP64.
template <class T>
__global__ __launch_bounds__(256) void k_adj_grad(Grid g, FieldPtrs<T> L, DesignArgs d, const T* __restrict__ h0,
                                                  const T* __restrict__ h1, const double* __restrict__ beta,
                                                  const double* __restrict__ gamma, double* grad) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= d.off[3]) return;
  int q;
  const int64_t idx = designIndex(g, d, t, q);
  const double lam = static_cast<double>(L.f[q][idx]);
  grad[t] += lam * (beta[t] * static_cast<double>(h1[t]) + gamma[t] * static_cast<double>(h0[t]));
}

// lambda += dJ/dF at the objective points for step n.
struct AdjPoint {
  int64_t idx;
  int32_t field;
  int32_t pad;
};

template <class T>
__global__ __launch_bounds__(64) void k_adj_source(FieldPtrs<T> L, const AdjPoint* __restrict__ pts, int n,
                                                   const double* __restrict__ src, int64_t nt, int64_t step) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  T* f = L.f[pts[q].field];
  f[pts[q].idx] += static_cast<T>(src[q * nt + step]);
}

template <class T>
struct AdjLaunch {
  static void adjH(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const CoefPtrs<T>& C, int ichunk);
  static void adjH(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const ClassCoef<T>& C, int ichunk);
  static void adjE(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const CoefPtrs<T>& C, int ichunk);
  static void adjE(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const ClassCoef<T>& C, int ichunk);
  static void adjH(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const WideCoef<T>& C, int ichunk);
  static void adjE(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const WideCoef<T>& C, int ichunk);
  static void cpmlMu(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& a);
  static void cpmlGather(hipStream_t s, bool is_e, const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& a);
  static void cpmlMu2(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& a0,
                      const SlabArgs<T>& a1);
  static void cpmlGather2(hipStream_t s, bool is_e, const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& a0,
                          const SlabArgs<T>& a1);
  static void hist(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const DesignArgs& d, T* hist);
  static void grad(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const DesignArgs& d, const T* h0, const T* h1,
                   const double* beta, const double* gamma, double* grad);
  static void source(hipStream_t s, const FieldPtrs<T>& L, const AdjPoint* pts, int n, const double* src, int64_t nt,
                     int64_t step);
  static void records(hipStream_t s, const FieldPtrs<T>& L, const Record* rec, const int64_t* seg, int64_t nseg,
                      T* lamJ);
  static void adePre(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const AdeArgs<T>& A, const AdeArgs<T>& LA,
                     const FieldPtrs<T>& X, int scheme);
  static void adePost(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const int16_t* mat, const FieldPtrs<T>& X);
};

}  // namespace xfd
