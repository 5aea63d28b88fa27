// Hot-path kernels for gfx950, templated on the field precision T.
//
// Every kernel here is an HBM-bound stencil or stream (about 0.1 flop per
// byte), so the design goal is the fewest bytes per cell, not MFMA: lanes run
// along k (the contiguous axis) so a 64-wide wavefront issues whole 128-byte
// lines, a workgroup is 64 (k) x 4 (j), and each thread marches a chunk of
// i-planes carrying the i-1 / i+1 neighbour of the previous plane in
// registers, so the only re-reads are the j +/- 1 and k +/- 1 neighbours,
// which hit in L1/L2.
//
// Arithmetic is written in the reference's evaluation order
// (src/include/updator/update_scheme.h:26-62: c1*f + c2*(p-q) + c3*(p-q),
// left to right).  The FP64 translation unit is compiled with
// -ffp-contract=off, which makes its results bitwise equal to the
// reference's CPU path; the FP32 unit lets the compiler fuse.
#pragma once

#include "layout.hpp"
#include "slab_inline.hpp"

namespace xfd {

// Sweep tile: kLanesK k-lanes (a multiple of the 64-wide wavefront) by
// kRowsJ j-rows; XFD_TILE_K / XFD_TILE_J override them at build time.
#ifndef XFD_TILE_K
#define XFD_TILE_K 64
#endif
#ifndef XFD_TILE_J
#define XFD_TILE_J 4
#endif
constexpr int kLanesK = XFD_TILE_K;
constexpr int kRowsJ = XFD_TILE_J;
constexpr int kTileThreads = kLanesK * kRowsJ;
static_assert(kLanesK % 64 == 0 && kTileThreads <= 1024, "sweep tile");

// The CPML corrections of one point, fused into a Yee sweep: for every slab
// of P containing the point (in boundary order), for its components A =
// (c+1)%3 and B = (c+2)%3 of this half-step,
//   psi = b psi + a d,  v += c_psi psi
// with d the dual difference along the slab axis c, which is the Yee curl
// difference the sweep already holds: dA[c] for component A, dB[c] for B.
// Same operations in the same order as k_cpml after the sweep, so results
// are unchanged; the field and its dual are not read again.

// Bit n set: slab n of P contains (i, j, k) for some i of [i0, i1).  Taken
// once per thread and i-chunk, so the sweep's inner loop only tests i.
template <class T>
__device__ __forceinline__ uint32_t cpmlChunkMask(const CpmlList<T>& P, int i0, int i1, int j, int k) {
  uint32_t m = 0;
  for (int n = 0; n < P.n; ++n) {
    const SlabArgs<T>& S = P.s[n];
    if (j >= S.lo[1] && j < S.lo[1] + S.ext[1] && k >= S.lo[2] && k < S.lo[2] + S.ext[2] && i1 > S.lo[0] &&
        i0 < S.lo[0] + S.ext[0])
      m |= 1u << n;
  }
  return m;
}

// The CPML corrections of one point, fused into a Yee sweep: for every slab
// of P containing the point (in boundary order), for its components A =
// (c+1)%3 and B = (c+2)%3 of this half-step,
//   psi = b psi + a d,  v += c_psi psi
// with d the dual difference along the slab axis c, which is the Yee curl
// difference the sweep already holds: dA[c] for component A, dB[c] for B.
// Same operations in the same order as k_cpml after the sweep, so results
// are unchanged; the field and its dual are not read again.
template <class T>
__device__ __forceinline__ void cpmlFused(const CpmlList<T>& P, uint32_t mask, int i, int j, int k, T& v0, T& v1,
                                          T& v2, const T dA0, const T dA1, const T dA2, const T dB0, const T dB1,
                                          const T dB2) {
  while (mask) {
    const int n = __builtin_ctz(mask);
    mask &= mask - 1;
    const SlabArgs<T>& S = P.s[n];
    const int u = i - S.lo[0];
    if (u < 0 || u >= S.ext[0]) continue;
    const int c = S.axis, a = c == 2 ? 0 : c + 1, b = c == 0 ? 2 : c - 1;
    const int64_t t = (static_cast<int64_t>(u) * S.ext[1] + (j - S.lo[1])) * S.ext[2] + (k - S.lo[2]);
    const int layer = pick3(c, i, j, k) - S.start;
    const int pa = pick3(a, i, j, k), pb = pick3(b, i, j, k);
    const T ca = S.ca[layer], cb = S.cb[layer];
    if (pa < S.extA[0] && pb < S.extA[1]) {
      const T d = c == 0 ? dA0 : (c == 1 ? dA1 : dA2);
      const T psi = cb * S.psiA[t] + ca * d;
      S.psiA[t] = psi;
      const T add = S.cpsiA[t] * psi;  // value selects: no address of v0..v2 is taken
      const T n0 = a == 0 ? v0 + add : v0, n1 = a == 1 ? v1 + add : v1, n2 = a == 2 ? v2 + add : v2;
      v0 = n0;
      v1 = n1;
      v2 = n2;
    }
    if (pa < S.extB[0] && pb < S.extB[1]) {
      const T d = c == 0 ? dB0 : (c == 1 ? dB1 : dB2);
      const T psi = cb * S.psiB[t] + ca * d;
      S.psiB[t] = psi;
      const T add = S.cpsiB[t] * psi;
      const T n0 = b == 0 ? v0 + add : v0, n1 = b == 1 ? v1 + add : v1, n2 = b == 2 ? v2 + add : v2;
      v0 = n0;
      v1 = n1;
      v2 = n2;
    }
  }
}

// A sweep lane's view of the z slabs (ZCpml): whether its k lies in one and
// in which, that layer's coefficients, and the box-local psi index of the
// lane's point, advanced by one x-plane per i.  The slab arrays are picked
// at use with value selects between the two uniform pointers (no
// runtime-indexed aggregate, which would go to scratch memory, and no
// long-lived per-lane pointers).
template <class T>
struct ZLane {
  bool in = false, sel = false;
  int t = 0, step = 0;
  T ca = T(0), cb = T(0);
};

template <class T>
__device__ __forceinline__ ZLane<T> zLane(const ZCpml<T>& Z, int i0, int j, int k) {
  ZLane<T> L;
  const ZSlab<T>& s0 = Z.z[0];
  const ZSlab<T>& s1 = Z.z[1];
  const bool in0 = Z.n > 0 && k >= s0.klo && k < s0.khi;
  const bool in1 = Z.n > 1 && k >= s1.klo && k < s1.khi;
  // FP64: a j row in a y slab is frame throughout (FP32 has no frame)
  L.in = (in0 || in1) && (sizeof(T) == 4 || !Z.fr.inY(j));
  L.sel = in1;
  if (!L.in) return L;
  const int klo = in1 ? s1.klo : s0.klo;
  const int ext1 = in1 ? s1.ext1 : s0.ext1, ext2 = in1 ? s1.ext2 : s0.ext2;
  const T* ca = in1 ? s1.ca : s0.ca;
  const T* cb = in1 ? s1.cb : s0.cb;
  L.ca = ca[k - klo];
  L.cb = cb[k - klo];
  L.step = ext1 * ext2;
  L.t = (i0 * ext1 + j) * ext2 + (k - klo);
  return L;
}

// The z-slab corrections of the sweep point (PMLCorrector order within the
// slab: component A, then B):
//   psi = b psi + a d,  v += c_psi psi,
// with d the dual difference along z, off the frame.  zLoad issues the
// point's psi and c_psi reads (and, for E, the k-1 duals) at the top of the
// plane, beside the Yee reads, so a slab lane waits for memory once per
// plane, not twice; zStore finishes the corrections once the Yee values are
// known.  psi and c_psi are plain (cached) accesses: one 128-byte line of
// them serves the 4 j-rows of a work-group (measured: the non-temporal hint
// costs 8 % of the H sweep).
template <class T>
struct ZVals {
  T pA, pB, cA, cB, dualA, dualB;
  bool act;
};

template <class T, bool IsE>
__device__ __forceinline__ ZVals<T> zLoad(const ZCpml<T>& Z, const ZLane<T>& L, int i, const T* __restrict__ duA,
                                          const T* __restrict__ duB, int64_t idx) {
  ZVals<T> v{T(0), T(0), T(0), T(0), T(0), T(0), false};
  v.act = L.in && (sizeof(T) == 4 || !Z.fr.inX(i));
  if (v.act) {
    const ZSlab<T>& s0 = Z.z[0];
    const ZSlab<T>& s1 = Z.z[1];
    v.pA = (L.sel ? s1.psiA : s0.psiA)[L.t];
    v.cA = (L.sel ? s1.cpsiA : s0.cpsiA)[L.t];
    v.pB = (L.sel ? s1.psiB : s0.psiB)[L.t];
    v.cB = (L.sel ? s1.cpsiB : s0.cpsiB)[L.t];
    if (IsE) {
      v.dualA = duA[idx - 1];
      v.dualB = duB[idx - 1];
    }
  }
  return v;
}

template <class T>
__device__ __forceinline__ void zStore(const ZCpml<T>& Z, const ZLane<T>& L, const ZVals<T>& v, T& vA, T& vB, T dA,
                                       T dB) {
  if (v.act) {
    const ZSlab<T>& s0 = Z.z[0];
    const ZSlab<T>& s1 = Z.z[1];
    const T psiA = L.cb * v.pA + L.ca * dA;
    const T psiB = L.cb * v.pB + L.ca * dB;
    (L.sel ? s1.psiA : s0.psiA)[L.t] = psiA;
    (L.sel ? s1.psiB : s0.psiB)[L.t] = psiB;
    vA += v.cA * psiA;
    vB += v.cB * psiB;
  }
}

// BasicUpdator::updateH (src/updator/basic_updator.cpp:23-41): Hx, Hy, Hz
// over [0,nx) x [0,ny) x [0,nz).  Coef is CoefPtrs (dense) or ClassCoef.
// One work-group tile: k-block bx, j-block by, i-chunk bz.  Fused: the CPML
// H corrections of the visited points in the same pass (the H dual
// differences along x, y, z are dual[idx + s_c] - dual[idx]).
template <class T, class Coef, bool Fused = false, bool ZF = false, bool All = false, int TK = kLanesK,
          int TJ = kRowsJ>
__device__ __forceinline__ void sweepH(const Grid& g, const FieldPtrs<T>& F, const Coef& C, int bx, int by, int bz,
                                       int ichunk, const CpmlList<T>& P = CpmlList<T>{},
                                       const ZCpml<T>& Z = ZCpml<T>{}, const StepSlabs<T>& S = StepSlabs<T>{}) {
  const int k = bx * TK + threadIdx.x;
  const int j = by * TJ + threadIdx.y;
  const int i0 = bz * ichunk;
  if (k >= g.nz || j >= g.ny || i0 >= g.nx) return;
  const int i1 = min(i0 + ichunk, g.nx);
  const int64_t pz = g.pz, pl = g.plane;
  int64_t idx = (static_cast<int64_t>(i0) * g.NY + j) * pz + k;
  const T* __restrict__ ex = F.f[0];
  const T* __restrict__ ey = F.f[1];
  const T* __restrict__ ez = F.f[2];
  T* __restrict__ hx = F.f[3];
  T* __restrict__ hy = F.f[4];
  T* __restrict__ hz = F.f[5];
  T ey_i = ey[idx];
  T ez_i = ez[idx];
  const uint32_t mask = Fused ? cpmlChunkMask(P, i0, i1, j, k) : 0u;
  ZLane<T> zl;
  if (ZF) zl = zLane(Z, i0, j, k);
  // All: every slab of S (its j and k masks, the k mask's wavefront union)
  const uint32_t sj = All ? axisMask(S, 1, j) : 0u, sk = All ? axisMask(S, 2, k) : 0u;
  const uint32_t sku = All ? waveUnion(sk) : 0u;
  for (int i = i0; i < i1; ++i, idx += pl) {
    ZVals<T> zv;
    if (ZF) zv = zLoad<T, false>(Z, zl, i, nullptr, nullptr, idx);
    const T ex0 = ex[idx], ex_j = ex[idx + pz], ex_k = ex[idx + 1];
    const T ey_k = ey[idx + 1], ey_n = ey[idx + pl];
    const T ez_j = ez[idx + pz], ez_n = ez[idx + pl];
    T s, a, b;
    C.getOwn(3, idx, s, a, b);
    T vx = s * ldOwn(hx + idx) + a * (ey_k - ey_i) + b * (ez_j - ez_i);
    C.getOwn(4, idx, s, a, b);
    T vy = s * ldOwn(hy + idx) + a * (ez_n - ez_i) + b * (ex_k - ex0);
    C.getOwn(5, idx, s, a, b);
    T vz = s * ldOwn(hz + idx) + a * (ex_j - ex0) + b * (ey_n - ey_i);
    if (Fused && mask)  // A: Hy(x) <- Ez, Hz(y) <- Ex, Hx(z) <- Ey;  B: Hz(x) <- Ey, Hx(y) <- Ez, Hy(z) <- Ex
      cpmlFused(P, mask, i, j, k, vx, vy, vz, ez_n - ez_i, ex_j - ex0, ey_k - ey_i, ey_n - ey_i, ez_j - ez_i,
                ex_k - ex0);
    if (ZF) {  // z slab: Hx <- Ey, Hy <- Ex
      zStore(Z, zl, zv, vx, vy, ey_k - ey_i, ex_k - ex0);
      zl.t += zl.step;
    }
    if (All) {
      const uint32_t si = axisMask(S, 0, i) & sj;
      if (si & sku)
        stepSlabs<T, true>(S, si & sku, si & sk, i, j, k, vx, vy, vz, ez_n - ez_i, ex_j - ex0, ey_k - ey_i,
                           ey_n - ey_i, ez_j - ez_i, ex_k - ex0);
    }
    stOwn(hx + idx, vx);
    stOwn(hy + idx, vy);
    stOwn(hz + idx, vz);
    ey_i = ey_n;
    ez_i = ez_n;
  }
}

template <class T, class Coef, bool Fused, bool ZF, bool All, int TK = kLanesK, int TJ = kRowsJ>
__global__ __launch_bounds__(TK * TJ) void k_update_h(Grid g, FieldPtrs<T> F, Coef C, SweepMap M, int ichunk,
                                             CpmlList<T> P, ZCpml<T> Z, StepSlabs<T> S) {
  int bx, by, bz;
  M.coords(bx, by, bz);
  sweepH<T, Coef, Fused, ZF, All, TK, TJ>(g, F, C, bx, by, bz, ichunk, P, Z, S);
}

// BasicUpdator3D::updateE (src/updator/basic_updator_3d.cpp:38-69):
// Ex over [0,nx) x [1,ny) x [1,nz), Ey over [1,nx) x [0,ny) x [1,nz),
// Ez over [1,nx) x [1,ny) x [0,nz).  Fused: the CPML E corrections of the
// visited points (i < nx, j < ny, k < nz) in the same pass, including the
// face points a Yee range leaves out; the dual differences along x, y, z are
// dual[idx] - dual[idx - s_c].
template <class T, class Coef, bool Fused = false, bool ZF = false, bool All = false, int TK = kLanesK,
          int TJ = kRowsJ>
__device__ __forceinline__ void sweepE(const Grid& g, const FieldPtrs<T>& F, const Coef& C, int bx, int by, int bz,
                                       int ichunk, const CpmlList<T>& P = CpmlList<T>{},
                                       const ZCpml<T>& Z = ZCpml<T>{}, const StepSlabs<T>& S = StepSlabs<T>{}) {
  const int k = bx * TK + threadIdx.x;
  const int j = by * TJ + threadIdx.y;
  const int i0 = bz * ichunk;
  if (k >= g.nz || j >= g.ny || i0 >= g.nx) return;
  const int i1 = min(i0 + ichunk, g.nx);
  const int64_t pz = g.pz, pl = g.plane;
  int64_t idx = (static_cast<int64_t>(i0) * g.NY + j) * pz + k;
  T* __restrict__ ex = F.f[0];
  T* __restrict__ ey = F.f[1];
  T* __restrict__ ez = F.f[2];
  const T* __restrict__ hx = F.f[3];
  const T* __restrict__ hy = F.f[4];
  const T* __restrict__ hz = F.f[5];
  T hz_p = i0 > 0 ? hz[idx - pl] : T(0);
  T hy_p = i0 > 0 ? hy[idx - pl] : T(0);
  const bool jk = j >= 1 && k >= 1;
  const int64_t dk = k >= 1 ? 1 : 0, dj = j >= 1 ? pz : 0;
  const uint32_t mask = Fused ? cpmlChunkMask(P, i0, i1, j, k) : 0u;
  ZLane<T> zl;
  if (ZF) zl = zLane(Z, i0, j, k);
  const uint32_t sj = All ? axisMask(S, 1, j) : 0u, sk = All ? axisMask(S, 2, k) : 0u;
  const uint32_t sku = All ? waveUnion(sk) : 0u;
  for (int i = i0; i < i1; ++i, idx += pl) {
    ZVals<T> zv;
    if (ZF) zv = zLoad<T, true>(Z, zl, i, hy, hx, idx);  // Ex <- Hy, Ey <- Hx (k >= 1 in a z slab)
    const T hx0 = hx[idx], hy0 = hy[idx], hz0 = hz[idx];
    T s, a, b;
    const uint32_t si = All ? axisMask(S, 0, i) & sj : 0u;
    if (All ? !(si & sku) : (!Fused || !mask)) {
      // every lane stores its three values (unchanged on the faces the Yee
      // ranges leave out), so each wavefront writes whole lines: skipping the
      // j = 0 / k = 0 lanes left partial lines and cost 6 % of the FP32 sweep.
      // The loads are unconditional (the k-1 / j-1 neighbours read at the
      // point itself on the k = 0 / j = 0 faces) and the Yee ranges select,
      // so all of a plane's loads are in flight together: loads under the
      // per-component range branches waited one round trip per component.
      T vx = ldOwn(ex + idx), vy = ldOwn(ey + idx), vz = ldOwn(ez + idx);
      const T hy_k = hy[idx - dk], hx_k = hx[idx - dk], hz_j = hz[idx - dj], hx_j = hx[idx - dj];
      T s1, a1, b1, s2, a2, b2;
      C.getOwn(0, idx, s, a, b);
      C.getOwn(1, idx, s1, a1, b1);
      C.getOwn(2, idx, s2, a2, b2);
      const T ux = s * vx + a * (hy0 - hy_k) + b * (hz0 - hz_j);
      const T uy = s1 * vy + a1 * (hz0 - hz_p) + b1 * (hx0 - hx_k);
      const T uz = s2 * vz + a2 * (hx0 - hx_j) + b2 * (hy0 - hy_p);
      vx = jk ? ux : vx;
      vy = (i >= 1 && k >= 1) ? uy : vy;
      vz = (i >= 1 && j >= 1) ? uz : vz;
      if (ZF) {
        zStore(Z, zl, zv, vx, vy, hy0 - zv.dualA, hx0 - zv.dualB);
        zl.t += zl.step;
      }
      stOwn(ex + idx, vx);
      stOwn(ey + idx, vy);
      stOwn(ez + idx, vz);
    } else {
      // curl differences: x uses the carried plane, y and z the j-1 / k-1
      // neighbours where they exist (a slab's axis never starts at 0)
      const T dzHy = k >= 1 ? hy0 - hy[idx - 1] : T(0);
      const T dzHx = k >= 1 ? hx0 - hx[idx - 1] : T(0);
      const T dyHz = j >= 1 ? hz0 - hz[idx - pz] : T(0);
      const T dyHx = j >= 1 ? hx0 - hx[idx - pz] : T(0);
      const T dxHz = hz0 - hz_p, dxHy = hy0 - hy_p;
      T vx = ex[idx], vy = ey[idx], vz = ez[idx];
      if (jk) {
        C.getOwn(0, idx, s, a, b);
        vx = s * vx + a * dzHy + b * dyHz;
      }
      if (i >= 1 && k >= 1) {
        C.getOwn(1, idx, s, a, b);
        vy = s * vy + a * dxHz + b * dzHx;
      }
      if (i >= 1 && j >= 1) {
        C.getOwn(2, idx, s, a, b);
        vz = s * vz + a * dyHx + b * dxHy;
      }
      // A: Ey(x) <- Hz, Ez(y) <- Hx, Ex(z) <- Hy;  B: Ez(x) <- Hy, Ex(y) <- Hz, Ey(z) <- Hx
      if (All)
        stepSlabs<T, true>(S, si & sku, si & sk, i, j, k, vx, vy, vz, dxHz, dyHx, dzHy, dxHy, dyHz, dzHx);
      else
        cpmlFused(P, mask, i, j, k, vx, vy, vz, dxHz, dyHx, dzHy, dxHy, dyHz, dzHx);
      ex[idx] = vx;
      ey[idx] = vy;
      ez[idx] = vz;
    }
    hz_p = hz0;
    hy_p = hy0;
  }
}

template <class T, class Coef, bool Fused, bool ZF, bool All, int TK = kLanesK, int TJ = kRowsJ>
__global__ __launch_bounds__(TK * TJ) void k_update_e(Grid g, FieldPtrs<T> F, Coef C, SweepMap M, int ichunk,
                                             CpmlList<T> P, ZCpml<T> Z, StepSlabs<T> S) {
  int bx, by, bz;
  M.coords(bx, by, bz);
  sweepE<T, Coef, Fused, ZF, All, TK, TJ>(g, F, C, bx, by, bz, ichunk, P, Z, S);
}

// 2D (TE) grid, nz = 1, in a layout of pitch pz = 2 along k (k = 1 holds
// the far face of Ex, Ey and Hz): lanes run along j, the contiguous axis
// once k has two entries, and work-group rows along i.
//
// BasicUpdator::updateH (src/updator/basic_updator.cpp:23-41) over
// [0,nx) x [0,ny) x {0}: all three components with every term, as the
// reference evaluates them (the Ex, Ey and Hz terms are zeros of a TE grid).
// TeZero: Ex, Ey and Hz are known to stay zero (no correction drives them),
// so the Ex / Ey terms are zeros whose addition changes no value and Hz keeps
// its zero: the half-step reads and writes Hx, Hy and Ez only.
template <class T, class Coef, bool TeZero = false>
__global__ __launch_bounds__(256) void k_te_h(Grid g, FieldPtrs<T> F, Coef C) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y * 4 + threadIdx.y;
  if (j >= g.ny || i >= g.nx) return;
  const int64_t pz = g.pz, pl = g.plane;
  const int64_t idx = (static_cast<int64_t>(i) * g.NY + j) * pz;
  const T* __restrict__ ex = F.f[0];
  const T* __restrict__ ey = F.f[1];
  const T* __restrict__ ez = F.f[2];
  T s, a, b;
  if (TeZero) {
    C.get(3, idx, s, a, b);
    const T vx = s * F.f[3][idx] + b * (ez[idx + pz] - ez[idx]);
    T s4, a4, b4;
    C.get(4, idx, s4, a4, b4);
    const T vy = s4 * F.f[4][idx] + a4 * (ez[idx + pl] - ez[idx]);
    F.f[3][idx] = vx;
    F.f[4][idx] = vy;
    return;
  }
  C.get(3, idx, s, a, b);
  const T vx = s * F.f[3][idx] + a * (ey[idx + 1] - ey[idx]) + b * (ez[idx + pz] - ez[idx]);
  C.get(4, idx, s, a, b);
  const T vy = s * F.f[4][idx] + a * (ez[idx + pl] - ez[idx]) + b * (ex[idx + 1] - ex[idx]);
  C.get(5, idx, s, a, b);
  const T vz = s * F.f[5][idx] + a * (ex[idx + pz] - ex[idx]) + b * (ey[idx + pl] - ey[idx]);
  F.f[3][idx] = vx;
  F.f[4][idx] = vy;
  F.f[5][idx] = vz;
}

// BasicUpdatorTE::updateE (src/updator/basic_updator_te.cpp:23-38): Ez over
// [1,nx) x [1,ny) x {0}; Ex and Ey (ranges [1,nz) along k) are not updated.
template <class T, class Coef>
__global__ __launch_bounds__(256) void k_te_e(Grid g, FieldPtrs<T> F, Coef C) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y * 4 + threadIdx.y;
  if (j < 1 || j >= g.ny || i < 1 || i >= g.nx) return;
  const int64_t pz = g.pz, pl = g.plane;
  const int64_t idx = (static_cast<int64_t>(i) * g.NY + j) * pz;
  const T* __restrict__ hx = F.f[3];
  const T* __restrict__ hy = F.f[4];
  T s, a, b;
  C.get(2, idx, s, a, b);
  F.f[2][idx] = s * F.f[2][idx] + a * (hx[idx] - hx[idx - pz]) + b * (hy[idx] - hy[idx - pl]);
}

// BasicUpdatorTEM::updateE (src/updator/basic_updator.cpp:90-109) on a 1D
// grid (nx = ny = 1): Ex(0,0,k) for k in [1,nz), its Hz term passed as
// zeros, eNext(cexe, ex, cexhy, hy(k), hy(k-1), 0, 0, 0).
template <class T, class Coef>
__global__ __launch_bounds__(256) void k_tem_e(Grid g, FieldPtrs<T> F, Coef C) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < 1 || k >= g.nz) return;
  const int64_t idx = k;
  const T* __restrict__ hy = F.f[4];
  T s, a, b;
  C.get(0, idx, s, a, b);
  F.f[0][idx] = s * F.f[0][idx] + a * (hy[idx] - hy[idx - 1]) + T(0) * (T(0) - T(0));
}

// Single-cell time series (FieldTimeMonitor sampling): one launch per step
// for all of them; probe p writes out[p * stride + t].
struct Probe {
  int64_t idx;
  int32_t field;
  int32_t pad;
};

template <class T>
__global__ __launch_bounds__(64) void k_probes(FieldPtrs<T> F, const Probe* __restrict__ p, int n, double* out,
                                               int64_t stride, int64_t t) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  out[q * stride + t] = static_cast<double>(F.f[p[q].field][p[q].idx]);
}

// Weighted-sum recordings (VoltageMonitor::update, voltage_monitor.cpp:
// 191-227; CurrentMonitor::update, current_monitor.cpp:272-385): monitor m
// sums, in term order, w * F[field][idx] into one of two partial sums and
// records scale * (g0 + g1) for step t, in FP64.  One thread per monitor
// keeps the reference's summation order.
struct LineTerm {
  int64_t idx;
  double w;
  int32_t field;
  int32_t group;
};

template <class T>
__device__ __forceinline__ void lineSum(const FieldPtrs<T>& F, const LineTerm* __restrict__ terms,
                                        const int64_t* __restrict__ off, const double* __restrict__ scale, int m,
                                        double* out, int64_t stride, int64_t t) {
  double g0 = 0.0, g1 = 0.0;
  for (int64_t q = off[m]; q < off[m + 1]; ++q) {
    const LineTerm L = terms[q];
    const double v = static_cast<double>(F.f[L.field][L.idx]) * L.w;
    if (L.group == 0)
      g0 = g0 + v;
    else
      g1 = g1 + v;
  }
  out[m * stride + t] = scale[m] * (g0 + g1);
}

template <class T>
__global__ __launch_bounds__(64) void k_lines(FieldPtrs<T> F, const LineTerm* __restrict__ terms,
                                              const int64_t* __restrict__ off, const double* __restrict__ scale,
                                              int n, double* out, int64_t stride, int64_t t) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n || t >= stride) return;
  lineSum(F, terms, off, scale, m, out, stride, t);
}

// correctPML (src/include/boundary/pml_corrector.h:14-52) for both
// components of one slab and one half-step.  IsE: psi = b psi + a (p - q)
// with q one plane below along the main axis; H: psi = b psi + a (q - p)
// with q one plane above.  field += c_psi * psi.
template <class T, bool IsE>
__device__ __forceinline__ void cpmlPoint(const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& s, int p0, int p1,
                                          int p2, int64_t t) {
  const int c = s.axis, a = c == 2 ? 0 : c + 1, b = c == 0 ? 2 : c - 1;
  const int layer = pick3(c, p0, p1, p2) - s.start;
  const T ca = s.ca[layer], cb = s.cb[layer];
  const int64_t idx = (static_cast<int64_t>(p0) * g.NY + p1) * g.pz + p2;
  const int64_t q = IsE ? idx - s.stride_c : idx + s.stride_c;
  const int pa = pick3(a, p0, p1, p2), pb = pick3(b, p0, p1, p2);
  if (pa < s.extA[0] && pb < s.extA[1]) {
    const T* dual = F.f[s.dualB];
    T psi = ldOwn(s.psiA + t);
    psi = IsE ? cb * psi + ca * (dual[idx] - dual[q]) : cb * psi + ca * (dual[q] - dual[idx]);
    stOwn(s.psiA + t, psi);
    T* f = F.f[s.fieldA];
    f[idx] += ldOwn(s.cpsiA + t) * psi;
  }
  if (pa < s.extB[0] && pb < s.extB[1]) {
    const T* dual = F.f[s.dualA];
    T psi = ldOwn(s.psiB + t);
    psi = IsE ? cb * psi + ca * (dual[idx] - dual[q]) : cb * psi + ca * (dual[q] - dual[idx]);
    stOwn(s.psiB + t, psi);
    T* f = F.f[s.fieldB];
    f[idx] += ldOwn(s.cpsiB + t) * psi;
  }
}

template <class T, bool IsE>
__device__ __forceinline__ void cpmlCell(const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& s, int64_t t) {
  const int64_t vol = static_cast<int64_t>(s.ext[0]) * s.ext[1] * s.ext[2];
  if (t >= vol) return;
  const int w = static_cast<int>(t % s.ext[2]);
  const int v = static_cast<int>((t / s.ext[2]) % s.ext[1]);
  const int u = static_cast<int>(t / (static_cast<int64_t>(s.ext[2]) * s.ext[1]));
  cpmlPoint<T, IsE>(g, F, s, s.lo[0] + u, s.lo[1] + v, s.lo[2] + w, t);
}

template <class T, bool IsE>
__global__ __launch_bounds__(256) void k_cpml(Grid g, FieldPtrs<T> F, SlabArgs<T> s) {
  cpmlCell<T, IsE>(g, F, s, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
}

// Two slabs that share no point (the two faces of one axis) in one launch:
// work-groups [0, nb0) take s0, the rest s1.  Each point still receives its
// slabs' corrections in boundary order, since no point belongs to both.
template <class T, bool IsE>
__global__ __launch_bounds__(256) void k_cpml2(Grid g, FieldPtrs<T> F, SlabArgs<T> s0, SlabArgs<T> s1,
                                               unsigned nb0) {
  if (blockIdx.x < nb0)
    cpmlCell<T, IsE>(g, F, s0, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
  else
    cpmlCell<T, IsE>(g, F, s1, static_cast<int64_t>(blockIdx.x - nb0) * blockDim.x + threadIdx.x);
}

// The points of a slab box that the fused sweeps do not visit: the far grid
// planes i = nx, j = ny, k = nz.  Point q of the union of the three plane
// pieces (X: i = nx; Y: j = ny, i < nx; Z: k = nz, i < nx, j < ny), or false.
struct FacePieces {
  int n0, n1;        // extents of the slab box clipped below nx / ny
  int64_t nX, nY, nZ;
};
template <class T>
__device__ __host__ __forceinline__ FacePieces facePieces(const Grid& g, const SlabArgs<T>& s) {
  FacePieces f;
  const int hi0 = s.lo[0] + s.ext[0], hi1 = s.lo[1] + s.ext[1], hi2 = s.lo[2] + s.ext[2];
  f.n0 = max(0, min(hi0, g.nx) - s.lo[0]);
  f.n1 = max(0, min(hi1, g.ny) - s.lo[1]);
  f.nX = (s.lo[0] <= g.nx && g.nx < hi0) ? static_cast<int64_t>(s.ext[1]) * s.ext[2] : 0;
  f.nY = (s.lo[1] <= g.ny && g.ny < hi1) ? static_cast<int64_t>(f.n0) * s.ext[2] : 0;
  f.nZ = (s.lo[2] <= g.nz && g.nz < hi2) ? static_cast<int64_t>(f.n0) * f.n1 : 0;
  return f;
}

template <class T, bool IsE>
__device__ __forceinline__ void cpmlFacePoint(const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& s,
                                              int64_t q) {
  const FacePieces f = facePieces(g, s);
  int p0, p1, p2;
  if (q < f.nX) {
    p0 = g.nx;
    p1 = s.lo[1] + static_cast<int>(q / s.ext[2]);
    p2 = s.lo[2] + static_cast<int>(q % s.ext[2]);
  } else if (q < f.nX + f.nY) {
    q -= f.nX;
    p0 = s.lo[0] + static_cast<int>(q / s.ext[2]);
    p1 = g.ny;
    p2 = s.lo[2] + static_cast<int>(q % s.ext[2]);
  } else if (q < f.nX + f.nY + f.nZ) {
    q -= f.nX + f.nY;
    p0 = s.lo[0] + static_cast<int>(q / f.n1);
    p1 = s.lo[1] + static_cast<int>(q % f.n1);
    p2 = g.nz;
  } else {
    return;
  }
  const int64_t t = (static_cast<int64_t>(p0 - s.lo[0]) * s.ext[1] + (p1 - s.lo[1])) * s.ext[2] + (p2 - s.lo[2]);
  cpmlPoint<T, IsE>(g, F, s, p0, p1, p2, t);
}

template <class T, bool IsE>
__global__ __launch_bounds__(256) void k_cpml_faces(Grid g, FieldPtrs<T> F, SlabArgs<T> s) {
  cpmlFacePoint<T, IsE>(g, F, s, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
}

// The points of a z slab the sweeps leave: its frame (ZFrame) and the far
// planes i = nx, j = ny, at the slab's place in boundary order.
template <class T, bool IsE>
__global__ __launch_bounds__(256) void k_cpml_zrest(Grid g, FieldPtrs<T> F, SlabArgs<T> s, ZFrame fr) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t vol = static_cast<int64_t>(s.ext[0]) * s.ext[1] * s.ext[2];
  if (t >= vol) return;
  const int w = static_cast<int>(t % s.ext[2]);
  const int v = static_cast<int>((t / s.ext[2]) % s.ext[1]);
  const int u = static_cast<int>(t / (static_cast<int64_t>(s.ext[2]) * s.ext[1]));
  const int p0 = s.lo[0] + u, p1 = s.lo[1] + v;
  if (p0 < g.nx && p1 < g.ny && !fr.inX(p0) && !fr.inY(p1)) return;
  cpmlPoint<T, IsE>(g, F, s, p0, p1, s.lo[2] + w, t);
}

enum AdeScheme { kAdeDrude = 0, kAdeDebye = 1, kAdeMLorentz = 2 };

// One dispersive E update after the plain curl update `en` of the old value
// `cur`: TemplateADEUpdateScheme::updateE (template_ade_update_scheme.h:
// 62-82) with DebyeADEUpdator / DrudeADEUpdator (debye_ade_updator.cpp:
// 31-49, drude_ade_updator.cpp:21-39), or MLorentzUpdator::updateE / updateJ
// (m_lor_ade_updator.cpp:66-135).  E = cee_prev * E_prev + en + cej * sum_J,
// the pole sum taken before the currents advance.
template <class T, int S>
__device__ __forceinline__ T adeUpdate(const AdeArgs<T>& A, int m, int q, int64_t idx, T cur, T en) {
  const T* c = A.tab + static_cast<int64_t>(m) * A.stride;
  T jsum = T(0);
  for (int p = 0; p < A.poles; ++p) {
    const T* cp = c + 2 + 7 * p;
    const int64_t o = p * A.total + idx;
    if (S == kAdeMLorentz)
      jsum += cp[5] * A.j[q][o] + cp[6] * A.jp[q][o];
    else
      jsum += cp[5] * A.j[q][o];
  }
  const T eprev = S == kAdeMLorentz ? A.eprev[q][idx] : T(0);
  const T ceep = S == kAdeMLorentz ? c[1] : T(0);
  const T enew = ceep * eprev + en + c[0] * jsum;
  for (int p = 0; p < A.poles; ++p) {
    const T* cp = c + 2 + 7 * p;
    const int64_t o = p * A.total + idx;
    const T jv = A.j[q][o];
    if (S == kAdeDebye) {
      A.j[q][o] = cp[3] * jv + cp[1] * (enew - cur);
    } else if (S == kAdeDrude) {
      A.j[q][o] = cp[3] * jv + cp[1] * (enew + cur);
    } else {
      const T jn = cp[0] * enew + cp[1] * cur + cp[2] * eprev + cp[3] * jv + cp[4] * A.jp[q][o];
      A.jp[q][o] = jv;
      A.j[q][o] = jn;
    }
  }
  if (S == kAdeMLorentz) A.eprev[q][idx] = cur;
  return enew;
}

// adeUpdate for the three components of one cell with every load issued
// first: the cell's pole currents (and, modified Lorentz, previous currents
// and previous E) of all components, then the arithmetic of adeUpdate in its
// order, component by component.  `r[q]`: component q is inside its Yee range
// (outside it the forward update leaves E and the currents alone).  Up to
// kAdeFastPoles poles (a uniform bound: the unused slots are never loaded).
constexpr int kAdeFastPoles = 4;
template <class T, int S>
__device__ __forceinline__ void adeCell(const AdeArgs<T>& A, int m, int64_t idx, const bool r[3], const T cur[3],
                                        const T en[3], T out[3]) {
  constexpr int P = kAdeFastPoles;
  const int np = A.poles;
  const T* c = A.tab + static_cast<int64_t>(m) * A.stride;
  T jv[3][P], jpv[3][P], ep[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    ep[q] = S == kAdeMLorentz ? A.eprev[q][idx] : T(0);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      jv[q][p] = T(0);
      jpv[q][p] = T(0);
      if (p < np) {
        jv[q][p] = A.j[q][p * A.total + idx];
        if (S == kAdeMLorentz) jpv[q][p] = A.jp[q][p * A.total + idx];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    out[q] = cur[q];
    if (!r[q]) continue;
    T jsum = T(0);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (p >= np) break;
      const T* cp = c + 2 + 7 * p;
      if (S == kAdeMLorentz)
        jsum += cp[5] * jv[q][p] + cp[6] * jpv[q][p];
      else
        jsum += cp[5] * jv[q][p];
    }
    const T eprev = S == kAdeMLorentz ? ep[q] : T(0);
    const T ceep = S == kAdeMLorentz ? c[1] : T(0);
    const T enew = ceep * eprev + en[q] + c[0] * jsum;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (p >= np) break;
      const T* cp = c + 2 + 7 * p;
      const int64_t o = p * A.total + idx;
      const T j0 = jv[q][p];
      if (S == kAdeDebye) {
        A.j[q][o] = cp[3] * j0 + cp[1] * (enew - cur[q]);
      } else if (S == kAdeDrude) {
        A.j[q][o] = cp[3] * j0 + cp[1] * (enew + cur[q]);
      } else {
        const T jn = cp[0] * enew + cp[1] * cur[q] + cp[2] * eprev + cp[3] * j0 + cp[4] * jpv[q][p];
        A.jp[q][o] = j0;
        A.j[q][o] = jn;
      }
    }
    if (S == kAdeMLorentz) A.eprev[q][idx] = cur[q];
    out[q] = enew;
  }
}

// The E sweep of k_update_e with the dispersive update on cells whose
// material id is set (the reference runs its ADE updator on every cell; on
// a cell without dispersion its coefficients are zero and the result is the
// plain update).  As in sweepE the curl terms' loads are unconditional (the
// k-1 / j-1 neighbours read at the point itself on the k = 0 / j = 0 faces)
// and the Yee ranges select, so a plane's loads are in flight together.
template <class T, class Coef, int S, int TK = kLanesK, int TJ = kRowsJ>
__device__ __forceinline__ void sweepEAde(const Grid& g, const FieldPtrs<T>& F, const Coef& C, const AdeArgs<T>& A,
                                          int bx, int by, int bz, int ichunk) {
  const int k = bx * TK + threadIdx.x;
  const int j = by * TJ + threadIdx.y;
  const int i0 = bz * ichunk;
  if (k >= g.nz || j >= g.ny || i0 >= g.nx) return;
  const int i1 = min(i0 + ichunk, g.nx);
  const int64_t pz = g.pz, pl = g.plane;
  int64_t idx = (static_cast<int64_t>(i0) * g.NY + j) * pz + k;
  T* __restrict__ ex = F.f[0];
  T* __restrict__ ey = F.f[1];
  T* __restrict__ ez = F.f[2];
  const T* __restrict__ hx = F.f[3];
  const T* __restrict__ hy = F.f[4];
  const T* __restrict__ hz = F.f[5];
  T hz_p = i0 > 0 ? hz[idx - pl] : T(0);
  T hy_p = i0 > 0 ? hy[idx - pl] : T(0);
  const bool jk = j >= 1 && k >= 1;
  const int64_t dk = k >= 1 ? 1 : 0, dj = j >= 1 ? pz : 0;
  const bool fast = A.poles <= kAdeFastPoles;
  for (int i = i0; i < i1; ++i, idx += pl) {
    const T hx0 = hx[idx], hy0 = hy[idx], hz0 = hz[idx];
    const int m = ldOwn(A.mat + idx);
    // whole-line stores as in sweepE (face values rewritten unchanged)
    const T cur[3] = {ldOwn(ex + idx), ldOwn(ey + idx), ldOwn(ez + idx)};
    const T hy_k = hy[idx - dk], hx_k = hx[idx - dk], hz_j = hz[idx - dj], hx_j = hx[idx - dj];
    T s0, a0, b0, s1, a1, b1, s2, a2, b2;
    C.getOwn(0, idx, s0, a0, b0);
    C.getOwn(1, idx, s1, a1, b1);
    C.getOwn(2, idx, s2, a2, b2);
    const T en[3] = {s0 * cur[0] + a0 * (hy0 - hy_k) + b0 * (hz0 - hz_j),
                     s1 * cur[1] + a1 * (hz0 - hz_p) + b1 * (hx0 - hx_k),
                     s2 * cur[2] + a2 * (hx0 - hx_j) + b2 * (hy0 - hy_p)};
    const bool r[3] = {jk, i >= 1 && k >= 1, i >= 1 && j >= 1};
    T v[3];
    if (m < 0) {
      for (int q = 0; q < 3; ++q) v[q] = r[q] ? en[q] : cur[q];
    } else if (fast) {
      adeCell<T, S>(A, m, idx, r, cur, en, v);
    } else {
      for (int q = 0; q < 3; ++q) v[q] = r[q] ? adeUpdate<T, S>(A, m, q, idx, cur[q], en[q]) : cur[q];
    }
    stOwn(ex + idx, v[0]);
    stOwn(ey + idx, v[1]);
    stOwn(ez + idx, v[2]);
    hz_p = hz0;
    hy_p = hy0;
  }
}

template <class T, class Coef, int S, int TK = kLanesK, int TJ = kRowsJ>
__global__ __launch_bounds__(TK * TJ) void k_update_e_ade(Grid g, FieldPtrs<T> F, Coef C, AdeArgs<T> A, SweepMap M,
                                                         int ichunk) {
  int bx, by, bz;
  M.coords(bx, by, bz);
  sweepEAde<T, Coef, S, TK, TJ>(g, F, C, A, bx, by, bz, ichunk);
}

// interpolate::interpolateSurfaceCenter (interpolate_scheme.h:18-138): field
// component f (0..5) at the centre of the face normal to axis c at the
// padded index idx.
template <class T>
__device__ __forceinline__ T faceCentre(const Grid& g, const FieldPtrs<T>& F, int c, int f, int64_t idx) {
  const T* a = F.f[f];
  const int64_t sx = g.plane, sy = g.pz, sz = 1;
  const T h = T(0.5), q = T(0.25);
  if (c == 0) {
    if (f == 5) return q * (a[idx + sz] + a[idx] + a[idx - sx + sz] + a[idx - sx]);
    if (f == 4) return q * (a[idx + sy] + a[idx] + a[idx - sx + sy] + a[idx - sx]);
    if (f == 2) return h * (a[idx + sy] + a[idx]);
    return h * (a[idx + sz] + a[idx]);  // f == 1
  }
  if (c == 1) {
    if (f == 5) return q * (a[idx + sz] + a[idx] + a[idx - sy + sz] + a[idx - sy]);
    if (f == 3) return q * (a[idx + sx] + a[idx] + a[idx + sx - sy] + a[idx - sy]);
    if (f == 2) return h * (a[idx + sx] + a[idx]);
    return h * (a[idx + sz] + a[idx]);  // f == 0
  }
  if (f == 3) return q * (a[idx + sx] + a[idx] + a[idx - sz] + a[idx + sx - sz]);
  if (f == 4) return q * (a[idx] + a[idx + sy] + a[idx - sz] + a[idx + sy - sz]);
  if (f == 0) return h * (a[idx + sy] + a[idx]);
  return h * (a[idx + sx] + a[idx]);  // f == 1
}

// NF2FF running DFT, one thread per surface cell: FDPlaneData::calculateJ
// and calculateM (nffft_fd_data_update.cpp:22-122) for every frequency,
// each accumulator advanced `repeat` times (nf reproduces
// NFFFTFrequencyDomain::update, which repeats the whole update once per
// frequency).  FP64 accumulators.
template <class T>
__device__ __forceinline__ void nf2ffCell(const Grid& g, const FieldPtrs<T>& F, const Nf2ffArgs& a, int64_t cell,
                                          int64_t t) {
  if (cell >= a.off[6]) return;
  int d = 0;
  while (cell >= a.off[d + 1]) ++d;
  const int64_t q = cell - a.off[d];
  const int e1 = a.ext[d][1], e2 = a.ext[d][2];
  const int k = a.lo[d][2] + static_cast<int>(q % e2);
  const int j = a.lo[d][1] + static_cast<int>((q / e2) % e1);
  const int i = a.lo[d][0] + static_cast<int>(q / (static_cast<int64_t>(e2) * e1));
  const int64_t idx = (static_cast<int64_t>(i) * g.NY + j) * g.pz + k;
  const int c = d / 2, A = (c + 1) % 3, B = (c + 2) % 3;
  const bool pos = d & 1;
  // coff_ja / coff_jb (calculateJ) and coff_ma / coff_mb (calculateM)
  const double xja = (pos ? -1.0 : 1.0) * static_cast<double>(faceCentre(g, F, c, 3 + B, idx));
  const double xjb = (pos ? 1.0 : -1.0) * static_cast<double>(faceCentre(g, F, c, 3 + A, idx));
  const double xma = (pos ? 1.0 : -1.0) * static_cast<double>(faceCentre(g, F, c, B, idx));
  const double xmb = (pos ? -1.0 : 1.0) * static_cast<double>(faceCentre(g, F, c, A, idx));
  const int64_t cells = a.off[6];
  for (int f = 0; f < a.nf; ++f) {
    const double2 th = a.th[f * a.nt + t], te = a.te[f * a.nt + t];
    const double2 ija = {xja * th.x, xja * th.y}, ijb = {xjb * th.x, xjb * th.y};
    const double2 ima = {xma * te.x, xma * te.y}, imb = {xmb * te.x, xmb * te.y};
    double2* p = a.acc + static_cast<int64_t>(f) * 4 * cells + cell;
    double2 ja = p[0], jb = p[cells], ma = p[2 * cells], mb = p[3 * cells];
    for (int r = 0; r < a.repeat; ++r) {
      ja.x = ja.x + ija.x;
      ja.y = ja.y + ija.y;
      jb.x = jb.x + ijb.x;
      jb.y = jb.y + ijb.y;
      ma.x = ma.x + ima.x;
      ma.y = ma.y + ima.y;
      mb.x = mb.x + imb.x;
      mb.y = mb.y + imb.y;
    }
    p[0] = ja;
    p[cells] = jb;
    p[2 * cells] = ma;
    p[3 * cells] = mb;
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_nf2ff(Grid g, FieldPtrs<T> F, Nf2ffArgs a, int64_t t) {
  nf2ffCell(g, F, a, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x, t);
}

// Far-field potentials of one frequency (FDPlaneData::calculatePotential,
// nffft_fd_data_process.cpp:44-160, rVector / ds :196-225): work-group n
// reduces over every surface cell
//   P(n) += (cur_a t_a(n) + cur_b t_b(n)) exp(j k r.rhat(n)) ds
// for A (ja, jb) and F (ma, mb), each with the theta and the phi rows of the
// Cartesian-to-spherical transform (transform.h:24-58).
struct FarArgs {
  const double* e_node[3];  // per axis, N + 1 nodes
  const double* h_node[3];  // per axis, N cell centres
  const double* e_size[3];  // per axis, N cell sizes
  const double* theta;
  const double* phi;
  double origin[3];
  double k;                 // 2 pi f / c
};

// Work-groups (direction, chunk): chunk c of `chunks` sums its share of the
// surface cells for direction n into part[(n * chunks + c) * 8 + w];
// k_far_reduce then adds the chunks in order (a fixed summation order, so
// repeated calls give the same bits).  Several chunks per direction keep the
// GPU filled when there are few directions.
// Templated on the field type only so that each precision's unit holds its own copy.
template <class T>
__global__ __launch_bounds__(256) void k_far_field(Nf2ffArgs a, int f, FarArgs fa, int chunks, double* part) {
  __shared__ double red[8][256];
  const int n = blockIdx.x;
  const int chunk = blockIdx.y;
  const double th = fa.theta[n], ph = fa.phi[n];
  const double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
  const double stcp = st * cp, stsp = st * sp;
  const double tt0 = ct * cp, tt1 = ct * sp, tt2 = -1 * st;  // theta row
  const double tp0 = -1 * sp, tp1 = cp;                      // phi row (z entry 0)
  const int64_t cells = a.off[6];
  const double2* acc = a.acc + static_cast<int64_t>(f) * 4 * cells;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // A_theta, A_phi, F_theta, F_phi (re, im)
  const int64_t per = (cells + chunks - 1) / chunks;
  const int64_t q0 = static_cast<int64_t>(chunk) * per, q1 = min(cells, q0 + per);
  for (int64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    int d = 0;
    while (q >= a.off[d + 1]) ++d;
    const int64_t l = q - a.off[d];
    const int e1 = a.ext[d][1], e2 = a.ext[d][2];
    const int kk = a.lo[d][2] + static_cast<int>(l % e2);
    const int j = a.lo[d][1] + static_cast<int>((l / e2) % e1);
    const int i = a.lo[d][0] + static_cast<int>(l / (static_cast<int64_t>(e2) * e1));
    const int c = d / 2;
    double r0, r1, r2, ds;
    if (c == 0) {
      r0 = fa.e_node[0][i];
      r1 = fa.h_node[1][j];
      r2 = fa.h_node[2][kk];
      ds = fa.e_size[1][j] * fa.e_size[2][kk];
    } else if (c == 1) {
      r0 = fa.h_node[0][i];
      r1 = fa.e_node[1][j];
      r2 = fa.h_node[2][kk];
      ds = fa.e_size[2][kk] * fa.e_size[0][i];
    } else {
      r0 = fa.h_node[0][i];
      r1 = fa.h_node[1][j];
      r2 = fa.e_node[2][kk];
      ds = fa.e_size[0][i] * fa.e_size[1][j];
    }
    const double rx = r0 - fa.origin[0], ry = r1 - fa.origin[1], rz = r2 - fa.origin[2];
    const double v = rx * stcp + ry * stsp + rz * ct;
    double sn, cs;
    sincos(fa.k * v, &sn, &cs);
    const int A = c == 2 ? 0 : c + 1, B = c == 0 ? 2 : c - 1;
    const double tta = A == 0 ? tt0 : (A == 1 ? tt1 : tt2), ttb = B == 0 ? tt0 : (B == 1 ? tt1 : tt2);
    const double tpa = A == 0 ? tp0 : (A == 1 ? tp1 : 0.0), tpb = B == 0 ? tp0 : (B == 1 ? tp1 : 0.0);
    const double2 ja = acc[q], jb = acc[cells + q], ma = acc[2 * cells + q], mb = acc[3 * cells + q];
    // (cur_a t_a + cur_b t_b) * phase * ds, phase = (cs, sn)
    auto add = [&](double* out2, const double2 ca, const double2 cb, double ta, double tb) {
      const double re = ca.x * ta + cb.x * tb, im = ca.y * ta + cb.y * tb;
      out2[0] += (re * cs - im * sn) * ds;
      out2[1] += (re * sn + im * cs) * ds;
    };
    add(s + 0, ja, jb, tta, ttb);
    add(s + 2, ja, jb, tpa, tpb);
    add(s + 4, ma, mb, tta, ttb);
    add(s + 6, ma, mb, tpa, tpb);
  }
  for (int w = 0; w < 8; ++w) red[w][threadIdx.x] = s[w];
  __syncthreads();
  for (int half = 128; half > 0; half >>= 1) {
    if (threadIdx.x < half)
      for (int w = 0; w < 8; ++w) red[w][threadIdx.x] += red[w][threadIdx.x + half];
    __syncthreads();
  }
  if (threadIdx.x < 8) part[(static_cast<int64_t>(n) * chunks + chunk) * 8 + threadIdx.x] = red[threadIdx.x][0];
}

template <class T>
__global__ __launch_bounds__(256) void k_far_reduce(int na, int chunks, const double* __restrict__ part,
                                                    double* __restrict__ out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(na) * 8) return;
  const int64_t n = t / 8, w = t % 8;
  double v = 0;
  for (int c = 0; c < chunks; ++c) v += part[(n * chunks + c) * 8 + w];
  out[t] = v;
}

// Lumped-element and TFSF corrections: one thread per corrected cell applies
// that cell's records in the reference's corrector order.  `state` holds the
// inductor currents (one slot per inductor edge).
template <class T>
__device__ __forceinline__ void recordSeg(const FieldPtrs<T>& F, const Record* __restrict__ rec,
                                          const int64_t* __restrict__ seg, int64_t nseg,
                                          const double* __restrict__ pool, T* __restrict__ state, int64_t s,
                                          int64_t t) {
  if (s >= nseg) return;
  const int64_t r0 = seg[s], r1 = seg[s + 1];
  T* f = F.f[rec[r0].field];
  const int64_t target = rec[r0].target;
  double acc = static_cast<double>(f[target]);
  for (int64_t r = r0; r < r1; ++r) {
    const Record R = rec[r];
    if (R.kind == 0) {
      // TFSFCorrector::exInc..hzInc (tfsf_corrector.cpp:16-122) times the
      // signed coefficient (tfsf_corrector.h:216-225)
      const double* row = pool + R.pool + t * R.row;
      const double inc = R.tr * (R.omw * row[0] + R.w * row[1]);
      acc = acc + inc * R.coef;
    } else if (R.kind == 1) {
      // Voltage/CurrentSourceCorrector: E += coff * w[t]
      acc = acc + R.coef * pool[R.pool + t];
    } else {
      // InductorCorrector::correctE: E += cecjc * J, then J += cjcec * E
      const double j = static_cast<double>(state[R.pool]);
      acc = acc + R.coef * j;
      if (sizeof(T) == sizeof(float)) acc = static_cast<double>(static_cast<T>(acc));
      state[R.pool] = static_cast<T>(j + R.w * acc);
    }
    if (sizeof(T) == sizeof(float)) acc = static_cast<double>(static_cast<T>(acc));
  }
  f[target] = static_cast<T>(acc);
}

template <class T>
__global__ __launch_bounds__(256) void k_records(FieldPtrs<T> F, const Record* __restrict__ rec,
                                                 const int64_t* __restrict__ seg, int64_t nseg,
                                                 const double* __restrict__ pool, T* __restrict__ state,
                                                 int64_t t) {
  recordSeg(F, rec, seg, nseg, pool, state, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x, t);
}

// Recording: copy the box [o, e) of one field into a double frame.
template <class T>
__global__ __launch_bounds__(256) void k_gather(Grid g, FieldPtrs<T> F, GatherArgs a, double* out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= a.frame_elems) return;
  const int n2 = a.e[2] - a.o[2], n1 = a.e[1] - a.o[1];
  const int w = static_cast<int>(t % n2);
  const int v = static_cast<int>((t / n2) % n1);
  const int u = static_cast<int>(t / (static_cast<int64_t>(n2) * n1));
  const int64_t idx = (static_cast<int64_t>(a.o[0] + u) * g.NY + (a.o[1] + v)) * g.pz + (a.o[2] + w);
  out[t] = static_cast<double>(F.f[a.field][idx]);
}

// Dense copy between the reference extent (k fastest, n0 x n1 x n2 doubles)
// and the padded layout, converting precision.
template <class T>
__global__ __launch_bounds__(256) void k_unpack(Grid g, const double* __restrict__ src, T* dst, int n0, int n1,
                                                int n2) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t vol = static_cast<int64_t>(n0) * n1 * n2;
  if (t >= vol) return;
  const int w = static_cast<int>(t % n2);
  const int v = static_cast<int>((t / n2) % n1);
  const int u = static_cast<int>(t / (static_cast<int64_t>(n2) * n1));
  dst[(static_cast<int64_t>(u) * g.NY + v) * g.pz + w] = static_cast<T>(src[t]);
}

template <class T>
__global__ __launch_bounds__(256) void k_pack(Grid g, const T* __restrict__ src, double* dst, int n0, int n1,
                                              int n2) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t vol = static_cast<int64_t>(n0) * n1 * n2;
  if (t >= vol) return;
  const int w = static_cast<int>(t % n2);
  const int v = static_cast<int>((t / n2) % n1);
  const int u = static_cast<int>(t / (static_cast<int64_t>(n2) * n1));
  dst[t] = static_cast<double>(src[(static_cast<int64_t>(u) * g.NY + v) * g.pz + w]);
}

// Time-domain NF2FF, phase 1, one thread per (group, cell):
// TDPlaneData::update (nffft_td_data.h:159-231) up to the potential adds --
// the face-centre tangential values, their change since the previous step,
// the bin nn = floor(s + t) with weight f = s + t - nn, and the four
// contributions coeff (1 - f) ds delta / coeff f ds delta to bins nn, nn + 1.
template <class T>
__device__ __forceinline__ void tdCell(const Grid& g, const FieldPtrs<T>& F, const TdNf2ffArgs& a, int64_t cell,
                                       int64_t t) {
  if (cell >= a.off[12]) return;
  int gr = 0;
  while (cell >= a.off[gr + 1]) ++gr;
  const int d = gr >> 1;
  const bool is_e = (gr & 1) == 0;
  const int64_t q = cell - a.off[gr];
  const int e1 = a.ext[d][1], e2 = a.ext[d][2];
  const int k = a.lo[d][2] + static_cast<int>(q % e2);
  const int j = a.lo[d][1] + static_cast<int>((q / e2) % e1);
  const int i = a.lo[d][0] + static_cast<int>(q / (static_cast<int64_t>(e2) * e1));
  const int64_t idx = (static_cast<int64_t>(i) * g.NY + j) * g.pz + k;
  const int c = d / 2, A = (c + 1) % 3, B = (c + 2) % 3;
  const bool neg = (d & 1) == 0;
  double coeff_a = is_e ? 1.0 : -1.0;
  if (neg) coeff_a *= -1.0;
  double coeff_b = is_e ? -1.0 : 1.0;
  if (neg) coeff_b *= -1.0;
  const int base = is_e ? 0 : 3;
  const double a_avg = static_cast<double>(faceCentre(g, F, c, base + A, idx));
  const double b_avg = static_cast<double>(faceCentre(g, F, c, base + B, idx));
  const double delta_a = a_avg - a.prev_a[cell];
  const double delta_b = b_avg - a.prev_b[cell];
  a.prev_a[cell] = a_avg;
  a.prev_b[cell] = b_avg;
  const double v = a.s[cell] + static_cast<double>(t);
  const int64_t nn = static_cast<int64_t>(floor(v));
  const double coeff_f = v - static_cast<double>(nn);
  const double ds = a.ds[cell];
  const double ca1 = coeff_a * (1 - coeff_f) * ds * delta_b;
  const double ca2 = coeff_a * coeff_f * ds * delta_b;
  const double cb1 = coeff_b * (1 - coeff_f) * ds * delta_a;
  const double cb2 = coeff_b * coeff_f * ds * delta_a;
  if (a.fixed) {
    const int64_t p1 = a.slot[2 * cell], p2 = a.slot[2 * cell + 1];
    a.va[p1] = ca1;
    a.va[p2] = ca2;
    a.vb[p1] = cb1;
    a.vb[p2] = cb2;
  } else {
    a.nn[cell] = nn;
    a.ca1[cell] = ca1;
    a.ca2[cell] = ca2;
    a.cb1[cell] = cb1;
    a.cb2[cell] = cb2;
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_td_cells(Grid g, FieldPtrs<T> F, TdNf2ffArgs a, int64_t t) {
  tdCell(g, F, a, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x, t);
}

// Phase 2, one thread per (group, bin of this step's window): bin n adds the
// contributions of the cells with nn == n (first) or nn + 1 == n (second),
// in cell order, to the running potential sums -- the reference's order.
__device__ __forceinline__ void tdBin(const TdNf2ffArgs& a, int64_t id, int64_t t) {
  if (id >= a.bin_off[12]) return;
  int gr = 0;
  while (id >= a.bin_off[gr + 1]) ++gr;
  const int64_t local = id - a.bin_off[gr];
  const int64_t n = t + a.omin[gr] + local;
  if (n < 0 || n >= a.num) return;
  if (a.fixed) {  // this bin's contributions, already in order
    const int d = gr >> 1;
    double* pa = a.pot + (static_cast<int64_t>(d) * 4 + ((gr & 1) ? 2 : 0)) * a.num;
    double* pb = pa + a.num;
    double va = pa[n], vb = pb[n];
    const int64_t q0 = a.lstart[a.loff[gr] + local], q1 = a.lstart[a.loff[gr] + local + 1];
    for (int64_t q = q0; q < q1; ++q) {
      va += a.va[q];
      vb += a.vb[q];
    }
    pa[n] = va;
    pb[n] = vb;
    return;
  }
  const int64_t* bs = a.bstart + a.boff[gr];
  const int64_t nb = a.nbk[gr];
  // buckets local - 2, local - 1, local (clamped to the group's range)
  int64_t p[3], e[3];
  for (int r = 0; r < 3; ++r) {
    const int64_t b = local - 2 + r;
    if (b < 0 || b >= nb) {
      p[r] = e[r] = 0;
    } else {
      p[r] = bs[b];
      e[r] = bs[b + 1];
    }
  }
  const int64_t c0 = a.off[gr];
  const int d = gr >> 1;
  const int w = (gr & 1) ? 2 : 0;
  double* pa = a.pot + (static_cast<int64_t>(d) * 4 + w) * a.num;
  double* pb = pa + a.num;
  double va = pa[n], vb = pb[n];
  while (true) {
    int r = -1;
    int64_t best = 0;
    for (int x = 0; x < 3; ++x)
      if (p[x] < e[x] && (r < 0 || a.order[p[x]] < best)) {
        r = x;
        best = a.order[p[x]];
      }
    if (r < 0) break;
    ++p[r];
    const int64_t cell = c0 + best;
    const int64_t nn = a.nn[cell];
    if (nn == n) {
      va += a.ca1[cell];
      vb += a.cb1[cell];
    } else if (nn + 1 == n) {
      va += a.ca2[cell];
      vb += a.cb2[cell];
    }
  }
  pa[n] = va;
  pb[n] = vb;
}

template <class T>  // one instance per precision's translation unit
__global__ __launch_bounds__(256) void k_td_bins(TdNf2ffArgs a, int64_t t) {
  tdBin(a, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x, t);
}

// Lane l's value of v (l a compile-time lane index after unrolling): two
// v_readlane of the 32-bit halves.
__device__ __forceinline__ double laneValue(double v, int l) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffull), l));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), l));
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// The fixed-order bins, one 64-lane wavefront per bin: the lanes read 64
// consecutive contributions of the bin's run at a time (one coalesced load)
// and the wavefront adds them in lane order, so the sum is the same
// sequence of additions as one thread walking the run.  Lanes past the run
// hold +0.0: a running sum that starts at +0.0 never becomes -0.0 (round to
// nearest), so adding +0.0 leaves it bitwise unchanged.
__device__ __forceinline__ void tdBinWave(const TdNf2ffArgs& a, int64_t id, int lane, int64_t t) {
  if (id >= a.bin_off[12]) return;  // uniform over the wavefront
  int gr = 0;
  while (id >= a.bin_off[gr + 1]) ++gr;
  const int64_t local = id - a.bin_off[gr];
  const int64_t n = t + a.omin[gr] + local;
  if (n < 0 || n >= a.num) return;
  const int d = gr >> 1;
  double* pa = a.pot + (static_cast<int64_t>(d) * 4 + ((gr & 1) ? 2 : 0)) * a.num;
  double* pb = pa + a.num;
  double va = pa[n], vb = pb[n];
  const int64_t q0 = a.lstart[a.loff[gr] + local], q1 = a.lstart[a.loff[gr] + local + 1];
  for (int64_t q = q0; q < q1; q += 64) {
    const bool in = q + lane < q1;
    const double xa = in ? a.va[q + lane] : 0.0;
    const double xb = in ? a.vb[q + lane] : 0.0;
#pragma unroll
    for (int l = 0; l < 64; ++l) {
      va += laneValue(xa, l);
      vb += laneValue(xb, l);
    }
  }
  if (lane == 0) {
    pa[n] = va;
    pb[n] = vb;
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_td_bins_wave(TdNf2ffArgs a, int64_t t) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  tdBinWave(a, g / 64, static_cast<int>(g % 64), t);
}

// Host-side launchers, instantiated once per precision.
template <class T>
struct Launch {
  // P.n > 0: the sweep applies the slabs of P in the same pass; Z.n > 0:
  // the sweep applies the z slabs of Z (then cpml_faces completes either on
  // the far grid planes)
  // S.n > 0: the sweep applies every slab of S (in boundary order; P and Z
  // unused)
  static void update_h(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const CoefPtrs<T>& C, int ichunk,
                       const CpmlList<T>& P = CpmlList<T>{}, const ZCpml<T>& Z = ZCpml<T>{},
                       const StepSlabs<T>& S = StepSlabs<T>{});
  static void update_e(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const CoefPtrs<T>& C, int ichunk,
                       const CpmlList<T>& P = CpmlList<T>{}, const ZCpml<T>& Z = ZCpml<T>{},
                       const StepSlabs<T>& S = StepSlabs<T>{});
  static void update_h(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const ClassCoef<T>& C, int ichunk,
                       const CpmlList<T>& P = CpmlList<T>{}, const ZCpml<T>& Z = ZCpml<T>{},
                       const StepSlabs<T>& S = StepSlabs<T>{});
  static void update_e(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const ClassCoef<T>& C, int ichunk,
                       const CpmlList<T>& P = CpmlList<T>{}, const ZCpml<T>& Z = ZCpml<T>{},
                       const StepSlabs<T>& S = StepSlabs<T>{});
  static void update_h(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const WideCoef<T>& C, int ichunk,
                       const CpmlList<T>& P = CpmlList<T>{}, const ZCpml<T>& Z = ZCpml<T>{},
                       const StepSlabs<T>& S = StepSlabs<T>{});
  static void update_e(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const WideCoef<T>& C, int ichunk,
                       const CpmlList<T>& P = CpmlList<T>{}, const ZCpml<T>& Z = ZCpml<T>{},
                       const StepSlabs<T>& S = StepSlabs<T>{});
  // the Yee half-steps of a 2D (TE, dim 2) or 1D (TEM, dim 1) grid; the 1D
  // H half-step is the 3D sweep (update_h)
  // te_zero (2D): Ex, Ey and Hz are known to stay zero, so the H half-step
  // leaves out Hz and the Ex / Ey terms
  static void update_low(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const CoefPtrs<T>& C, int dim,
                         bool is_h, bool te_zero = false);
  static void update_low(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const ClassCoef<T>& C, int dim,
                         bool is_h, bool te_zero = false);
  static void cpml_faces(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& a, bool is_e);
  static void cpml_zrest(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& a,
                         const ZFrame& fr, bool is_e);
  static void update_e_ade(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const CoefPtrs<T>& C,
                           const AdeArgs<T>& A, int scheme, int ichunk);
  static void update_e_ade(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const ClassCoef<T>& C,
                           const AdeArgs<T>& A, int scheme, int ichunk);
  static void update_e_ade(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const WideCoef<T>& C,
                           const AdeArgs<T>& A, int scheme, int ichunk);
  static void probes(hipStream_t s, const FieldPtrs<T>& F, const Probe* p, int n, double* out, int64_t stride,
                     int64_t t);
  static void cpml(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& a, bool is_e);
  static void cpml2(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& a0, const SlabArgs<T>& a1,
                    bool is_e);
  static void records(hipStream_t s, const FieldPtrs<T>& F, const Record* rec, const int64_t* seg, int64_t nseg,
                      const double* pool, T* state, int64_t t);
  static void gather(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const GatherArgs& a, double* out);
  static void lines(hipStream_t s, const FieldPtrs<T>& F, const LineTerm* terms, const int64_t* off,
                    const double* scale, int n, double* out, int64_t stride, int64_t t);
  static void nf2ff(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const Nf2ffArgs& a, int64_t t);
  // time-domain NF2FF: both phases of step t
  static void tdNf2ff(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const TdNf2ffArgs& a, int64_t t);
  // na directions, out: na x 8 doubles (A_theta, A_phi, F_theta, F_phi as re, im)
  // chunks per direction for na directions, and the partial sums it needs
  // (doubles) beside the na x 8 output
  static int farChunks(int na);
  static void far_field(hipStream_t s, const Nf2ffArgs& a, int f, const FarArgs& fa, int na, double* part,
                        double* out);
  static void unpack(hipStream_t s, const Grid& g, const double* src, T* dst, int n0, int n1, int n2);
  static void pack(hipStream_t s, const Grid& g, const T* src, double* dst, int n0, int n1, int n2);
};

}  // namespace xfd
