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
__global__ __launch_bounds__(kTileThreads) void k_adj_h(Grid g, FieldPtrs<T> L, Coef C) {
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
__global__ __launch_bounds__(kTileThreads) void k_adj_e(Grid g, FieldPtrs<T> L, Coef C) {
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

// CPML adjoint, first half: for every psi point of one slab and half-step,
// mu = b mu + c_psi lambda_field (the slab's psi arrays in `s` hold mu, and
// the field pointers are the adjoint fields).
template <class T>
__global__ __launch_bounds__(256) void k_adj_cpml_mu(Grid g, FieldPtrs<T> L, SlabArgs<T> s) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
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
__global__ __launch_bounds__(256) void k_adj_cpml_gather(Grid g, FieldPtrs<T> L, SlabArgs<T> s) {
  const int c = s.axis, a = (c + 1) % 3, b = (c + 2) % 3;
  int wext[3] = {s.ext[0], s.ext[1], s.ext[2]};
  wext[c] += 1;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
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
  static void adjH(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const CoefPtrs<T>& C);
  static void adjH(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const ClassCoef<T>& C);
  static void adjE(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const CoefPtrs<T>& C);
  static void adjE(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const ClassCoef<T>& C);
  static void cpmlMu(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& a);
  static void cpmlGather(hipStream_t s, bool is_e, const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& a);
  static void hist(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const DesignArgs& d, T* hist);
  static void grad(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const DesignArgs& d, const T* h0, const T* h1,
                   const double* beta, const double* gamma, double* grad);
  static void source(hipStream_t s, const FieldPtrs<T>& L, const AdjPoint* pts, int n, const double* src, int64_t nt,
                     int64_t step);
};

}  // namespace xfd
