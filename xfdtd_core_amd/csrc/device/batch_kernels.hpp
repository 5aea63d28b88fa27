// Batched launches: many independent simulations stepped together.
//
// Each phase of a time step (Yee H, TFSF/lumped records, CPML, Yee E, ...)
// is one persistent launch over the work of every case: a table lists the
// entries (a case's sweep, a CPML slab, a record list, an NF2FF) with the
// first work-group id each entry owns, and a fixed grid of work-groups walks
// the global work-group ids with a grid stride, finding its entry by binary
// search.  Cases of uneven size share the launch, so small cases no longer
// leave the GPU idle, and one step is a handful of launches that a hipGraph
// replays.  The time step index lives in device memory (advanced by k_tick at
// the end of every step), so the captured step is independent of t.
#pragma once

#include "kernels.hpp"

namespace xfd {

// Entry e owns global work-groups [start[e], start[e+1]).
__device__ __forceinline__ int findEntry(const int64_t* __restrict__ start, int n, int64_t b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Work-items of an entry per global work-group id for the point-wise batched
// launches (CPML, far planes, records, DFT): one entry lookup serves this many
// 256-lane passes.
constexpr int kBatchPasses = 4;
constexpr int64_t kBatchGroupItems = 256 * kBatchPasses;

template <class T>
struct BatchCase {
  Grid g;
  FieldPtrs<T> F;
  ClassCoef<T> K;
  CpmlList<T> ph, pe;  // slabs applied inside the sweeps (empty: separate CPML launches)
  int nbx, nby, nbz, ichunk;
  AdeArgs<T> A;        // dispersive cases (k_batch_sweep_ade): state and material tables
  int scheme;          // AdeScheme, or -1
};

template <class T>
struct BatchSlab {
  Grid g;
  FieldPtrs<T> F;
  SlabArgs<T> s;
};

template <class T>
struct BatchRec {
  FieldPtrs<T> F;
  const Record* rec;
  const int64_t* seg;
  int64_t nseg;
  const double* pool;
  T* state;
};

template <class T>
struct BatchNf2ff {
  Grid g;
  FieldPtrs<T> F;
  Nf2ffArgs a;
};

// The sweep tiles the batched launches can use (TK k-lanes by TJ rows); the
// autotuner times each (xfd_batch_tune) and keeps the fastest.
constexpr int kBatchTiles = 5;
constexpr int kBatchTileK[kBatchTiles] = {64, 64, 128, 128, 64};
constexpr int kBatchTileJ[kBatchTiles] = {4, 8, 2, 4, 2};

template <class T, bool IsH, bool Fused, int TK, int TJ>
__global__ __launch_bounds__(TK * TJ) void k_batch_sweep(const BatchCase<T>* __restrict__ cases,
                                                     const int64_t* __restrict__ start, int n, int64_t total) {
  for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const int e = findEntry(start, n, b);
    const BatchCase<T>& C = cases[e];
    const int64_t local = b - start[e];
    const int bx = static_cast<int>(local % C.nbx);
    const int64_t r = local / C.nbx;
    const int by = static_cast<int>(r % C.nby);
    const int bz = static_cast<int>(r / C.nby);
    if (IsH)
      sweepH<T, ClassCoef<T>, Fused, false, false, TK, TJ>(C.g, C.F, C.K, bx, by, bz, C.ichunk, C.ph);
    else
      sweepE<T, ClassCoef<T>, Fused, false, false, TK, TJ>(C.g, C.F, C.K, bx, by, bz, C.ichunk, C.pe);
  }
}

// The E sweep of the dispersive cases (their H sweep is k_batch_sweep's):
// sweepEAde with the case's scheme.
template <class T, int TK, int TJ>
__global__ __launch_bounds__(TK * TJ) void k_batch_sweep_ade(const BatchCase<T>* __restrict__ cases,
                                                         const int64_t* __restrict__ start, int n, int64_t total) {
  for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const int e = findEntry(start, n, b);
    const BatchCase<T>& C = cases[e];
    const int64_t local = b - start[e];
    const int bx = static_cast<int>(local % C.nbx);
    const int64_t r = local / C.nbx;
    const int by = static_cast<int>(r % C.nby);
    const int bz = static_cast<int>(r / C.nby);
    if (C.scheme == kAdeDrude)
      sweepEAde<T, ClassCoef<T>, kAdeDrude, TK, TJ>(C.g, C.F, C.K, C.A, bx, by, bz, C.ichunk);
    else if (C.scheme == kAdeDebye)
      sweepEAde<T, ClassCoef<T>, kAdeDebye, TK, TJ>(C.g, C.F, C.K, C.A, bx, by, bz, C.ichunk);
    else
      sweepEAde<T, ClassCoef<T>, kAdeMLorentz, TK, TJ>(C.g, C.F, C.K, C.A, bx, by, bz, C.ichunk);
  }
}

template <class T, bool IsE>
__global__ __launch_bounds__(256) void k_batch_cpml(const BatchSlab<T>* __restrict__ slabs,
                                                    const int64_t* __restrict__ start, int n, int64_t total) {
  for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const int e = findEntry(start, n, b);
    const BatchSlab<T>& S = slabs[e];
    const int64_t base = (b - start[e]) * kBatchGroupItems + threadIdx.x;
    for (int r = 0; r < kBatchPasses; ++r) cpmlCell<T, IsE>(S.g, S.F, S.s, base + r * 256);
  }
}

// The far-plane points of every slab entry (the fused sweeps' remainder).
template <class T, bool IsE>
__global__ __launch_bounds__(256) void k_batch_cpml_faces(const BatchSlab<T>* __restrict__ slabs,
                                                          const int64_t* __restrict__ start, int n, int64_t total) {
  for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const int e = findEntry(start, n, b);
    const BatchSlab<T>& S = slabs[e];
    const int64_t base = (b - start[e]) * kBatchGroupItems + threadIdx.x;
    for (int r = 0; r < kBatchPasses; ++r) cpmlFacePoint<T, IsE>(S.g, S.F, S.s, base + r * 256);
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_batch_records(const BatchRec<T>* __restrict__ recs,
                                                       const int64_t* __restrict__ start, int n, int64_t total,
                                                       const int64_t* __restrict__ tdev) {
  const int64_t t = *tdev;
  for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const int e = findEntry(start, n, b);
    const BatchRec<T>& R = recs[e];
    const int64_t base = (b - start[e]) * kBatchGroupItems + threadIdx.x;
    for (int r = 0; r < kBatchPasses; ++r) recordSeg(R.F, R.rec, R.seg, R.nseg, R.pool, R.state, base + r * 256, t);
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_batch_nf2ff(const BatchNf2ff<T>* __restrict__ nfs,
                                                     const int64_t* __restrict__ start, int n, int64_t total,
                                                     const int64_t* __restrict__ tdev) {
  const int64_t t = *tdev;
  for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const int e = findEntry(start, n, b);
    const BatchNf2ff<T>& N = nfs[e];
    const int64_t base = (b - start[e]) * kBatchGroupItems + threadIdx.x;
    for (int r = 0; r < kBatchPasses; ++r) nf2ffCell(N.g, N.F, N.a, base + r * 256, t);
  }
}

// Recordings with the schedule evaluated on the device: frame
// (t - first) / interval of a box, when t is on the schedule.
template <class T>
__global__ __launch_bounds__(256) void k_gather_at(Grid g, FieldPtrs<T> F, GatherArgs a, double* out, int64_t first,
                                                   int64_t interval, int64_t count, const int64_t* __restrict__ tdev) {
  const int64_t t = *tdev;
  if (t < first || (t - first) % interval != 0) return;
  const int64_t frame = (t - first) / interval;
  if (frame >= count) return;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= a.frame_elems) return;
  const int n2 = a.e[2] - a.o[2], n1 = a.e[1] - a.o[1];
  const int w = static_cast<int>(q % n2);
  const int v = static_cast<int>((q / n2) % n1);
  const int u = static_cast<int>(q / (static_cast<int64_t>(n2) * n1));
  const int64_t idx = (static_cast<int64_t>(a.o[0] + u) * g.NY + (a.o[1] + v)) * g.pz + (a.o[2] + w);
  out[frame * a.frame_elems + q] = static_cast<double>(F.f[a.field][idx]);
}

template <class T>
__global__ __launch_bounds__(64) void k_probes_at(FieldPtrs<T> F, const Probe* __restrict__ p, int n, double* out,
                                                  int64_t stride, const int64_t* __restrict__ tdev) {
  const int64_t t = *tdev;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n || t >= stride) return;
  out[q * stride + t] = static_cast<double>(F.f[p[q].field][p[q].idx]);
}

template <class T>
__global__ __launch_bounds__(64) void k_lines_at(FieldPtrs<T> F, const LineTerm* __restrict__ terms,
                                                 const int64_t* __restrict__ off, const double* __restrict__ scale,
                                                 int n, double* out, int64_t stride, const int64_t* __restrict__ tdev) {
  const int64_t t = *tdev;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n || t >= stride) return;
  lineSum(F, terms, off, scale, m, out, stride, t);
}

// Advances the device-side step counter (one lane writes it).  Templated on
// the field type only so that each precision's unit holds its own copy.
template <class T>
__global__ void k_tick(int64_t* t) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *t = *t + 1;
}

template <class T>
struct BatchLaunch {
  // fused: some case applies leading CPML slabs inside its sweeps (ph / pe);
  // tile: index into kBatchTileK / kBatchTileJ
  static void sweep(hipStream_t s, bool is_h, const BatchCase<T>* cases, const int64_t* start, int n, int64_t total,
                    int grid, bool fused, int tile = 0);
  // the dispersive cases' E sweep (k_batch_sweep_ade)
  static void sweepAde(hipStream_t s, const BatchCase<T>* cases, const int64_t* start, int n, int64_t total, int grid,
                       int tile = 0);
  static void cpml(hipStream_t s, bool is_e, const BatchSlab<T>* slabs, const int64_t* start, int n, int64_t total,
                   int grid);
  static void cpmlFaces(hipStream_t s, bool is_e, const BatchSlab<T>* slabs, const int64_t* start, int n,
                        int64_t total, int grid);
  static void records(hipStream_t s, const BatchRec<T>* recs, const int64_t* start, int n, int64_t total, int grid,
                      const int64_t* tdev);
  static void nf2ff(hipStream_t s, const BatchNf2ff<T>* nfs, const int64_t* start, int n, int64_t total, int grid,
                    const int64_t* tdev);
  static void gatherAt(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const GatherArgs& a, double* out,
                       int64_t first, int64_t interval, int64_t count, const int64_t* tdev);
  static void probesAt(hipStream_t s, const FieldPtrs<T>& F, const Probe* p, int n, double* out, int64_t stride,
                       const int64_t* tdev);
  static void linesAt(hipStream_t s, const FieldPtrs<T>& F, const LineTerm* terms, const int64_t* off,
                      const double* scale, int n, double* out, int64_t stride, const int64_t* tdev);
  static void tick(hipStream_t s, int64_t* t);
};

}  // namespace xfd
