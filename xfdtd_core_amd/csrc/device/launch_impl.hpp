// Launcher definitions; included by exactly one translation unit per
// precision (kernels_f64.hip, kernels_f32.hip), which differ only in their
// floating-point contraction flags.
#pragma once

#include <algorithm>
#include <cstdlib>

#include "adjoint_kernels.hpp"
#include "batch_kernels.hpp"
#include "kernels.hpp"
#include "step_kernels.hpp"

namespace xfd {

inline unsigned blocks_for(int64_t n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

// The sweep's work-group map.  Measured on the 384^3 box the dispatcher's
// own order is ~4 % faster than the per-XCD contiguous runs, so the remap is
// off unless XFD_XCD_REMAP=1.
inline SweepMap sweepMap(const Grid& g, int ichunk, int tk = kLanesK, int tj = kRowsJ) {
  static const int xcd = [] {
    const char* e = std::getenv("XFD_XCD_REMAP");
    return (e && e[0] == '1') ? 8 : 0;
  }();
  return SweepMap{static_cast<int>(blocks_for(g.nz, tk)), static_cast<int>(blocks_for(g.ny, tj)),
                  static_cast<int>(blocks_for(g.nx, ichunk)), xcd};
}

// The plain and z-in-sweep Yee sweeps take 128 k-lanes x 2 j-rows where nz
// is a multiple of 128 (no idle lanes; half the k-halo loads per point of a
// 64-lane row): measured on the 384^3 box, ms per step FP32 / FP64, 64 x 4
// 1.115 / 2.216, 128 x 2 1.088 / 2.183.  Elsewhere 64 x 4, whose rows waste
// fewer lanes on ragged nz.  XFD_SWEEP_TILE=64 / 128 forces one of them.
inline bool sweepTile128(const Grid& g) {
  const char* e = std::getenv("XFD_SWEEP_TILE");
  if (e && e[0]) return std::atoi(e) == 128;
  return g.nz % 128 == 0;
}

template <class T, class Coef, bool Fused, bool ZF, bool All, int TK, int TJ>
void launchSweep(bool is_h, hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const Coef& C, int ichunk,
                 const CpmlList<T>& P, const ZCpml<T>& Z, const StepSlabs<T>& S) {
  const SweepMap m = sweepMap(g, ichunk, TK, TJ);
  const dim3 grid(static_cast<unsigned>(m.nbx * m.nby * m.nbz)), block(TK, TJ);
  if (is_h)
    hipLaunchKernelGGL((k_update_h<T, Coef, Fused, ZF, All, TK, TJ>), grid, block, 0, s, g, F, C, m, ichunk, P, Z, S);
  else
    hipLaunchKernelGGL((k_update_e<T, Coef, Fused, ZF, All, TK, TJ>), grid, block, 0, s, g, F, C, m, ichunk, P, Z, S);
}

template <class T, class Coef>
void launchUpdate(bool is_h, hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const Coef& C, int ichunk,
                  const CpmlList<T>& P, const ZCpml<T>& Z, const StepSlabs<T>& S = StepSlabs<T>{}) {
  const bool all = S.n > 0, fused = !all && P.n > 0, zf = !all && !fused && Z.n > 0;
  const bool t128 = !all && !fused && sweepTile128(g);
  if (all)
    launchSweep<T, Coef, false, false, true, kLanesK, kRowsJ>(is_h, s, g, F, C, ichunk, P, Z, S);
  else if (fused)
    launchSweep<T, Coef, true, false, false, kLanesK, kRowsJ>(is_h, s, g, F, C, ichunk, P, Z, S);
  else if (zf && t128)
    launchSweep<T, Coef, false, true, false, 128, 2>(is_h, s, g, F, C, ichunk, P, Z, S);
  else if (zf)
    launchSweep<T, Coef, false, true, false, kLanesK, kRowsJ>(is_h, s, g, F, C, ichunk, P, Z, S);
  else if (t128)
    launchSweep<T, Coef, false, false, false, 128, 2>(is_h, s, g, F, C, ichunk, P, Z, S);
  else
    launchSweep<T, Coef, false, false, false, kLanesK, kRowsJ>(is_h, s, g, F, C, ichunk, P, Z, S);
}

template <class T>
void Launch<T>::update_h(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const CoefPtrs<T>& C, int ichunk,
                         const CpmlList<T>& P, const ZCpml<T>& Z, const StepSlabs<T>& S) {
  launchUpdate<T>(true, s, g, F, C, ichunk, P, Z, S);
}
template <class T>
void Launch<T>::update_e(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const CoefPtrs<T>& C, int ichunk,
                         const CpmlList<T>& P, const ZCpml<T>& Z, const StepSlabs<T>& S) {
  launchUpdate<T>(false, s, g, F, C, ichunk, P, Z, S);
}
template <class T>
void Launch<T>::update_h(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const ClassCoef<T>& C, int ichunk,
                         const CpmlList<T>& P, const ZCpml<T>& Z, const StepSlabs<T>& S) {
  launchUpdate<T>(true, s, g, F, C, ichunk, P, Z, S);
}
template <class T>
void Launch<T>::update_e(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const ClassCoef<T>& C, int ichunk,
                         const CpmlList<T>& P, const ZCpml<T>& Z, const StepSlabs<T>& S) {
  launchUpdate<T>(false, s, g, F, C, ichunk, P, Z, S);
}
template <class T>
void Launch<T>::update_h(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const WideCoef<T>& C, int ichunk,
                         const CpmlList<T>& P, const ZCpml<T>& Z, const StepSlabs<T>& S) {
  launchUpdate<T>(true, s, g, F, C, ichunk, P, Z, S);
}
template <class T>
void Launch<T>::update_e(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const WideCoef<T>& C, int ichunk,
                         const CpmlList<T>& P, const ZCpml<T>& Z, const StepSlabs<T>& S) {
  launchUpdate<T>(false, s, g, F, C, ichunk, P, Z, S);
}
template <class T, class Coef>
void launchLow(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const Coef& C, int dim, bool is_h,
               bool te_zero) {
  if (dim == 2) {
    const dim3 grid(blocks_for(g.ny, 64), blocks_for(g.nx, 4)), block(64, 4);
    if (is_h && te_zero)
      hipLaunchKernelGGL((k_te_h<T, Coef, true>), grid, block, 0, s, g, F, C);
    else if (is_h)
      hipLaunchKernelGGL((k_te_h<T, Coef>), grid, block, 0, s, g, F, C);
    else
      hipLaunchKernelGGL((k_te_e<T, Coef>), grid, block, 0, s, g, F, C);
  } else if (!is_h) {
    hipLaunchKernelGGL((k_tem_e<T, Coef>), dim3(blocks_for(g.nz, 256)), dim3(256), 0, s, g, F, C);
  } else {
    launchUpdate<T>(true, s, g, F, C, 1, CpmlList<T>{}, ZCpml<T>{});
  }
}
template <class T>
void Launch<T>::update_low(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const CoefPtrs<T>& C, int dim,
                           bool is_h, bool te_zero) {
  launchLow<T>(s, g, F, C, dim, is_h, te_zero);
}
template <class T>
void Launch<T>::update_low(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const ClassCoef<T>& C, int dim,
                           bool is_h, bool te_zero) {
  launchLow<T>(s, g, F, C, dim, is_h, te_zero);
}
template <class T>
void Launch<T>::cpml_faces(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& a, bool is_e) {
  const FacePieces f = facePieces(g, a);
  const int64_t n = f.nX + f.nY + f.nZ;
  if (n <= 0) return;
  if (is_e)
    hipLaunchKernelGGL((k_cpml_faces<T, true>), dim3(blocks_for(n, 256)), dim3(256), 0, s, g, F, a);
  else
    hipLaunchKernelGGL((k_cpml_faces<T, false>), dim3(blocks_for(n, 256)), dim3(256), 0, s, g, F, a);
}

template <class T>
void Launch<T>::cpml_zrest(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& a,
                           const ZFrame& fr, bool is_e) {
  const int64_t n = static_cast<int64_t>(a.ext[0]) * a.ext[1] * a.ext[2];
  if (n <= 0) return;
  if (is_e)
    hipLaunchKernelGGL((k_cpml_zrest<T, true>), dim3(blocks_for(n, 256)), dim3(256), 0, s, g, F, a, fr);
  else
    hipLaunchKernelGGL((k_cpml_zrest<T, false>), dim3(blocks_for(n, 256)), dim3(256), 0, s, g, F, a, fr);
}

template <class T, class Coef, int TK, int TJ>
void launchUpdateAdeTile(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const Coef& C, const AdeArgs<T>& A,
                         int scheme, int ichunk) {
  const SweepMap m = sweepMap(g, ichunk, TK, TJ);
  const dim3 grid(static_cast<unsigned>(m.nbx * m.nby * m.nbz)), block(TK, TJ);
  switch (scheme) {
    case kAdeDrude:
      hipLaunchKernelGGL((k_update_e_ade<T, Coef, kAdeDrude, TK, TJ>), grid, block, 0, s, g, F, C, A, m, ichunk);
      break;
    case kAdeDebye:
      hipLaunchKernelGGL((k_update_e_ade<T, Coef, kAdeDebye, TK, TJ>), grid, block, 0, s, g, F, C, A, m, ichunk);
      break;
    default:
      hipLaunchKernelGGL((k_update_e_ade<T, Coef, kAdeMLorentz, TK, TJ>), grid, block, 0, s, g, F, C, A, m, ichunk);
      break;
  }
}
// the dispersive E sweep takes the plain sweeps' tile (sweepTile128)
template <class T, class Coef>
void launchUpdateAde(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const Coef& C, const AdeArgs<T>& A,
                     int scheme, int ichunk) {
  if (sweepTile128(g))
    launchUpdateAdeTile<T, Coef, 128, 2>(s, g, F, C, A, scheme, ichunk);
  else
    launchUpdateAdeTile<T, Coef, kLanesK, kRowsJ>(s, g, F, C, A, scheme, ichunk);
}
template <class T>
void Launch<T>::update_e_ade(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const CoefPtrs<T>& C,
                             const AdeArgs<T>& A, int scheme, int ichunk) {
  launchUpdateAde<T>(s, g, F, C, A, scheme, ichunk);
}
template <class T>
void Launch<T>::update_e_ade(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const ClassCoef<T>& C,
                             const AdeArgs<T>& A, int scheme, int ichunk) {
  launchUpdateAde<T>(s, g, F, C, A, scheme, ichunk);
}
template <class T>
void Launch<T>::update_e_ade(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const WideCoef<T>& C,
                             const AdeArgs<T>& A, int scheme, int ichunk) {
  launchUpdateAde<T>(s, g, F, C, A, scheme, ichunk);
}

template <class T>
void Launch<T>::nf2ff(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const Nf2ffArgs& a, int64_t t) {
  if (a.off[6] <= 0) return;
  hipLaunchKernelGGL(k_nf2ff<T>, dim3(blocks_for(a.off[6], 256)), dim3(256), 0, s, g, F, a, t);
}

template <class T>
void Launch<T>::tdNf2ff(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const TdNf2ffArgs& a, int64_t t) {
  if (a.off[12] <= 0) return;
  hipLaunchKernelGGL(k_td_cells<T>, dim3(blocks_for(a.off[12], 256)), dim3(256), 0, s, g, F, a, t);
  if (a.fixed)
    hipLaunchKernelGGL(k_td_bins_wave<T>, dim3(blocks_for(a.bin_off[12] * 64, 256)), dim3(256), 0, s, a, t);
  else
    hipLaunchKernelGGL(k_td_bins<T>, dim3(blocks_for(a.bin_off[12], 256)), dim3(256), 0, s, a, t);
}

template <class T>
void Launch<T>::probes(hipStream_t s, const FieldPtrs<T>& F, const Probe* p, int n, double* out, int64_t stride,
                       int64_t t) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_probes<T>, dim3(blocks_for(n, 64)), dim3(64), 0, s, F, p, n, out, stride, t);
}

template <class T>
void Launch<T>::cpml(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& a, bool is_e) {
  const int64_t vol = static_cast<int64_t>(a.ext[0]) * a.ext[1] * a.ext[2];
  if (vol <= 0) return;
  if (is_e)
    hipLaunchKernelGGL((k_cpml<T, true>), dim3(blocks_for(vol, 256)), dim3(256), 0, s, g, F, a);
  else
    hipLaunchKernelGGL((k_cpml<T, false>), dim3(blocks_for(vol, 256)), dim3(256), 0, s, g, F, a);
}

template <class T>
void Launch<T>::cpml2(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const SlabArgs<T>& a0,
                      const SlabArgs<T>& a1, bool is_e) {
  const unsigned nb0 = blocks_for(static_cast<int64_t>(a0.ext[0]) * a0.ext[1] * a0.ext[2], 256);
  const unsigned nb1 = blocks_for(static_cast<int64_t>(a1.ext[0]) * a1.ext[1] * a1.ext[2], 256);
  if (nb0 + nb1 == 0) return;
  if (is_e)
    hipLaunchKernelGGL((k_cpml2<T, true>), dim3(nb0 + nb1), dim3(256), 0, s, g, F, a0, a1, nb0);
  else
    hipLaunchKernelGGL((k_cpml2<T, false>), dim3(nb0 + nb1), dim3(256), 0, s, g, F, a0, a1, nb0);
}

template <class T>
void Launch<T>::records(hipStream_t s, const FieldPtrs<T>& F, const Record* rec, const int64_t* seg, int64_t nseg,
                        const double* pool, T* state, int64_t t) {
  if (nseg <= 0) return;
  hipLaunchKernelGGL(k_records<T>, dim3(blocks_for(nseg, 256)), dim3(256), 0, s, F, rec, seg, nseg, pool, state,
                     t);
}

template <class T>
void Launch<T>::gather(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const GatherArgs& a, double* out) {
  if (a.frame_elems <= 0) return;
  hipLaunchKernelGGL(k_gather<T>, dim3(blocks_for(a.frame_elems, 256)), dim3(256), 0, s, g, F, a, out);
}

template <class T>
void Launch<T>::unpack(hipStream_t s, const Grid& g, const double* src, T* dst, int n0, int n1, int n2) {
  const int64_t vol = static_cast<int64_t>(n0) * n1 * n2;
  if (vol <= 0) return;
  hipLaunchKernelGGL(k_unpack<T>, dim3(blocks_for(vol, 256)), dim3(256), 0, s, g, src, dst, n0, n1, n2);
}

template <class T>
void Launch<T>::pack(hipStream_t s, const Grid& g, const T* src, double* dst, int n0, int n1, int n2) {
  const int64_t vol = static_cast<int64_t>(n0) * n1 * n2;
  if (vol <= 0) return;
  hipLaunchKernelGGL(k_pack<T>, dim3(blocks_for(vol, 256)), dim3(256), 0, s, g, src, dst, n0, n1, n2);
}

// ------------------------------------------------------------- batched --
template <class T, int TK, int TJ>
void batchSweepTile(hipStream_t s, bool is_h, const BatchCase<T>* cases, const int64_t* start, int n, int64_t total,
                    int grid, bool fused) {
  dim3 block(TK, TJ);
  dim3 g(static_cast<unsigned>(std::min<int64_t>(total, grid)));
  if (is_h && fused)
    hipLaunchKernelGGL((k_batch_sweep<T, true, true, TK, TJ>), g, block, 0, s, cases, start, n, total);
  else if (is_h)
    hipLaunchKernelGGL((k_batch_sweep<T, true, false, TK, TJ>), g, block, 0, s, cases, start, n, total);
  else if (fused)
    hipLaunchKernelGGL((k_batch_sweep<T, false, true, TK, TJ>), g, block, 0, s, cases, start, n, total);
  else
    hipLaunchKernelGGL((k_batch_sweep<T, false, false, TK, TJ>), g, block, 0, s, cases, start, n, total);
}
template <class T>
void BatchLaunch<T>::sweep(hipStream_t s, bool is_h, const BatchCase<T>* cases, const int64_t* start, int n,
                           int64_t total, int grid, bool fused, int tile) {
  if (total <= 0) return;
  static_assert(kBatchTiles == 5, "tile list");
  switch (tile) {
    case 1: batchSweepTile<T, kBatchTileK[1], kBatchTileJ[1]>(s, is_h, cases, start, n, total, grid, fused); break;
    case 2: batchSweepTile<T, kBatchTileK[2], kBatchTileJ[2]>(s, is_h, cases, start, n, total, grid, fused); break;
    case 3: batchSweepTile<T, kBatchTileK[3], kBatchTileJ[3]>(s, is_h, cases, start, n, total, grid, fused); break;
    case 4: batchSweepTile<T, kBatchTileK[4], kBatchTileJ[4]>(s, is_h, cases, start, n, total, grid, fused); break;
    default: batchSweepTile<T, kBatchTileK[0], kBatchTileJ[0]>(s, is_h, cases, start, n, total, grid, fused); break;
  }
}
template <class T, int TK, int TJ>
void batchSweepAdeTile(hipStream_t s, const BatchCase<T>* cases, const int64_t* start, int n, int64_t total,
                       int grid) {
  hipLaunchKernelGGL((k_batch_sweep_ade<T, TK, TJ>), dim3(static_cast<unsigned>(std::min<int64_t>(total, grid))),
                     dim3(TK, TJ), 0, s, cases, start, n, total);
}
template <class T>
void BatchLaunch<T>::sweepAde(hipStream_t s, const BatchCase<T>* cases, const int64_t* start, int n, int64_t total,
                              int grid, int tile) {
  if (total <= 0) return;
  switch (tile) {
    case 1: batchSweepAdeTile<T, kBatchTileK[1], kBatchTileJ[1]>(s, cases, start, n, total, grid); break;
    case 2: batchSweepAdeTile<T, kBatchTileK[2], kBatchTileJ[2]>(s, cases, start, n, total, grid); break;
    case 3: batchSweepAdeTile<T, kBatchTileK[3], kBatchTileJ[3]>(s, cases, start, n, total, grid); break;
    case 4: batchSweepAdeTile<T, kBatchTileK[4], kBatchTileJ[4]>(s, cases, start, n, total, grid); break;
    default: batchSweepAdeTile<T, kBatchTileK[0], kBatchTileJ[0]>(s, cases, start, n, total, grid); break;
  }
}
template <class T>
void BatchLaunch<T>::cpml(hipStream_t s, bool is_e, const BatchSlab<T>* slabs, const int64_t* start, int n,
                          int64_t total, int grid) {
  if (total <= 0) return;
  dim3 g(static_cast<unsigned>(std::min<int64_t>(total, grid)));
  if (is_e)
    hipLaunchKernelGGL((k_batch_cpml<T, true>), g, dim3(256), 0, s, slabs, start, n, total);
  else
    hipLaunchKernelGGL((k_batch_cpml<T, false>), g, dim3(256), 0, s, slabs, start, n, total);
}
template <class T>
void BatchLaunch<T>::cpmlFaces(hipStream_t s, bool is_e, const BatchSlab<T>* slabs, const int64_t* start, int n,
                               int64_t total, int grid) {
  if (total <= 0) return;
  dim3 g(static_cast<unsigned>(std::min<int64_t>(total, grid)));
  if (is_e)
    hipLaunchKernelGGL((k_batch_cpml_faces<T, true>), g, dim3(256), 0, s, slabs, start, n, total);
  else
    hipLaunchKernelGGL((k_batch_cpml_faces<T, false>), g, dim3(256), 0, s, slabs, start, n, total);
}
template <class T>
void BatchLaunch<T>::records(hipStream_t s, const BatchRec<T>* recs, const int64_t* start, int n, int64_t total,
                             int grid, const int64_t* tdev) {
  if (total <= 0) return;
  dim3 g(static_cast<unsigned>(std::min<int64_t>(total, grid)));
  hipLaunchKernelGGL(k_batch_records<T>, g, dim3(256), 0, s, recs, start, n, total, tdev);
}
template <class T>
void BatchLaunch<T>::nf2ff(hipStream_t s, const BatchNf2ff<T>* nfs, const int64_t* start, int n, int64_t total,
                           int grid, const int64_t* tdev) {
  if (total <= 0) return;
  dim3 g(static_cast<unsigned>(std::min<int64_t>(total, grid)));
  hipLaunchKernelGGL(k_batch_nf2ff<T>, g, dim3(256), 0, s, nfs, start, n, total, tdev);
}
template <class T>
void BatchLaunch<T>::gatherAt(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const GatherArgs& a, double* out,
                              int64_t first, int64_t interval, int64_t count, const int64_t* tdev) {
  if (a.frame_elems <= 0 || count <= 0) return;
  hipLaunchKernelGGL(k_gather_at<T>, dim3(blocks_for(a.frame_elems, 256)), dim3(256), 0, s, g, F, a, out, first,
                     interval, count, tdev);
}
template <class T>
void BatchLaunch<T>::probesAt(hipStream_t s, const FieldPtrs<T>& F, const Probe* p, int n, double* out,
                              int64_t stride, const int64_t* tdev) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_probes_at<T>, dim3(blocks_for(n, 64)), dim3(64), 0, s, F, p, n, out, stride, tdev);
}
template <class T>
void BatchLaunch<T>::linesAt(hipStream_t s, const FieldPtrs<T>& F, const LineTerm* terms, const int64_t* off,
                             const double* scale, int n, double* out, int64_t stride, const int64_t* tdev) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lines_at<T>, dim3(blocks_for(n, 64)), dim3(64), 0, s, F, terms, off, scale, n, out, stride,
                     tdev);
}
template <class T>
void Launch<T>::lines(hipStream_t s, const FieldPtrs<T>& F, const LineTerm* terms, const int64_t* off,
                      const double* scale, int n, double* out, int64_t stride, int64_t t) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lines<T>, dim3(blocks_for(n, 64)), dim3(64), 0, s, F, terms, off, scale, n, out, stride, t);
}
template <class T>
void BatchLaunch<T>::tick(hipStream_t s, int64_t* t) {
  hipLaunchKernelGGL(k_tick<T>, dim3(1), dim3(64), 0, s, t);
}

template <class T>
int Launch<T>::farChunks(int na) {
  // about 2048 work-groups in all, at most 32 per direction
  return static_cast<int>(std::max(1, std::min(32, (2048 + std::max(na, 1) - 1) / std::max(na, 1))));
}
template <class T>
void Launch<T>::far_field(hipStream_t s, const Nf2ffArgs& a, int f, const FarArgs& fa, int na, double* part,
                          double* out) {
  if (na <= 0) return;
  const int chunks = farChunks(na);
  hipLaunchKernelGGL(k_far_field<T>, dim3(na, chunks), dim3(256), 0, s, a, f, fa, chunks, part);
  hipLaunchKernelGGL(k_far_reduce<T>, dim3(blocks_for(static_cast<int64_t>(na) * 8, 256)), dim3(256), 0, s, na,
                     chunks, part, out);
}

// ---------------------------------------------------------------- adjoint --
inline dim3 adjGrid(const Grid& g) {
  return dim3(blocks_for(g.nz + 1, kLanesK), blocks_for(g.ny + 1, kRowsJ), static_cast<unsigned>(g.nx + 1));
}
// The i-marching sweeps, or with XFD_ADJ_POINT=1 the one-point-per-thread
// kernels (the transposes term by term, kept as their cross-check: in FP64
// both give the same bits).  Read at every launch.
inline bool adjPoint() {
  const char* e = std::getenv("XFD_ADJ_POINT");
  return e && e[0] == '1';
}
template <class T, class Coef>
void launchAdj(bool is_h, hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const Coef& C, int ichunk) {
  const dim3 block(kLanesK, kRowsJ);
  if (adjPoint()) {
    if (is_h)
      hipLaunchKernelGGL((k_adj_h_point<T, Coef>), adjGrid(g), block, 0, s, g, L, C);
    else
      hipLaunchKernelGGL((k_adj_e_point<T, Coef>), adjGrid(g), block, 0, s, g, L, C);
    return;
  }
  ichunk = std::max(1, ichunk);
  // H: the R_H tiles; E: every point up to the far faces
  const int e = is_h ? 0 : 1;
  const SweepMap m{static_cast<int>(blocks_for(g.nz + e, kLanesK)), static_cast<int>(blocks_for(g.ny + e, kRowsJ)),
                   static_cast<int>(blocks_for(g.nx + e, ichunk)), 0};
  const dim3 grid(static_cast<unsigned>(m.nbx * m.nby * m.nbz));
  if (is_h)
    hipLaunchKernelGGL((k_adj_h<T, Coef>), grid, block, 0, s, g, L, C, m, ichunk);
  else
    hipLaunchKernelGGL((k_adj_e<T, Coef>), grid, block, 0, s, g, L, C, m, ichunk);
}
template <class T>
void AdjLaunch<T>::adjH(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const CoefPtrs<T>& C, int ichunk) {
  launchAdj<T>(true, s, g, L, C, ichunk);
}
template <class T>
void AdjLaunch<T>::adjH(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const ClassCoef<T>& C, int ichunk) {
  launchAdj<T>(true, s, g, L, C, ichunk);
}
template <class T>
void AdjLaunch<T>::adjH(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const WideCoef<T>& C, int ichunk) {
  launchAdj<T>(true, s, g, L, C, ichunk);
}
template <class T>
void AdjLaunch<T>::adjE(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const CoefPtrs<T>& C, int ichunk) {
  launchAdj<T>(false, s, g, L, C, ichunk);
}
template <class T>
void AdjLaunch<T>::adjE(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const ClassCoef<T>& C, int ichunk) {
  launchAdj<T>(false, s, g, L, C, ichunk);
}
template <class T>
void AdjLaunch<T>::adjE(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const WideCoef<T>& C, int ichunk) {
  launchAdj<T>(false, s, g, L, C, ichunk);
}
template <class T>
void AdjLaunch<T>::records(hipStream_t s, const FieldPtrs<T>& L, const Record* rec, const int64_t* seg, int64_t nseg,
                           T* lamJ) {
  if (nseg <= 0) return;
  hipLaunchKernelGGL(k_adj_records<T>, dim3(blocks_for(nseg, 256)), dim3(256), 0, s, L, rec, seg, nseg, lamJ);
}
template <class T>
void AdjLaunch<T>::adePre(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const AdeArgs<T>& A,
                          const AdeArgs<T>& LA, const FieldPtrs<T>& X, int scheme) {
  const int64_t n = static_cast<int64_t>(g.nx) * g.ny * g.nz;
  const dim3 grid(blocks_for(n, 256));
  if (scheme == kAdeDrude)
    hipLaunchKernelGGL((k_adj_ade_pre<T, kAdeDrude>), grid, dim3(256), 0, s, g, L, A, LA, X);
  else if (scheme == kAdeDebye)
    hipLaunchKernelGGL((k_adj_ade_pre<T, kAdeDebye>), grid, dim3(256), 0, s, g, L, A, LA, X);
  else
    hipLaunchKernelGGL((k_adj_ade_pre<T, kAdeMLorentz>), grid, dim3(256), 0, s, g, L, A, LA, X);
}
template <class T>
void AdjLaunch<T>::adePost(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const int16_t* mat,
                           const FieldPtrs<T>& X) {
  const int64_t n = static_cast<int64_t>(g.nx) * g.ny * g.nz;
  hipLaunchKernelGGL(k_adj_ade_post<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, g, L, mat, X);
}
template <class T>
void AdjLaunch<T>::cpmlMu(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& a) {
  const int64_t vol = static_cast<int64_t>(a.ext[0]) * a.ext[1] * a.ext[2];
  if (vol <= 0) return;
  hipLaunchKernelGGL(k_adj_cpml_mu<T>, dim3(blocks_for(vol, 256)), dim3(256), 0, s, g, L, a);
}
template <class T>
void AdjLaunch<T>::cpmlGather(hipStream_t s, bool is_e, const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& a) {
  int64_t vol = static_cast<int64_t>(a.ext[0]) * a.ext[1] * a.ext[2];
  if (vol <= 0) return;
  vol = vol / a.ext[a.axis] * (a.ext[a.axis] + 1);
  if (is_e)
    hipLaunchKernelGGL((k_adj_cpml_gather<T, true>), dim3(blocks_for(vol, 256)), dim3(256), 0, s, g, L, a);
  else
    hipLaunchKernelGGL((k_adj_cpml_gather<T, false>), dim3(blocks_for(vol, 256)), dim3(256), 0, s, g, L, a);
}
template <class T>
void AdjLaunch<T>::cpmlMu2(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const SlabArgs<T>& a0,
                           const SlabArgs<T>& a1) {
  const unsigned nb0 = blocks_for(static_cast<int64_t>(a0.ext[0]) * a0.ext[1] * a0.ext[2], 256);
  const unsigned nb1 = blocks_for(static_cast<int64_t>(a1.ext[0]) * a1.ext[1] * a1.ext[2], 256);
  if (nb0 + nb1 == 0) return;
  hipLaunchKernelGGL(k_adj_cpml_mu2<T>, dim3(nb0 + nb1), dim3(256), 0, s, g, L, a0, a1, nb0);
}
template <class T>
void AdjLaunch<T>::cpmlGather2(hipStream_t s, bool is_e, const Grid& g, const FieldPtrs<T>& L,
                               const SlabArgs<T>& a0, const SlabArgs<T>& a1) {
  auto widened = [](const SlabArgs<T>& a) {
    const int64_t vol = static_cast<int64_t>(a.ext[0]) * a.ext[1] * a.ext[2];
    return vol <= 0 ? int64_t(0) : vol / a.ext[a.axis] * (a.ext[a.axis] + 1);
  };
  const unsigned nb0 = blocks_for(widened(a0), 256), nb1 = blocks_for(widened(a1), 256);
  if (nb0 + nb1 == 0) return;
  if (is_e)
    hipLaunchKernelGGL((k_adj_cpml_gather2<T, true>), dim3(nb0 + nb1), dim3(256), 0, s, g, L, a0, a1, nb0);
  else
    hipLaunchKernelGGL((k_adj_cpml_gather2<T, false>), dim3(nb0 + nb1), dim3(256), 0, s, g, L, a0, a1, nb0);
}
template <class T>
void AdjLaunch<T>::hist(hipStream_t s, const Grid& g, const FieldPtrs<T>& F, const DesignArgs& d, T* h) {
  if (d.off[3] <= 0) return;
  hipLaunchKernelGGL(k_adj_hist<T>, dim3(blocks_for(d.off[3], 256)), dim3(256), 0, s, g, F, d, h);
}
template <class T>
void AdjLaunch<T>::grad(hipStream_t s, const Grid& g, const FieldPtrs<T>& L, const DesignArgs& d, const T* h0,
                        const T* h1, const double* beta, const double* gamma, double* gr) {
  if (d.off[3] <= 0) return;
  hipLaunchKernelGGL(k_adj_grad<T>, dim3(blocks_for(d.off[3], 256)), dim3(256), 0, s, g, L, d, h0, h1, beta, gamma,
                     gr);
}
template <class T>
void AdjLaunch<T>::source(hipStream_t s, const FieldPtrs<T>& L, const AdjPoint* pts, int n, const double* src,
                          int64_t nt, int64_t step) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_adj_source<T>, dim3(blocks_for(n, 64)), dim3(64), 0, s, L, pts, n, src, nt, step);
}

template <class T, class Coef>
void launchStep(hipStream_t s, const Grid& g, const FieldPtrs<T>& A, const FieldPtrs<T>& B, const Coef& C,
                const StepSlabs<T>& SH, const StepSlabs<T>& SE, int ichunk) {
  const SweepMap m = sweepMap(g, ichunk);
  hipLaunchKernelGGL((k_step<T, Coef>), dim3(static_cast<unsigned>(m.nbx * m.nby * m.nbz)),
                     dim3(kLanesK, kRowsJ), 0, s, g, A, B, C, SH, SE, m, ichunk);
}
template <class T>
void StepLaunch<T>::step(hipStream_t s, const Grid& g, const FieldPtrs<T>& A, const FieldPtrs<T>& B,
                         const CoefPtrs<T>& C, const StepSlabs<T>& SH, const StepSlabs<T>& SE, int ichunk) {
  launchStep<T>(s, g, A, B, C, SH, SE, ichunk);
}
template <class T>
void StepLaunch<T>::step(hipStream_t s, const Grid& g, const FieldPtrs<T>& A, const FieldPtrs<T>& B,
                         const ClassCoef<T>& C, const StepSlabs<T>& SH, const StepSlabs<T>& SE, int ichunk) {
  launchStep<T>(s, g, A, B, C, SH, SE, ichunk);
}

}  // namespace xfd
