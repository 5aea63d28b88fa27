// Device-side data layout shared by the kernels and the C-ABI.
//
// All six field arrays, the 18 coefficient arrays and the recordings' source
// fields live in one padded 3D layout of (nx+1) x (ny+1) x pz elements with k
// fastest: element (i, j, k) of every component is at (i*(ny+1) + j)*pz + k.
// pz rounds nz+1 up to a whole number of 128-byte lines, so every k-row
// starts on a line boundary and a 64-lane wavefront reading 64 consecutive k
// touches the minimum number of lines.  A component's own extent (e.g. Ex is
// nx x (ny+1) x (nz+1)) is a sub-box of the common one; the padding cells are
// zero and never written.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace xfd {

struct Grid {
  int nx, ny, nz;
  int NY;          // ny + 1
  int64_t pz;      // padded k pitch (elements)
  int64_t plane;   // NY * pz
  int64_t total;   // (nx + 1) * plane
};

// Work-group order of the Yee sweeps.  The sweep grid is nbx (k) x nby (j)
// x nbz (i-chunk) work-groups launched as a 1D grid.  The dispatcher hands
// consecutive work-group ids to the 8 XCDs in turn, and each XCD has its own
// L2; with `xcd` set, hardware id h is renumbered so that each XCD receives
// one contiguous run of logical ids (k fastest, then j, then i), i.e. whole
// neighbourhoods of tiles, and the rows/planes that neighbouring tiles share
// are served from one L2 instead of being fetched from HBM once per XCD.
struct SweepMap {
  int nbx, nby, nbz;
  int xcd;  // 0: identity order, else the number of XCDs
  __device__ __forceinline__ void coords(int& bx, int& by, int& bz) const {
    const int n = nbx * nby * nbz;
    const int h = blockIdx.x;
    int logical = h;
    if (xcd > 1) {
      const int per = n / xcd;
      if (h < per * xcd) logical = (h % xcd) * per + h / xcd;
    }
    bx = logical % nbx;
    const int r = logical / nbx;
    by = r % nby;
    bz = r / nby;
  }
};

template <class T>
struct FieldPtrs {
  T* f[6];  // Ex Ey Ez Hx Hy Hz
};

// Streamed per-cell data (own fields, class ids, CPML psi: read and written
// once per pass) carries the non-temporal hint, so it does not displace the
// neighbour rows other work-groups still read from L2 (measured on the 384^3
// box: FP32 +5 %, FP64 +9 %; batched sweep +13 %).  XFD_STREAMING_OWN=0 at
// build time turns it off.
#ifndef XFD_STREAMING_OWN
#define XFD_STREAMING_OWN 1
#endif
template <class T>
__device__ __forceinline__ T ldOwn(const T* p) {
#if XFD_STREAMING_OWN
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
template <class T>
__device__ __forceinline__ void stOwn(T* p, T v) {
#if XFD_STREAMING_OWN
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// Coefficient access for component q (0..5 = Ex Ey Ez Hx Hy Hz): the
// triple (self, tangential-A dual, tangential-B dual) of the reference's
// per-cell arrays.  Dense: 18 padded arrays.  Classed: one uint16 class id
// per cell and component into a table of distinct triples -- the same FP64
// values, 2 bytes per cell instead of 3*sizeof(T).
template <class T>
struct CoefPtrs {
  const T* c[18];  // cexe cexhy cexhz ceye ceyhz ceyhx ceze cezhx cezhy chxh ... chzey
  __device__ __forceinline__ void get(int q, int64_t idx, T& s, T& a, T& b) const {
    s = c[3 * q][idx];
    a = c[3 * q + 1][idx];
    b = c[3 * q + 2][idx];
  }
  __device__ __forceinline__ void getOwn(int q, int64_t idx, T& s, T& a, T& b) const {
    s = ldOwn(c[3 * q] + idx);
    a = ldOwn(c[3 * q + 1] + idx);
    b = ldOwn(c[3 * q + 2] + idx);
  }
  // split access (a load issued ahead of its use, then its values): the
  // three coefficients themselves
  struct Raw {
    T s, a, b;
  };
  __device__ __forceinline__ Raw raw(int q, int64_t idx) const {
    return Raw{ldOwn(c[3 * q] + idx), ldOwn(c[3 * q + 1] + idx), ldOwn(c[3 * q + 2] + idx)};
  }
  __device__ __forceinline__ void resolve(int, const Raw& r, T& s, T& a, T& b) const {
    s = r.s;
    a = r.a;
    b = r.b;
  }
};

template <class T, class Id = uint16_t>
struct ClassCoef {
  const Id* id[6];
  const T* tab[6];  // 3 values per class
  __device__ __forceinline__ void get(int q, int64_t idx, T& s, T& a, T& b) const {
    const T* e = tab[q] + 3 * static_cast<int>(id[q][idx]);
    s = e[0];
    a = e[1];
    b = e[2];
  }
  // the same for the cell a sweep owns (its id is read once per pass)
  __device__ __forceinline__ void getOwn(int q, int64_t idx, T& s, T& a, T& b) const {
    const T* e = tab[q] + 3 * static_cast<int>(ldOwn(id[q] + idx));
    s = e[0];
    a = e[1];
    b = e[2];
  }
  // split access (a load issued ahead of its use, then its values): the
  // class id, then its table entry
  struct Raw {
    int id;
  };
  __device__ __forceinline__ Raw raw(int q, int64_t idx) const { return Raw{static_cast<int>(ldOwn(id[q] + idx))}; }
  __device__ __forceinline__ void resolve(int q, const Raw& r, T& s, T& a, T& b) const {
    const T* e = tab[q] + 3 * r.id;
    s = e[0];
    a = e[1];
    b = e[2];
  }
};

// 32-bit class ids, for grids with more distinct triples per component than
// 16-bit ids hold (per-cell permittivity maps, design regions): 4 bytes per
// cell and component instead of 2, still far below the dense layout's three
// coefficients, and the same branch-free reads.
template <class T>
using WideCoef = ClassCoef<T, uint32_t>;

// ADE state and per-material coefficients for the dispersive E update.
// `mat` is a per-cell material id in the padded layout (-1: plain update);
// material m's coefficients start at tab + m * stride:
//   [0] cej_sum  [1] cee_prev  then per pole p at 2 + 7p:
//   cj_en, cj_e, cj_ep, cj_j, cj_jp, cjsum_j, cjsum_jp.
// j[q] / jp[q] hold `poles` padded arrays each (pole p at offset p * total);
// jp and eprev are used by the modified-Lorentz scheme only.
template <class T>
struct AdeArgs {
  int poles;
  int stride;
  int64_t total;
  const int16_t* mat;
  const T* tab;
  T* j[3];
  T* jp[3];
  T* eprev[3];
};

// NF2FF running DFT: the six surfaces' cells flattened (surface d owns
// [off[d], off[d+1])), each surface a box [lo, hi) with k fastest.
// acc holds nf x 4 x cells complex values: (f * 4 + w) * cells + cell, w =
// ja, jb, ma, mb.  te / th: nf x nt complex phase tables.
struct Nf2ffArgs {
  int lo[6][3], ext[6][3];
  int64_t off[7];
  int nf;
  int repeat;
  int64_t nt;
  const double2* te;
  const double2* th;
  double2* acc;
};

// Time-domain NF2FF (NFFFTTimeDomain): 12 groups g = 2 d + attr of surface
// d (XN..ZP) and attribute attr (0: E -> potentials ua, ub; 1: H -> wa, wb),
// each the cells of the surface's box in the reference's loop order (k
// fastest).  Group g owns cells [off[g], off[g + 1]); per group cell s (the
// time step delay plus the attribute's offset: the bin of step t is
// floor(s + t)) and ds (the face area).  Per step, k_td_cells computes every cell's bin nn
// and its four contributions; k_td_bins then adds them bin by bin.  A cell's
// bin at step t is o + t with o = floor(s) or floor(s) + 1 (rounding of
// s + t), so the cells are bucketed by floor(s) (`order`, sorted by bucket
// then cell; group g's buckets at bstart + boff[g], its nbk[g] buckets
// starting at omin[g]); bin n of step t draws on buckets n - t - 2 .. n - t,
// merged in cell order, so every bin sums in the reference's order.
// When no s + t of the run can round up to the next integer (`fixed`, the
// usual case; the host checks it), a cell's bin is exactly floor(s) + t, so
// the order of the contributions to every bin is the same each step: bin L
// of group g (L = n - t - omin[g], 0..nbk[g]) takes bucket L's first and
// bucket L - 1's second contributions merged by cell.  `slot` gives each
// cell's two positions in that order (va / vb, one run per bin starting at
// lstart[loff[g] + L]); k_td_cells writes there and k_td_bins streams each
// run, nbk[g] + 1 bins per group and step.
// pot holds per surface d: ua, ub, wa, wb of num values (FP64).
struct TdNf2ffArgs {
  int lo[6][3], ext[6][3];
  int64_t off[13];
  int64_t omin[12];
  int64_t nbk[12];
  int64_t boff[13];      // group g's bucket starts: bstart[boff[g] .. boff[g] + nbk[g]]
  int64_t bin_off[13];   // thread ids of group g's bins: nbk[g] + 2 bins per step
  int64_t num;
  const double* s;
  const double* ds;
  const int64_t* bstart;
  const int64_t* order;  // cell ids within the group
  double* prev_a;
  double* prev_b;
  int64_t* nn;
  double* ca1;
  double* ca2;
  double* cb1;
  double* cb2;
  double* pot;
  int fixed;
  const int64_t* slot;    // 2 per cell: first, second contribution
  const int64_t* lstart;  // per group, nbk + 2 run starts
  int64_t loff[12];
  double* va;
  double* vb;
};

// One CPML slab for one half-step.  The thread space is the union box of the
// two corrected components; `ext_a` / `ext_b` give each component's extent
// along the tangential axes a and b (the other component's is one larger).
template <class T>
struct SlabArgs {
  int axis;          // main axis c
  int n;             // layers
  int start;         // first corrected plane along c
  int lo[3], ext[3]; // union box in grid indices
  int extA[2];       // component A extent along (a, b)
  int extB[2];       // component B extent along (a, b)
  int fieldA, fieldB, dualA, dualB;  // field ids (0..5)
  int64_t stride_c;  // padded-layout stride of the main axis
  const T* cpsiA;    // box-local layout, k fastest
  const T* cpsiB;
  T* psiA;
  T* psiB;
  const T* ca;       // per-layer recursive-convolution coefficients
  const T* cb;
};

// The CPML slabs of one half-step for the sweeps that apply them in the same
// pass (device array of `n` slabs in boundary add order).
template <class T>
struct CpmlList {
  const SlabArgs<T>* s = nullptr;
  int n = 0;
};

// A z-normal CPML slab applied inside a Yee sweep.  A sweep lane keeps one k
// for its whole i-march, so it lies in at most one such slab and its layer
// coefficients are loaded once; only psi and c_psi are read per cell, and the
// field and dual values come from the sweep's registers.  The slab spans the
// whole x/y range (box-local psi index t = (i * ext1 + j) * ext2 + k - klo)
// and both of its components cover every sweep point (i < nx, j < ny) of its
// layers [klo, khi); the far planes i = nx, j = ny (k_cpml_faces) and, in
// FP64, the frame (ZFrame, k_cpml_zrest) are left to a launch of their own.
template <class T>
struct ZSlab {
  int klo, khi;  // corrected k planes (layer = k - klo)
  int ext1, ext2;
  const T* cpsiA;
  const T* cpsiB;
  T* psiA;
  T* psiB;
  const T* ca;
  const T* cb;
};

// The points of a z slab that x or y slabs also correct (the frame): i in
// one of the x slabs' ranges or j in one of the y slabs'.  In FP64 the sweep
// applies a z slab only off the frame, where it is the one correction of the
// point, so the reference's order is kept; the frame follows the x/y slabs
// in a launch of its own (k_cpml_zrest).  FP32 leaves the frame empty.
struct ZFrame {
  int nx = 0, ny = 0;
  int xr[4][2] = {}, yr[4][2] = {};  // [lo, hi)
  __device__ __forceinline__ bool inX(int i) const {
    bool r = false;
    for (int q = 0; q < nx; ++q) r = r || (i >= xr[q][0] && i < xr[q][1]);
    return r;
  }
  __device__ __forceinline__ bool inY(int j) const {
    bool r = false;
    for (int q = 0; q < ny; ++q) r = r || (j >= yr[q][0] && j < yr[q][1]);
    return r;
  }
};

// The z slabs (at most two, disjoint in k) of one half-step.
template <class T>
struct ZCpml {
  ZSlab<T> z[2];
  int n = 0;
  ZFrame fr;
};

// Additive correction applied after a Yee half-step, in the reference's
// corrector order (objects, then TFSF; src/simulation/simulation.cpp:437-487).
struct Record {
  int64_t target;     // padded linear index
  int64_t pool;       // table offset of column j at t = 0
  int64_t row;        // table row stride
  double w, omw;      // interpolation weights (kind 0)
  double tr, coef;    // transform and coefficient (kind 0); coef (kind 1)
  int32_t field;      // 0..5
  int32_t kind;       // 0: TFSF interpolated incident field, 1: waveform * coef
};

struct GatherArgs {
  int field;
  int o[3], e[3];
  int64_t frame_elems;
};

}  // namespace xfd
