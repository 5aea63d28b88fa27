// C-ABI of the device library: owns device memory, turns the descriptor into
// device-friendly tables and runs the step loop on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../../../include/xfdtd_core_amd.h"
#include "adjoint_kernels.hpp"
#include "batch_kernels.hpp"
#include "kernels.hpp"
#include "step_kernels.hpp"

namespace {

thread_local std::string g_err;

struct XfdError : std::runtime_error {
  int code;
  XfdError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define XFD_HIP(x)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess)                                                                   \
      throw XfdError(e_ == hipErrorOutOfMemory ? XFD_ENOMEM : XFD_EDEVICE,                  \
                     std::string(#x) + ": " + hipGetErrorString(e_));                       \
  } while (0)

void require(bool ok, const std::string& msg) {
  if (!ok) throw XfdError(XFD_EINVAL, msg);
}

template <class F>
int guard(F&& f) {
  try {
    f();
    return XFD_OK;
  } catch (const XfdError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return XFD_EINVAL;
  }
}

// Reference extent of field `id` (src/simulation/simulation.cpp:514-528).
void extentOf(int id, int nx, int ny, int nz, int out[3]) {
  const int n[3] = {nx, ny, nz};
  const int c = id % 3;
  const bool h = id >= 3;
  for (int a = 0; a < 3; ++a) out[a] = h ? (a == c ? n[a] + 1 : n[a]) : (a == c ? n[a] : n[a] + 1);
}

// coefficient slot -> field id it multiplies (self) for extent purposes
int coefField(int slot) { return slot / 3; }

// Distinct (self, dual-A, dual-B) coefficient triples of one component, by
// bit pattern, and the per-cell class id in the padded layout.  A uniform
// grid with a few materials has few classes (the 384^3 random-sphere box:
// about 1.6k per E and 700 per H component), so 2 bytes per cell replace
// three coefficients.  Returns false when a component has more than 65535
// classes (the dense layout is used instead).
struct Triple {
  uint64_t v[3];
  bool operator==(const Triple& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
};
struct TripleHash {
  size_t operator()(const Triple& t) const {
    uint64_t h = t.v[0] * 0x9E3779B97F4A7C15ull;
    h ^= t.v[1] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h ^= t.v[2] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return static_cast<size_t>(h);
  }
};

// Id: uint16_t (at most 65535 classes) or uint32_t (at most max_classes).
template <class Id>
bool classify(const double* A, const double* B, const double* C, const int ext[3], const xfd::Grid& g,
              std::vector<Id>& ids, std::vector<double>& table, size_t max_classes = 65535) {
  const int64_t plane = static_cast<int64_t>(ext[1]) * ext[2];
  const unsigned hw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  const int parts = static_cast<int>(std::min<int64_t>(hw, ext[0]));
  auto key = [&](int64_t p) {
    Triple t;
    std::memcpy(&t.v[0], A + p, 8);
    std::memcpy(&t.v[1], B + p, 8);
    std::memcpy(&t.v[2], C + p, 8);
    return t;
  };
  // pass 1: distinct triples per thread, merged in thread order
  std::vector<std::vector<Triple>> found(parts);
  std::atomic<bool> overflow{false};
  auto rangeOf = [&](int p, int& lo, int& hi) {
    const int step = (ext[0] + parts - 1) / parts;
    lo = p * step;
    hi = std::min(ext[0], lo + step);
  };
  {
    std::vector<std::thread> th;
    for (int p = 0; p < parts; ++p)
      th.emplace_back([&, p] {
        int lo, hi;
        rangeOf(p, lo, hi);
        std::unordered_map<Triple, int, TripleHash> seen;
        Triple last{};
        bool have = false;
        for (int64_t q = lo * plane; q < hi * plane; ++q) {
          Triple t = key(q);
          if (have && t == last) continue;
          last = t;
          have = true;
          if (seen.emplace(t, 0).second) found[p].push_back(t);
          if (seen.size() > max_classes) {
            overflow = true;
            return;
          }
        }
      });
    for (auto& t : th) t.join();
  }
  if (overflow) return false;
  std::unordered_map<Triple, Id, TripleHash> id;
  table.clear();
  for (auto& f : found)
    for (auto& t : f)
      if (id.find(t) == id.end()) {
        if (id.size() >= max_classes) return false;
        id.emplace(t, static_cast<Id>(id.size()));
        for (int q = 0; q < 3; ++q) {
          double d;
          std::memcpy(&d, &t.v[q], 8);
          table.push_back(d);
        }
      }
  // pass 2: ids in the padded layout (padding stays class 0, never read)
  ids.assign(static_cast<size_t>(g.total), 0);
  std::vector<std::thread> th;
  for (int p = 0; p < parts; ++p)
    th.emplace_back([&, p] {
      int lo, hi;
      rangeOf(p, lo, hi);
      Triple last{};
      Id last_id = 0;
      bool have = false;
      for (int64_t i = lo; i < hi; ++i)
        for (int64_t j = 0; j < ext[1]; ++j)
          for (int64_t k = 0; k < ext[2]; ++k) {
            const Triple t = key((i * ext[1] + j) * ext[2] + k);
            if (!have || !(t == last)) {
              last = t;
              last_id = id.at(t);
              have = true;
            }
            ids[(i * g.NY + j) * g.pz + k] = last_id;
          }
    });
  for (auto& t : th) t.join();
  return true;
}

// Set while xfd_batch_create builds its cases.
thread_local bool g_batch_create = false;

enum KernelClass {
  K_UPDATE_H = 0, K_UPDATE_E, K_CPML_H, K_CPML_E, K_REC_H, K_REC_E, K_GATHER, K_NF2FF, K_STEP, K_COUNT
};
const char* kKernelNames[K_COUNT] = {"update_h",  "update_e", "cpml_h", "cpml_e", "records_h",
                                     "records_e", "gather",   "nf2ff",  "step"};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    reset();
    bytes = b;
    if (b == 0) return;
    XFD_HIP(hipMalloc(&p, b));
    XFD_HIP(hipMemset(p, 0, b));
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace

struct xfd_ctx {
  int dtype = XFD_F64;
  xfd::Grid g{};
  hipStream_t stream = nullptr;
  size_t elem = 8;
  int ichunk = 16;
  DevBuf fields[6];
  DevBuf coefs[18];   // dense coefficient layout
  bool classed = false;
  // wide class ids (classed too): uint32 ids, when a component has more
  // than 65535 distinct triples
  bool wide = false;
  DevBuf ids[6];      // classed layout: uint16 (uint32 when wide) class per cell and component
  DevBuf tabs[6];     // and the triple table of each component
  int64_t classes[6] = {0, 0, 0, 0, 0, 0};

  struct Slab {
    int axis, n, e_start, h_start;
    int lo_e[3], lo_h[3], ext[3] = {0, 0, 0}, ext_h[3] = {0, 0, 0};
    int extA_e[2], extB_e[2], extA_h[2], extB_h[2];
    DevBuf cpsi[4], psi[4];  // ea_hb, eb_ha, ha_eb, hb_ea (box layout)
    DevBuf psi_h2[2];        // the single-pass step's second set of ha_eb, hb_ea
    DevBuf ae, be, ah, bh;
  };
  std::vector<std::unique_ptr<Slab>> slabs;

  DevBuf pool;
  int64_t pool_rows = 0;  // smallest table row count: steps must stay below it
  DevBuf rec_e, seg_e, rec_h, seg_h;
  DevBuf state;  // inductor currents, one per inductor edge
  // dispersive materials (ADE): scheme or -1, per-cell material ids,
  // per-material coefficient table, pole currents, previous currents and
  // previous E (full padded arrays per component and pole)
  // NF2FF running DFTs: surfaces, phase tables, FP64 accumulators
  struct Nf2ff {
    xfd::Nf2ffArgs a{};
    DevBuf te, th, acc;
  };
  std::vector<std::unique_ptr<Nf2ff>> nf2ffs;
  // time-domain NF2FFs: per group cell s, ds, previous values, this step's
  // bins and contributions; delay buckets; FP64 potentials
  struct TdNf2ff {
    xfd::TdNf2ffArgs a{};
    int64_t omax = 0;  // largest bucket over the groups (bins reach t + omax + 2)
    DevBuf s, ds, bstart, order, prev_a, prev_b, nn, ca1, ca2, cb1, cb2, pot;
    DevBuf slot, lstart, va, vb;  // fixed-order runs
  };
  std::vector<std::unique_ptr<TdNf2ff>> tdnfs;
  int ade_scheme = -1;
  int ade_poles = 0;
  DevBuf ade_mat, ade_tab, ade_j[3], ade_jp[3], ade_eprev[3];
  int64_t nseg_e = 0, nseg_h = 0;

  struct Gather {
    xfd_gather_desc d;
    xfd::GatherArgs a;
    DevBuf out;
    int probe = -1;  // row in the batched probe buffer, or -1
    // Staged recording (movie-size frames): frames go to a ring of two
    // device slots and each is copied to pinned host memory on the copy
    // stream as soon as it is gathered; a slot is reused once its previous
    // frame's copy is done.
    bool staged = false;
    double* host = nullptr;  // count x frame_elems, pinned
    hipEvent_t copied[2] = {nullptr, nullptr};
    hipEvent_t gathered = nullptr;
    ~Gather() {
      for (auto e : copied)
        if (e) (void)hipEventDestroy(e);
      if (gathered) (void)hipEventDestroy(gathered);
      if (host) (void)hipHostFree(host);
    }
  };
  hipStream_t copy_stream = nullptr;  // staged frames' device-to-host copies
  std::vector<std::unique_ptr<Gather>> gathers;
  // weighted-sum recordings (voltage / current monitors): one launch per step
  DevBuf line_terms, line_off, line_scale, line_out;
  int n_line = 0;
  int64_t line_count = 0;
  // single-cell, every-step recordings share one launch per step
  DevBuf probe_list, probe_out;
  int nprobe = 0;
  int64_t probe_count = 0;

  // CPML applied inside the Yee sweeps (see cpmlFused): the slab arguments
  // of each half-step as device arrays
  // grid dimension: 3, 2 (TE, nz = 1) or 1 (TEM, nx = ny = 1)
  int dim = 3;
  bool te_zero = false;  // 2D: Ex, Ey and Hz stay zero (k_te_h<..., true>)
  int device = 0;  // the HIP device the context lives on
  // The single-pass step (k_step): a second set of the six fields and of
  // the H-side CPML psi; `parity` says which set holds the current state
  // (0: fields / psi, 1: fields_b / psi_h2).
  bool fstep = false;
  int fstep_chunk = 4;
  // every CPML slab applied inside the two Yee sweeps (kernel-argument slab
  // list), the far grid planes by k_cpml_faces
  bool allslabs = false;
  int parity = 0;
  DevBuf fields_b[6];
  bool fused = false;
  int n_fused = 0;  // slabs [0, n_fused) run inside the sweeps, the rest separately
  // the z slabs run inside the sweeps (ZCpml), the rest separately
  bool zfused = false;
  std::vector<char> zslab;  // per slab: applied inside the sweeps
  DevBuf cpml_list_e, cpml_list_h;

  // checkpoints of a forward run for the adjoint (xfd_forward_checkpointed)
  DevBuf ckpt;
  int64_t ck_interval = 0, ck_nt = 0, ck_bytes = 0;

  bool prof = false;
  uint32_t prof_mask = 0;  // kernel classes timed when prof
  int64_t launches[K_COUNT] = {};
  double total_ms[K_COUNT] = {};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> event_pool;

  ~xfd_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    gathers.clear();  // their events and pinned buffers before the streams
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    for (auto& p : pending) {
      event_pool.push_back(p.second.first);
      event_pool.push_back(p.second.second);
    }
    for (auto e : event_pool) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int64_t deviceBytes() const {
    int64_t b = 0;
    for (auto& f : fields) b += f.bytes;
    for (auto& c : coefs) b += c.bytes;
    for (int q = 0; q < 6; ++q) b += ids[q].bytes + tabs[q].bytes;
    b += probe_list.bytes + probe_out.bytes;
    for (auto& s : slabs) {
      for (auto& x : s->cpsi) b += x.bytes;
      for (auto& x : s->psi) b += x.bytes;
      b += s->ae.bytes + s->be.bytes + s->ah.bytes + s->bh.bytes;
    }
    b += pool.bytes + rec_e.bytes + seg_e.bytes + rec_h.bytes + seg_h.bytes + state.bytes;
    b += ade_mat.bytes + ade_tab.bytes;
    for (auto& n : nf2ffs) b += n->te.bytes + n->th.bytes + n->acc.bytes;
    for (auto& n : tdnfs)
      for (DevBuf* x : {&n->s, &n->ds, &n->bstart, &n->order, &n->prev_a, &n->prev_b, &n->nn, &n->ca1, &n->ca2,
                        &n->cb1, &n->cb2, &n->pot, &n->slot, &n->lstart, &n->va, &n->vb})
        b += x->bytes;
    for (int q = 0; q < 3; ++q) b += ade_j[q].bytes + ade_jp[q].bytes + ade_eprev[q].bytes;
    for (auto& gth : gathers) b += gth->out.bytes;
    b += ckpt.bytes;
    b += line_terms.bytes + line_off.bytes + line_scale.bytes + line_out.bytes;
    return b;
  }

  hipEvent_t takeEvent() {
    if (!event_pool.empty()) {
      auto e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    XFD_HIP(hipEventCreate(&e));
    return e;
  }

  template <class F>
  void timed(int cls, F&& launch) {
    if (!prof || !(prof_mask & (1u << cls))) {
      launch();
      return;
    }
    hipEvent_t a = takeEvent(), b = takeEvent();
    XFD_HIP(hipEventRecord(a, stream));
    launch();
    XFD_HIP(hipEventRecord(b, stream));
    pending.push_back({cls, {a, b}});
  }

  void drainEvents() {
    for (auto& p : pending) {
      float ms = 0;
      XFD_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
      launches[p.first] += 1;
      total_ms[p.first] += ms;
      event_pool.push_back(p.second.first);
      event_pool.push_back(p.second.second);
    }
    pending.clear();
  }
};

namespace {

// The set of field arrays holding the current state (set 1 only after an
// odd number of single-pass steps).
DevBuf* curFields(xfd_ctx& c) { return c.parity ? c.fields_b : c.fields; }

template <class T>
xfd::FieldPtrs<T> fieldPtrs(xfd_ctx& c) {
  xfd::FieldPtrs<T> F;
  for (int i = 0; i < 6; ++i) F.f[i] = curFields(c)[i].as<T>();
  return F;
}

template <class T>
xfd::FieldPtrs<T> otherFieldPtrs(xfd_ctx& c) {
  xfd::FieldPtrs<T> F;
  for (int i = 0; i < 6; ++i) F.f[i] = (c.parity ? c.fields : c.fields_b)[i].as<T>();
  return F;
}

// Brings the current state back to set 0 (fields, psi) and leaves the
// single-pass step: for entry points that work on set 0 directly
// (checkpoints, the adjoint, writes of whole fields, batches).
void leaveStep(xfd_ctx& c) {
  if (!c.fstep) return;
  if (c.parity) {
    for (int f = 0; f < 6; ++f)
      XFD_HIP(hipMemcpyAsync(c.fields[f].p, c.fields_b[f].p, c.fields[f].bytes, hipMemcpyDeviceToDevice, c.stream));
    for (auto& sl : c.slabs)
      for (int q = 0; q < 2; ++q)
        XFD_HIP(hipMemcpyAsync(sl->psi[2 + q].p, sl->psi_h2[q].p, sl->psi[2 + q].bytes, hipMemcpyDeviceToDevice,
                               c.stream));
    XFD_HIP(hipStreamSynchronize(c.stream));
  }
  c.parity = 0;
  c.fstep = false;
  for (auto& f : c.fields_b) f.reset();
  for (auto& sl : c.slabs)
    for (auto& b : sl->psi_h2) b.reset();
}

// The dispersive update's arguments (zero when the context has no ADE).
template <class T>
xfd::AdeArgs<T> adeArgs(xfd_ctx& c) {
  xfd::AdeArgs<T> A{};
  if (c.ade_scheme < 0) return A;
  A.poles = c.ade_poles;
  A.stride = 2 + 7 * c.ade_poles;
  A.total = c.g.total;
  A.mat = c.ade_mat.as<int16_t>();
  A.tab = c.ade_tab.as<T>();
  for (int q = 0; q < 3; ++q) {
    A.j[q] = c.ade_j[q].as<T>();
    A.jp[q] = c.ade_jp[q].as<T>();
    A.eprev[q] = c.ade_eprev[q].as<T>();
  }
  return A;
}

template <class T>
xfd::WideCoef<T> wideCoef(xfd_ctx& c) {
  xfd::WideCoef<T> M{};
  for (int q = 0; q < 6; ++q) {
    M.id[q] = c.ids[q].as<uint32_t>();
    M.tab[q] = c.tabs[q].as<T>();
  }
  return M;
}

template <class T>
xfd::CoefPtrs<T> coefPtrs(xfd_ctx& c) {
  xfd::CoefPtrs<T> C;
  for (int i = 0; i < 18; ++i) C.c[i] = c.coefs[i].as<T>();
  return C;
}

// Uploads a double array of the reference extent into a padded device array.
template <class T>
void uploadPadded(xfd_ctx& c, const double* host, const int ext[3], T* dst, DevBuf& staging) {
  const size_t n = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  XFD_HIP(hipMemcpyAsync(staging.p, host, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  xfd::Launch<T>::unpack(c.stream, c.g, staging.as<double>(), dst, ext[0], ext[1], ext[2]);
  XFD_HIP(hipStreamSynchronize(c.stream));
}

template <class T>
void uploadVector(const std::vector<T>& v, DevBuf& b) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) XFD_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

int64_t padIndex(const xfd::Grid& g, int64_t i, int64_t j, int64_t k) { return (i * g.NY + j) * g.pz + k; }

// Expands the TFSF corrector loops (tfsf_corrector.h:89-229, call order
// tfsf_corrector.cpp:177-232) into per-cell records for one half-step.
void tfsfRecords(const xfd::Grid& g, const xfd_tfsf_desc& d, int64_t e_off, int64_t h_off, bool is_h,
                 std::vector<xfd::Record>& out) {
  const int64_t G0[3] = {d.origin[0], d.origin[1], d.origin[2]};
  const int64_t G1[3] = {d.origin[0] + d.size[0], d.origin[1] + d.size[1], d.origin[2] + d.size[2]};
  const double* pint[3] = {d.px_int, d.py_int, d.pz_int};
  const double* phalf[3] = {d.px_half, d.py_half, d.pz_half};
  const int64_t nint[3] = {d.size[0] + 1, d.size[1] + 1, d.size[2] + 1};
  const int64_t nhalf[3] = {d.size[0] + 2, d.size[1] + 2, d.size[2] + 2};

  // incident component `comp` of E (inc_h = false) or H at global (i,j,k):
  // projection sum, then truncating index (tfsf_corrector.cpp:16-122)
  auto lookup = [&](bool inc_h, int comp, int64_t p[3], xfd::Record& r) {
    double proj = 0;
    for (int a = 0; a < 3; ++a) {
      int64_t rel = p[a] - G0[a];
      // E_comp uses the half projection on its own axis, H_comp on the two
      // other axes; a half projection is indexed one past the integer one.
      bool half = inc_h ? (a != comp) : (a == comp);
      int64_t idx = half ? rel + 1 : rel;
      require(idx >= 0 && idx < (half ? nhalf[a] : nint[a]), "TFSF projection index out of range");
      double v = half ? phalf[a][idx] : pint[a][idx];
      proj = a == 0 ? v : proj + v;
    }
    if (inc_h) proj = proj - 0.5;
    require(proj >= 0, "TFSF projection is negative");
    const uint64_t idx = static_cast<uint64_t>(proj);
    const double w = proj - static_cast<double>(idx);
    const int64_t cols = inc_h ? d.aux - 1 : d.aux;
    require(static_cast<int64_t>(idx) + 1 < cols, "TFSF incident index exceeds the auxiliary line");
    r.w = w;
    r.omw = 1 - w;
    r.tr = inc_h ? d.transform_h[comp] : d.transform_e[comp];
    r.pool = (inc_h ? h_off : e_off) + static_cast<int64_t>(idx);
    r.row = cols;
    r.kind = 0;
  };

  struct Call {
    int dir;  // 0..5 = XN XP YN YP ZN ZP
    int comp;
  };
  static const Call e_calls[12] = {{0, 2}, {1, 2}, {2, 2}, {3, 2}, {4, 1}, {5, 1},
                                   {0, 1}, {1, 1}, {4, 0}, {5, 0}, {2, 0}, {3, 0}};
  static const Call h_calls[12] = {{0, 1}, {1, 1}, {2, 0}, {3, 0}, {4, 0}, {5, 0},
                                   {0, 2}, {1, 2}, {2, 2}, {3, 2}, {4, 1}, {5, 1}};
  // TFSF2DCorrector (tfsf_corrector.cpp:153-175): the four x/y faces, Ez
  // and Hy / Hx; TFSF1DCorrector (:124-151): the ZN (forward) or ZP face,
  // Ex and Hy
  static const Call e_calls_2d[4] = {{0, 2}, {1, 2}, {2, 2}, {3, 2}};
  static const Call h_calls_2d[4] = {{0, 1}, {1, 1}, {2, 0}, {3, 0}};
  const Call e_call_1d[1] = {{d.forward ? 4 : 5, 0}};
  const Call h_call_1d[1] = {{d.forward ? 4 : 5, 1}};
  const Call* calls = is_h ? h_calls : e_calls;
  int n_calls = 12;
  if (d.kind == 2) {
    calls = is_h ? h_calls_2d : e_calls_2d;
    n_calls = 4;
  } else if (d.kind == 1) {
    calls = is_h ? h_call_1d : e_call_1d;
    n_calls = 1;
  } else {
    require(d.kind == 0 || d.kind == 3, "bad TFSF kind");
  }
  for (int ci = 0; ci < n_calls; ++ci) {
    const int dir = calls[ci].dir, fcomp = calls[ci].comp;
    const int xyz = dir / 2;
    const bool neg = dir % 2 == 0;
    const int da = (xyz + 1) % 3, db = (xyz + 2) % 3;
    const int dual_comp = fcomp == da ? db : da;
    // task in abc coordinates, extended on the node's far end
    int64_t lo[3] = {G0[0], G0[1], G0[2]}, hi[3] = {G1[0], G1[1], G1[2]};
    const int offset_a = ((!is_h && fcomp == db) || (is_h && fcomp == da)) ? 1 : 0;
    if (offset_a == 1) hi[da] += 1;
    else hi[db] += 1;
    const int64_t cs = neg ? G0[xyz] : G1[xyz];
    lo[xyz] = cs;
    hi[xyz] = cs + 1;
    const int flag = (is_h ? -1 : 1) * (fcomp == da ? -1 : 1) * (neg ? -1 : 1);
    const double coef = (is_h ? d.cb[xyz] : d.ca[xyz]) * flag;
    const int field = (is_h ? 3 : 0) + fcomp;
    for (int64_t i = lo[0]; i < hi[0]; ++i)
      for (int64_t j = lo[1]; j < hi[1]; ++j)
        for (int64_t k = lo[2]; k < hi[2]; ++k) {
          int64_t p[3] = {i, j, k};
          int64_t tgt[3] = {i, j, k};
          int64_t src[3] = {i, j, k};
          if (neg) {
            if (!is_h) src[xyz] -= 1;  // E: incident H one cell outside
            else tgt[xyz] -= 1;        // H: the outside H is corrected
          }
          (void)p;
          xfd::Record r{};
          lookup(!is_h, dual_comp, src, r);
          r.coef = coef;
          r.field = field;
          int ext[3];
          extentOf(field, g.nx, g.ny, g.nz, ext);
          require(tgt[0] >= 0 && tgt[1] >= 0 && tgt[2] >= 0 && tgt[0] < ext[0] && tgt[1] < ext[1] &&
                      tgt[2] < ext[2],
                  "TFSF box touches the grid edge");
          r.target = padIndex(g, tgt[0], tgt[1], tgt[2]);
          out.push_back(r);
        }
  }
}

// Orders records by target (stable, keeping the corrector order per cell)
// and returns segment starts.
void segment(std::vector<xfd::Record>& rec, std::vector<int64_t>& seg) {
  std::stable_sort(rec.begin(), rec.end(), [](const xfd::Record& a, const xfd::Record& b) {
    return a.field != b.field ? a.field < b.field : a.target < b.target;
  });
  seg.clear();
  for (size_t r = 0; r < rec.size(); ++r)
    if (r == 0 || rec[r].field != rec[r - 1].field || rec[r].target != rec[r - 1].target)
      seg.push_back(static_cast<int64_t>(r));
  seg.push_back(static_cast<int64_t>(rec.size()));
}

// True when a record (TFSF or lumped correction) writes a component that a
// CPML slab of the same half-step also corrects at that point: the
// reference applies records before CPML there, so such a context keeps the
// separate CPML launches.
bool recordsMeetSlabs(const xfd_ctx& c, const std::vector<xfd::Record>& recs, bool is_e) {
  for (const auto& r : recs) {
    const int64_t i = r.target / c.g.plane, rem = r.target % c.g.plane;
    const int64_t p[3] = {i, rem / c.g.pz, rem % c.g.pz};
    for (const auto& sl : c.slabs) {
      const int* lo = is_e ? sl->lo_e : sl->lo_h;
      bool in = true;
      for (int q = 0; q < 3; ++q) in = in && p[q] >= lo[q] && p[q] < lo[q] + sl->ext[q];
      if (in && r.field % 3 != sl->axis) return true;
    }
  }
  return false;
}

template <class T>
xfd::SlabArgs<T> slabArgs(xfd_ctx::Slab& s, bool is_e);

template <class T>
void createTyped(xfd_ctx& c, const xfd_desc& d) {
  const xfd::Grid& g = c.g;
  for (auto& f : c.fields) f.alloc(g.total * sizeof(T));
  // coefficients
  size_t max_ext = 0;
  for (int s = 0; s < 18; ++s) {
    int ext[3];
    extentOf(coefField(s), g.nx, g.ny, g.nz, ext);
    max_ext = std::max(max_ext, static_cast<size_t>(ext[0]) * ext[1] * ext[2]);
  }
  for (int s = 0; s < 18; ++s) require(d.coeff[s] != nullptr, "coefficient array missing");
  // Classed layout with 16-bit ids; when a component has more than 65535
  // distinct triples, 32-bit ids for every component (3D grids) while each
  // component's distinct triples are at most a quarter of its cells;
  // otherwise, or with XFD_COEF_MODE=dense, the 18-array dense layout.
  // XFD_COEF_MODE=classed rules out the 32-bit ids.
  const char* mode = std::getenv("XFD_COEF_MODE");
  const bool force_dense = mode && std::string(mode) == "dense";
  const bool no_wide = mode && std::string(mode) == "classed";
  c.classed = !force_dense;
  c.wide = false;
  for (int pass = 0; pass < 2 && c.classed; ++pass) {
    const bool wide = pass == 1;
    if (wide && (no_wide || c.dim != 3)) {
      c.classed = false;
      break;
    }
    bool ok = true;
    for (int q = 0; q < 6 && ok; ++q) {
      int ext[3];
      extentOf(q, g.nx, g.ny, g.nz, ext);
      std::vector<double> tab;
      if (wide) {
        std::vector<uint32_t> ids;
        const size_t cells = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
        ok = classify(d.coeff[3 * q], d.coeff[3 * q + 1], d.coeff[3 * q + 2], ext, g, ids, tab, cells / 4);
        if (ok) uploadVector(ids, c.ids[q]);
      } else {
        std::vector<uint16_t> ids;
        ok = classify(d.coeff[3 * q], d.coeff[3 * q + 1], d.coeff[3 * q + 2], ext, g, ids, tab);
        if (ok) uploadVector(ids, c.ids[q]);
      }
      if (!ok) break;
      std::vector<T> tt(tab.begin(), tab.end());
      uploadVector(tt, c.tabs[q]);
      c.classes[q] = static_cast<int64_t>(tab.size() / 3);
    }
    if (ok) {
      c.wide = wide;
      break;
    }
    if (wide) c.classed = false;
  }
  if (!c.classed) {
    for (int q = 0; q < 6; ++q) {
      c.ids[q].reset();
      c.tabs[q].reset();
      c.classes[q] = 0;
    }
    DevBuf staging;
    staging.alloc(max_ext * sizeof(double));
    for (int s3 = 0; s3 < 18; ++s3) {
      int ext[3];
      extentOf(coefField(s3), g.nx, g.ny, g.nz, ext);
      c.coefs[s3].alloc(g.total * sizeof(T));
      uploadPadded<T>(c, d.coeff[s3], ext, c.coefs[s3].as<T>(), staging);
    }
  }

  // CPML slabs
  const int N[3] = {g.nx, g.ny, g.nz};
  for (int si = 0; si < d.n_cpml; ++si) {
    const xfd_cpml_desc& p = d.cpml[si];
    require(p.axis >= 0 && p.axis < 3 && p.n > 0, "bad CPML slab");
    const int cax = p.axis, a = (cax + 1) % 3, b = (cax + 2) % 3;
    // E reads the dual H one plane below its first corrected plane, H reads
    // the dual E one plane above its last
    require(p.e_start >= 1 && p.e_start + p.n <= N[cax] && p.h_start >= 0 && p.h_start + p.n <= N[cax],
            "CPML slab outside the grid");
    auto s = std::make_unique<xfd_ctx::Slab>();
    s->axis = cax;
    s->n = p.n;
    s->e_start = p.e_start;
    s->h_start = p.h_start;
    for (int q = 0; q < 3; ++q) {
      s->lo_e[q] = 0;
      s->lo_h[q] = 0;
      s->ext[q] = N[q] + 1;
    }
    s->lo_e[cax] = p.e_start;
    s->lo_h[cax] = p.h_start;
    s->ext[cax] = p.n;
    // pml_corrector.h:219-243 / 257-282 / 316-339 / 353-376
    s->extA_e[0] = N[a];
    s->extA_e[1] = N[b] + 1;
    s->extB_e[0] = N[a] + 1;
    s->extB_e[1] = N[b];
    s->extA_h[0] = N[a] + 1;
    s->extA_h[1] = N[b];
    s->extB_h[0] = N[a];
    s->extB_h[1] = N[b] + 1;
    const int64_t vol = static_cast<int64_t>(s->ext[0]) * s->ext[1] * s->ext[2];
    const double* src[4] = {p.c_ea_psi_hb, p.c_eb_psi_ha, p.c_ha_psi_eb, p.c_hb_psi_ea};
    // psi extents in (a, b) per array
    const int pe[4][2] = {{N[a], N[b] + 1}, {N[a] + 1, N[b]}, {N[a] + 1, N[b]}, {N[a], N[b] + 1}};
    for (int q = 0; q < 4; ++q) {
      require(src[q] != nullptr, "CPML c_psi array missing");
      std::vector<T> box(vol, T(0));
      // reference psi arrays are xyz-ordered with the layer index on the
      // main axis (pml.cpp:297-337)
      int ex[3];
      ex[cax] = p.n;
      ex[a] = pe[q][0];
      ex[b] = pe[q][1];
      for (int64_t x = 0; x < ex[0]; ++x)
        for (int64_t y = 0; y < ex[1]; ++y)
          for (int64_t z = 0; z < ex[2]; ++z) {
            const int64_t lin = (x * ex[1] + y) * ex[2] + z;
            // box-local: the main axis starts at the slab start in both boxes
            const int64_t t = (x * s->ext[1] + y) * s->ext[2] + z;
            box[t] = static_cast<T>(src[q][lin]);
          }
      uploadVector(box, s->cpsi[q]);
      s->psi[q].alloc(vol * sizeof(T));
    }
    auto up = [&](const double* v, DevBuf& b) {
      require(v != nullptr, "CPML profile missing");
      std::vector<T> t(p.n);
      for (int q = 0; q < p.n; ++q) t[q] = static_cast<T>(v[q]);
      uploadVector(t, b);
    };
    up(p.a_e, s->ae);
    up(p.b_e, s->be);
    up(p.a_h, s->ah);
    up(p.b_h, s->bh);
    c.slabs.push_back(std::move(s));
  }

  // table pool: lumped waveforms, then TFSF e_inc / h_inc
  std::vector<double> pool;
  std::vector<int64_t> lumped_off(d.n_lumped), e_off(d.n_tfsf), h_off(d.n_tfsf);
  int64_t min_rows = INT64_MAX;
  for (int li = 0; li < d.n_lumped; ++li) {
    const auto& L = d.lumped[li];
    require(L.kind == XFD_LUMPED_SOURCE || L.kind == XFD_LUMPED_INDUCTOR, "bad lumped element kind");
    if (L.kind != XFD_LUMPED_SOURCE) continue;
    require(L.waveform != nullptr && L.nt > 0, "lumped waveform missing");
    lumped_off[li] = static_cast<int64_t>(pool.size());
    pool.insert(pool.end(), L.waveform, L.waveform + L.nt);
    min_rows = std::min(min_rows, L.nt);
  }
  for (int ti = 0; ti < d.n_tfsf; ++ti) {
    const auto& Tf = d.tfsf[ti];
    require(Tf.e_inc && Tf.h_inc && Tf.nt > 0 && Tf.aux > 1, "TFSF tables missing");
    e_off[ti] = static_cast<int64_t>(pool.size());
    pool.insert(pool.end(), Tf.e_inc, Tf.e_inc + Tf.nt * Tf.aux);
    h_off[ti] = static_cast<int64_t>(pool.size());
    pool.insert(pool.end(), Tf.h_inc, Tf.h_inc + Tf.nt * (Tf.aux - 1));
    min_rows = std::min(min_rows, Tf.nt);
  }
  c.pool_rows = pool.empty() ? INT64_MAX : min_rows;
  uploadVector(pool, c.pool);

  std::vector<xfd::Record> re, rh;
  int64_t nstate = 0;
  for (int li = 0; li < d.n_lumped; ++li) {  // object correctors first
    const auto& L = d.lumped[li];
    require(L.component >= 0 && L.component < 3 && L.coff != nullptr, "bad lumped element");
    const bool inductor = L.kind == XFD_LUMPED_INDUCTOR;
    require(!inductor || L.coff_j != nullptr, "inductor cjcec missing");
    int ext[3];
    extentOf(L.component, g.nx, g.ny, g.nz, ext);
    for (int q = 0; q < 3; ++q)
      require(L.origin[q] >= 0 && L.size[q] >= 0 && L.origin[q] + L.size[q] <= ext[q],
              "lumped element outside the field extent");
    for (int64_t i = 0; i < L.size[0]; ++i)
      for (int64_t j = 0; j < L.size[1]; ++j)
        for (int64_t k = 0; k < L.size[2]; ++k) {
          xfd::Record r{};
          const int64_t q = (i * L.size[1] + j) * L.size[2] + k;
          r.field = L.component;
          r.target = padIndex(g, L.origin[0] + i, L.origin[1] + j, L.origin[2] + k);
          r.coef = L.coff[q];
          if (inductor) {
            r.kind = 2;
            r.w = L.coff_j[q];
            r.pool = nstate++;  // this edge's J slot
            r.row = 0;
          } else {
            r.kind = 1;
            r.pool = lumped_off[li];
            r.row = 1;
          }
          re.push_back(r);
        }
  }
  for (int ti = 0; ti < d.n_tfsf; ++ti) {
    tfsfRecords(g, d.tfsf[ti], e_off[ti], h_off[ti], false, re);
    tfsfRecords(g, d.tfsf[ti], e_off[ti], h_off[ti], true, rh);
  }
  if (nstate) {  // inductor currents start at zero
    std::vector<T> z(static_cast<size_t>(nstate), T(0));
    uploadVector(z, c.state);
  }
  std::vector<int64_t> se, sh;
  segment(re, se);
  segment(rh, sh);
  uploadVector(re, c.rec_e);
  uploadVector(se, c.seg_e);
  uploadVector(rh, c.rec_h);
  uploadVector(sh, c.seg_h);
  c.nseg_e = re.empty() ? 0 : static_cast<int64_t>(se.size()) - 1;
  c.nseg_h = rh.empty() ? 0 : static_cast<int64_t>(sh.size()) - 1;
  const bool records_meet_slabs = recordsMeetSlabs(c, re, true) || recordsMeetSlabs(c, rh, false);
  // A 2D grid whose Ex, Ey and Hz no correction drives keeps them at zero
  // (their CPML psi are driven by each other only), so its H half-step can
  // leave out Hz and the Ex / Ey terms (adding a zero changes no value).
  c.te_zero = c.dim == 2;
  for (const auto& r : re) c.te_zero = c.te_zero && r.field != 0 && r.field != 1;
  for (const auto& r : rh) c.te_zero = c.te_zero && r.field != 5;

  require(!d.ade || c.dim == 3, "ADE materials need a 3D grid");
  require((d.n_nf2ff == 0 && d.n_tdnf == 0) || c.dim == 3, "NF2FF needs a 3D grid");
  if (d.ade) {
    const auto& A = *d.ade;
    require(A.scheme >= XFD_ADE_DRUDE && A.scheme <= XFD_ADE_MLORENTZ, "bad ADE scheme");
    require(A.poles > 0 && A.n_materials > 0 && A.n_materials <= 32767, "bad ADE pole or material count");
    require(A.coeff != nullptr && A.cell_material != nullptr, "ADE arrays missing");
    const int stride = 2 + 7 * A.poles;
    std::vector<T> tab(static_cast<size_t>(A.n_materials) * stride);
    for (size_t q = 0; q < tab.size(); ++q) tab[q] = static_cast<T>(A.coeff[q]);
    uploadVector(tab, c.ade_tab);
    std::vector<int16_t> mat(static_cast<size_t>(g.total), int16_t(-1));
    for (int64_t i = 0; i < g.nx; ++i)
      for (int64_t j = 0; j < g.ny; ++j)
        for (int64_t k = 0; k < g.nz; ++k) {
          const int16_t m = A.cell_material[(i * g.ny + j) * g.nz + k];
          require(m >= -1 && m < A.n_materials, "ADE cell material out of range");
          mat[padIndex(g, i, j, k)] = m;
        }
    uploadVector(mat, c.ade_mat);
    const bool lor = A.scheme == XFD_ADE_MLORENTZ;
    for (int q = 0; q < 3; ++q) {
      c.ade_j[q].alloc(static_cast<size_t>(A.poles) * g.total * sizeof(T));
      if (lor) {
        c.ade_jp[q].alloc(static_cast<size_t>(A.poles) * g.total * sizeof(T));
        c.ade_eprev[q].alloc(static_cast<size_t>(g.total) * sizeof(T));
      }
    }
    c.ade_scheme = A.scheme;
    c.ade_poles = A.poles;
  }

  for (int ni = 0; ni < d.n_nf2ff; ++ni) {
    const auto& N = d.nf2ff[ni];
    require(N.nf > 0 && N.nt > 0 && N.te != nullptr && N.th != nullptr, "NF2FF tables missing");
    auto nn = std::make_unique<xfd_ctx::Nf2ff>();
    const int n3[3] = {g.nx, g.ny, g.nz};
    nn->a.off[0] = 0;
    for (int s6 = 0; s6 < 6; ++s6) {
      int64_t cells = 1;
      for (int q = 0; q < 3; ++q) {
        const int lo = N.lo[s6][q], hi = N.hi[s6][q];
        // the face-centre stencils reach one cell down and one up
        require(lo >= 1 && lo <= hi && hi <= n3[q], "NF2FF surface outside the grid interior");
        nn->a.lo[s6][q] = lo;
        nn->a.ext[s6][q] = hi - lo;
        cells *= hi - lo;
      }
      nn->a.off[s6 + 1] = nn->a.off[s6] + cells;
    }
    require(N.repeat >= 1, "NF2FF repeat must be at least 1");
    nn->a.nf = N.nf;
    nn->a.repeat = N.repeat;
    nn->a.nt = N.nt;
    const size_t tab = static_cast<size_t>(N.nf) * N.nt * 2;
    uploadVector(std::vector<double>(N.te, N.te + tab), nn->te);
    uploadVector(std::vector<double>(N.th, N.th + tab), nn->th);
    nn->acc.alloc(static_cast<size_t>(N.nf) * 4 * nn->a.off[6] * 2 * sizeof(double));
    nn->a.te = nn->te.as<double2>();
    nn->a.th = nn->th.as<double2>();
    nn->a.acc = nn->acc.as<double2>();
    c.pool_rows = std::min(c.pool_rows, N.nt);
    c.nf2ffs.push_back(std::move(nn));
  }

  for (int ni = 0; ni < d.n_tdnf; ++ni) {
    const auto& N = d.tdnf[ni];
    require(N.s_e && N.s_h && N.ds && N.num > 0, "time-domain NF2FF arrays missing");
    auto td = std::make_unique<xfd_ctx::TdNf2ff>();
    auto& A = td->a;
    const int n3[3] = {g.nx, g.ny, g.nz};
    int64_t cells[6];
    int64_t total = 0;
    for (int s6 = 0; s6 < 6; ++s6) {
      cells[s6] = 1;
      for (int q = 0; q < 3; ++q) {
        const int lo = N.lo[s6][q], hi = N.hi[s6][q];
        require(lo >= 1 && lo <= hi && hi <= n3[q], "NF2FF surface outside the grid interior");
        A.lo[s6][q] = lo;
        A.ext[s6][q] = hi - lo;
        cells[s6] *= hi - lo;
      }
      total += cells[s6];
    }
    // groups g = 2 d + attr, each a copy of surface d's cells
    std::vector<double> s(2 * total), ds(2 * total);
    A.off[0] = 0;
    int64_t src = 0;
    for (int s6 = 0; s6 < 6; ++s6) {
      for (int at = 0; at < 2; ++at) {
        const int gr = 2 * s6 + at;
        A.off[gr + 1] = A.off[gr] + cells[s6];
        for (int64_t q = 0; q < cells[s6]; ++q) {
          s[A.off[gr] + q] = (at == 0 ? N.s_e : N.s_h)[src + q];
          ds[A.off[gr] + q] = N.ds[src + q];
        }
      }
      src += cells[s6];
    }
    // fixed order: no s + t of the run (t < num) can round up to the next
    // integer when 1 - frac(s) exceeds an ulp of s + num, so every cell's
    // bin is floor(s) + t
    double smax = 0;
    for (double v : s) smax = std::max(smax, v);
    const double ulp = std::ldexp(1.0, std::ilogb(smax + static_cast<double>(N.num)) - 52);
    bool fixed = std::getenv("XFD_TDNF_MERGE") == nullptr;  // =1: always the merging bins
    for (double v : s) fixed = fixed && (std::floor(v) + 1 - v) > 2 * ulp;
    A.fixed = fixed ? 1 : 0;
    // delay buckets floor(s), cells sorted by (bucket, cell) per group
    std::vector<int64_t> bstart, order;
    std::vector<int64_t> slot(2 * A.off[12], 0), lstart;  // two per group cell
    A.boff[0] = 0;
    A.bin_off[0] = 0;
    for (int gr = 0; gr < 12; ++gr) {
      const int64_t c0 = A.off[gr], nc = A.off[gr + 1] - c0;
      int64_t omin = 0, omax = -1;
      std::vector<int64_t> o(nc);
      for (int64_t q = 0; q < nc; ++q) {
        require(s[c0 + q] >= 0 && s[c0 + q] < 1e15, "time-domain NF2FF delay out of range");
        o[q] = static_cast<int64_t>(std::floor(s[c0 + q]));
        omin = q == 0 ? o[q] : std::min(omin, o[q]);
        omax = q == 0 ? o[q] : std::max(omax, o[q]);
      }
      const int64_t nb = nc ? omax - omin + 1 : 0;
      A.omin[gr] = omin;
      A.nbk[gr] = nb;
      td->omax = std::max(td->omax, omax);
      std::vector<int64_t> cnt(nb + 1, 0);
      for (int64_t q = 0; q < nc; ++q) ++cnt[o[q] - omin + 1];
      const int64_t base = static_cast<int64_t>(order.size());
      for (int64_t b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
      for (int64_t b = 0; b <= nb; ++b) bstart.push_back(base + cnt[b]);
      A.boff[gr + 1] = static_cast<int64_t>(bstart.size());
      order.resize(base + nc);
      std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
      for (int64_t q = 0; q < nc; ++q) order[base + fill[o[q] - omin]++] = q;  // ascending q within a bucket
      // fixed-order runs: bin L merges bucket L (first contributions) and
      // bucket L - 1 (second) by cell
      A.loff[gr] = static_cast<int64_t>(lstart.size());
      int64_t pos = 2 * c0;
      for (int64_t L = 0; nc && L <= nb; ++L) {
        lstart.push_back(pos);
        int64_t p1 = L < nb ? bstart[A.boff[gr] + L] : 0, e1 = L < nb ? bstart[A.boff[gr] + L + 1] : 0;
        int64_t p2 = L > 0 ? bstart[A.boff[gr] + L - 1] : 0, e2 = L > 0 ? bstart[A.boff[gr] + L] : 0;
        while (p1 < e1 || p2 < e2) {
          const bool first = p2 >= e2 || (p1 < e1 && order[p1] < order[p2]);
          const int64_t q = first ? order[p1++] : order[p2++];
          slot[2 * (c0 + q) + (first ? 0 : 1)] = pos++;
        }
      }
      if (nc) lstart.push_back(pos);
      A.bin_off[gr + 1] = A.bin_off[gr] + (nc ? nb + (fixed ? 1 : 2) : 0);
    }
    A.num = N.num;
    uploadVector(slot, td->slot);
    uploadVector(lstart.empty() ? std::vector<int64_t>{0} : lstart, td->lstart);
    td->va.alloc(static_cast<size_t>(std::max<int64_t>(2 * A.off[12], 1)) * sizeof(double));
    td->vb.alloc(static_cast<size_t>(std::max<int64_t>(2 * A.off[12], 1)) * sizeof(double));
    A.slot = td->slot.as<int64_t>();
    A.lstart = td->lstart.as<int64_t>();
    A.va = td->va.as<double>();
    A.vb = td->vb.as<double>();
    uploadVector(s, td->s);
    uploadVector(ds, td->ds);
    uploadVector(bstart, td->bstart);
    uploadVector(order, td->order);
    const size_t nb = static_cast<size_t>(std::max<int64_t>(A.off[12], 1));
    for (DevBuf* x : {&td->prev_a, &td->prev_b, &td->ca1, &td->ca2, &td->cb1, &td->cb2}) {
      x->alloc(nb * sizeof(double));
      XFD_HIP(hipMemsetAsync(x->p, 0, x->bytes, c.stream));
    }
    td->nn.alloc(nb * sizeof(int64_t));
    td->pot.alloc(static_cast<size_t>(24) * N.num * sizeof(double));
    XFD_HIP(hipMemsetAsync(td->pot.p, 0, td->pot.bytes, c.stream));
    A.s = td->s.as<double>();
    A.ds = td->ds.as<double>();
    A.bstart = td->bstart.as<int64_t>();
    A.order = td->order.as<int64_t>();
    A.prev_a = td->prev_a.as<double>();
    A.prev_b = td->prev_b.as<double>();
    A.nn = td->nn.as<int64_t>();
    A.ca1 = td->ca1.as<double>();
    A.ca2 = td->ca2.as<double>();
    A.cb1 = td->cb1.as<double>();
    A.cb2 = td->cb2.as<double>();
    A.pot = td->pot.as<double>();
    c.tdnfs.push_back(std::move(td));
  }

  for (int gi = 0; gi < d.n_gather; ++gi) {
    const auto& G = d.gather[gi];
    require(G.field >= 0 && G.field < 6 && G.interval > 0 && G.count >= 0, "bad recording");
    int ext[3];
    extentOf(G.field, g.nx, g.ny, g.nz, ext);
    for (int q = 0; q < 3; ++q)
      require(G.origin[q] >= 0 && G.origin[q] <= G.end[q] && G.end[q] <= ext[q], "recording box outside field");
    auto gg = std::make_unique<xfd_ctx::Gather>();
    gg->d = G;
    gg->a.field = G.field;
    for (int q = 0; q < 3; ++q) {
      gg->a.o[q] = G.origin[q];
      gg->a.e[q] = G.end[q];
    }
    gg->a.frame_elems = static_cast<int64_t>(G.end[0] - G.origin[0]) * (G.end[1] - G.origin[1]) *
                        (G.end[2] - G.origin[2]);
    const bool probe = gg->a.frame_elems == 1 && G.first == 0 && G.interval == 1 && G.count > 0 &&
                       (c.nprobe == 0 || G.count == c.probe_count);
    // Movies (four or more frames of 4 MiB or more): staged to the host at
    // their cadence, so device memory holds two frames whatever the run
    // length, and the copies overlap the stepping (XFD_STAGE=0 keeps every
    // frame on the device, =1 stages any recording of several frames);
    // batched cases record on the device.
    const char* senv = std::getenv("XFD_STAGE");
    const size_t fbytes = static_cast<size_t>(gg->a.frame_elems) * sizeof(double);
    const bool stage = !g_batch_create && !probe && G.count >= 2 && fbytes > 0 &&
                       ((senv && senv[0] == '1') || (!(senv && senv[0] == '0') && G.count >= 4 &&
                                                     fbytes >= (size_t(4) << 20)));
    if (probe) {
      gg->probe = c.nprobe++;
      c.probe_count = G.count;
    } else if (stage) {
      gg->staged = true;
      gg->out.alloc(2 * fbytes);
      XFD_HIP(hipHostMalloc(reinterpret_cast<void**>(&gg->host), fbytes * static_cast<size_t>(G.count)));
      std::memset(gg->host, 0, fbytes * static_cast<size_t>(G.count));
      for (auto& e : gg->copied) XFD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      XFD_HIP(hipEventCreateWithFlags(&gg->gathered, hipEventDisableTiming));
      if (!c.copy_stream) XFD_HIP(hipStreamCreate(&c.copy_stream));
    } else {
      gg->out.alloc(static_cast<size_t>(gg->a.frame_elems * G.count) * sizeof(double));
    }
    c.gathers.push_back(std::move(gg));
  }
  if (c.nprobe) {
    std::vector<xfd::Probe> pl(c.nprobe);
    for (auto& gg : c.gathers)
      if (gg->probe >= 0)
        pl[gg->probe] = {padIndex(g, gg->a.o[0], gg->a.o[1], gg->a.o[2]), gg->a.field, 0};
    uploadVector(pl, c.probe_list);
    c.probe_out.alloc(static_cast<size_t>(c.nprobe * c.probe_count) * sizeof(double));
  }
  if (d.n_line > 0) {  // weighted-sum recordings
    require(d.line != nullptr, "line recordings missing");
    std::vector<xfd::LineTerm> terms;
    std::vector<int64_t> off{0};
    std::vector<double> scale;
    int64_t count = -1;
    for (int m = 0; m < d.n_line; ++m) {
      const auto& L = d.line[m];
      require(L.n_terms >= 0 && (L.n_terms == 0 || (L.terms && L.weights)), "bad line recording");
      require(count < 0 || L.count == count, "line recordings share one length");
      count = L.count;
      for (int q = 0; q < L.n_terms; ++q) {
        const int32_t* tq = L.terms + 5 * q;
        require(tq[0] >= 0 && tq[0] < 6 && (tq[4] == 0 || tq[4] == 1), "bad line term");
        int ext[3];
        extentOf(tq[0], g.nx, g.ny, g.nz, ext);
        require(tq[1] >= 0 && tq[2] >= 0 && tq[3] >= 0 && tq[1] < ext[0] && tq[2] < ext[1] && tq[3] < ext[2],
                "line term outside the field");
        terms.push_back({padIndex(g, tq[1], tq[2], tq[3]), L.weights[q], tq[0], tq[4]});
      }
      off.push_back(static_cast<int64_t>(terms.size()));
      scale.push_back(L.scale);
    }
    uploadVector(terms, c.line_terms);
    uploadVector(off, c.line_off);
    uploadVector(scale, c.line_scale);
    c.n_line = d.n_line;
    c.line_count = std::max<int64_t>(0, count);
    c.line_out.alloc(static_cast<size_t>(c.n_line * c.line_count) * sizeof(double));
  }
  {  // CPML inside the sweeps when XFD_CPML_FUSED=1, without ADE or records in a slab.
    // Off by default: measured on MI355X the separate slab launches are faster
    // (384^3 step FP32 1.192 ms vs 1.454 ms fused, FP64 2.246 vs 2.352 ms),
    // because the fused sweep reads the slab parameters per cell on the slab
    // tiles.
    const char* env = std::getenv("XFD_CPML_FUSED");
    const bool off = !(env && env[0] == '1');
    // Only x and y slabs: their membership is uniform over a wavefront (one i,
    // one j row), while a z slab splits the 64 k-lanes of a wave and would
    // make it run both paths.  The fused slabs are the leading x/y slabs in
    // boundary order, so the slabs' relative order is kept.
    int prefix = 0;
    while (prefix < static_cast<int>(c.slabs.size()) && c.slabs[prefix]->axis != 2) ++prefix;
    c.fused = !off && prefix > 0 && c.ade_scheme < 0 && !records_meet_slabs && c.dim == 3;
    c.n_fused = c.fused ? prefix : 0;
    if (c.fused) {
      const int64_t stride[3] = {g.plane, g.pz, 1};
      std::vector<xfd::SlabArgs<T>> le, lh;
      for (int n = 0; n < prefix; ++n) {
        auto& sl = c.slabs[n];
        le.push_back(slabArgs<T>(*sl, true));
        lh.push_back(slabArgs<T>(*sl, false));
        le.back().stride_c = stride[le.back().axis];
        lh.back().stride_c = stride[lh.back().axis];
      }
      uploadVector(le, c.cpml_list_e);
      uploadVector(lh, c.cpml_list_h);
    }
    // The z slabs inside the sweeps.  A z slab holds 8-12 cells of each
    // 128-byte k-row, so a separate launch moves 2-2.5x its bytes in partial
    // lines, while inside the sweep the field and dual values are already in
    // registers and only psi and c_psi are read.  The slab lanes' extra work
    // slows the wavefronts that carry them, so it pays only when few do:
    // measured, the 384^3 box (6 k-tiles per row) steps 1.192 ms with it and
    // 1.244 ms without, while the batched design sweep (nz 48-160, 1-3
    // k-tiles) steps 3.66 ms with it and 3.41 ms without.  FP32 applies the
    // whole slab in the sweep (its corrections then precede the x/y slabs'
    // where they overlap: FP32 rounding only); default on rows of at least 3
    // k-tiles (FP32 ms per step, separate / inside: n = 128 0.109 / 0.125,
    // 160 0.163 / 0.149, 192 0.201 / 0.198, 384 ~1.18 / 1.12) and 16-bit
    // class ids or dense coefficients; with 32-bit class ids the sweeps lose
    // more than the z launches cost (160^3 adjoint design run, 400 steps:
    // sweeps 37.5 -> 43.5 us, 206.7 ms separate against 214.0 inside).  FP64 keeps the reference's order of corrections: the sweeps
    // apply a z slab only where no x/y slab corrects the point, and the frame
    // where they overlap follows the x/y slabs at the z slab's place
    // (k_cpml_zrest), bitwise; the frame's launch costs more than the fusion
    // saves (384^3: 2.40 ms against 2.26 ms), so FP64 takes it only when
    // forced.  XFD_CPML_ZFUSED=1 forces it, =0 turns it off.  Never together
    // with XFD_CPML_FUSED=1.
    const char* zenv = std::getenv("XFD_CPML_ZFUSED");
    const bool zoff = zenv && zenv[0] == '0';
    const bool zforce = zenv && zenv[0] == '1';
    const bool zwide = (g.nz + xfd::kLanesK - 1) / xfd::kLanesK >= 3;
    std::vector<int> zq;
    for (size_t q = 0; q < c.slabs.size(); ++q)
      if (c.slabs[q]->axis == 2) zq.push_back(static_cast<int>(q));
    bool disjoint = true;
    if (zq.size() == 2) {
      const auto& s0 = *c.slabs[zq[0]];
      const auto& s1 = *c.slabs[zq[1]];
      disjoint = (s0.e_start + s0.n <= s1.e_start || s1.e_start + s1.n <= s0.e_start) &&
                 (s0.h_start + s0.n <= s1.h_start || s1.h_start + s1.n <= s0.h_start);
    }
    int per_axis[2] = {0, 0};
    for (auto& sl : c.slabs)
      if (sl->axis < 2) ++per_axis[sl->axis];
    c.zfused = !zoff && (zforce || (zwide && c.elem == 4 && !c.wide)) && off && !zq.empty() && zq.size() <= 2 && disjoint &&
               per_axis[0] <= 4 && per_axis[1] <= 4 && c.ade_scheme < 0 && !records_meet_slabs && c.dim == 3;
    c.zslab.assign(c.slabs.size(), 0);
    if (c.zfused)
      for (int q : zq) c.zslab[q] = 1;
  }
  {  // The single-pass step (k_step) where every correction between the H and
    // E half-steps is a CPML slab: no TFSF H records, no ADE, no record on a
    // point a slab corrects or on a grid boundary plane (the pass leaves the
    // far planes alone), at most kMaxStepSlabs slabs.  Opt-in with
    // XFD_STEP_FUSED=1 (measured slower than the two sweeps on the 384^3 box);
    // XFD_STEP_CHUNK sets the i-planes per work-item.
    const char* env = std::getenv("XFD_STEP_FUSED");
    const char* fenv = std::getenv("XFD_CPML_FUSED");
    const char* zenv1 = std::getenv("XFD_CPML_ZFUSED");
    // an explicit request for slabs inside the two sweeps keeps the sweeps
    const bool off = !(env && env[0] == '1') || (fenv && fenv[0] == '1') || (zenv1 && zenv1[0] == '1');
    bool edge = false;
    for (const auto& r : re) {
      const int64_t i = r.target / g.plane, rem = r.target % g.plane, j = rem / g.pz, k = rem % g.pz;
      edge = edge || i == 0 || i >= g.nx || j == 0 || j >= g.ny || k == 0 || k >= g.nz;
    }
    c.fstep = !off && c.dim == 3 && c.ade_scheme < 0 && c.nseg_h == 0 && !records_meet_slabs && !edge &&
              !c.wide && c.slabs.size() <= static_cast<size_t>(xfd::kMaxStepSlabs);
    if (c.fstep) {
      c.fused = false;
      c.n_fused = 0;
      c.zfused = false;
      std::fill(c.zslab.begin(), c.zslab.end(), 0);
      if (const char* e = std::getenv("XFD_STEP_CHUNK")) {
        const int v = std::atoi(e);
        require(v >= 1 && v <= xfd::kMaxStepChunk, "XFD_STEP_CHUNK out of range");
        c.fstep_chunk = v;
      }
      for (int f = 0; f < 6; ++f) c.fields_b[f].alloc(c.fields[f].bytes);
      for (auto& sl : c.slabs)
        for (int q = 0; q < 2; ++q) sl->psi_h2[q].alloc(sl->psi[2 + q].bytes);
      c.parity = 0;
    }
  }
  {  // Every slab inside the two sweeps (boundary order per point, so FP64
    // stays bitwise; the far planes follow in k_cpml_faces), for 3D grids
    // without ADE, without records on points a slab corrects and with at
    // most kMaxStepSlabs slabs.  Opt-in with XFD_CPML_INSWEEP=1: measured on
    // the 384^3 box it is slower than the separate slab launches (FP32 1.41
    // against 1.24 ms per step, FP64 2.27 against 2.24 ms).
    const char* env = std::getenv("XFD_CPML_INSWEEP");
    const char* fenv = std::getenv("XFD_CPML_FUSED");
    const char* zenv1 = std::getenv("XFD_CPML_ZFUSED");
    const bool other = (fenv && fenv[0] == '1') || (zenv1 && zenv1[0] == '1');
    c.allslabs = (env && env[0] == '1') && !other && !c.fstep && c.dim == 3 && c.ade_scheme < 0 &&
                 !records_meet_slabs && !c.slabs.empty() && c.slabs.size() <= static_cast<size_t>(xfd::kMaxStepSlabs);
    if (c.allslabs) {
      c.fused = false;
      c.n_fused = 0;
      c.zfused = false;
      std::fill(c.zslab.begin(), c.zslab.end(), 0);
    }
  }
  XFD_HIP(hipStreamSynchronize(c.stream));
}

// The slabs of one half-step as the sweeps read them: in place, or for the
// single-pass step's H side (pingpong) from the current psi set into the
// other.
template <class T>
xfd::StepSlabs<T> stepSlabs(xfd_ctx& c, bool is_e, bool pingpong = true) {
  xfd::StepSlabs<T> L{};
  for (auto& sp : c.slabs) {
    const xfd::SlabArgs<T> a = slabArgs<T>(*sp, is_e);
    xfd::StepSlab<T>& S = L.s[L.n++];
    S.axis = a.axis;
    S.start = a.start;
    for (int q = 0; q < 3; ++q) {
      S.lo[q] = a.lo[q];
      S.ext[q] = a.ext[q];
    }
    for (int q = 0; q < 2; ++q) {
      S.extA[q] = a.extA[q];
      S.extB[q] = a.extB[q];
    }
    S.cpsiA = a.cpsiA;
    S.cpsiB = a.cpsiB;
    S.ca = a.ca;
    S.cb = a.cb;
    if (is_e || !pingpong) {
      S.rdA = S.wrA = a.psiA;
      S.rdB = S.wrB = a.psiB;
    } else {
      T* cur[2] = {sp->psi[2].as<T>(), sp->psi[3].as<T>()};
      T* alt[2] = {sp->psi_h2[0].as<T>(), sp->psi_h2[1].as<T>()};
      S.rdA = c.parity ? alt[0] : cur[0];
      S.rdB = c.parity ? alt[1] : cur[1];
      S.wrA = c.parity ? cur[0] : alt[0];
      S.wrB = c.parity ? cur[1] : alt[1];
    }
  }
  return L;
}

template <class T>
xfd::SlabArgs<T> slabArgs(xfd_ctx::Slab& s, bool is_e) {
  xfd::SlabArgs<T> a{};
  const int c = s.axis, ta = (c + 1) % 3, tb = (c + 2) % 3;
  a.axis = c;
  a.n = s.n;
  a.start = is_e ? s.e_start : s.h_start;
  for (int q = 0; q < 3; ++q) {
    a.lo[q] = is_e ? s.lo_e[q] : s.lo_h[q];
    a.ext[q] = s.ext[q];
  }
  if (is_e) {
    a.extA[0] = s.extA_e[0];
    a.extA[1] = s.extA_e[1];
    a.extB[0] = s.extB_e[0];
    a.extB[1] = s.extB_e[1];
    a.fieldA = ta;
    a.fieldB = tb;
    a.dualA = 3 + ta;
    a.dualB = 3 + tb;
    a.cpsiA = s.cpsi[0].as<T>();
    a.cpsiB = s.cpsi[1].as<T>();
    a.psiA = s.psi[0].as<T>();
    a.psiB = s.psi[1].as<T>();
    a.ca = s.ae.as<T>();
    a.cb = s.be.as<T>();
  } else {
    a.extA[0] = s.extA_h[0];
    a.extA[1] = s.extA_h[1];
    a.extB[0] = s.extB_h[0];
    a.extB[1] = s.extB_h[1];
    a.fieldA = 3 + ta;
    a.fieldB = 3 + tb;
    a.dualA = ta;
    a.dualB = tb;
    a.cpsiA = s.cpsi[2].as<T>();
    a.cpsiB = s.cpsi[3].as<T>();
    a.psiA = s.psi[2].as<T>();
    a.psiB = s.psi[3].as<T>();
    a.ca = s.ah.as<T>();
    a.cb = s.bh.as<T>();
  }
  return a;
}

// The z slabs a context applies inside its FP32 sweeps, per half-step.  The
// sweep's z path assumes a slab spanning the whole x/y range whose two
// components cover every sweep point of its layers; the slabs built from an
// xfd_cpml_desc always do (box lo 0 and extent N + 1 across, component
// extents N or N + 1), and this checks it.
template <class T>
xfd::ZSlab<T> zSlab(const xfd::SlabArgs<T>& a, const xfd::Grid& g) {
  require(a.axis == 2 && a.lo[0] == 0 && a.lo[1] == 0 && a.ext[0] == g.nx + 1 && a.ext[1] == g.ny + 1 &&
              a.extA[0] >= g.nx && a.extA[1] >= g.ny && a.extB[0] >= g.nx && a.extB[1] >= g.ny &&
              a.lo[2] == a.start && a.ext[2] == a.n,
          "z CPML slab shape");
  require(static_cast<int64_t>(a.ext[0]) * a.ext[1] * a.ext[2] < (int64_t{1} << 31), "z CPML slab too large");
  xfd::ZSlab<T> z{};
  z.klo = a.start;
  z.khi = a.start + a.n;
  z.ext1 = a.ext[1];
  z.ext2 = a.ext[2];
  z.cpsiA = a.cpsiA;
  z.cpsiB = a.cpsiB;
  z.psiA = a.psiA;
  z.psiB = a.psiB;
  z.ca = a.ca;
  z.cb = a.cb;
  return z;
}

template <class T>
void zCpml(xfd_ctx& c, xfd::ZCpml<T>& ze, xfd::ZCpml<T>& zh) {
  ze = {};
  zh = {};
  for (size_t q = 0; q < c.slabs.size(); ++q) {
    if (c.zslab[q]) {
      ze.z[ze.n++] = zSlab<T>(slabArgs<T>(*c.slabs[q], true), c.g);
      zh.z[zh.n++] = zSlab<T>(slabArgs<T>(*c.slabs[q], false), c.g);
      continue;
    }
    const int ax = c.slabs[q]->axis;  // FP64, an x or y slab: its range joins the frame
    if (ax == 2 || sizeof(T) == 4) continue;
    for (int is_e = 0; is_e < 2; ++is_e) {
      const auto a = slabArgs<T>(*c.slabs[q], is_e != 0);
      xfd::ZFrame& fr = is_e ? ze.fr : zh.fr;
      int& n = ax == 0 ? fr.nx : fr.ny;
      require(n < 4, "more than four CPML slabs along one axis");
      auto& r = ax == 0 ? fr.xr[n] : fr.yr[n];
      r[0] = a.lo[ax];
      r[1] = a.lo[ax] + a.ext[ax];
      ++n;
    }
  }
}

template <class T>
void stepTyped(xfd_ctx& c, int64_t t0, int64_t n, bool record = true) {
  require(t0 >= 0 && n >= 0, "bad step range");
  require(t0 + n <= c.pool_rows, "step beyond the source tables (run length given at init)");
  auto F = fieldPtrs<T>(c);
  const auto C = coefPtrs<T>(c);
  xfd::ClassCoef<T> K{};
  for (int q = 0; q < 6; ++q) {
    K.id[q] = c.ids[q].as<uint16_t>();
    K.tab[q] = c.tabs[q].as<T>();
  }
  const auto M = wideCoef<T>(c);
  const xfd::AdeArgs<T> A = adeArgs<T>(c);
  std::vector<xfd::SlabArgs<T>> se, sh;
  for (auto& s : c.slabs) {
    se.push_back(slabArgs<T>(*s, true));
    sh.push_back(slabArgs<T>(*s, false));
  }
  const int64_t stride[3] = {c.g.plane, c.g.pz, 1};
  for (auto& a : se) a.stride_c = stride[a.axis];
  for (auto& a : sh) a.stride_c = stride[a.axis];
  hipStream_t st = c.stream;
  xfd::CpmlList<T> PE{}, PH{};
  if (c.fused) {
    PE = {c.cpml_list_e.as<xfd::SlabArgs<T>>(), c.n_fused};
    PH = {c.cpml_list_h.as<xfd::SlabArgs<T>>(), c.n_fused};
  }
  xfd::ZCpml<T> ZE{}, ZH{};
  if (c.zfused) zCpml<T>(c, ZE, ZH);
  xfd::StepSlabs<T> AH{}, AE{};
  if (c.allslabs) {
    AH = stepSlabs<T>(c, false, false);
    AE = stepSlabs<T>(c, true, false);
  }
  auto inSweep = [&](size_t q) { return c.allslabs || static_cast<int>(q) < c.n_fused; };
  // Slabs q and q + 1 share one launch when both run separately, lie on one
  // axis and share no point (XFD_CPML_PAIR=0: one launch per slab).
  const char* pair_env = std::getenv("XFD_CPML_PAIR");
  const bool pair_on = !(pair_env && pair_env[0] == '0');
  auto separate = [&](size_t q) { return !inSweep(q) && !(c.zfused && c.zslab[q]); };
  auto pairs = [&](const std::vector<xfd::SlabArgs<T>>& v, size_t q) {
    if (!pair_on || q + 1 >= v.size() || !separate(q) || !separate(q + 1)) return false;
    const auto &a = v[q], &b = v[q + 1];
    if (a.axis != b.axis) return false;
    const int x = a.axis;
    return a.lo[x] + a.ext[x] <= b.lo[x] || b.lo[x] + b.ext[x] <= a.lo[x];
  };
  for (int64_t t = t0; t < t0 + n; ++t) {
    if (c.fstep) {
      // H, its CPML, E and its CPML in one pass from the current set into
      // the other, which then becomes current
      const auto FB = otherFieldPtrs<T>(c);
      const auto SH = stepSlabs<T>(c, false), SE = stepSlabs<T>(c, true);
      c.timed(K_STEP, [&] {
        if (c.classed)
          xfd::StepLaunch<T>::step(st, c.g, F, FB, K, SH, SE, c.fstep_chunk);
        else
          xfd::StepLaunch<T>::step(st, c.g, F, FB, C, SH, SE, c.fstep_chunk);
      });
      c.parity ^= 1;
      F = FB;
    } else {
    c.timed(K_UPDATE_H, [&] {
      if (c.dim != 3) {
        if (c.classed)
          xfd::Launch<T>::update_low(st, c.g, F, K, c.dim, true, c.te_zero);
        else
          xfd::Launch<T>::update_low(st, c.g, F, C, c.dim, true, c.te_zero);
      } else if (c.wide) {
        xfd::Launch<T>::update_h(st, c.g, F, M, c.ichunk, PH, ZH, AH);
      } else if (c.classed) {
        xfd::Launch<T>::update_h(st, c.g, F, K, c.ichunk, PH, ZH, AH);
      } else {
        xfd::Launch<T>::update_h(st, c.g, F, C, c.ichunk, PH, ZH, AH);
      }
    });
    if (c.nseg_h)
      c.timed(K_REC_H, [&] {
        xfd::Launch<T>::records(st, F, c.rec_h.as<xfd::Record>(), c.seg_h.as<int64_t>(), c.nseg_h,
                                c.pool.as<double>(), c.state.as<T>(), t);
      });
    for (size_t q = 0; q < sh.size(); ++q)
      c.timed(K_CPML_H, [&] {
        if (pairs(sh, q)) {
          xfd::Launch<T>::cpml2(st, c.g, F, sh[q], sh[q + 1], false);
          ++q;
        } else if (c.zfused && c.zslab[q] && sizeof(T) == 8)
          xfd::Launch<T>::cpml_zrest(st, c.g, F, sh[q], ZH.fr, false);
        else if (inSweep(q) || (c.zfused && c.zslab[q]))
          xfd::Launch<T>::cpml_faces(st, c.g, F, sh[q], false);
        else
          xfd::Launch<T>::cpml(st, c.g, F, sh[q], false);
      });
    c.timed(K_UPDATE_E, [&] {
      if (c.dim != 3) {
        if (c.classed)
          xfd::Launch<T>::update_low(st, c.g, F, K, c.dim, false);
        else
          xfd::Launch<T>::update_low(st, c.g, F, C, c.dim, false);
      } else if (c.ade_scheme >= 0) {
        if (c.wide)
          xfd::Launch<T>::update_e_ade(st, c.g, F, M, A, c.ade_scheme, c.ichunk);
        else if (c.classed)
          xfd::Launch<T>::update_e_ade(st, c.g, F, K, A, c.ade_scheme, c.ichunk);
        else
          xfd::Launch<T>::update_e_ade(st, c.g, F, C, A, c.ade_scheme, c.ichunk);
      } else if (c.wide) {
        xfd::Launch<T>::update_e(st, c.g, F, M, c.ichunk, PE, ZE, AE);
      } else if (c.classed) {
        xfd::Launch<T>::update_e(st, c.g, F, K, c.ichunk, PE, ZE, AE);
      } else {
        xfd::Launch<T>::update_e(st, c.g, F, C, c.ichunk, PE, ZE, AE);
      }
    });
    }
    if (c.nseg_e)
      c.timed(K_REC_E, [&] {
        xfd::Launch<T>::records(st, F, c.rec_e.as<xfd::Record>(), c.seg_e.as<int64_t>(), c.nseg_e,
                                c.pool.as<double>(), c.state.as<T>(), t);
      });
    for (size_t q = 0; q < se.size() && !c.fstep; ++q)
      c.timed(K_CPML_E, [&] {
        if (pairs(se, q)) {
          xfd::Launch<T>::cpml2(st, c.g, F, se[q], se[q + 1], true);
          ++q;
        } else if (c.zfused && c.zslab[q] && sizeof(T) == 8)
          xfd::Launch<T>::cpml_zrest(st, c.g, F, se[q], ZE.fr, true);
        else if (inSweep(q) || (c.zfused && c.zslab[q]))
          xfd::Launch<T>::cpml_faces(st, c.g, F, se[q], true);
        else
          xfd::Launch<T>::cpml(st, c.g, F, se[q], true);
      });
    if (!record) {
      XFD_HIP(hipGetLastError());
      continue;
    }
    if (c.nprobe && t < c.probe_count)
      c.timed(K_GATHER, [&] {
        xfd::Launch<T>::probes(st, F, c.probe_list.as<xfd::Probe>(), c.nprobe, c.probe_out.as<double>(),
                               c.probe_count, t);
      });
    for (auto& gp : c.gathers) {
      const auto& d = gp->d;
      if (gp->probe >= 0) continue;
      if (t < d.first || (t - d.first) % d.interval != 0) continue;
      const int64_t frame = (t - d.first) / d.interval;
      if (frame >= d.count) continue;
      if (gp->staged) {
        const int slot = static_cast<int>(frame & 1);
        double* ring = gp->out.as<double>() + slot * gp->a.frame_elems;
        const size_t fb = static_cast<size_t>(gp->a.frame_elems) * sizeof(double);
        if (frame >= 2) XFD_HIP(hipStreamWaitEvent(st, gp->copied[slot], 0));  // the slot's last copy is done
        c.timed(K_GATHER, [&] { xfd::Launch<T>::gather(st, c.g, F, gp->a, ring); });
        XFD_HIP(hipEventRecord(gp->gathered, st));
        XFD_HIP(hipStreamWaitEvent(c.copy_stream, gp->gathered, 0));
        XFD_HIP(hipMemcpyAsync(gp->host + frame * gp->a.frame_elems, ring, fb, hipMemcpyDeviceToHost,
                               c.copy_stream));
        XFD_HIP(hipEventRecord(gp->copied[slot], c.copy_stream));
        continue;
      }
      c.timed(K_GATHER, [&] {
        xfd::Launch<T>::gather(st, c.g, F, gp->a, gp->out.as<double>() + frame * gp->a.frame_elems);
      });
    }
    if (c.n_line && t < c.line_count)
      c.timed(K_GATHER, [&] {
        xfd::Launch<T>::lines(st, F, c.line_terms.as<xfd::LineTerm>(), c.line_off.as<int64_t>(),
                              c.line_scale.as<double>(), c.n_line, c.line_out.as<double>(), c.line_count, t);
      });
    for (auto& nn : c.nf2ffs)  // Domain::record: monitors, then NF2FF (domain.cpp:176-184)
      c.timed(K_NF2FF, [&] { xfd::Launch<T>::nf2ff(st, c.g, F, nn->a, t); });
    for (auto& td : c.tdnfs) {
      // TDPlaneData::update writes potential.at(nn + 1): the arrays end
      // before this step's last bin
      require(t + td->omax + 2 < td->a.num, "time-domain NF2FF potential too short for this step");
      c.timed(K_NF2FF, [&] { xfd::Launch<T>::tdNf2ff(st, c.g, F, td->a, t); });
    }
    XFD_HIP(hipGetLastError());
    if (c.prof && c.pending.size() > 4096) {
      XFD_HIP(hipStreamSynchronize(st));
      c.drainEvents();
    }
  }
  XFD_HIP(hipStreamSynchronize(st));
  if (c.prof) c.drainEvents();
}

// Zeroes every piece of time-stepping state of a case: fields, CPML psi,
// inductor currents, ADE state, NF2FF accumulators and recordings.
void resetState(xfd_ctx& c) {
  auto zero = [&](DevBuf& b) {
    if (b.p && b.bytes) XFD_HIP(hipMemsetAsync(b.p, 0, b.bytes, c.stream));
  };
  for (auto& f : c.fields) zero(f);
  for (auto& f : c.fields_b) zero(f);
  for (auto& sl : c.slabs) {
    for (auto& p : sl->psi) zero(p);
    for (auto& p : sl->psi_h2) zero(p);
  }
  c.parity = 0;
  zero(c.state);
  for (int q = 0; q < 3; ++q) {
    zero(c.ade_j[q]);
    zero(c.ade_jp[q]);
    zero(c.ade_eprev[q]);
  }
  for (auto& n : c.nf2ffs) zero(n->acc);
  for (auto& n : c.tdnfs) {
    zero(n->prev_a);
    zero(n->prev_b);
    zero(n->pot);
  }
  if (c.copy_stream) XFD_HIP(hipStreamSynchronize(c.copy_stream));
  for (auto& gg : c.gathers) {
    zero(gg->out);
    if (gg->staged) std::memset(gg->host, 0, static_cast<size_t>(gg->a.frame_elems * gg->d.count) * sizeof(double));
  }
  zero(c.probe_out);
  zero(c.line_out);
  XFD_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace

namespace {

// The arrays that make up the time-stepping state of a context (checkpoints
// save and restore them).
std::vector<DevBuf*> stateBuffers(xfd_ctx& c) {
  std::vector<DevBuf*> v;
  for (auto& f : c.fields) v.push_back(&f);
  for (auto& sl : c.slabs)
    for (auto& p : sl->psi) v.push_back(&p);
  v.push_back(&c.state);
  for (int q = 0; q < 3; ++q) {
    v.push_back(&c.ade_j[q]);
    v.push_back(&c.ade_jp[q]);
    v.push_back(&c.ade_eprev[q]);
  }
  return v;
}

void checkpointCopy(xfd_ctx& c, int64_t slot, bool save) {
  char* base = c.ckpt.as<char>() + slot * c.ck_bytes;
  for (DevBuf* b : stateBuffers(c)) {
    if (!b->bytes) continue;
    if (save)
      XFD_HIP(hipMemcpyAsync(base, b->p, b->bytes, hipMemcpyDeviceToDevice, c.stream));
    else
      XFD_HIP(hipMemcpyAsync(b->p, base, b->bytes, hipMemcpyDeviceToDevice, c.stream));
    base += b->bytes;
  }
}

template <class T>
void forwardCheckpointed(xfd_ctx& c, int64_t nt, int64_t interval) {
  require(nt > 0 && interval > 0, "bad run length or checkpoint interval");
  int64_t bytes = 0;
  for (DevBuf* b : stateBuffers(c)) bytes += static_cast<int64_t>(b->bytes);
  const int64_t n_ck = (nt + interval - 1) / interval;
  c.ckpt.alloc(static_cast<size_t>(bytes * n_ck));
  c.ck_bytes = bytes;
  c.ck_interval = interval;
  c.ck_nt = nt;
  resetState(c);
  for (int64_t k = 0; k < n_ck; ++k) {
    checkpointCopy(c, k, true);
    const int64_t t0 = k * interval;
    stepTyped<T>(c, t0, std::min(interval, nt - t0));
  }
}

// Design edges of the cells [o, o + n): Ex o x (n0, n1+1, n2+1), Ey
// (n0+1, n1, n2+1), Ez (n0+1, n1+1, n2).
xfd::DesignArgs designArgs(const int32_t o[3], const int32_t n[3]) {
  xfd::DesignArgs d{};
  int64_t off = 0;
  for (int q = 0; q < 3; ++q) {
    for (int a = 0; a < 3; ++a) {
      d.o[q][a] = o[a];
      d.e[q][a] = a == q ? n[a] : n[a] + 1;
    }
    d.off[q] = off;
    off += static_cast<int64_t>(d.e[q][0]) * d.e[q][1] * d.e[q][2];
  }
  d.off[3] = off;
  return d;
}

template <class T>
void backwardTyped(xfd_ctx& c, const xfd_adjoint_desc& D, double* grad_out) {
  require(c.ck_nt > 0 && c.ckpt.p, "no checkpointed forward run (xfd_forward_checkpointed)");
  require(D.nt == c.ck_nt, "objective length differs from the forward run");
  require(c.dim == 3, "the adjoint needs a 3D grid");
  for (int a = 0; a < 3; ++a) {
    const int n = a == 0 ? c.g.nx : (a == 1 ? c.g.ny : c.g.nz);
    require(D.size[a] > 0 && D.origin[a] >= 1 && D.origin[a] + D.size[a] <= n - 1, "design box outside the grid");
  }
  require(D.n_points >= 0 && (D.n_points == 0 || (D.points && D.source)), "bad objective points");
  const xfd::DesignArgs d = designArgs(D.origin, D.size);
  const int64_t nd = d.off[3];
  {  // The gradient reads each design edge's E after the Yee update: no E
    // record (lumped element, TFSF) and no dispersive update may touch one.
    std::vector<char> edge(static_cast<size_t>(c.g.total), 0);
    for (int q = 0; q < 3; ++q)
      for (int u = 0; u < d.e[q][0]; ++u)
        for (int v = 0; v < d.e[q][1]; ++v)
          for (int w = 0; w < d.e[q][2]; ++w)
            edge[padIndex(c.g, d.o[q][0] + u, d.o[q][1] + v, d.o[q][2] + w)] |= static_cast<char>(1 << q);
    if (c.nseg_e > 0) {
      std::vector<xfd::Record> rec(c.rec_e.bytes / sizeof(xfd::Record));
      XFD_HIP(hipMemcpy(rec.data(), c.rec_e.p, c.rec_e.bytes, hipMemcpyDeviceToHost));
      for (const auto& r : rec)
        require(r.field > 2 || !(edge[r.target] & (1 << r.field)), "an E record (lumped element or TFSF) touches a design edge");
    }
    for (const auto& sl : c.slabs) {  // CPML E corrections along the slab axis
      const int a = sl->axis;
      require(D.origin[a] + D.size[a] + 1 <= sl->e_start || D.origin[a] >= sl->e_start + sl->n,
              "a CPML slab corrects a design edge");
    }
    if (c.ade_scheme >= 0) {
      std::vector<int16_t> mat(static_cast<size_t>(c.g.total));
      XFD_HIP(hipMemcpy(mat.data(), c.ade_mat.p, c.ade_mat.bytes, hipMemcpyDeviceToHost));
      for (size_t q = 0; q < mat.size(); ++q)
        require(!(mat[q] >= 0 && edge[q]), "a dispersive (ADE) cell holds a design edge");
    }
  }
  require(D.beta && D.gamma && grad_out, "bad design coefficients");
  hipStream_t st = c.stream;
  const auto F = fieldPtrs<T>(c);
  const auto C = coefPtrs<T>(c);
  xfd::ClassCoef<T> K{};
  for (int q = 0; q < 6; ++q) {
    K.id[q] = c.ids[q].as<uint16_t>();
    K.tab[q] = c.tabs[q].as<T>();
  }
  const auto Mx = wideCoef<T>(c);
  // adjoint fields, adjoint psi, history, design coefficients, objective
  DevBuf lam[6];
  xfd::FieldPtrs<T> L;
  for (int q = 0; q < 6; ++q) {
    lam[q].alloc(c.fields[q].bytes);
    L.f[q] = lam[q].as<T>();
  }
  std::vector<std::unique_ptr<DevBuf[]>> mu;
  std::vector<xfd::SlabArgs<T>> me, mh;
  const int64_t stride[3] = {c.g.plane, c.g.pz, 1};
  for (auto& sl : c.slabs) {
    mu.emplace_back(new DevBuf[4]);
    for (int q = 0; q < 4; ++q) mu.back()[q].alloc(sl->psi[q].bytes);
    auto e = slabArgs<T>(*sl, true), h = slabArgs<T>(*sl, false);
    e.stride_c = stride[e.axis];
    h.stride_c = stride[h.axis];
    e.psiA = mu.back()[0].as<T>();
    e.psiB = mu.back()[1].as<T>();
    h.psiA = mu.back()[2].as<T>();
    h.psiB = mu.back()[3].as<T>();
    me.push_back(e);
    mh.push_back(h);
  }
  // adjoint inductor currents; adjoint ADE state (the forward state's
  // layout) and the deferred part of lambdaE on dispersive cells
  DevBuf lam_state, lam_ade[9], ade_x[3];
  xfd::AdeArgs<T> Ad{}, LA{};
  xfd::FieldPtrs<T> X{};
  if (c.state.bytes) {
    lam_state.alloc(c.state.bytes);
    XFD_HIP(hipMemset(lam_state.p, 0, c.state.bytes));
  }
  if (c.ade_scheme >= 0) {
    Ad.poles = LA.poles = c.ade_poles;
    Ad.stride = LA.stride = 2 + 7 * c.ade_poles;
    Ad.total = LA.total = c.g.total;
    Ad.mat = LA.mat = c.ade_mat.as<int16_t>();
    Ad.tab = LA.tab = c.ade_tab.as<T>();
    for (int q = 0; q < 3; ++q) {
      DevBuf* src3[3] = {&c.ade_j[q], &c.ade_jp[q], &c.ade_eprev[q]};
      T** dst3[3] = {&LA.j[q], &LA.jp[q], &LA.eprev[q]};
      for (int w = 0; w < 3; ++w) {
        DevBuf& b = lam_ade[3 * q + w];
        if (src3[w]->bytes) {
          b.alloc(src3[w]->bytes);
          XFD_HIP(hipMemset(b.p, 0, b.bytes));
        }
        *dst3[w] = b.as<T>();
      }
      ade_x[q].alloc(c.fields[q].bytes);
      X.f[q] = ade_x[q].as<T>();
    }
  }
  const int64_t K_int = c.ck_interval;
  DevBuf hist, beta, gamma, grad, pts, src;
  hist.alloc(static_cast<size_t>((K_int + 1) * nd) * sizeof(T));
  uploadVector(std::vector<double>(D.beta, D.beta + nd), beta);
  uploadVector(std::vector<double>(D.gamma, D.gamma + nd), gamma);
  grad.alloc(static_cast<size_t>(nd) * sizeof(double));
  std::vector<xfd::AdjPoint> hp;
  for (int p = 0; p < D.n_points; ++p) {
    const int32_t* q = D.points + 4 * p;
    require(q[0] >= 0 && q[0] < 6, "bad objective field");
    const int ext_ok = q[1] >= 0 && q[2] >= 0 && q[3] >= 0 && q[1] <= c.g.nx && q[2] <= c.g.ny && q[3] <= c.g.nz;
    require(ext_ok, "objective point outside the grid");
    hp.push_back({padIndex(c.g, q[1], q[2], q[3]), q[0], 0});
  }
  uploadVector(hp, pts);
  if (D.n_points)
    uploadVector(std::vector<double>(D.source, D.source + static_cast<int64_t>(D.n_points) * D.nt), src);
  using A = xfd::AdjLaunch<T>;
  // The transposed slabs two per launch (XFD_CPML_PAIR=0: one each): the mu
  // updates of any two slabs; the gathers of two slabs of one axis whose
  // boxes, widened by the gathered plane, share no point.
  const char* pe = std::getenv("XFD_CPML_PAIR");
  const bool pair_on = !(pe && pe[0] == '0');
  auto muAll = [&](const std::vector<xfd::SlabArgs<T>>& v) {
    size_t q = 0;
    for (; pair_on && q + 1 < v.size(); q += 2) A::cpmlMu2(st, c.g, L, v[q], v[q + 1]);
    for (; q < v.size(); ++q) A::cpmlMu(st, c.g, L, v[q]);
  };
  auto gatherAll = [&](const std::vector<xfd::SlabArgs<T>>& v, bool is_e) {
    for (size_t q = 0; q < v.size(); ++q) {
      if (pair_on && q + 1 < v.size() && v[q].axis == v[q + 1].axis) {
        const auto &a = v[q], &b = v[q + 1];
        const int x = a.axis;
        if (a.lo[x] + a.ext[x] + 1 <= b.lo[x] || b.lo[x] + b.ext[x] + 1 <= a.lo[x]) {
          A::cpmlGather2(st, is_e, c.g, L, a, b);
          ++q;
          continue;
        }
      }
      A::cpmlGather(st, is_e, c.g, L, v[q]);
    }
  };
  const int64_t n_ck = (c.ck_nt + K_int - 1) / K_int;
  T* H0 = hist.as<T>();
  for (int64_t seg = n_ck - 1; seg >= 0; --seg) {
    const int64_t t0 = seg * K_int, t1 = std::min(t0 + K_int, c.ck_nt);
    checkpointCopy(c, seg, false);  // recompute the segment, keeping the design edges
    A::hist(st, c.g, F, d, H0);
    for (int64_t t = t0; t < t1; ++t) {
      stepTyped<T>(c, t, 1, false);
      A::hist(st, c.g, F, d, H0 + (t - t0 + 1) * nd);
    }
    for (int64_t t = t1 - 1; t >= t0; --t) {
      A::source(st, L, pts.as<xfd::AdjPoint>(), D.n_points, src.as<double>(), D.nt, t);
      A::grad(st, c.g, L, d, H0 + (t - t0) * nd, H0 + (t - t0 + 1) * nd, beta.as<double>(), gamma.as<double>(),
              grad.as<double>());
      muAll(me);
      if (c.state.bytes)
        A::records(st, L, c.rec_e.as<xfd::Record>(), c.seg_e.as<int64_t>(), c.nseg_e, lam_state.as<T>());
      if (c.ade_scheme >= 0) A::adePre(st, c.g, L, Ad, LA, X, c.ade_scheme);
      if (c.wide)
        A::adjH(st, c.g, L, Mx, c.ichunk);
      else if (c.classed)
        A::adjH(st, c.g, L, K, c.ichunk);
      else
        A::adjH(st, c.g, L, C, c.ichunk);
      gatherAll(me, true);
      muAll(mh);
      if (c.wide)
        A::adjE(st, c.g, L, Mx, c.ichunk);
      else if (c.classed)
        A::adjE(st, c.g, L, K, c.ichunk);
      else
        A::adjE(st, c.g, L, C, c.ichunk);
      if (c.ade_scheme >= 0) A::adePost(st, c.g, L, Ad.mat, X);
      gatherAll(mh, false);
      XFD_HIP(hipGetLastError());
    }
  }
  XFD_HIP(hipMemcpyAsync(grad_out, grad.p, static_cast<size_t>(nd) * sizeof(double), hipMemcpyDeviceToHost, st));
  XFD_HIP(hipStreamSynchronize(st));
}

}  // namespace

// A batch of independent cases stepped by shared persistent launches (see
// batch_kernels.hpp).  The cases are ordinary contexts (the batch owns them),
// so fields and recordings are read through xfd_batch_case.
struct xfd_batch {
  int dtype = XFD_F64;
  std::vector<std::unique_ptr<xfd_ctx>> cases;
  bool shared = true;  // every case class-indexed (16-bit ids): batched launches

  hipStream_t stream = nullptr;
  DevBuf tdev;
  int ichunk = 16;
  int tile = 0;     // sweep tile: index into kBatchTileK / kBatchTileJ
  int grid = 2048;  // work-groups of each persistent launch
  int cus = 256;
  // One table per launch: entries and their first work-group ids.  CPML
  // slabs of one case overlap at edges and corners, where the reference adds
  // their corrections in boundary order, so slab o of every case goes into
  // CPML launch o.
  struct Table {
    DevBuf tab, start;
    int n = 0;
    int64_t total = 0;
  };
  // sweep: every case's sweep (the H sweeps; the E sweeps too when no case
  // is dispersive); sweep_e / sweep_ade: the E sweeps of the other and of the
  // dispersive cases when some case is
  Table sweep, sweep_e, sweep_ade, rec_e, rec_h, nf;
  bool any_ade = false;
  std::vector<Table> cpml_e, cpml_h;            // slabs stepped by separate launches
  std::vector<Table> faces_e, faces_h;          // far planes of slabs fused into the sweeps
  bool any_fused = false;  // some case applies leading slabs inside its sweeps
  int64_t pool_rows = INT64_MAX;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0;
  ~xfd_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void dropGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

namespace {

template <class E>
void fillTable(xfd_batch::Table& t, const std::vector<E>& v, const std::vector<int64_t>& start, int64_t total) {
  uploadVector(v, t.tab);
  uploadVector(start, t.start);
  t.n = static_cast<int>(v.size());
  t.total = total;
}

// Collects entries with their first work-group ids.
template <class E>
struct Collect {
  std::vector<E> v;
  std::vector<int64_t> start;
  int64_t total = 0;
  void add(const E& e, int64_t groups) {
    if (groups <= 0) return;
    v.push_back(e);
    start.push_back(total);
    total += groups;
  }
  void to(xfd_batch::Table& t) const { fillTable(t, v, start, total); }
  void append(const Collect& o) {
    for (size_t i = 0; i < o.v.size(); ++i)
      add(o.v[i], (i + 1 < o.v.size() ? o.start[i + 1] : o.total) - o.start[i]);
  }
};

// Global work-group ids an entry of `items` point-wise work-items owns.
int64_t groupsOf(int64_t items) { return (items + xfd::kBatchGroupItems - 1) / xfd::kBatchGroupItems; }

// Builds the entry tables of every batched launch for the current ichunk.
template <class T>
void batchTables(xfd_batch& b) {
  b.any_fused = false;
  Collect<xfd::BatchCase<T>> sw, swe, swa;
  b.any_ade = false;
  std::vector<Collect<xfd::BatchSlab<T>>> ce, ch, fe, fh;
  Collect<xfd::BatchRec<T>> re, rh;
  Collect<xfd::BatchNf2ff<T>> nf;
  // ordinal o and o + 1 share a launch when, in every case holding both, the
  // two slabs run separately, lie on one axis and share no point
  std::vector<char> pairable;
  for (auto& up : b.cases) {
    const xfd_ctx& c = *up;
    if (pairable.size() < c.slabs.size()) pairable.resize(c.slabs.size(), 1);
    for (size_t o = 0; o + 1 < c.slabs.size(); ++o) {
      const auto &u = *c.slabs[o], &w = *c.slabs[o + 1];
      const bool ok = static_cast<int>(o) >= c.n_fused && u.axis == w.axis &&
                      (u.e_start + u.n <= w.e_start || w.e_start + w.n <= u.e_start) &&
                      (u.h_start + u.n <= w.h_start || w.h_start + w.n <= u.h_start);
      if (!ok) pairable[o] = 0;
    }
  }
  for (auto& up : b.cases) {
    xfd_ctx& c = *up;
    const auto F = fieldPtrs<T>(c);
    xfd::BatchCase<T> bc{};
    bc.g = c.g;
    bc.F = F;
    for (int q = 0; q < 6; ++q) {
      bc.K.id[q] = c.ids[q].as<uint16_t>();
      bc.K.tab[q] = c.tabs[q].as<T>();
    }
    bc.ichunk = b.ichunk;
    if (c.fused) {
      b.any_fused = true;
      bc.ph = {c.cpml_list_h.as<xfd::SlabArgs<T>>(), c.n_fused};
      bc.pe = {c.cpml_list_e.as<xfd::SlabArgs<T>>(), c.n_fused};
    }
    require(!c.zfused, "batched cases keep their z slabs separate");
    const int tk = xfd::kBatchTileK[b.tile], tj = xfd::kBatchTileJ[b.tile];
    bc.nbx = (c.g.nz + tk - 1) / tk;
    bc.nby = (c.g.ny + tj - 1) / tj;
    bc.nbz = (c.g.nx + b.ichunk - 1) / b.ichunk;
    bc.scheme = c.ade_scheme;
    if (c.ade_scheme >= 0) {
      b.any_ade = true;
      bc.A = adeArgs<T>(c);
    }
    const int64_t tiles = static_cast<int64_t>(bc.nbx) * bc.nby * bc.nbz;
    sw.add(bc, tiles);
    (c.ade_scheme >= 0 ? swa : swe).add(bc, tiles);
    const int64_t stride[3] = {c.g.plane, c.g.pz, 1};
    for (size_t o = 0; o < c.slabs.size(); ++o) {
      if (ce.size() <= o) {
        ce.resize(o + 1);
        ch.resize(o + 1);
        fe.resize(o + 1);
        fh.resize(o + 1);
      }
      for (int is_e = 0; is_e < 2; ++is_e) {
        xfd::BatchSlab<T> e{c.g, F, slabArgs<T>(*c.slabs[o], is_e != 0)};
        e.s.stride_c = stride[e.s.axis];
        if (static_cast<int>(o) < c.n_fused) {  // only the far-plane remainder
          const auto fp = xfd::facePieces(c.g, e.s);
          (is_e ? fe : fh)[o].add(e, groupsOf(fp.nX + fp.nY + fp.nZ));
        } else {
          const int64_t vol = static_cast<int64_t>(e.s.ext[0]) * e.s.ext[1] * e.s.ext[2];
          (is_e ? ce : ch)[o].add(e, groupsOf(vol));
        }
      }
    }
    if (c.nseg_h)
      rh.add({F, c.rec_h.as<xfd::Record>(), c.seg_h.as<int64_t>(), c.nseg_h, c.pool.as<double>(), c.state.as<T>()},
             groupsOf(c.nseg_h));
    if (c.nseg_e)
      re.add({F, c.rec_e.as<xfd::Record>(), c.seg_e.as<int64_t>(), c.nseg_e, c.pool.as<double>(), c.state.as<T>()},
             groupsOf(c.nseg_e));
    for (auto& n : c.nf2ffs) nf.add({c.g, F, n->a}, groupsOf(n->a.off[6]));
    b.pool_rows = std::min(b.pool_rows, c.pool_rows);
  }
  sw.to(b.sweep);
  swe.to(b.sweep_e);
  swa.to(b.sweep_ade);
  re.to(b.rec_e);
  rh.to(b.rec_h);
  nf.to(b.nf);
  const char* pair_env = std::getenv("XFD_CPML_PAIR");
  const bool pair_on = !(pair_env && pair_env[0] == '0');
  std::vector<size_t> launch_of(ce.size());
  size_t nl = 0;
  for (size_t o = 0; o < ce.size(); ++o) {
    launch_of[o] = nl++;
    if (pair_on && o + 1 < ce.size() && pairable[o] && fe[o].v.empty() && fh[o].v.empty() && fe[o + 1].v.empty() &&
        fh[o + 1].v.empty()) {
      launch_of[o + 1] = launch_of[o];
      ce[o].append(ce[o + 1]);
      ch[o].append(ch[o + 1]);
      ++o;
    }
  }
  for (auto* v : {&b.cpml_e, &b.cpml_h, &b.faces_e, &b.faces_h}) {
    v->clear();
    v->resize(nl);
  }
  for (size_t o = 0; o < ce.size(); ++o) {
    if (o > 0 && launch_of[o] == launch_of[o - 1]) continue;  // appended to the previous ordinal's launch
    const size_t l = launch_of[o];
    ce[o].to(b.cpml_e[l]);
    ch[o].to(b.cpml_h[l]);
    fe[o].to(b.faces_e[l]);
    fh[o].to(b.faces_h[l]);
  }
}

// One time step of every case, in Domain::run order, reading the step index
// from b.tdev and advancing it at the end.
template <class T>
void batchEnqueueStep(xfd_batch& b) {
  using L = xfd::BatchLaunch<T>;
  hipStream_t st = b.stream;
  const int64_t* t = b.tdev.as<int64_t>();
  auto sweep = [&](bool is_h) {
    if (is_h || !b.any_ade) {
      L::sweep(st, is_h, b.sweep.tab.as<xfd::BatchCase<T>>(), b.sweep.start.as<int64_t>(), b.sweep.n,
               b.sweep.total, b.grid, b.any_fused, b.tile);
      return;
    }
    L::sweep(st, false, b.sweep_e.tab.as<xfd::BatchCase<T>>(), b.sweep_e.start.as<int64_t>(), b.sweep_e.n,
             b.sweep_e.total, b.grid, b.any_fused, b.tile);
    L::sweepAde(st, b.sweep_ade.tab.as<xfd::BatchCase<T>>(), b.sweep_ade.start.as<int64_t>(), b.sweep_ade.n,
                b.sweep_ade.total, b.grid, b.tile);
  };
  auto records = [&](xfd_batch::Table& r) {
    if (r.n) L::records(st, r.tab.as<xfd::BatchRec<T>>(), r.start.as<int64_t>(), r.n, r.total, b.grid, t);
  };
  auto cpml = [&](bool is_e) {  // slab ordinal by ordinal: boundary order within each case
    auto& full = is_e ? b.cpml_e : b.cpml_h;
    auto& faces = is_e ? b.faces_e : b.faces_h;
    for (size_t o = 0; o < full.size(); ++o) {
      auto& f = faces[o];
      if (f.n) L::cpmlFaces(st, is_e, f.tab.as<xfd::BatchSlab<T>>(), f.start.as<int64_t>(), f.n, f.total, b.grid);
      auto& c = full[o];
      if (c.n) L::cpml(st, is_e, c.tab.as<xfd::BatchSlab<T>>(), c.start.as<int64_t>(), c.n, c.total, b.grid);
    }
  };
  sweep(true);
  records(b.rec_h);
  cpml(false);
  sweep(false);
  records(b.rec_e);
  cpml(true);
  for (auto& up : b.cases) {  // recordings, on the device's schedule
    xfd_ctx& c = *up;
    const auto F = fieldPtrs<T>(c);
    if (c.nprobe)
      L::probesAt(st, F, c.probe_list.as<xfd::Probe>(), c.nprobe, c.probe_out.as<double>(), c.probe_count, t);
    for (auto& gp : c.gathers)
      if (gp->probe < 0)
        L::gatherAt(st, c.g, F, gp->a, gp->out.as<double>(), gp->d.first, gp->d.interval, gp->d.count, t);
    if (c.n_line)
      L::linesAt(st, F, c.line_terms.as<xfd::LineTerm>(), c.line_off.as<int64_t>(), c.line_scale.as<double>(),
                 c.n_line, c.line_out.as<double>(), c.line_count, t);
  }
  if (b.nf.n) L::nf2ff(st, b.nf.tab.as<xfd::BatchNf2ff<T>>(), b.nf.start.as<int64_t>(), b.nf.n, b.nf.total, b.grid, t);
  L::tick(st, b.tdev.as<int64_t>());
}

template <class T>
void batchStep(xfd_batch& b, int64_t t0, int64_t n) {
  require(t0 >= 0 && n >= 0, "bad step range");
  if (!b.shared) {  // cases the batched launches do not cover: step them one by one
    for (auto& c : b.cases) stepTyped<T>(*c, t0, n);
    return;
  }
  require(t0 + n <= b.pool_rows, "step beyond the source tables (run length given at init)");
  XFD_HIP(hipMemcpy(b.tdev.p, &t0, sizeof(int64_t), hipMemcpyHostToDevice));
  XFD_HIP(hipEventRecord(b.ev0, b.stream));
  if (b.use_graph) {
    if (!b.exec) {
      XFD_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
      batchEnqueueStep<T>(b);
      XFD_HIP(hipStreamEndCapture(b.stream, &b.graph));
      XFD_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    }
    for (int64_t s = 0; s < n; ++s) XFD_HIP(hipGraphLaunch(b.exec, b.stream));
  } else {
    for (int64_t s = 0; s < n; ++s) batchEnqueueStep<T>(b);
  }
  XFD_HIP(hipGetLastError());
  XFD_HIP(hipEventRecord(b.ev1, b.stream));
  XFD_HIP(hipStreamSynchronize(b.stream));
  XFD_HIP(hipEventElapsedTime(&b.last_ms, b.ev0, b.ev1));
}

// Times `reps` pairs of batched Yee sweeps for each candidate (sweep tile
// shape, i-chunk, work-groups per launch) and keeps the fastest.
template <class T>
void batchTune(xfd_batch& b, int reps) {
  const int chunks[] = {4, 8, 16, 32};
  const int mults[] = {4, 8, 16, 32};
  using L = xfd::BatchLaunch<T>;
  // one step's sweeps and CPML launches (the grid size is every persistent
  // launch's), without records or DFTs: the state only runs through linear
  // updates and is zeroed afterwards
  auto timePair = [&](int n) {
    auto pair = [&] {
      for (bool is_h : {true, false}) {
        if (is_h || !b.any_ade) {
          L::sweep(b.stream, is_h, b.sweep.tab.as<xfd::BatchCase<T>>(), b.sweep.start.as<int64_t>(), b.sweep.n,
                   b.sweep.total, b.grid, b.any_fused, b.tile);
        } else {
          L::sweep(b.stream, false, b.sweep_e.tab.as<xfd::BatchCase<T>>(), b.sweep_e.start.as<int64_t>(),
                   b.sweep_e.n, b.sweep_e.total, b.grid, b.any_fused, b.tile);
          L::sweepAde(b.stream, b.sweep_ade.tab.as<xfd::BatchCase<T>>(), b.sweep_ade.start.as<int64_t>(),
                      b.sweep_ade.n, b.sweep_ade.total, b.grid, b.tile);
        }
        auto& full = is_h ? b.cpml_h : b.cpml_e;
        auto& faces = is_h ? b.faces_h : b.faces_e;
        for (size_t o = 0; o < full.size(); ++o) {
          auto& f = faces[o];
          if (f.n)
            L::cpmlFaces(b.stream, !is_h, f.tab.as<xfd::BatchSlab<T>>(), f.start.as<int64_t>(), f.n, f.total, b.grid);
          auto& c = full[o];
          if (c.n) L::cpml(b.stream, !is_h, c.tab.as<xfd::BatchSlab<T>>(), c.start.as<int64_t>(), c.n, c.total, b.grid);
        }
      }
    };
    pair();  // warm-up
    XFD_HIP(hipEventRecord(b.ev0, b.stream));
    for (int r = 0; r < n; ++r) pair();
    XFD_HIP(hipEventRecord(b.ev1, b.stream));
    XFD_HIP(hipStreamSynchronize(b.stream));
    float ms = 0;
    XFD_HIP(hipEventElapsedTime(&ms, b.ev0, b.ev1));
    return ms;
  };
  struct Cand {
    float ms;
    int tile, ichunk, grid;
  };
  // coarse pass over every candidate, then the four fastest timed again,
  // three times each with twice the repetitions, keeping each one's best
  std::vector<Cand> all;
  for (int tile = 0; tile < xfd::kBatchTiles; ++tile)
    for (int ch : chunks) {
      b.tile = tile;
      b.ichunk = ch;
      batchTables<T>(b);
      for (int m : mults) {
        b.grid = b.cus * m;
        all.push_back({timePair(reps), tile, ch, b.grid});
      }
    }
  std::sort(all.begin(), all.end(), [](const Cand& x, const Cand& y) { return x.ms < y.ms; });
  all.resize(std::min<size_t>(all.size(), 4));
  for (auto& c : all) {
    b.tile = c.tile;
    b.ichunk = c.ichunk;
    b.grid = c.grid;
    batchTables<T>(b);
    float best = -1;
    for (int trial = 0; trial < 3; ++trial) {
      const float ms = timePair(2 * reps) / 2;
      if (best < 0 || ms < best) best = ms;
    }
    c.ms = best;
  }
  const Cand w = *std::min_element(all.begin(), all.end(), [](const Cand& x, const Cand& y) { return x.ms < y.ms; });
  b.tile = w.tile;
  b.ichunk = w.ichunk;
  b.grid = w.grid;
  batchTables<T>(b);
  b.dropGraph();
  for (auto& c : b.cases) resetState(*c);
}

}  // namespace

extern "C" {

const char* xfd_last_error(void) { return g_err.c_str(); }

const char* xfd_version(void) { return "xfdtd_core_amd 0.1 (gfx950)"; }

int xfd_device_count(int32_t* count) {
  return guard([&] {
    int n = 0;
    XFD_HIP(hipGetDeviceCount(&n));
    *count = n;
  });
}

int xfd_device_synchronize(void) {
  return guard([&] { XFD_HIP(hipDeviceSynchronize()); });
}

int xfd_device_info(char* name, int32_t name_len, int64_t* total_bytes) {
  return guard([&] {
    int dev = 0;
    XFD_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    XFD_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) std::snprintf(name, static_cast<size_t>(name_len), "%s (%s)", p.name, p.gcnArchName);
    if (total_bytes) *total_bytes = static_cast<int64_t>(p.totalGlobalMem);
  });
}

int xfd_create(const xfd_desc* d, xfd_ctx** out) {
  return guard([&] {
    require(d != nullptr && out != nullptr, "null argument");
    *out = nullptr;
    const int dim = d->dim == 0 ? 3 : d->dim;
    require(dim >= 1 && dim <= 3, "bad grid dimension");
    if (dim == 3) require(d->nx > 1 && d->ny > 1 && d->nz > 1, "grid too small");
    if (dim == 2) require(d->nx > 1 && d->ny > 1 && d->nz == 1, "a 2D grid has nx, ny > 1 and nz = 1");
    if (dim == 1) require(d->nx == 1 && d->ny == 1 && d->nz > 1, "a 1D grid has nx = ny = 1 and nz > 1");
    require(d->dtype == XFD_F64 || d->dtype == XFD_F32, "bad dtype");
    require(d->n_cpml >= 0 && d->n_lumped >= 0 && d->n_tfsf >= 0 && d->n_gather >= 0, "bad counts");
    auto c = std::make_unique<xfd_ctx>();
    c->dim = dim;
    XFD_HIP(hipGetDevice(&c->device));
    c->dtype = d->dtype;
    c->elem = d->dtype == XFD_F64 ? 8 : 4;
    xfd::Grid& g = c->g;
    g.nx = d->nx;
    g.ny = d->ny;
    g.nz = d->nz;
    g.NY = d->ny + 1;
    const int64_t line = 128 / static_cast<int64_t>(c->elem);
    // a 2D grid keeps k's two entries unpadded, so its rows along j are
    // contiguous (pitch 2) for the TE kernels' lanes
    g.pz = dim == 2 ? 2 : ((d->nz + 1) + line - 1) / line * line;
    g.plane = static_cast<int64_t>(g.NY) * g.pz;
    g.total = static_cast<int64_t>(d->nx + 1) * g.plane;
    const int64_t bxy = ((g.nz + xfd::kLanesK - 1) / xfd::kLanesK) * ((g.ny + xfd::kRowsJ - 1) / xfd::kRowsJ);
    // i-planes marched per sweep work-item: measured on the 384^3 box (FP32 /
    // FP64 ms per step) 1: 1.405 / 2.403, 2: 1.336 / 2.376, 3: 1.324 / 2.373,
    // 4: 1.325 / 2.381, 8: 1.340 / 2.395, 16: 1.363 / 2.425, 32: 1.430 / 2.459
    c->ichunk = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(4, g.nx * bxy / 2048)));
    if (const char* e = std::getenv("XFD_ICHUNK")) {  // i-planes per sweep work-item (measurement knob)
      const int v = std::atoi(e);
      require(v >= 1 && v <= 4096, "XFD_ICHUNK out of range");
      c->ichunk = v;
    }
    XFD_HIP(hipStreamCreate(&c->stream));
    if (c->dtype == XFD_F64)
      createTyped<double>(*c, *d);
    else
      createTyped<float>(*c, *d);
    *out = c.release();
  });
}

void xfd_destroy(xfd_ctx* c) { delete c; }

int xfd_step(xfd_ctx* c, int64_t t0, int64_t n) {
  return guard([&] {
    require(c != nullptr, "null context");
    if (c->dtype == XFD_F64)
      stepTyped<double>(*c, t0, n);
    else
      stepTyped<float>(*c, t0, n);
  });
}

int xfd_read_field(xfd_ctx* c, int32_t field, double* dst) {
  return guard([&] {
    require(c && dst && field >= 0 && field < 6, "bad argument");
    int ext[3];
    extentOf(field, c->g.nx, c->g.ny, c->g.nz, ext);
    const size_t n = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
    DevBuf tmp;
    tmp.alloc(n * sizeof(double));
    DevBuf& cur = curFields(*c)[field];
    if (c->dtype == XFD_F64)
      xfd::Launch<double>::pack(c->stream, c->g, cur.as<double>(), tmp.as<double>(), ext[0], ext[1], ext[2]);
    else
      xfd::Launch<float>::pack(c->stream, c->g, cur.as<float>(), tmp.as<double>(), ext[0], ext[1], ext[2]);
    XFD_HIP(hipMemcpyAsync(dst, tmp.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    XFD_HIP(hipStreamSynchronize(c->stream));
  });
}

int xfd_write_field(xfd_ctx* c, int32_t field, const double* src) {
  return guard([&] {
    require(c && src && field >= 0 && field < 6, "bad argument");
    if (field == 0 || field == 1 || field == 5) c->te_zero = false;  // no longer known to stay zero
    int ext[3];
    extentOf(field, c->g.nx, c->g.ny, c->g.nz, ext);
    DevBuf tmp;
    tmp.alloc(static_cast<size_t>(ext[0]) * ext[1] * ext[2] * sizeof(double));
    DevBuf& cur = curFields(*c)[field];
    if (c->dtype == XFD_F64)
      uploadPadded<double>(*c, src, ext, cur.as<double>(), tmp);
    else
      uploadPadded<float>(*c, src, ext, cur.as<float>(), tmp);
    if (c->fstep) {  // both sets hold the points the single-pass step does not visit
      DevBuf& alt = (c->parity ? c->fields : c->fields_b)[field];
      XFD_HIP(hipMemcpyAsync(alt.p, cur.p, cur.bytes, hipMemcpyDeviceToDevice, c->stream));
      XFD_HIP(hipStreamSynchronize(c->stream));
    }
  });
}

int xfd_read_gather(xfd_ctx* c, int32_t gi, double* dst) {
  return guard([&] {
    require(c && dst && gi >= 0 && gi < static_cast<int>(c->gathers.size()), "bad argument");
    auto& g = *c->gathers[gi];
    if (g.staged) {  // the frames copied so far, from pinned host memory
      XFD_HIP(hipStreamSynchronize(c->stream));
      XFD_HIP(hipStreamSynchronize(c->copy_stream));
      std::memcpy(dst, g.host, static_cast<size_t>(g.a.frame_elems * g.d.count) * sizeof(double));
      return;
    }
    if (g.probe >= 0)
      XFD_HIP(hipMemcpyAsync(dst, c->probe_out.as<double>() + g.probe * c->probe_count,
                             static_cast<size_t>(c->probe_count) * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    else if (g.out.bytes)
      XFD_HIP(hipMemcpyAsync(dst, g.out.p, g.out.bytes, hipMemcpyDeviceToHost, c->stream));
    XFD_HIP(hipStreamSynchronize(c->stream));
  });
}

int xfd_set_profiling(xfd_ctx* c, int32_t on) { return xfd_set_profiling_classes(c, on ? ~0u : 0u); }

int xfd_set_profiling_classes(xfd_ctx* c, uint32_t mask) {
  return guard([&] {
    require(c != nullptr, "null context");
    c->prof_mask = mask & ((1u << K_COUNT) - 1);
    c->prof = c->prof_mask != 0;
    for (int k = 0; k < K_COUNT; ++k) {
      c->launches[k] = 0;
      c->total_ms[k] = 0;
    }
  });
}

int xfd_kernel_stats(xfd_ctx* c, int32_t k, const char** name, int64_t* launches, double* total_ms) {
  return guard([&] {
    require(c != nullptr && k >= 0 && k < K_COUNT, "bad kernel class");
    if (name) *name = kKernelNames[k];
    if (launches) *launches = c->launches[k];
    if (total_ms) *total_ms = c->total_ms[k];
  });
}

int xfd_forward_checkpointed(xfd_ctx* c, int64_t nt, int64_t interval) {
  return guard([&] {
    require(c != nullptr, "null context");
    require(c->dim == 3, "the adjoint form needs a 3D grid");
    leaveStep(*c);  // checkpoints save and restore set 0
    if (c->dtype == XFD_F64)
      forwardCheckpointed<double>(*c, nt, interval);
    else
      forwardCheckpointed<float>(*c, nt, interval);
  });
}

int xfd_design_edges(const int32_t origin[3], const int32_t size[3], int64_t counts[3]) {
  return guard([&] {
    require(origin && size && counts, "bad argument");
    const auto d = designArgs(origin, size);
    for (int q = 0; q < 3; ++q) counts[q] = d.off[q + 1] - d.off[q];
  });
}

int xfd_backward(xfd_ctx* c, const xfd_adjoint_desc* d, double* grad) {
  return guard([&] {
    require(c != nullptr && d != nullptr, "bad argument");
    require(c->dim == 3, "the adjoint form needs a 3D grid");
    leaveStep(*c);
    if (c->dtype == XFD_F64)
      backwardTyped<double>(*c, *d, grad);
    else
      backwardTyped<float>(*c, *d, grad);
  });
}

int xfd_nf2ff_far_field(xfd_ctx* c, int32_t n, int32_t f, int32_t na, const double* theta, const double* phi,
                        const double* origin, double k, const double* const* e_node, const double* const* h_node,
                        const double* const* e_size, double* dst) {
  return guard([&] {
    require(c != nullptr && dst != nullptr && theta != nullptr && phi != nullptr && origin != nullptr,
            "bad argument");
    require(n >= 0 && n < static_cast<int32_t>(c->nf2ffs.size()), "bad NF2FF index");
    auto& N = *c->nf2ffs[n];
    require(f >= 0 && f < N.a.nf, "bad frequency index");
    require(na >= 0, "bad direction count");
    if (na == 0) return;
    const int dims[3] = {c->g.nx, c->g.ny, c->g.nz};
    DevBuf geo[9], ang[2], out;
    xfd::FarArgs fa{};
    for (int a = 0; a < 3; ++a) {
      require(e_node[a] && h_node[a] && e_size[a], "grid arrays missing");
      uploadVector(std::vector<double>(e_node[a], e_node[a] + dims[a] + 1), geo[a]);
      uploadVector(std::vector<double>(h_node[a], h_node[a] + dims[a]), geo[3 + a]);
      uploadVector(std::vector<double>(e_size[a], e_size[a] + dims[a]), geo[6 + a]);
      fa.e_node[a] = geo[a].as<double>();
      fa.h_node[a] = geo[3 + a].as<double>();
      fa.e_size[a] = geo[6 + a].as<double>();
      fa.origin[a] = origin[a];
    }
    uploadVector(std::vector<double>(theta, theta + na), ang[0]);
    uploadVector(std::vector<double>(phi, phi + na), ang[1]);
    fa.theta = ang[0].as<double>();
    fa.phi = ang[1].as<double>();
    fa.k = k;
    out.alloc(static_cast<size_t>(na) * 8 * sizeof(double));
    DevBuf part;
    part.alloc(static_cast<size_t>(na) * xfd::Launch<double>::farChunks(na) * 8 * sizeof(double));
    xfd::Launch<double>::far_field(c->stream, N.a, f, fa, na, part.as<double>(), out.as<double>());
    XFD_HIP(hipGetLastError());
    XFD_HIP(hipMemcpyAsync(dst, out.p, out.bytes, hipMemcpyDeviceToHost, c->stream));
    XFD_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"

// Device memory of xfd_nf2ff_far_field_noalloc: the staged grid geometry and
// directions and the kernel's output, in one file-scope array per device
// (held until the process ends).  In host code `scratch` names a host-side
// stand-in, not device memory: host code reaches each device's copy only
// through the address hipGetSymbolAddress returns with that device current,
// taken once per device on the entry point's first call there.
constexpr size_t kScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratchPtr(int dev) {
  require(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    XFD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

size_t alignUp(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" {

int xfd_nf2ff_far_field_noalloc(xfd_ctx* c, int32_t n, int32_t f, int32_t na, const double* theta,
                                const double* phi, const double* origin, double k, const double* const* e_node,
                                const double* const* h_node, const double* const* e_size, double* dst) {
  return guard([&] {
    require(c != nullptr && dst != nullptr && theta != nullptr && phi != nullptr && origin != nullptr,
            "bad argument");
    require(n >= 0 && n < static_cast<int32_t>(c->nf2ffs.size()), "bad NF2FF index");
    auto& N = *c->nf2ffs[n];
    require(f >= 0 && f < N.a.nf, "bad frequency index");
    require(na >= 0, "bad direction count");
    if (na == 0) return;
    int cur = 0;
    XFD_HIP(hipGetDevice(&cur));
    require(cur == c->device, "the context's device must be current");
    const int dims[3] = {c->g.nx, c->g.ny, c->g.nz};
    // layout: per axis e_node (N+1), h_node (N), e_size (N); theta, phi;
    // the output (na x 8); the per-chunk partial sums (na x chunks x 8)
    size_t need = 0;
    size_t off_geo[9];
    for (int a = 0; a < 3; ++a) {
      const size_t cnt[3] = {static_cast<size_t>(dims[a]) + 1, static_cast<size_t>(dims[a]),
                             static_cast<size_t>(dims[a])};
      for (int w = 0; w < 3; ++w) {
        off_geo[3 * w + a] = need;
        need += alignUp(cnt[w] * sizeof(double));
      }
    }
    const size_t off_th = need;
    need += alignUp(static_cast<size_t>(na) * sizeof(double));
    const size_t off_ph = need;
    need += alignUp(static_cast<size_t>(na) * sizeof(double));
    const size_t off_out = need;
    need += alignUp(static_cast<size_t>(na) * 8 * sizeof(double));
    const size_t off_part = need;  // per-chunk partial sums
    need += alignUp(static_cast<size_t>(na) * xfd::Launch<double>::farChunks(na) * 8 * sizeof(double));
    if (need > kScratchBytes)
      throw XfdError(XFD_EINVAL, "xfd_nf2ff_far_field_noalloc: grid and " + std::to_string(na) +
                                     " directions need " + std::to_string(need) + " bytes of scratch, above its " +
                                     std::to_string(kScratchBytes) + "-byte maximum");
    unsigned char* base = static_cast<unsigned char*>(scratchPtr(c->device));
    xfd::FarArgs fa{};
    const double* const* src[3] = {e_node, h_node, e_size};
    for (int a = 0; a < 3; ++a) {
      require(e_node[a] && h_node[a] && e_size[a], "grid arrays missing");
      const size_t cnt[3] = {static_cast<size_t>(dims[a]) + 1, static_cast<size_t>(dims[a]),
                             static_cast<size_t>(dims[a])};
      for (int w = 0; w < 3; ++w)
        XFD_HIP(hipMemcpyAsync(base + off_geo[3 * w + a], src[w][a], cnt[w] * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
      fa.e_node[a] = reinterpret_cast<const double*>(base + off_geo[a]);
      fa.h_node[a] = reinterpret_cast<const double*>(base + off_geo[3 + a]);
      fa.e_size[a] = reinterpret_cast<const double*>(base + off_geo[6 + a]);
      fa.origin[a] = origin[a];
    }
    XFD_HIP(hipMemcpyAsync(base + off_th, theta, static_cast<size_t>(na) * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    XFD_HIP(hipMemcpyAsync(base + off_ph, phi, static_cast<size_t>(na) * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    fa.theta = reinterpret_cast<const double*>(base + off_th);
    fa.phi = reinterpret_cast<const double*>(base + off_ph);
    fa.k = k;
    double* out = reinterpret_cast<double*>(base + off_out);
    xfd::Launch<double>::far_field(c->stream, N.a, f, fa, na, reinterpret_cast<double*>(base + off_part), out);
    XFD_HIP(hipGetLastError());
    XFD_HIP(hipMemcpyAsync(dst, out, static_cast<size_t>(na) * 8 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    XFD_HIP(hipStreamSynchronize(c->stream));
  });
}

int xfd_cpml_fused(xfd_ctx* c, int32_t* fused) {
  return guard([&] {
    require(c && fused, "bad argument");
    *fused = c->fstep ? 3 : (c->allslabs ? 4 : (c->fused ? 1 : (c->zfused ? 2 : 0)));
  });
}

int xfd_reset(xfd_ctx* c) {
  return guard([&] {
    require(c != nullptr, "null context");
    resetState(*c);
  });
}

int xfd_batch_create(int32_t n, const xfd_desc* descs, xfd_batch** out) {
  struct BatchFlag {  // batched cases record on the device (no staging)
    BatchFlag() { g_batch_create = true; }
    ~BatchFlag() { g_batch_create = false; }
  } flag;
  return guard([&] {
    require(n > 0 && descs != nullptr && out != nullptr, "bad argument");
    *out = nullptr;
    auto b = std::make_unique<xfd_batch>();
    b->dtype = descs[0].dtype;
    for (int i = 0; i < n; ++i) require(descs[i].dim == 0 || descs[i].dim == 3, "batched cases must be 3D grids");
    for (int i = 0; i < n; ++i) {
      require(descs[i].dtype == b->dtype, "every case of a batch has the same dtype");
      xfd_ctx* c = nullptr;
      if (xfd_create(&descs[i], &c) != XFD_OK) throw XfdError(XFD_EINVAL, std::string("case ") + std::to_string(i) + ": " + g_err);
      b->cases.emplace_back(c);
      leaveStep(*c);  // the batched launches step set 0 with their own tables
      c->allslabs = false;
      // batched cases keep their z slabs as separate launches: with most
      // wavefronts of these small boxes carrying slab lanes, the sweeps were
      // measured slower with them inside
      c->zfused = false;
      std::fill(c->zslab.begin(), c->zslab.end(), 0);
      b->shared = b->shared && c->classed && !c->wide && c->tdnfs.empty();

    }
    XFD_HIP(hipStreamCreate(&b->stream));
    XFD_HIP(hipEventCreate(&b->ev0));
    XFD_HIP(hipEventCreate(&b->ev1));
    b->tdev.alloc(sizeof(int64_t));
    int dev = 0;
    XFD_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    XFD_HIP(hipGetDeviceProperties(&prop, dev));
    b->cus = std::max(1, prop.multiProcessorCount);
    b->grid = b->cus * 8;
    if (const char* e = std::getenv("XFD_BATCH_TILE")) {  // a compiled sweep tile before any tuning
      const int v = std::atoi(e);
      require(v >= 0 && v < xfd::kBatchTiles, "XFD_BATCH_TILE out of range");
      b->tile = v;
    }
    if (b->shared) {
      if (b->dtype == XFD_F64)
        batchTables<double>(*b);
      else
        batchTables<float>(*b);
    }
    *out = b.release();
  });
}

void xfd_batch_destroy(xfd_batch* b) { delete b; }

int xfd_batch_case(xfd_batch* b, int32_t i, xfd_ctx** ctx) {
  return guard([&] {
    require(b && ctx && i >= 0 && i < static_cast<int32_t>(b->cases.size()), "bad case index");
    *ctx = b->cases[i].get();
  });
}

int xfd_batch_tile(xfd_batch* b, int32_t* lanes_k, int32_t* rows_j) {
  return guard([&] {
    require(b != nullptr, "null batch");
    if (lanes_k) *lanes_k = xfd::kBatchTileK[b->tile];
    if (rows_j) *rows_j = xfd::kBatchTileJ[b->tile];
  });
}

int xfd_batch_info(xfd_batch* b, int32_t* shared, int32_t* ichunk, int32_t* grid) {
  return guard([&] {
    require(b != nullptr, "null batch");
    if (shared) *shared = b->shared ? 1 : 0;
    if (ichunk) *ichunk = b->ichunk;
    if (grid) *grid = b->grid;
  });
}

int xfd_batch_set_graph(xfd_batch* b, int32_t on) {
  return guard([&] {
    require(b != nullptr, "null batch");
    b->use_graph = on != 0;
    b->dropGraph();
  });
}

int xfd_batch_tune(xfd_batch* b, int32_t reps) {
  return guard([&] {
    require(b != nullptr && reps > 0, "bad argument");
    if (!b->shared) return;
    if (b->dtype == XFD_F64)
      batchTune<double>(*b, reps);
    else
      batchTune<float>(*b, reps);
  });
}

int xfd_batch_step(xfd_batch* b, int64_t t0, int64_t n) {
  return guard([&] {
    require(b != nullptr, "null batch");
    if (b->dtype == XFD_F64)
      batchStep<double>(*b, t0, n);
    else
      batchStep<float>(*b, t0, n);
  });
}

int xfd_batch_last_ms(xfd_batch* b, double* ms) {
  return guard([&] {
    require(b && ms, "bad argument");
    *ms = b->last_ms;
  });
}

int xfd_read_line(xfd_ctx* c, int32_t m, double* dst) {
  return guard([&] {
    require(c && dst && m >= 0 && m < c->n_line, "bad line recording index");
    if (!c->line_count) return;
    XFD_HIP(hipMemcpy(dst, c->line_out.as<double>() + m * c->line_count, c->line_count * sizeof(double),
                      hipMemcpyDeviceToHost));
  });
}

int xfd_read_nf2ff(xfd_ctx* c, int32_t n, int32_t f, int32_t d, int32_t which, double* dst) {
  return guard([&] {
    require(c && dst && n >= 0 && n < static_cast<int32_t>(c->nf2ffs.size()), "bad NF2FF index");
    auto& nn = *c->nf2ffs[n];
    require(f >= 0 && f < nn.a.nf && d >= 0 && d < 6 && which >= 0 && which < 4, "bad NF2FF slice");
    const int64_t cells = nn.a.off[6], cnt = nn.a.off[d + 1] - nn.a.off[d];
    if (cnt == 0) return;
    const double2* src = nn.acc.as<double2>() + (static_cast<int64_t>(f) * 4 + which) * cells + nn.a.off[d];
    XFD_HIP(hipMemcpy(dst, src, cnt * sizeof(double2), hipMemcpyDeviceToHost));
  });
}

int xfd_read_td_nf2ff(xfd_ctx* c, int32_t n, int32_t d, int32_t which, double* dst) {
  return guard([&] {
    require(c && dst && n >= 0 && n < static_cast<int32_t>(c->tdnfs.size()), "bad time-domain NF2FF index");
    require(d >= 0 && d < 6 && which >= 0 && which < 4, "bad time-domain NF2FF slice");
    auto& td = *c->tdnfs[n];
    XFD_HIP(hipMemcpy(dst, td.pot.as<double>() + (static_cast<int64_t>(d) * 4 + which) * td.a.num,
                      td.a.num * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int xfd_coef_layout(xfd_ctx* c, int32_t* classed, int64_t* classes) {
  return guard([&] {
    require(c && classed, "bad argument");
    *classed = c->wide ? 2 : (c->classed ? 1 : 0);
    if (classes)
      for (int q = 0; q < 6; ++q) classes[q] = c->classes[q];
  });
}

int xfd_device_bytes(xfd_ctx* c, int64_t* bytes) {
  return guard([&] {
    require(c && bytes, "bad argument");
    *bytes = c->deviceBytes();
  });
}

}  // extern "C"
