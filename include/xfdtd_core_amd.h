/* xfdtd_core_amd.h — C-ABI of the MI355X FDTD time-stepping library.
 *
 * This is the device boundary behind Simulation::run(): the host performs
 * the reference's FP64 precompute (grid, material averaging, the 18 Yee
 * coefficient arrays, CPML profiles, TFSF incident tables, lumped-element
 * patches) and hands the result to xfd_create(); xfd_step() then runs the
 * per-step loop that Domain::run() runs in the reference
 * (src/domain/domain.cpp:82-135): updateH -> correctH -> updateE -> correctE
 * -> record.  Plain pointers and sizes only; every entry point returns 0 on
 * success and a negative code on failure (message in xfd_last_error()).
 *
 * Array conventions follow the reference: row-major, k (z) fastest
 * (include/xfdtd/common/type_define.h:16-31), staggered extents as in
 * Simulation::generateEMF (src/simulation/simulation.cpp:514-528):
 *   Ex nx x (ny+1) x (nz+1)   Ey (nx+1) x ny x (nz+1)   Ez (nx+1) x (ny+1) x nz
 *   Hx (nx+1) x ny x nz       Hy nx x (ny+1) x nz       Hz nx x ny x (nz+1)
 * Coefficient arrays have their field's extent.  Host buffers are borrowed
 * only for the duration of the call; the library owns all device memory and
 * its HIP stream.
 */
#ifndef XFDTD_CORE_AMD_H_
#define XFDTD_CORE_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XFD_OK 0
#define XFD_EINVAL -1   /* bad argument or inconsistent descriptor */
#define XFD_EDEVICE -2  /* HIP runtime error */
#define XFD_ENOMEM -3   /* device allocation failed */

enum xfd_dtype { XFD_F64 = 0, XFD_F32 = 1 };
enum xfd_field { XFD_EX = 0, XFD_EY, XFD_EZ, XFD_HX, XFD_HY, XFD_HZ };

/* One CPML slab: PML::init / calRecursiveConvolutionCoeff / correctCoefficient*
 * (src/boundary/pml.cpp:98-196, 286-633) and the PMLCorrector ranges
 * (src/include/boundary/pml_corrector.h:178-382) for a single domain.
 * Tangential axes of `axis` are a = (axis+1)%3, b = (axis+2)%3.  The four
 * c_psi arrays have the reference's psi extents: the slab box
 * (n along axis, grid size along a and b) extended by one along b for
 * c_ea_psi_hb / c_hb_psi_ea and along a for c_eb_psi_ha / c_ha_psi_eb. */
typedef struct {
  int32_t axis;     /* 0 = x, 1 = y, 2 = z */
  int32_t n;        /* layers */
  int32_t e_start;  /* first corrected E plane (PML::nodeENodeStartIndexMainAxis) */
  int32_t h_start;  /* first corrected H plane (PML::nodeHNodeStartIndexMainAxis) */
  const double *a_e, *b_e, *a_h, *b_h;  /* n each */
  const double *c_ea_psi_hb, *c_eb_psi_ha, *c_ha_psi_eb, *c_hb_psi_ea;
} xfd_cpml_desc;

/* TFSF3D plane-wave source: the TFSFCorrector constructor arguments
 * (src/waveform_source/tfsf_3d.cpp:21-45) for a single domain. */
typedef struct {
  int32_t origin[3], size[3];            /* TFSF box (TFSF::globalBox) */
  const double *px_int, *py_int, *pz_int;    /* size[a] + 1 each */
  const double *px_half, *py_half, *pz_half; /* size[a] + 2 each */
  int64_t nt;                            /* rows of the incident tables */
  int64_t aux;                           /* e_inc columns; h_inc has aux-1 */
  const double *e_inc, *h_inc;           /* nt x aux, nt x (aux-1) */
  double transform_e[3], transform_h[3];
  double ca[3], cb[3];                   /* TFSF::cax..caz, cbx..cbz */
  /* corrector calls per half-step: 0 or 3 TFSF3DCorrector (the 12 faces x
   * components of tfsf_corrector.cpp:177-232), 2 TFSF2DCorrector (:153-175),
   * 1 TFSF1DCorrector (:124-151) on the ZN face (forward) or ZP */
  int32_t kind;
  int32_t forward;
} xfd_tfsf_desc;

/* Per-step lumped-element correction after the E update, in object add
 * order (src/object/lumped_element/lumped_element_corrrector.cpp):
 *   XFD_LUMPED_SOURCE   voltage or current source, E += coff * w[t] (:8-38);
 *   XFD_LUMPED_INDUCTOR E += coff * J, then J += coff_j * E, with J a state
 *                       per edge starting at zero (:40-99, inductor.cpp). */
enum xfd_lumped_kind { XFD_LUMPED_SOURCE = 0, XFD_LUMPED_INDUCTOR = 1 };

typedef struct {
  int32_t component;        /* 0..2: Ex, Ey, Ez */
  int32_t origin[3], size[3];  /* edge box in grid indices */
  const double* coff;       /* size[0] x size[1] x size[2]: coff_v / coff_i, or cecjc */
  int64_t nt;
  const double* waveform;   /* source: nt samples, indexed by the time step */
  int32_t kind;             /* enum xfd_lumped_kind */
  const double* coff_j;     /* inductor: cjcec, same shape as coff */
} xfd_lumped_desc;

/* A recording: `count` copies of the box [origin, end) of one field, taken
 * after correctE of steps first, first + interval, ... (MovieMonitor,
 * FieldMonitor, FieldTimeMonitor sampling in src/monitor/). */
typedef struct {
  int32_t field;
  int32_t origin[3], end[3];
  int64_t first, interval, count;
} xfd_gather_desc;

/* Linear dispersion by the auxiliary differential equation method: the
 * reference's ADEMethodStorage (include/xfdtd/material/ade_method/
 * ade_method.h) and E updators (src/include/updator/ade_updator/
 * template_ade_update_scheme.h:14-88, src/updator/m_lor_ade_updator.cpp).
 * The reference keeps the ADE coefficients per cell; they only depend on the
 * cell's material, so they are passed once per material plus a per-cell
 * material id.  The E coefficients of dispersive cells (cece, cecha, cechb)
 * are already patched in xfd_desc.coeff, as after buildDispersiveSpace.
 * Per material, `coeff` holds 2 + 7 * poles doubles:
 *   cej_sum, cee_prev, then per pole
 *   cj_en, cj_e, cj_ep, cj_j, cj_jp, cjsum_j, cjsum_jp
 * (coeffEJSum, coeffEEPrev, coeffJENext, coeffJE, coeffJEPrev, coeffJJ,
 * coeffJJPrev, coeffJSumJ, coeffJSumJPrev; entries a scheme does not use are
 * ignored).  Pole currents, previous currents and previous E start at zero. */
enum xfd_ade_scheme { XFD_ADE_DRUDE = 0, XFD_ADE_DEBYE = 1, XFD_ADE_MLORENTZ = 2 };

typedef struct {
  int32_t scheme;            /* enum xfd_ade_scheme (simulation.cpp:806-818) */
  int32_t poles;             /* the storage's pole count */
  int32_t n_materials;
  const double* coeff;       /* n_materials x (2 + 7 * poles) */
  const int16_t* cell_material;  /* nx x ny x nz cells; -1: not dispersive */
} xfd_ade_desc;

/* Frequency-domain NF2FF running DFT (NFFFTFrequencyDomain::update,
 * src/nffft/nffft_frequency_domain.cpp:45-53, FDPlaneData::calculateJ/M,
 * src/nffft/nffft_fd_data_update.cpp:22-122), run after the recordings of
 * every step.  Per frequency f and surface d the library accumulates four
 * complex arrays on the surface's cells (k fastest):
 *   ja, jb += (+-1) * face-centre(H_b, H_a) * th[f][t]
 *   ma, mb += (+-1) * face-centre(E_b, E_a) * te[f][t]
 * in FP64, `repeat` times per step (the reference's update loop repeats the
 * whole update once per frequency, so its currents carry a factor nf). */
typedef struct {
  int32_t lo[6][3], hi[6][3];  /* surfaces XN XP YN YP ZN ZP; hi == lo: unused */
  int32_t nf;                  /* frequencies */
  int32_t repeat;              /* accumulations per step: nf reproduces the
                                  reference's loop, 1 is the plain DFT */
  int64_t nt;                  /* rows of the phase tables */
  const double* te;            /* nf x nt complex (re, im): dt exp(-j 2 pi f (t + 1) dt) */
  const double* th;            /* nf x nt complex: dt exp(-j 2 pi f (t + 1/2) dt) */
} xfd_nf2ff_desc;

/* One time-domain NF2FF (NFFFTTimeDomain, src/nffft/nfffft_time_domain.cpp;
 * TDPlaneData::update, src/include/nffft/nffft_td_data.h:159-231), run
 * after the recordings of every step t.  For each used surface d (XN XP YN
 * YP ZN ZP) and each of its cells (i, j, k; k fastest) the library keeps the
 * previous face-centre tangential E (a, b) and H (a, b) and adds, with
 * f = s + t - floor(s + t) and nn = floor(s + t),
 *   E: ua[nn] += ca (1 - f) ds dEb, ua[nn + 1] += ca f ds dEb,
 *      ub likewise with cb and dEa;   H: wa, wb with dHb, dHa,
 * in FP64 and in the reference's order (ca, cb = +-1 by attribute and side).
 * s_e / s_h: per cell of the used surfaces in order, the time-step delay
 * (dmax - r . u) / (c dt) plus 0.5 (E) or 0 (H); ds: the face areas. */
typedef struct {
  int32_t lo[6][3], hi[6][3];  /* surfaces XN XP YN YP ZN ZP; hi == lo: unused */
  const double* s_e;
  const double* s_h;
  const double* ds;
  int64_t num;                 /* potential length */
} xfd_tdnf_desc;

/* A weighted-sum recording: VoltageMonitor / CurrentMonitor
 * (voltage_monitor.cpp:191-227, current_monitor.cpp:272-385).  After the
 * corrections of every step t < count the library records, in FP64,
 *   scale * (g0 + g1),  g_q = sum over the terms of group q, in term order,
 *   of weight * F[field](i, j, k)
 * (the reference accumulates its integrals term by term; a current monitor
 * keeps its two loop sides in the two groups). */
typedef struct {
  int32_t n_terms;
  const int32_t* terms;    /* n_terms x 5: field (enum xfd_field), i, j, k, group (0 or 1) */
  const double* weights;   /* n_terms */
  double scale;
  int64_t count;
} xfd_line_desc;

typedef struct {
  int32_t nx, ny, nz;
  int32_t dtype;               /* enum xfd_dtype: field and update precision */
  /* cexe cexhy cexhz ceye ceyhz ceyhx ceze cezhx cezhy
   * chxh chxey chxez chyh chyez chyex chzh chzex chzey */
  const double* coeff[18];
  int32_t n_cpml;
  const xfd_cpml_desc* cpml;   /* in boundary add order */
  int32_t n_lumped;
  const xfd_lumped_desc* lumped;  /* in object add order */
  int32_t n_tfsf;
  const xfd_tfsf_desc* tfsf;
  int32_t n_gather;
  const xfd_gather_desc* gather;
  const xfd_ade_desc* ade;     /* NULL: no dispersive material */
  int32_t n_nf2ff;
  const xfd_nf2ff_desc* nf2ff;
  int32_t n_line;
  const xfd_line_desc* line;
  int32_t n_tdnf;
  const xfd_tdnf_desc* tdnf;
  /* grid dimension (GridSpace::Dimension): 0 or 3 a 3D grid; 2 a 2D grid
   * (nz = 1: BasicUpdatorTE, Ez from Hx/Hy); 1 a 1D grid (nx = ny = 1:
   * BasicUpdatorTEM, Ex from Hy) */
  int32_t dim;
} xfd_desc;

typedef struct xfd_ctx xfd_ctx;

const char* xfd_last_error(void);
const char* xfd_version(void);
int xfd_device_count(int32_t* count);
/* Waits for all work on the current device (hipDeviceSynchronize). */
int xfd_device_synchronize(void);
/* Name of the current device and its HBM capacity in bytes. */
int xfd_device_info(char* name, int32_t name_len, int64_t* total_bytes);

/* Replaces Simulation::generateDomain + the Domain objects
 * (src/simulation/simulation.cpp:401-494). */
int xfd_create(const xfd_desc* desc, xfd_ctx** out);
void xfd_destroy(xfd_ctx* ctx);

/* Runs steps t0 .. t0+n-1 (Domain::run, src/domain/domain.cpp:82-135) and
 * waits for them.  Time step t indexes the TFSF tables, lumped waveforms and
 * the recording schedule. */
int xfd_step(xfd_ctx* ctx, int64_t t0, int64_t n);

/* Copies a field out to / in from a host array of the reference extent,
 * converting to / from double. */
int xfd_read_field(xfd_ctx* ctx, int32_t field, double* dst);
int xfd_write_field(xfd_ctx* ctx, int32_t field, const double* src);

/* Copies recording `g` (count x box volume doubles) to the host.  Movies
 * (four or more frames of 4 MiB or more) are staged while the steps run:
 * each frame is gathered into one of two device slots and copied to pinned
 * host memory on a second stream (MovieMonitor::update's cadence,
 * movie_monitor.cpp:26-34), so device memory holds two frames whatever the
 * run length; smaller recordings stay on the device until read.
 * XFD_STAGE=0 keeps every recording on the device, =1 stages any recording
 * of two or more frames. */
int xfd_read_gather(xfd_ctx* ctx, int32_t g, double* dst);

/* Copies weighted-sum recording `m` (count doubles) to the host. */
int xfd_read_line(xfd_ctx* ctx, int32_t m, double* dst);

/* Copies the accumulated surface current `which` (0 ja, 1 jb, 2 ma, 3 mb)
 * of NF2FF `n`, frequency `f`, surface `d` (XN XP YN YP ZN ZP) to the host
 * as interleaved (re, im) doubles in the surface's (i, j, k) order. */
int xfd_read_nf2ff(xfd_ctx* ctx, int32_t n, int32_t f, int32_t d, int32_t which, double* dst);

/* Copies potential `which` (0 ua, 1 ub, 2 wa, 3 wb) of surface `d` of
 * time-domain NF2FF `n` (num doubles) to the host. */
int xfd_read_td_nf2ff(xfd_ctx* ctx, int32_t n, int32_t d, int32_t which, double* dst);

/* Far-field potentials of NF2FF `n`, frequency `f`, from its accumulated
 * surface currents, on the device (FDPlaneData::getPotential /
 * calculatePotential, nffft_fd_data_process.cpp:9-26, 44-160): for each of
 * the `na` directions (theta[q], phi[q]) the sums over all surface cells of
 * (cur_a t_a + cur_b t_b) exp(j k r.rhat) ds for A (ja, jb) and F (ma, mb),
 * with the theta and the phi rows of the transform (transform.h:24-58).
 * k = 2 pi f / c.  e_node[a] (N_a + 1 values), h_node[a] and e_size[a] (N_a
 * values each) are the grid's node coordinates, cell centres and cell sizes
 * (GridSpace).  dst: na x 4 complex (A_theta, A_phi, F_theta, F_phi) as
 * interleaved (re, im) doubles. */
int xfd_nf2ff_far_field(xfd_ctx* ctx, int32_t n, int32_t f, int32_t na, const double* theta, const double* phi,
                        const double* origin, double k, const double* const* e_node, const double* const* h_node,
                        const double* const* e_size, double* dst);

/* xfd_nf2ff_far_field without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The staged grid arrays, directions, the kernels' partial sums and output
 * live in the library's file-scope device array of 4 MiB (one per GPU, held
 * until the process ends); a call whose grid and `na` need more (about
 * 29 000 directions on a 384^3 grid) fails with a message and does nothing.  Same
 * arguments, kernel and results (bitwise) as xfd_nf2ff_far_field.  Every
 * call on one GPU shares that array, so calls that overlap on one GPU (from
 * several host threads or streams) race; the context's device must be
 * current. */
int xfd_nf2ff_far_field_noalloc(xfd_ctx* ctx, int32_t n, int32_t f, int32_t na, const double* theta,
                                const double* phi, const double* origin, double k, const double* const* e_node,
                                const double* const* h_node, const double* const* e_size, double* dst);

/* Kernel timing with HIP events on the library's stream.  When enabled,
 * xfd_step records an event pair around every launch; xfd_kernel_stats then
 * reports, for kernel class `k`, its name, launch count and summed time. */
int xfd_set_profiling(xfd_ctx* ctx, int32_t on);
/* The same for the kernel classes whose bit is set in `mask` (bit k: class
 * k of xfd_kernel_stats) -- an event pair costs about 10 us of idle GPU time
 * between launches, so timing only the class of interest keeps the step
 * close to its unprofiled time.  No reference counterpart. */
int xfd_set_profiling_classes(xfd_ctx* ctx, uint32_t mask);
int xfd_kernel_stats(xfd_ctx* ctx, int32_t k, const char** name, int64_t* launches,
                     double* total_ms);

/* Device bytes held by the context. */
int xfd_device_bytes(xfd_ctx* ctx, int64_t* bytes);

/* Coefficient storage chosen by xfd_create: *classed = 1 when every
 * component's distinct (self, dual-A, dual-B) triples fit a 16-bit class id
 * (classes[q] = table size of Ex..Hz); 2 (3D grids) when some component has
 * more distinct triples than that but none more than a quarter of its
 * cells: 32-bit class ids; 0 for the 18 dense arrays.  The environment
 * variable XFD_COEF_MODE=dense forces the dense layout, XFD_COEF_MODE=classed
 * rules out the 32-bit ids. */
int xfd_coef_layout(xfd_ctx* ctx, int32_t* classed, int64_t* classes);

/* Which CPML slabs the Yee sweeps apply in the same pass (the far grid
 * planes always stay separate launches); no reference counterpart, the
 * reference runs every slab as its own corrector (pml_corrector.h:14-52).
 *   *fused = 0: none, every slab is a separate launch;
 *   *fused = 1: the leading x/y slabs (same operations in the same order, so
 *     FP64 stays bitwise); opt-in with XFD_CPML_FUSED=1 (slower on MI355X);
 *   *fused = 2: the z slabs.  FP32: by default when a k-row spans at
 *     least 4 sweep tiles (256 cells), where it is faster; the z
 *     corrections then precede the x/y slabs' where they overlap (FP32
 *     rounding only).  FP64: only when forced; the sweeps correct only the
 *     points in no x or y slab, and the frame where slabs overlap follows
 *     the x/y slabs in one launch, so the reference's order holds (bitwise).
 *     XFD_CPML_ZFUSED=1 forces it, =0 turns it off.
 *   *fused = 3: the single-pass step: one launch per step computes H, its
 *     CPML slabs, E and its CPML slabs (every slab, in boundary order, so
 *     FP64 stays bitwise), reading one set of fields and writing a second.
 *     Opt-in with XFD_STEP_FUSED=1, for 3D grids without TFSF (its H
 *     corrections would have to reach recomputed points), ADE, records on a
 *     grid boundary plane or more than 8 slabs; XFD_CPML_FUSED=1 or
 *     XFD_CPML_ZFUSED=1 keep the two sweeps.  XFD_STEP_CHUNK (1..1024)
 *     sets the i-planes each work-item marches.
 *   *fused = 4: every slab inside the two sweeps, in boundary order per
 *     point (FP64 bitwise), each slab's parameters as kernel arguments and a
 *     point's membership from per-axis masks.  Opt-in with
 *     XFD_CPML_INSWEEP=1 (slower than modes 0 and 2 on MI355X).
 * Modes 1, 2 and 4 never with ADE materials or when a TFSF or lumped
 * correction lands inside a slab (the reference applies those first). */
int xfd_cpml_fused(xfd_ctx* ctx, int32_t* fused);

/* Zeroes a context's time-stepping state: fields, CPML auxiliaries, inductor
 * currents, ADE state, NF2FF accumulators and recordings (a fresh
 * Simulation::run without rebuilding the coefficients). */
int xfd_reset(xfd_ctx* ctx);

/* Reverse-time discrete adjoint (no reference counterpart): the gradient of
 * an objective J(F) of recorded field values with respect to the
 * permittivity of the E edges of a design box of cells.
 *
 * xfd_forward_checkpointed zeroes the state and runs steps 0 .. nt-1 like
 * xfd_step (recordings included), saving the whole state (fields, CPML psi)
 * at steps 0, K, 2K, ... in device memory.  xfd_backward then walks the
 * segments from last to first: it restores a checkpoint, recomputes the
 * segment while keeping the design edges' E, and runs the transposed step
 * (adjoint Yee sweeps, CPML, inductor records and dispersive ADE updates)
 * backwards over it, accumulating in FP64
 *   dJ/d eps_e = sum_n lambda_e^{n+1} (beta_e E_e^{n+1} + gamma_e E_e^n),
 * where, for E = s E + a dH_a + b dH_b with s = (2 eps - sigma dt) /
 * (2 eps + sigma dt) and a, b proportional to 1 / (2 eps + sigma dt),
 *   beta = -2 / (2 eps + sigma dt),  gamma = 4 sigma dt / (2 eps + sigma dt)^2
 *   - beta s.
 * The objective enters as `source[p * nt + n]` = dJ/dF_p after step n (F_p
 * the field value at point p once step n's corrections are applied).  The
 * design edges must be updated by the plain Yee sweep only: a design box whose
 * edges a CPML slab, an E record (TFSF, lumped element) or a dispersive cell
 * touches is refused with XFD_EINVAL.  After xfd_backward the fields hold a
 * recomputed intermediate state. */
typedef struct {
  int32_t origin[3], size[3];  /* design cells [origin, origin + size) */
  const double* beta;          /* per design edge (xfd_design_edges order) */
  const double* gamma;
  int32_t n_points;
  const int32_t* points;       /* n_points x 4: field (enum xfd_field), i, j, k */
  int64_t nt;                  /* steps of the checkpointed forward run */
  const double* source;        /* n_points x nt */
} xfd_adjoint_desc;

int xfd_forward_checkpointed(xfd_ctx* ctx, int64_t nt, int64_t interval);
/* Edge counts of the design box per component: Ex edges are the box
 * [origin, origin + (size0, size1 + 1, size2 + 1)) of Ex indices, Ey and Ez
 * likewise; the gradient lists Ex, then Ey, then Ez edges, each box k
 * fastest. */
int xfd_design_edges(const int32_t origin[3], const int32_t size[3], int64_t counts[3]);
int xfd_backward(xfd_ctx* ctx, const xfd_adjoint_desc* desc, double* grad);

/* Batched stepping of many independent simulations.  No reference
 * counterpart: the reference runs one Simulation per process (an MPI domain
 * split of one grid, src/simulation/simulation.cpp:401-494).  A batch owns
 * one context per descriptor (all of one dtype) and steps them together:
 * every phase of Domain::run (updateH, H correctors, CPML H, updateE, E
 * correctors, CPML E, recordings, NF2FF) is one persistent launch over the
 * work of all cases, whatever their sizes, with the step index kept on the
 * device, so one step is a short fixed sequence of launches that can be
 * replayed as a HIP graph.  Cases with dense coefficients or dispersive
 * materials are stepped one by one instead (xfd_batch_info: shared = 0).
 * Results are bitwise those of stepping each context with xfd_step. */
typedef struct xfd_batch xfd_batch;

int xfd_batch_create(int32_t n, const xfd_desc* descs, xfd_batch** out);
void xfd_batch_destroy(xfd_batch* batch);
/* Borrowed pointer to case i (read fields, recordings, NF2FF through it). */
int xfd_batch_case(xfd_batch* batch, int32_t i, xfd_ctx** ctx);
/* shared: 1 when the batched launches are used; ichunk and grid: the current
 * i-planes per work-group and work-groups per persistent launch. */
int xfd_batch_info(xfd_batch* batch, int32_t* shared, int32_t* ichunk, int32_t* grid);
/* The current sweep tile of the batched launches: lanes along k by rows
 * along j (64 x 4 until xfd_batch_tune picks another compiled shape). */
int xfd_batch_tile(xfd_batch* batch, int32_t* lanes_k, int32_t* rows_j);
/* on = 1: the first xfd_batch_step captures one step as a HIP graph and
 * every step replays it. */
int xfd_batch_set_graph(xfd_batch* batch, int32_t on);
/* Autotuner: times `reps` steps' Yee sweeps and CPML launches (no records,
 * no DFTs) for each compiled sweep tile
 * (64x4, 64x8, 128x2, 128x4, 64x2 k-lanes by j-rows), i-chunk (4, 8, 16,
 * 32) and persistent grid size (4, 8, 16, 32 work-groups per CU), times the
 * four fastest again (three trials of 2 x reps steps, each one's best) and
 * keeps the fastest of those, then zeroes every case's state (xfd_reset). */
int xfd_batch_tune(xfd_batch* batch, int32_t reps);
/* Runs steps t0 .. t0+n-1 of every case and waits for them. */
int xfd_batch_step(xfd_batch* batch, int64_t t0, int64_t n);
/* Device time of the last xfd_batch_step, in milliseconds. */
int xfd_batch_last_ms(xfd_batch* batch, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* XFDTD_CORE_AMD_H_ */
