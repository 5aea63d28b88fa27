/* xfdtd_core_amd_sim.h — scenario-level C-ABI over the C++ Simulation API.
 *
 * For callers without a C++ toolchain (ctypes, cgo, JNI).  A scenario is the
 * line-oriented spec text documented in xfdtd_core_amd/spec.py; each line is
 * the constructor call of one reference object (Simulation, Object, PML,
 * TFSF3D, lumped elements, monitors), so a scenario reads like the
 * reference's example programs (test/src/examples/*.cpp).  These calls wrap
 * xfdtd::Simulation (include/xfdtd_core_amd.hpp), which replaces the
 * reference's Simulation::init/run (src/simulation/simulation.cpp:284-389).
 * Return codes follow xfdtd_core_amd.h; messages via xfd_sim_last_error().
 */
#ifndef XFDTD_CORE_AMD_SIM_H_
#define XFDTD_CORE_AMD_SIM_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* xfd_sim_last_error(void);
/* Builds a Simulation from spec text; NULL on a parse error. */
void* xfd_sim_create(const char* spec);
void xfd_sim_destroy(void* sim);
/* 0 = FP64 (default), 1 = FP32 fields and updates. */
int xfd_sim_set_precision(void* sim, int32_t precision);
int xfd_sim_set_profiling(void* sim, int32_t on);
/* Times only the kernel classes whose bit is set (xfd_set_profiling_classes). */
int xfd_sim_set_profiling_classes(void* sim, uint32_t mask);
/* Host precompute only (no GPU): Simulation::init without the device. */
int xfd_sim_precompute(void* sim, int64_t steps);
/* Simulation::init(steps): precompute and upload to the GPU. */
int xfd_sim_init(void* sim, int64_t steps);
/* The time loop after xfd_sim_init, then monitor outputs and final fields. */
int xfd_sim_run_loop(void* sim);
/* The loop in pieces: steps t0..t0+n-1 (returns when done), then outputs. */
int xfd_sim_advance(void* sim, int64_t t0, int64_t n);
int xfd_sim_finish(void* sim);
/* Simulation::run(steps) = init + loop. */
int xfd_sim_run(void* sim, int64_t steps);
double xfd_sim_dt(void* sim);
double xfd_sim_loop_seconds(void* sim);
int xfd_sim_grid(void* sim, int64_t* n3);
int64_t xfd_sim_device_bytes(void* sim);
/* Coefficient storage as xfd_coef_layout: 1 16-bit class ids, 2 32-bit
 * class ids, 0 dense arrays. */
int32_t xfd_sim_coef_classed(void* sim);
/* Which CPML slabs run inside the Yee sweeps: 0, 1 or 2 as xfd_cpml_fused. */
int32_t xfd_sim_cpml_fused(void* sim);

/* NF2FF (NFFFTFrequencyDomain) results after run() or finish(), in the
 * order the scenario adds them: surface boxes (d = XN XP YN YP ZN ZP),
 * accumulated currents (which = 0 ja, 1 jb, 2 ma, 3 mb; interleaved complex),
 * the potentials aTheta / aPhi (a_pot = 1) and fTheta / fPhi (a_pot = 0) over
 * a theta or phi cut (nffft_fd_data_process.cpp:9-26), the radiated power,
 * and the reference's output files (processFarField, outputRadiationPower:
 * nffft_frequency_domain.cpp:56-160; nt == 1 selects the phi cut). */
int32_t xfd_sim_nf2ff_count(void* sim);

/* S-parameter networks (spec line `network`; Network / Port, network.cpp,
 * port.cpp) in spec order.  xfd_sim_network_output runs Network::output
 * (network.cpp:47-84) on the recordings of the last run: every port's power
 * waves, S(j, i) = b_j / a_i for each source port i keyed 10 j + i, and the
 * .npy files when the network has an output directory.  keys: up to cap
 * keys, returns their count (-1 on error).  xfd_sim_network_s copies one
 * S-parameter as nf interleaved (re, im) doubles. */
/* Time-domain NF2FFs (spec line `nf2ff_td`) in spec order, after a run:
 * the potential length, potential `which` (0 ua, 1 ub, 2 wa, 3 wb) of
 * surface d, and the far field (NFFFTTimeDomain::processFarField: w_theta,
 * w_phi, u_theta, u_phi, 4 x num doubles; the .npy files too when the
 * NF2FF has an output directory). */
int32_t xfd_sim_nf2ff_td_count(void* sim);
int64_t xfd_sim_nf2ff_td_num(void* sim, int32_t n);
int xfd_sim_nf2ff_td_potential(void* sim, int32_t n, int32_t d, int32_t which, double* dst);
int xfd_sim_nf2ff_td_far_field(void* sim, int32_t n, double* dst);

int32_t xfd_sim_network_count(void* sim);
int32_t xfd_sim_network_nf(void* sim, int32_t n); /* frequency count, -1 if n is out of range */
int xfd_sim_network_output(void* sim, int32_t n);
int32_t xfd_sim_network_keys(void* sim, int32_t n, int32_t* keys, int32_t cap);
int xfd_sim_network_s(void* sim, int32_t n, int32_t key, double* dst);
int xfd_sim_nf2ff_surface(void* sim, int32_t n, int32_t d, int64_t* lo, int64_t* hi);
int xfd_sim_nf2ff_current(void* sim, int32_t n, int32_t f, int32_t d, int32_t which, double* dst);
int xfd_sim_nf2ff_potential(void* sim, int32_t n, int32_t f, int32_t a_pot, int32_t theta_comp, int32_t nt,
                            const double* theta, int32_t np, const double* phi, const double* origin, double* dst);
int xfd_sim_nf2ff_power(void* sim, int32_t n, int32_t f, double* dst);
int xfd_sim_nf2ff_process_far_field(void* sim, int32_t n, int32_t nt, const double* theta, int32_t np,
                                    const double* phi, const char* sub_dir, const double* origin);
int xfd_sim_nf2ff_output_power(void* sim, int32_t n);
/* which = 0: e-node coordinates, otherwise h sizes; returns the length. */
int64_t xfd_sim_grid_array(void* sim, int32_t which, int32_t axis, double* dst);
/* Arrays of the reference extent; dst may be NULL to query the shape. */
int xfd_sim_field(void* sim, int32_t field, int64_t* shape3, double* dst);
int xfd_sim_coeff(void* sim, const char* name, int64_t* shape3, double* dst);
int xfd_sim_monitor_count(void* sim);
/* The device context (xfd_ctx*, include/xfdtd_core_amd.h) the simulation
 * steps, for direct C-ABI calls; NULL before init.  Borrowed. */
void* xfd_sim_device_context(void* sim);
/* Monitor::output() of monitor m (monitor.h:65): writes its .npy file(s) in
 * its output directory; a movie's frames are written by the run itself. */
int xfd_sim_monitor_output(void* sim, int32_t monitor);
int xfd_sim_monitor_records(void* sim, int32_t monitor);
int xfd_sim_monitor_record(void* sim, int32_t monitor, int32_t record, int64_t* shape3, double* dst);
int xfd_sim_kernel_count(void* sim);
int xfd_sim_kernel_stat(void* sim, int32_t k, char* name, int32_t name_len, int64_t* launches,
                        double* total_ms);

/* Adjoint (no reference counterpart; spec lines `design_region` and
 * `dft_objective`, see xfdtd::Simulation::setDesignRegion/setDftObjective).
 * forward: init + checkpointed loop, returns J in *value; backward: dJ/d eps_r
 * per design cell (k fastest) into grad; objective_dft: the objective points'
 * DFT values as interleaved complex. */
int xfd_sim_set_design(void* sim, const int64_t* origin3, const int64_t* size3, const double* eps_r);
int xfd_sim_forward(void* sim, int64_t steps, int64_t interval, double* value);
int xfd_sim_backward(void* sim, double* grad);
int xfd_sim_objective_dft(void* sim, double* dst);

/* Batched stepping of several scenarios on one GPU (xfdtd::SimulationBatch;
 * no reference counterpart).  The handles stay owned by the caller and must
 * outlive the batch; results are read from each handle as after run().
 * info: shared = 1 when the batched launches are used, the i-chunk and
 * persistent grid size, host seconds and device milliseconds of the last
 * advance. */
void* xfd_sim_batch_create(int32_t n, void* const* sims);
void xfd_sim_batch_destroy(void* batch);
int xfd_sim_batch_init(void* batch, int64_t steps);
int xfd_sim_batch_set_graph(void* batch, int32_t on);
int xfd_sim_batch_tune(void* batch, int32_t reps);
int xfd_sim_batch_advance(void* batch, int64_t t0, int64_t n);
int xfd_sim_batch_finish(void* batch);
int xfd_sim_batch_info(void* batch, int32_t* shared, int32_t* ichunk, int32_t* grid, double* loop_s,
                       double* device_ms);
/* The batched sweep tile (lanes along k, rows along j). */
int xfd_sim_batch_tile(void* batch, int32_t* lanes_k, int32_t* rows_j);

#ifdef __cplusplus
}
#endif

#endif /* XFDTD_CORE_AMD_SIM_H_ */
