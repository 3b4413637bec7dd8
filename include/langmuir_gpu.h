/*
 * langmuir_gpu.h -- C-ABI of the MI355X (gfx950) implementation of Langmuir's
 * data-parallel hot path.
 *
 * It replaces the C++ class LangmuirCore::OpenClHelper
 * (src/langmuirCore/include/openclhelper.h:22-241, openclhelper.cpp:17-583)
 * and the per-step sweep it served (src/langmuirCore/simulation.cpp:29-111).
 * Plain pointers and sizes only; no C++ types, no exceptions and no abort
 * cross this boundary.  Every function returns an lmg_status; the message of
 * the last failure is lmg_last_error(ctx) (ctx may be NULL for failures of
 * lmg_create).
 *
 * Threading: a context is single-threaded (callers serialise); distinct
 * contexts are independent.  Ownership: the caller owns every host array for
 * the duration of a call; the context owns all device memory.  Calls taking
 * host pointers are synchronous; the *_dev variants take device pointers, are
 * ordered on the context's stream and return before the device finishes
 * (lmg_sync waits).
 *
 * Site ids follow the reference: s = X*(y + z*Y) + x
 * (src/langmuirCore/cubicgrid.cpp:162-165); ids V..V+34 are the special-agent
 * slots of a grid (cubicgrid.cpp:558-589).
 */
#ifndef LANGMUIR_GPU_H
#define LANGMUIR_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t lmg_status;
#define LMG_OK             0
#define LMG_ERR_ARG       -1   /* invalid argument or size                        */
#define LMG_ERR_NO_DEVICE -2   /* no usable GPU: the OpenClHelper okCL=false case  */
#define LMG_ERR_HIP       -3   /* HIP runtime failure                              */
#define LMG_ERR_CAPACITY  -4   /* input larger than the context was sized for     */
#define LMG_ERR_STATE     -5   /* call out of order (e.g. step before load)       */
#define LMG_ERR_DEVICE    -6   /* device-side consistency check failed            */

/* pair-sum flags */
#define LMG_GAUSS      1  /* terms times erf(r / (sqrt(2) sigma)): gauss1/gauss2        */
#define LMG_IMAGE_LEFT 2  /* left-electrode image, dx = |xi + xj| + 1, sign -1
                             (Potential::coulombImage*, potential.cpp:384-415)          */
#define LMG_NO_CUTOFF  4  /* ignore the cutoff (map only)                               */
#define LMG_MAP_PAIRS  8  /* map: every charge at every site (no culling)             */
#define LMG_MAP_CULLED 16 /* map: only charges in 8^3-site bricks within the cutoff of
                             each 64-site tile (ignored with LMG_NO_CUTOFF).  Neither
                             flag: culled when the cutoff is well under the lattice */

typedef struct lmg_ctx lmg_ctx;

/* Constant inputs of the hot path; the role of the World/SimulationParameters
 * fields OpenClHelper::initializeOpenCL pre-binds (openclhelper.cpp:124-168). */
typedef struct lmg_params {
    int32_t grid_x, grid_y, grid_z;  /* grid.x/y/z                                 */
    int32_t cutoff;                  /* electrostatic.cutoff (sites)               */
    int32_t hopping_range;           /* hopping.range, 1 or 2                      */
    int32_t defects_charge;          /* defects.charge                             */
    int32_t device_id;               /* opencl.device.id analogue (HIP ordinal)    */
    int32_t reserved;
    double prefactor;                /* electrostatic.prefactor, parameters.h:386  */
    double inverse_kt;               /* inverse.kt, parameters.h:387               */
    double gaussian_sigma;           /* coulomb.gaussian.sigma (0 = plain 1/r)     */
    double exciton_binding;          /* exciton.binding                            */
} lmg_params;

/* Create / destroy: replaces initializeOpenCL + toggleOpenCL
 * (openclhelper.cpp:17-218).  Fails with LMG_ERR_NO_DEVICE when no GPU is
 * present, which a host maps onto the reference's CPU fallback rule. */
lmg_status lmg_create(const lmg_params* params, lmg_ctx** out);
void lmg_destroy(lmg_ctx* ctx);
const char* lmg_last_error(const lmg_ctx* ctx);
int32_t lmg_device_count(void);
lmg_status lmg_sync(lmg_ctx* ctx);
/* The context's HIP stream (hipStream_t), for callers that order their own work. */
void* lmg_stream(lmg_ctx* ctx);
/* Copy of the parameters the context was created with (the World fields
 * OpenClHelper reads back through m_world, openclhelper.cpp:124-168). */
lmg_status lmg_get_params(const lmg_ctx* ctx, lmg_params* out);

/* ---------------------------------------------------------------------------
 * Pair sums at an explicit target list: the coulomb2/gauss2 kernels with
 * launchCoulombKernel2/launchGaussKernel2 + getOutputHost
 * (kernel.cl:222-369, openclhelper.cpp:372-544,567-583), and the image sums
 * Potential::coulombImageE/H/D.  out[t] = pref * sum_j q_j f(r_tj) over
 * 0 < r^2 < cutoff^2 (times -1 for LMG_IMAGE_LEFT; every r > 0 with
 * LMG_NO_CUTOFF), FP32 terms with compensated accumulation, delivered as double.
 * ------------------------------------------------------------------------- */
lmg_status lmg_pair_potential(lmg_ctx* ctx, const int32_t* charge_site,
                              const int32_t* charge_q, int32_t n_charges,
                              const int32_t* target_site, int32_t n_targets,
                              int32_t flags, double* out);
lmg_status lmg_pair_potential_dev(lmg_ctx* ctx, const int32_t* d_charge_site,
                                  const int32_t* d_charge_q, int32_t n_charges,
                                  const int32_t* d_target_site, int32_t n_targets,
                                  int32_t flags, double* d_out);

/* lmg_pair_potential_dev that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its working
 * memory (target and charge staging, per-chunk partials) is one fixed array per
 * GPU, about 7.6 MiB, held until the process ends; it takes at most 65,536
 * targets and 262,144 charges (LMG_ERR_CAPACITY beyond).  Results are the same
 * bits as lmg_pair_potential_dev.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (any context, any stream) race. */
lmg_status lmg_pair_potential_noalloc_dev(lmg_ctx* ctx, const int32_t* d_charge_site,
                                          const int32_t* d_charge_q, int32_t n_charges,
                                          const int32_t* d_target_site, int32_t n_targets,
                                          int32_t flags, double* d_out);

/* ---------------------------------------------------------------------------
 * Full-lattice potential map: the coulomb1/gauss1 kernels with
 * launchCoulombKernel1 (kernel.cl:56-220, openclhelper.cpp:220-370).  Writes
 * all V sites (the reference read back only n values).  out is float[V].
 * ------------------------------------------------------------------------- */
lmg_status lmg_potential_map(lmg_ctx* ctx, const int32_t* charge_site,
                             const int32_t* charge_q, int32_t n_charges,
                             int32_t flags, float* out);
lmg_status lmg_potential_map_dev(lmg_ctx* ctx, const int32_t* d_charge_site,
                                 const int32_t* d_charge_q, int32_t n_charges,
                                 int32_t flags, float* d_out);
/* lmg_potential_map_dev that runs only if *d_run != 0 when the device reaches
 * it (a condition computed earlier on the stream); d_out is untouched
 * otherwise.  No counterpart in the reference: lets the surrogate's error gate
 * fall back to the exact map without a host round trip. */
lmg_status lmg_potential_map_if_dev(lmg_ctx* ctx, const int32_t* d_charge_site,
                                    const int32_t* d_charge_q, int32_t n_charges,
                                    int32_t flags, const int32_t* d_run, float* d_out);

/* ---------------------------------------------------------------------------
 * KMC sweep on the device (Simulation::performIterations with the Coulomb
 * branch, simulation.cpp:29-111, and the agents it drives).
 * ------------------------------------------------------------------------- */

/* flux agents, World order (world.cpp:961-1122) */
#define LMG_FLUX_ESL 0
#define LMG_FLUX_ESR 1
#define LMG_FLUX_HSL 2
#define LMG_FLUX_HSR 3
#define LMG_FLUX_XS  4
#define LMG_FLUX_EDL 5
#define LMG_FLUX_EDR 6
#define LMG_FLUX_HDL 7
#define LMG_FLUX_HDR 8
#define LMG_FLUX_XD  9

typedef struct lmg_sim_config {
    int32_t simulation_type;        /* 0 transistor, 1 solarcell                   */
    int32_t coulomb_carriers;       /* coulomb.carriers                            */
    int32_t source_metropolis;      /* source.metropolis                           */
    int32_t source_coulomb;         /* source.coulomb                              */
    int32_t balance_charges;        /* balance.charges                             */
    int32_t recombination_range;    /* recombination.range                         */
    int32_t max_electrons;          /* World::maxElectronAgents (world.cpp:598)    */
    int32_t max_holes;              /* World::maxHoleAgents (world.cpp:595)        */
    int32_t e_drain_slot[2];        /* slot ids of the left/right electron drains  */
    int32_t h_drain_slot[2];        /* slot ids of the left/right hole drains      */
    double voltage_left, voltage_right;
    double rate[10];                /* flux-agent probabilities (m_probability)     */
    double coupling[27];            /* couplingConstants, (h+1)^3, [dx][dy][dz]    */
} lmg_sim_config;

/* Host-side state image.  For lmg_sim_load every array holds n_* entries; for
 * lmg_sim_store the caller provides arrays of at least the capacities
 * (max_electrons, max_holes) and the callee sets the n_* fields. */
typedef struct lmg_sim_state {
    int32_t n_electrons, n_holes, n_defects, reserved;
    int32_t* e_site;  int32_t* e_lifetime;  int32_t* e_pathlength;
    int32_t* h_site;  int32_t* h_lifetime;  int32_t* h_pathlength;
    int32_t* defect_site;           /* load only                                   */
    double* potential;              /* load only: background potential, V sites    */
    uint32_t rng_window[624];       /* raw MT19937 words x[k .. k+623]             */
    uint64_t rng_consumed;          /* k = engine outputs consumed since seeding    */
    uint64_t flux_attempts[10];
    uint64_t flux_successes[10];
    int64_t current_step;
} lmg_sim_state;

lmg_status lmg_sim_load(lmg_ctx* ctx, const lmg_sim_config* config,
                        const lmg_sim_state* state);
lmg_status lmg_sim_store(lmg_ctx* ctx, lmg_sim_state* state);
/* n KMC steps, enqueued on the context stream; returns once enqueued unless
 * the context is in synchronous mode (lmg_sim_set_sync). */
lmg_status lmg_sim_step(lmg_ctx* ctx, int32_t n_steps);
lmg_status lmg_sim_set_sync(lmg_ctx* ctx, int32_t synchronous);
/* lmg_sim_step replays one captured step (a HIP graph of its kernels, made at
 * the first step after a load) when on (the default), or launches the kernels
 * one by one.  Same results either way.  No counterpart in the reference. */
lmg_status lmg_sim_set_graphs(lmg_ctx* ctx, int32_t on);
/* The ΔE kernel's shape after the load, picked by timing each candidate on the
 * loaded state, and its measured time in microseconds: on the all-pairs path
 * the number of charge chunks; on the culled path the number of work items (1
 * to 3 device-resident waves of workgroups, four items each; the time covers
 * the list building too).  A forced shape (lmg_sim_set_de_chunks) reports
 * *kernel_us = -1.  The sums do not depend on either.  No counterpart in the
 * reference. */
lmg_status lmg_sim_get_tuning(const lmg_ctx* ctx, int32_t* shape, double* kernel_us);
/* Fixes the ΔE kernel's shape for the following loads instead of choosing it
 * (0, the default): on the all-pairs path the charge-chunk count (1..128), on
 * the culled path the number of workgroups (at least one per target group;
 * at most 65536).  The sums do not depend on it.  No counterpart in the
 * reference. */
lmg_status lmg_sim_set_de_chunks(lmg_ctx* ctx, int32_t chunks);
/* dE path for the following lmg_sim_load: 0 = automatic (culled when a group
 * of targets plus the cutoff covers well under the lattice and the pair count
 * is large), 1 = all pairs, 2 = culled (8^3-site brick lists).  Both give the
 * same pair sums to FP32 rounding.  No counterpart in the reference, whose one
 * path is the all-pairs loop of kernel.cl / ChargeAgent::coulombCPU
 * (chargeagent.cpp:215-271).  lmg_sim_get_de_path reports the loaded path
 * (1 or 2). */
lmg_status lmg_sim_set_de_path(lmg_ctx* ctx, int32_t mode);
lmg_status lmg_sim_get_de_path(const lmg_ctx* ctx, int32_t* mode);
/* lmg_sim_step with HIP events around the phases of every step on the
 * context stream; returns once the steps are done.  kernel_ms[0] = summed dE
 * phase time, kernel_ms[1] = summed commit-phase time, kernel_ms[2] =
 * first-to-last event span, kernel_ms[3] = summed time of the pair-sum kernel
 * alone (de_kernel, or cull_de_kernel without the list building before it).
 * No counterpart in the reference: its only timing is the real:time column of
 * out.dat, writer.cpp:362,376. */
lmg_status lmg_sim_step_timed(lmg_ctx* ctx, int32_t n_steps, double* kernel_ms);
/* Proposed sites and dE of the pending step for every carrier (electrons then
 * holes) without consuming the decision draws: the values
 * ChargeAgent::coulombCPU/coulombGPU would hold (chargeagent.cpp:215-296). */
lmg_status lmg_sim_delta_e(lmg_ctx* ctx, int32_t* fsite_e, int32_t* fsite_h,
                           double* de_e, double* de_h);
/* *code = 0 when healthy; otherwise a device-side error code:
 * 1 = random-word ring overrun, 2 = carrier list above capacity,
 * 3 = invalid site in a carrier list, 4 = balanceCharges gave up (the
 * reference's qFatal "can not balance"). */
lmg_status lmg_sim_device_error(lmg_ctx* ctx, int32_t* code);

/* ---------------------------------------------------------------------------
 * Batched replicas: n independent simulations (each with its own lattice,
 * cutoff, carriers, traps and seed) advanced together, one launch per kernel
 * per step for all of them.  The reference ran independent jobs as separate
 * processes (LangmuirPython/batch/frank/gpu.batch:1-18, LangmuirPython/utils/
 * scan.py:13); there is no in-process counterpart.  Replicas must agree on
 * coulomb_carriers and on Gaussian screening (sigma > 0 or not).
 * ------------------------------------------------------------------------- */
typedef struct lmg_batch lmg_batch;

lmg_status lmg_batch_create(int32_t n, const lmg_params* params, const lmg_sim_config* configs,
                            const lmg_sim_state* states, int32_t device_id, lmg_batch** out);
/* lmg_batch_create with a choice of ΔE path: 0 automatic (the replicas whose
 * cutoff box around a target group covers well under their lattice are
 * candidates for the culled path; the autotuner times them culled against all
 * of the batch on the all-pairs path and keeps the faster), 1 every replica on
 * the all-pairs path, 2 every replica that can be culled (Coulomb on and the
 * lattice within the culled path's limits) on the culled path.  The culled
 * replicas' list building and sums run as one launch each for all of them. */
lmg_status lmg_batch_create_with_path(int32_t n, const lmg_params* params, const lmg_sim_config* configs,
                                      const lmg_sim_state* states, int32_t device_id, int32_t de_path,
                                      lmg_batch** out);
void lmg_batch_destroy(lmg_batch* batch);
const char* lmg_batch_last_error(const lmg_batch* batch);
int32_t lmg_batch_size(const lmg_batch* batch);
lmg_status lmg_batch_step(lmg_batch* batch, int32_t n_steps);
/* as lmg_sim_set_graphs, for the batch step (graph replay is the default) */
lmg_status lmg_batch_set_graphs(lmg_batch* batch, int32_t on);
/* the common charge-chunk count of the batch's all-pairs dE kernel (0 when
 * every replica is culled), autotuned at creation on the loaded replicas, and
 * the time of the whole dE phase for all replicas (us) */
lmg_status lmg_batch_get_tuning(const lmg_batch* batch, int32_t* chunks, double* kernel_us);
/* paths[r] = 1 (all pairs) or 2 (culled): the ΔE path replica r runs */
lmg_status lmg_batch_get_de_paths(const lmg_batch* batch, int32_t* paths);
/* every candidate the batch's autotuner timed, as a JSON object: lists of
 * [shape, microseconds] for the culled workgroup totals ("cull_candidates"),
 * the all-pairs chunk counts ("pairs_candidates", and "all_pairs_candidates"
 * with every replica on that path), and the choice.  Valid until destroy. */
const char* lmg_batch_tuning_report(const lmg_batch* batch);
/* n steps timed with HIP events on the batch stream; *ms = event span */
lmg_status lmg_batch_step_timed(lmg_batch* batch, int32_t n_steps, double* ms);
/* n steps launched kernel by kernel with HIP events around each replica-wide dE
 * kernel: kernel_ms[0] = summed dE-kernel time, kernel_ms[1] = summed time of the
 * rest of the step, kernel_ms[2] = first-to-last event span.  Same results as
 * lmg_batch_step. */
lmg_status lmg_batch_step_phases(lmg_batch* batch, int32_t n_steps, double* kernel_ms);
lmg_status lmg_batch_sync(lmg_batch* batch);
lmg_status lmg_batch_store(lmg_batch* batch, int32_t replica, lmg_sim_state* state);
/* lmg_sim_delta_e of one replica: its pending proposals and their dE (waits for
 * the batch stream first; consumes no draws). */
lmg_status lmg_batch_delta_e(lmg_batch* batch, int32_t replica, int32_t* fsite_e, int32_t* fsite_h,
                             double* de_e, double* de_h);

/* ---------------------------------------------------------------------------
 * Learned surrogate of the potential map.  No counterpart in the reference
 * (its only map is the exact coulomb1/gauss1 kernel, kernel.cl:56-220): a
 * spectral operator map = window(irfft3(W . rfft3(pad(q)))) with complex
 * weights W on the half spectrum of a zero-padded grid (padded[] per axis,
 * >= N + cutoff - 1, or >= 2N - 1 with LMG_NO_CUTOFF), trained on the GPU from
 * (occupancy, exact map) pairs.  It approximates lmg_potential_map_dev with the
 * flags given at creation (LMG_GAUSS, LMG_NO_CUTOFF) for the context's lattice
 * and uses the context's stream.
 *
 * Occupancy grids are int8 charge sums per site (V bytes per grid, grids
 * contiguous); a single grid written by lmg_surrogate_occupancy_dev must start
 * 4-byte aligned and hold ceil(V / 4) * 4 bytes.  Weights are n_weights
 * interleaved (re, im) float pairs in [kz][ky][kx] order, kx < padded[0]/2 + 1.
 * ------------------------------------------------------------------------- */
typedef struct lmg_surrogate lmg_surrogate;

/* max_batch bounds the batch of infer/gradient/train calls (device memory is
 * sized for it).  Weights start at zero. */
lmg_status lmg_surrogate_create(lmg_ctx* ctx, int32_t flags, int32_t max_batch, lmg_surrogate** out);
void lmg_surrogate_destroy(lmg_surrogate* s);
const char* lmg_surrogate_last_error(const lmg_surrogate* s);
lmg_status lmg_surrogate_shape(const lmg_surrogate* s, int32_t* padded_xyz, int64_t* n_weights);
lmg_status lmg_surrogate_get_weights(lmg_surrogate* s, float* w);
/* w == NULL zeroes the weights; either way the momentum is reset. */
lmg_status lmg_surrogate_set_weights(lmg_surrogate* s, const float* w);
/* int8 occupancy grid from a charge list (sites < V, |q| <= 127; a bad entry is
 * reported by the next lmg_surrogate_map_dev or device sync). */
lmg_status lmg_surrogate_occupancy_dev(lmg_surrogate* s, const int32_t* d_site, const int32_t* d_q, int32_t n,
                                       int8_t* d_occ);
/* batch maps (float[V] each) from batch occupancy grids */
lmg_status lmg_surrogate_infer_dev(lmg_surrogate* s, const int8_t* d_occ, int32_t batch, float* d_map);
/* Gradient of L = sum_b |map_b - target_b|^2 / sum_b |target_b|^2 with respect
 * to the weights (2 n_weights floats, device) at the current weights; *loss =
 * L.  Synchronous. */
lmg_status lmg_surrogate_gradient_dev(lmg_surrogate* s, const int8_t* d_occ, const float* d_target, int32_t batch,
                                      float* d_grad, double* loss);
/* iterations steps of momentum SGD on random minibatches of `batch` of the
 * n_snapshots (occupancy, target) pairs, starting from the current weights.
 * Each step adds a real update that is even in ky and kz (the symmetry of the
 * exact operator under lattice reflections), preconditioned per frequency by
 * the inverse mean spectral power of the first min(n, 256) occupancies; the
 * step size is 0.9 (1 - momentum) / lambda_max with lambda_max from a power
 * iteration on the preconditioned Hessian.  loss_history (host, may be NULL)
 * receives each step's minibatch loss.  Synchronous. */
lmg_status lmg_surrogate_train_dev(lmg_surrogate* s, const int8_t* d_occ, const float* d_target, int32_t n_snapshots,
                                   int32_t batch, int32_t iterations, double momentum, uint64_t seed,
                                   float* loss_history);
/* Gated map of a charge list: surrogate map, exact pair sums at gate_sites
 * (1..8192) sampled sites (seeded), gate error = relative RMS error over them; when it
 * exceeds tolerance the exact map (lmg_potential_map_if_dev) overwrites d_map.
 * The decision is taken on the device.  With gate_error and used_exact both
 * NULL the call is asynchronous on the context stream (lmg_surrogate_last_gate
 * reads the outcome later); otherwise it waits and reports the error and
 * whether the exact map was used. */
lmg_status lmg_surrogate_map_dev(lmg_surrogate* s, const int32_t* d_site, const int32_t* d_q, int32_t n, float* d_map,
                                 double tolerance, int32_t gate_sites, uint64_t gate_seed, double* gate_error,
                                 int32_t* used_exact);
/* Outcome of the latest lmg_surrogate_map_dev (waits for the stream). */
lmg_status lmg_surrogate_last_gate(lmg_surrogate* s, double* gate_error, int32_t* used_exact);
/* A batch of gated maps in one call (1 <= batch <= max_batch, grid_z * batch
 * <= 65535): map b is that of charges [o_b, o_b + counts[b]) of d_site / d_q,
 * o_b = counts[0] + ... + counts[b-1] (counts on the host), written to
 * d_map + b V; each map has its own gate over the same gate sites and its own
 * exact fallback.  gate_errors / used_exact (host, batch entries each, may be
 * NULL) as for lmg_surrogate_map_dev.  lmg_surrogate_map_dev is this call with
 * batch 1. */
lmg_status lmg_surrogate_map_batch_dev(lmg_surrogate* s, const int32_t* d_site, const int32_t* d_q,
                                       const int32_t* counts, int32_t batch, float* d_map, double tolerance,
                                       int32_t gate_sites, uint64_t gate_seed, double* gate_errors,
                                       int32_t* used_exact);
/* Outcomes of the latest gated call's first `batch` maps (waits for the stream). */
lmg_status lmg_surrogate_last_gates(lmg_surrogate* s, int32_t batch, double* gate_errors, int32_t* used_exact);
/* lmg_surrogate_map_dev records its launches once as a HIP graph and replays
 * it while the pointers, sizes and tolerance stay the same (default on); off
 * launches every kernel each call. */
lmg_status lmg_surrogate_set_graphs(lmg_surrogate* s, int32_t on);

#ifdef __cplusplus
}
#endif

#endif /* LANGMUIR_GPU_H */
