/*
 * slsf.h -- C ABI of the MI355X (gfx950) Monte Carlo backend for the 3D
 * layered-disorder XY model of SLSF-rs.
 *
 * The reference binds its hot path in two places, and this ABI replaces both:
 *   1. the carlo-mc plugin `impl MonteCarlo for ThetaModel`
 *      (src/model.rs:308-477: new / init / sweep / measure / finalize_estimates /
 *      task_metadata / build_tasks), and
 *   2. the free functions it calls (src/updates.rs:440,551;
 *      src/observables.rs:80,166; src/simulation.rs:13).
 * The lattice, its trig cache and its RNG state live on the device inside an
 * opaque batch handle; a single task is a batch of one.  Only FP64 scalars and
 * bins cross back to the host.  All entry points return 0 on success and a
 * nonzero SLSF_E_* code otherwise; slsf_last_error() then returns the message,
 * which for every validation the reference performs is the reference's exact
 * string (e.g. "temperature T must be positive and finite", src/types.rs:375).
 *
 * Threading: a batch handle is used by one host thread at a time; each handle
 * owns one HIP stream.  Error messages are thread-local.
 */
#ifndef SLSF_H
#define SLSF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLSF_OK 0
#define SLSF_E_INVALID 1   /* bad argument; message is the reference's string */
#define SLSF_E_DEVICE 2    /* HIP runtime error, or no usable GPU */
#define SLSF_E_ADMISSION 3 /* does not fit the configured device-memory budget */
#define SLSF_E_STATE 4     /* call not valid in the current state */

/* Message of the last failed call on this thread ("" if none). */
const char* slsf_last_error(void);
/* ABI version, bumped on any incompatible change. */
int32_t slsf_abi_version(void);

/* ======================================================================== */
/* Replicas and batches                                                      */
/* ======================================================================== */

/* Arithmetic of a batch.
 * SLSF_FP64_REPLAY: FP64 angles and trig cache in HBM, keyed uniforms, the
 *   library's own FP64 sincos/exp; reproduces the oracle's accept/reject
 *   decisions and Wolff clusters bit-for-bit.
 * SLSF_FP32: FP32 sweeps (hardware sin/cos/exp2), FP64 accumulators.  Cubic
 *   replicas with L in {8,12,16,24,32} run in the register/LDS-resident batched
 *   kernel, larger ones in the HBM kernels. */
#define SLSF_FP64_REPLAY 0
#define SLSF_FP32 1

/* One replica: the GPU counterpart of ThetaLattice + Parameters + proposal
 * width + wolff_steps + task seed (src/types.rs:84-161; src/model.rs:20-47).
 * j_xy[z] couples the in-plane bonds of layer z, j_z[z] the bond z -> z+1
 * (src/updates.rs:131-167).  `seed` keys every uniform of the replica. */
typedef struct {
  int32_t l_x, l_y, l_z;
  int32_t wolff_steps;
  double temperature;
  double proposal_width;
  const double* j_xy; /* [l_z] */
  const double* j_z;  /* [l_z] */
  uint64_t seed;
} slsf_replica_desc;

typedef struct slsf_batch slsf_batch;

/* Replaces ThetaModel::new (src/model.rs:314-371) for n replicas at once.
 * Validates like the reference: "L_x must be positive" (and y/z), "temperature
 * T must be positive and finite", "proposal_width must be nonnegative and
 * finite", "theta simulation requires nonnegative layer couplings", and for the
 * checkerboard "red-black Metropolis sweep requires even lattice dimensions". */
int slsf_batch_create(const slsf_replica_desc* replicas, int32_t n, int32_t precision,
                      slsf_batch** out);
void slsf_batch_destroy(slsf_batch* batch);
int32_t slsf_batch_size(const slsf_batch* batch);
int32_t slsf_batch_precision(const slsf_batch* batch);
/* Device bytes held by the handle. */
int64_t slsf_batch_device_bytes(const slsf_batch* batch);

/* Replaces ThetaModel::init / initialize_angles (src/model.rs:373-377;
 * src/initialization.rs:48-62).  Random: theta = 2 pi u, u keyed by (seed, site);
 * cold: theta = 0.  Resets the sweep and cluster counters to 0. */
int slsf_init_angles(slsf_batch* batch, int32_t cold);

/* Host <-> device copies of one replica's angles (FP64 on the host side), for
 * checkpoints and tests.  set_theta refreshes the trig cache. */
int slsf_get_theta(slsf_batch* batch, int32_t replica, double* theta_out);
int slsf_set_theta(slsf_batch* batch, int32_t replica, const double* theta);

/* Checkpoint image of a batch's whole run state: angles and trig cache as the
 * kernels hold them, keyed-stream counters, acceptance and cluster tallies,
 * and the measurement state (open-bin accumulators, bin counters, completed
 * bins).  Replaces the serde checkpoint of ThetaModel (src/model.rs:255-271)
 * and the RNG position of FastRng (src/types.rs:27-36): a batch created from
 * the same replica descriptors and loaded from the image continues with
 * exactly the sweeps, uniforms and bins the saved one would have produced.
 * The image is versioned bytes; get its size, then save into a buffer of at
 * least that size.  load checks the version, precision, replica count and
 * each replica's size, seed, temperature, proposal width and wolff steps. */
int slsf_checkpoint_size(slsf_batch* batch, int64_t* bytes);
int slsf_checkpoint_save(slsf_batch* batch, void* buffer, int64_t bytes);
int slsf_checkpoint_load(slsf_batch* batch, const void* buffer, int64_t bytes);

/* Counters of keyed uniforms: Metropolis sweeps and Wolff clusters done so far.
 * Together with theta they are the whole Markov-chain state. */
int slsf_get_counters(const slsf_batch* batch, uint32_t* metro_sweeps, uint32_t* wolff_steps);
int slsf_set_counters(slsf_batch* batch, uint32_t metro_sweeps, uint32_t wolff_steps);

/* Replaces metropolis_sweep_with_scratch (src/updates.rs:440-486): one red/black
 * sweep of every replica.  acceptance[n] (may be NULL) receives accepted/V;
 * decisions (may be NULL) receives one byte per site of every replica
 * (concatenated in replica order) -- 1 where the proposal was accepted. */
int slsf_metropolis_sweep(slsf_batch* batch, double* acceptance, uint8_t* decisions);

/* Replaces wolff_cluster_step_with_theta_scratch (src/updates.rs:551-620): one
 * Wolff cluster per replica.  cluster_size[n] may be NULL; members (may be NULL)
 * receives one byte per site, 1 for the sites that were flipped. */
int slsf_wolff_cluster_step(slsf_batch* batch, int64_t* cluster_size, uint8_t* members);

/* Replaces ThetaModel::sweep (src/model.rs:379-400): `nsweeps` times one
 * Metropolis sweep + wolff_steps clusters.  acceptance_sum[n] (may be NULL)
 * accumulates the per-sweep acceptance like ThetaModel::acceptance_sum. */
int slsf_sweep(slsf_batch* batch, int32_t nsweeps, double* acceptance_sum);

/* ThetaObservables (src/types.rs:253-263) plus m^4 = (m^2)^2. */
typedef struct {
  double energy;                /* E / V */
  double magnetization_squared; /* |sum e^{i theta}|^2 / V^2 */
  double cos_x, cos_y, cos_z;   /* sum J cos(theta_i - theta_{i+a}) */
  double sin_x, sin_y, sin_z;   /* sum J sin(theta_i - theta_{i+a}) */
} slsf_observables;

/* Replaces measure_theta_observables_with_scratch (src/observables.rs:80-154). */
int slsf_measure_observables(slsf_batch* batch, slsf_observables* out /* [n] */);

/* Effective correlation radii for a replica (src/observables.rs:177-181);
 * rmax arguments < 0 mean "None". */
int slsf_correlation_rmax(const slsf_batch* batch, int32_t replica, int32_t rmax,
                          int32_t rmax_xy, int32_t rmax_z, int32_t* rxy, int32_t* rz);

/* Replaces measure_theta_correlations_with_scratch (src/observables.rs:166-265)
 * for one replica.  Buffers: corr_x, corr_y, corr_xy [rxy]; corr_xy_by_z
 * [l_z * rxy] indexed [z][r-1]; corr_z [rz]. */
int slsf_measure_correlations(slsf_batch* batch, int32_t replica, int32_t rmax,
                              int32_t rmax_xy, int32_t rmax_z, double* corr_x,
                              double* corr_y, double* corr_xy, double* corr_xy_by_z,
                              double* corr_z);

/* slsf_measure_correlations without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, same values bit for bit; the row sums and the fold live in
 * one fixed 16 MiB device array per GPU (2^21 doubles) that the library holds
 * until the process ends.  A replica needs rxy*ly*lz*3 + rz*ly*lz + 3*rxy +
 * lz*rxy + rz doubles of it (L = 64 cubic at the default radii: 0.53 M); larger
 * requests fail with SLSF_E_INVALID and a message.  Every call on a GPU uses
 * that GPU's one array: the library serialises its own calls per device until
 * their work has finished, so calls from several host threads never overlap
 * on it. */
int slsf_measure_correlations_noalloc(slsf_batch* batch, int32_t replica, int32_t rmax,
                                      int32_t rmax_xy, int32_t rmax_z, double* corr_x,
                                      double* corr_y, double* corr_xy,
                                      double* corr_xy_by_z, double* corr_z);

/* ======================================================================== */
/* Plain driver                                                              */
/* ======================================================================== */

/* ThetaSimulationOptions (src/types.rs:265-296); rmax < 0 means None. */
typedef struct {
  int32_t thermal_sweeps, measure_sweeps, measure_interval;
  double proposal_width;
  int32_t wolff_steps, correlation_interval;
  int32_t correlation_rmax, correlation_rmax_xy, correlation_rmax_z;
  int32_t cold_init;
} slsf_simulation_options;

/* ThetaSimulationResult (src/types.rs:298-320).  The correlation vectors are
 * caller buffers with capacity >= l_x/2 (corr_x, corr_y, corr_xy), l_z * l_x/2
 * (corr_xy_by_z, [z][r-1]) and l_z/2 (corr_z); rxy / rz receive their lengths. */
typedef struct {
  double energy, rho_sx, rho_sy, rho_sz, std_rho_sx, std_rho_sy, std_rho_sz;
  double magnetization_squared, chi, acceptance;
  int32_t num_measurements, num_correlation_measurements, rxy, rz;
  double* corr_x;
  double* corr_y;
  double* corr_xy;
  double* corr_xy_by_z;
  double* corr_z;
} slsf_simulation_result;

/* Replaces run_theta_simulation (src/simulation.rs:13-195) on the GPU.  Same
 * validation order and messages.  Disorder: j_xy_array / j_z_array when given
 * (length l_z), else the reference's initialize_disorder (uniform J_xy, two-point
 * J_z) drawn from SplitMix64(seed).  Angles and updates use the keyed uniforms of
 * `seed`.  precision: SLSF_FP64_REPLAY or SLSF_FP32. */
int slsf_run_theta_simulation(int32_t l_x, int32_t l_y, int32_t l_z, double j_xy,
                              double j_z_mean, double delta_j_z, double temperature,
                              const double* j_xy_array, const double* j_z_array,
                              const slsf_simulation_options* options, uint64_t seed,
                              int32_t precision, slsf_simulation_result* result);

/* ======================================================================== */
/* Production runs with on-device binning                                    */
/* ======================================================================== */

/* Per-measurement series, in the reference's names (src/model.rs:402-442,
 * 847-855) plus the m^4 extension used for the Binder ratio. */
#define SLSF_SERIES_ENERGY 0          /* "Energy" */
#define SLSF_SERIES_ENERGY2 1         /* "EnergySquared" */
#define SLSF_SERIES_M2 2              /* "MagnetizationSquared" */
#define SLSF_SERIES_M4 3              /* "MagnetizationFourth" (extension) */
#define SLSF_SERIES_COS_X 4           /* "_helicity_cos_x" */
#define SLSF_SERIES_COS_Y 5
#define SLSF_SERIES_COS_Z 6
#define SLSF_SERIES_SIN_X 7           /* "_helicity_sin_x" */
#define SLSF_SERIES_SIN_Y 8
#define SLSF_SERIES_SIN_Z 9
#define SLSF_SERIES_SIN2_X 10         /* "_helicity_sin2_x" */
#define SLSF_SERIES_SIN2_Y 11
#define SLSF_SERIES_SIN2_Z 12
#define SLSF_NUM_SERIES 13

/* Name of a series as the reference's results use it. */
const char* slsf_series_name(int32_t series);

typedef struct {
  int64_t thermalization;   /* sweeps before measuring (src/model.rs:29) */
  int64_t sweeps;           /* measured sweeps, one measurement each */
  int32_t binsize;          /* measurements per bin (carlo-mc binsize) */
  int32_t sweeps_per_launch;/* 0 = autotuned/default */
  int32_t use_graph;        /* capture launches in a hipGraph and replay */
  int32_t stream;           /* FP32 resident kernel: SLSF_STREAM_PRODUCTION (one
                               keyed Philox call per two site updates) or
                               SLSF_STREAM_REPLAY (the FP64 replay path's per-site
                               uniforms, for decision-mismatch statistics) */
  int32_t hbm_form;         /* replicas on the HBM kernels: 0 = by launch size,
                               1 = slab sweeps (three launches a sweep), 2 =
                               per-phase sweeps (seven); the same chains bit for
                               bit, the same bins up to FP64 summation order */
  int32_t reserved;         /* 0 */
} slsf_run_options;

#define SLSF_STREAM_PRODUCTION 0
#define SLSF_STREAM_REPLAY 1

typedef struct {
  double seconds;           /* wall time of the run (device-synchronised) */
  double kernel_ms;         /* sum of HIP-event kernel times */
  int64_t site_updates;     /* sum over replicas of V * sweeps */
  int64_t launches;         /* kernel launches issued */
  double acceptance_mean;   /* mean Metropolis acceptance over replicas */
  double mean_cluster_size; /* mean Wolff cluster size over replicas */
} slsf_run_stats;

/* Runs thermalization + sweeps for every replica, accumulating FP64 bin means
 * of every series on the device.  May be called repeatedly; bins append, and a
 * bin left open by one run (fewer than binsize measurements) is continued by
 * the next. */
int slsf_run(slsf_batch* batch, const slsf_run_options* opt, slsf_run_stats* stats);

/* Autotuner of the FP32 batched kernel.  For every replica size of the batch
 * it times each candidate number of z-columns per thread (i.e. waves per
 * replica) on the batch's own replicas, trial_sweeps sweeps in one launch per
 * candidate, twice in alternation, keeps the fastest, and restores the
 * replicas' state and counters, so tuning does not change any chain.  Runs keep
 * the default launch length (the longest, which only gains; sweeps_per_launch
 * of the records).  records (may be NULL) receives up to cap records, one per
 * (size, candidate); *count the number of records. */
typedef struct {
  int32_t l;                  /* replica edge of the group */
  int32_t columns_per_thread; /* 1 or 2 */
  int32_t waves_per_replica;
  int32_t sweeps_per_launch;
  double site_updates_per_s;  /* trial rate */
  int32_t chosen;             /* 1 for the configuration kept */
  int32_t reserved;
} slsf_tune_record;
int slsf_autotune(slsf_batch* batch, int32_t trial_sweeps, slsf_tune_record* records,
                  int32_t cap, int32_t* count);

/* Reserves device storage for nbins completed bins per (replica, series), so
 * that measured runs up to that many bins allocate nothing (the analogue of
 * reserving a Vec's capacity; no reference counterpart).  Completed bins are
 * kept. */
int slsf_reserve_bins(slsf_batch* batch, int32_t nbins);

/* Completed bins of one series of one replica: writes min(cap, nbins) values
 * and returns the count in *nbins. */
int slsf_get_bins(slsf_batch* batch, int32_t replica, int32_t series, double* out,
                  int32_t cap, int32_t* nbins);

/* Mean Metropolis acceptance per replica since init (task_metadata,
 * src/model.rs:460-467). */
int slsf_get_acceptance(slsf_batch* batch, double* acceptance /* [n] */);

/* ======================================================================== */
/* Estimators (host, FP64)                                                   */
/* ======================================================================== */

/* BinnedEstimate::from_samples as the reference test pins it
 * (src/tests.rs:436-446): bins are means of `binsize` consecutive samples,
 * mean = mean of bins, stderr = sample std of bins / sqrt(nbins).  Error:
 * "binsize is larger than the sample series". */
int slsf_binned_estimate(const double* samples, int64_t n, int32_t binsize, double* mean,
                         double* stderr_out, double* bins /* [n/binsize] */,
                         int64_t* nbins);

/* blocking_stderr (src/simulation.rs:210-226). */
double slsf_blocking_stderr(const double* x, int64_t n);

/* Derived observables of the reference's Evaluator (src/model.rs:857-888) plus
 * the Binder ratio, each with a jackknife error over bins.  bins is laid out
 * [SLSF_NUM_SERIES][nbins]. */
#define SLSF_EVAL_MAGNETIZATION 0  /* sqrt(<m2>) */
#define SLSF_EVAL_CHI 1            /* beta V <m2> */
#define SLSF_EVAL_SPECIFIC_HEAT 2  /* beta^2 V (<E2> - <E>^2) */
#define SLSF_EVAL_RHO_XY 3         /* (rho_x + rho_y) / 2 */
#define SLSF_EVAL_RHO_Z 4
#define SLSF_EVAL_BINDER 5         /* <m4> / <m2>^2 (extension) */
#define SLSF_EVAL_ENERGY 6         /* <E> */
#define SLSF_NUM_EVALS 7
const char* slsf_eval_name(int32_t eval);
int slsf_evaluate(const double* bins, int64_t nbins, double volume, double beta,
                  double* value /* [SLSF_NUM_EVALS] */, double* error /* [SLSF_NUM_EVALS] */);

/* Jackknife of f(mean of columns) for a user function given as values of the
 * leave-one-out estimates: loo[nbins], full estimate f_all.  Returns the
 * bias-corrected estimate and the jackknife standard error. */
int slsf_jackknife(const double* loo, int64_t nbins, double f_all, double* estimate,
                   double* stderr_out);

/* Disorder averages over realizations (extension; the reference lists every
 * task and averages nothing across samples).  means[r * SLSF_NUM_SERIES + s] is
 * realization r's mean of series s (e.g. the mean of its bins).  value[k] is the
 * mean over realizations of evaluable k of each realization
 * (src/model.rs:857-888), except the Binder ratio, which is the ratio of
 * disorder averages [<m^4>] / [<m^2>]^2; error[k] is the jackknife error over
 * realizations (each realization left out in turn).  Errors: fewer than two
 * realizations. */
int slsf_disorder_average(const double* means /* [n_real][SLSF_NUM_SERIES] */,
                          int32_t n_real, double volume, double beta,
                          double* value /* [SLSF_NUM_EVALS] */,
                          double* error /* [SLSF_NUM_EVALS] */);

/* Crossing temperature of the disorder average of evaluable `eval`, scaled by
 * scale_a and scale_b (1 for the Binder ratio, L for the helicity moduli, whose
 * L rho curves cross at a 3D XY transition), between two lattice sizes a and b
 * (extension): means_a[(t * n_real + r) * SLSF_NUM_SERIES + s] as in
 * slsf_disorder_average for temperature temps[t] (increasing), and the same for
 * b.  The first t with d = scale_a f_a - scale_b f_b changing sign (or zero)
 * between temps[t] and temps[t + 1] gives the crossing by linear interpolation;
 * *found = 0 when the curves do not cross.  The error is the jackknife over
 * realizations: realization r is left out of both sizes at every temperature
 * and the crossing found again (NaN when some leave-one-out curves do not
 * cross).  Errors: fewer than two temperatures or realizations. */
int slsf_crossing(const double* temps, int32_t n_temps, int32_t n_real, int32_t eval,
                  const double* means_a, double volume_a, double scale_a,
                  const double* means_b, double volume_b, double scale_b, int32_t* found,
                  double* t_cross, double* t_error);

/* ======================================================================== */
/* Serving: continuous batching of independent jobs                          */
/* ======================================================================== */
/* No reference counterpart: the reference runs each task in its own process
 * (src/results.rs:187-239).  A server keeps one slot pool per lattice size on
 * the device; a job occupies a slot from admission to completion, its state
 * and FP64 accumulators stay resident between chunks, and every step runs one
 * chunk of sweeps for all active slots of every pool (pools on their own
 * streams, so their work overlaps).  Jobs arriving while others run are
 * admitted as soon as a slot of their size is free and the pool fits the
 * device-memory budget.  A job's result does not depend on what else runs:
 * its uniforms are keyed by its own seed and counters. */
typedef struct {
  int32_t l;                 /* cubic edge, even */
  int32_t wolff_steps;
  double temperature;
  double proposal_width;
  double j_xy_min, j_xy_max; /* J_xy[z] ~ U[j_xy_min, j_xy_max] iid per layer */
  double j_perp;             /* J_z[z] for every layer */
  uint64_t disorder_seed;    /* SplitMix64 seed of the layer couplings */
  uint64_t seed;             /* key of the job's uniforms */
  int64_t thermalization, sweeps;
  int32_t binsize;           /* raised if sweeps/binsize exceeds the pool's bin capacity */
  /* Correlation series (src/model.rs:360-365,410-435): with a radius above 0
   * the job measures C(r) on the measurements m with m % interval == 0
   * (interval below 1 counts as 1) for r <= min(rmax_xy, l/2) in plane and
   * r <= min(rmax_z, l/2) along z, and bins the samples binsize to a bin. */
  int32_t correlation_interval;
  int32_t correlation_rmax_xy, correlation_rmax_z;
} slsf_job_request;

typedef struct {
  int64_t job_id;
  int32_t l, nbins;
  double submit_s, start_s, finish_s; /* server clock (seconds since create) */
  double acceptance;
  double value[SLSF_NUM_EVALS], error[SLSF_NUM_EVALS]; /* slsf_evaluate of the bins */
  double series_mean[SLSF_NUM_SERIES];
  int64_t site_updates;
  /* correlation series: radii, complete bins and series count
   * (3 rxy + l rxy + rz); the bins are taken with slsf_server_correlations */
  int32_t corr_rxy, corr_rz, corr_nbins, corr_nseries;
} slsf_job_result;

typedef struct {
  int64_t queued, active, finished, device_bytes, site_updates;
  double clock_s;
} slsf_server_stats;

typedef struct slsf_server slsf_server;

/* device_budget_bytes: admission limit for all pools; max_slots_per_size:
 * slots of one lattice size; chunk_sweeps: sweeps per step. */
int slsf_server_create(int64_t device_budget_bytes, int32_t max_slots_per_size,
                       int32_t chunk_sweeps, slsf_server** out);
void slsf_server_destroy(slsf_server* server);
/* Validates and queues a job (reference messages for bad L, T, width,
 * couplings); SLSF_E_ADMISSION if one slot of its size can never fit. */
int slsf_server_submit(slsf_server* server, const slsf_job_request* job, int64_t* job_id);
/* One scheduling step: admit, run one chunk everywhere, retire finished jobs. */
int slsf_server_step(slsf_server* server, int32_t* finished);
/* Drains finished results (oldest first). */
int slsf_server_poll(slsf_server* server, slsf_job_result* out, int32_t cap, int32_t* n);
int slsf_server_get_stats(const slsf_server* server, slsf_server_stats* out);
/* Bin means of a finished job's correlation series, [corr_nseries][corr_nbins],
 * series in the order CorrX_r1.., CorrY_r1.., CorrXY_r1.., CorrXY_z0_r1..
 * (layer-major), CorrZ_r1.. (src/model.rs:413-435).  out == NULL: *count
 * receives the number of doubles and the series stay; otherwise they are
 * copied (cap >= that number) and released.  The server holds a finished
 * job's series until this call takes them. */
int slsf_server_correlations(slsf_server* server, int64_t job_id, double* out, int64_t cap,
                             int64_t* count);
/* Bytes one slot of size l occupies (the admission footprint); a pool of size
 * l adds (l^2 + l) doubles for the correlation sums of its samples. */
int64_t slsf_server_slot_bytes(int32_t l);
/* Free and total memory of the current device (hipMemGetInfo), for choosing a
 * server budget. */
int slsf_device_memory(int64_t* free_bytes, int64_t* total_bytes);

/* ======================================================================== */
/* Device ceilings (measurement helper; no counterpart in the reference)     */
/* ======================================================================== */

/* Microbenchmarks of the ceilings the kernels run against, timed with HIP
 * events on the current device: LDS read bandwidth over all CUs, and the issue
 * rates (wave instructions per second) of independent FP32 FMAs and of the
 * hardware v_sin and v_exp. */
typedef struct {
  double lds_bytes_per_s;
  double fma_wave_instr_per_s;
  double sin_wave_instr_per_s;
  double exp_wave_instr_per_s;
  int32_t compute_units, reserved;
} slsf_ceilings;
int slsf_measure_ceilings(slsf_ceilings* out);

/* Device self-check: the number of floats x in {-0} U [+0, +inf] for which the
 * hardware v_exp_f32 (2^x) returns less than 1, over all of them.  The FP32
 * resident kernel decides "dE <= 0 or u < exp(-beta dE)" (src/updates.rs:255-264)
 * by the second test alone, which is exact only when this count is 0. */
int slsf_check_exp2_floor(uint64_t* below);

/* ======================================================================== */
/* Job configuration (TOML) and task expansion                               */
/* ======================================================================== */

/* A task as ThetaTask (src/model.rs:20-47).  Couplings are per layer. */
typedef struct {
  char name[128];
  int32_t l, l_x, l_y, l_z;
  double temperature, j_xy, delta_j_xy, j_z_mean, delta_j_z;
  uint64_t disorder_seed, seed;
  int32_t sample;
  int64_t sweeps, thermalization;
  int32_t binsize;
  double proposal_width;
  int32_t wolff_steps;
  int32_t correlation_rmax, correlation_rmax_xy, correlation_rmax_z;
  int32_t correlation_interval;
  int32_t disorder_kind; /* 0 two-point (reference), 1 uniform (extension) */
} slsf_task;

typedef struct slsf_job slsf_job;

/* Parses a job TOML with the reference's schema and aliases
 * (src/model.rs:159-213,549-644) and expands it into tasks with the
 * reference's names, seeds and disorder draws (src/model.rs:697-810).
 * Extension keys: [model] disorder = "two_point" | "uniform",
 * j_xy_min/j_xy_max (uniform bounds).  Errors carry the reference's messages. */
int slsf_job_from_toml(const char* toml_text, slsf_job** out);
void slsf_job_destroy(slsf_job* job);
const char* slsf_job_name(const slsf_job* job);
int64_t slsf_job_num_tasks(const slsf_job* job);
int slsf_job_task(const slsf_job* job, int64_t index, slsf_task* out);
/* Layer couplings of a task: j_xy[l_z], j_z[l_z]. */
int slsf_job_task_couplings(const slsf_job* job, int64_t index, double* j_xy, double* j_z);
/* Durations run_time / checkpoint_time in seconds (src/model.rs:1033-1060). */
int slsf_job_durations(const slsf_job* job, int64_t* run_time_s, int64_t* checkpoint_time_s);
/* Round-robin selection idx % world == rank (src/model.rs:62-69). */
int64_t slsf_job_selected_tasks(const slsf_job* job, int32_t rank, int32_t world,
                                int64_t* indices, int64_t cap);
/* "[D-][[HH:]MM:]SS" -> seconds (src/model.rs:1033-1060). */
int slsf_parse_duration(const char* text, int64_t* seconds);

#ifdef __cplusplus
}
#endif
#endif /* SLSF_H */
