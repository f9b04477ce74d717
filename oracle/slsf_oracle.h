/*
 * slsf_oracle.h -- CPU restatement of the SLSF-rs Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product in
 * slsf-rs_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product never links, calls or falls back to it.
 *
 * Two uniform sources sit behind the same update rules:
 *   - reference mode: SplitMix64 FastRng with the reference's draw order
 *     (src/types.rs:9-82, src/updates.rs:238-486,551-620), libm sin/cos/exp in
 *     place of sleef u10 (src/fast_math.rs:3-25);
 *   - replay mode: uniforms keyed by (seed, counter) through Philox4x32-10, and
 *     the project's own FP64 sincos/exp, so that the GPU FP64 kernels reproduce
 *     every accept/reject decision and every Wolff cluster bit-for-bit.
 *
 * Every function is pinned by known-answer vectors, closed forms or brute
 * force in the oracle tests, except the rand 0.10 adapters (or_rng_f64,
 * or_rng_bool, or_rng_range): parity unpinned, as no vector for them exists
 * outside the crate.
 */
#ifndef SLSF_ORACLE_H
#define SLSF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: SplitMix64 FastRng (src/types.rs:9-82) ---------------------- */
typedef struct {
  uint64_t seed;
  uint64_t state;
  uint64_t draws; /* reference keeps u128; draws never exceed 2^64 here */
} or_fastrng;

void or_rng_seed(or_fastrng* r, uint64_t seed);
uint64_t or_rng_next_u64(or_fastrng* r);
uint32_t or_rng_next_u32(or_fastrng* r);
void or_rng_set_position(or_fastrng* r, uint64_t position);
/* rand 0.10 adapters (third-party; parity unpinned): */
double or_rng_f64(or_fastrng* r);              /* (u >> 11) * 2^-53 */
int or_rng_bool(or_fastrng* r);                /* sign bit of next_u32 */
uint64_t or_rng_range(or_fastrng* r, uint64_t n); /* Canon widening multiply */

/* ---- Philox4x32-10 (Salmon et al., SC'11) ------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- FP64 math shared (as an algorithm) with the device library -------- */
void or_sincos(double x, double* s, double* c);
double or_exp(double x);

/* ---- lattice + scratch (src/types.rs:109-214) -------------------------- */
typedef struct {
  int lx, ly, lz;
  double* theta;   /* [V], x fastest: idx = x + lx*(y + ly*z) */
  double* j_xy;    /* [lz] */
  double* j_z;     /* [lz] bond z -> z+1 */
  double* cos_t;   /* ThetaScratch cos cache [V] */
  double* sin_t;   /* ThetaScratch sin cache [V] */
  uint8_t* in_cluster; /* WolffScratch [V] */
  int* stack;      /* [V] site indices (LIFO) */
  int* cluster;    /* [V] member list */
} or_lattice;

/* returns 0 or an error code; message via or_last_error() */
int or_lattice_new(or_lattice* lat, int lx, int ly, int lz);
void or_lattice_free(or_lattice* lat);
void or_scratch_refresh(or_lattice* lat, int replay_math);
const char* or_last_error(void);

/* ---- initialisation (src/initialization.rs) ---------------------------- */
int or_two_point_disorder(double* values, int n, double mean, double delta,
                          or_fastrng* rng, const char* coupling_name);
int or_uniform_disorder(double* values, int n, double lo, double hi, or_fastrng* rng);
int or_initialize_disorder(or_lattice* lat, double j_xy, double j_z_mean,
                           double delta_j_z, or_fastrng* rng);
void or_initialize_angles(or_lattice* lat, int cold, or_fastrng* rng);
void or_initialize_angles_keyed(or_lattice* lat, uint64_t seed);

/* ---- Metropolis -------------------------------------------------------- */
/* acceptance probability helper (src/updates.rs:21-53), lanes 4 or 8 */
int or_vector_exp_min_uphill(int lanes);
void or_metropolis_accept_probabilities(const double* delta, int lanes, double beta,
                                        double* prob);
/* reference mode: FastRng, batch8 when use_batch8 (AVX-512 host), then batch4,
 * then scalar tail (src/updates.rs:440-486).  Returns 0 on success. */
int or_metropolis_sweep_ref(or_lattice* lat, double temperature, double width,
                            int use_batch8, or_fastrng* rng, double* acceptance);
/* random-site single update (src/updates.rs:410-427) */
int or_local_metropolis_step_ref(or_lattice* lat, double temperature, double width,
                                 or_fastrng* rng, int* accepted);
/* replay mode: keyed uniforms, own FP64 math */
int or_metropolis_sweep_replay(or_lattice* lat, double temperature, double width,
                               uint64_t seed, uint32_t sweep, uint8_t* decisions,
                               double* acceptance);

/* ---- Wolff ------------------------------------------------------------- */
double or_wolff_add_probability(double beta, double coupling, double ri, double rj);
double or_wolff_reflect_angle(double theta, double phi);
/* reference mode (use_scratch selects the ThetaScratch path) */
int or_wolff_step_ref(or_lattice* lat, double temperature, int use_scratch,
                      or_fastrng* rng, int64_t* cluster_size);
/* replay mode: seed/phi and canonical-bond uniforms keyed by (seed, step);
 * membership written to members[V] (0/1) when non-NULL. */
int or_wolff_step_replay(or_lattice* lat, double temperature, uint64_t seed,
                         uint32_t step, uint8_t* members, int64_t* cluster_size);

/* ---- observables (src/observables.rs) ---------------------------------- */
typedef struct {
  double energy, magnetization_squared;
  double cos_x, cos_y, cos_z, sin_x, sin_y, sin_z;
} or_observables;

void or_measure_observables(const or_lattice* lat, or_observables* out);
double or_measure_energy_direct(const or_lattice* lat);
double or_measure_magnetization_squared_direct(const or_lattice* lat);
void or_helicity_sums_direct(const or_lattice* lat, double out[6]);
int or_measure_helicity_modulus(const or_lattice* lat, double temperature, double out[3]);

/* correlations (src/observables.rs:166-265).  rmax_* < 0 means "None".
 * Output sizes: corr_x/corr_y/corr_xy [rxy], corr_xy_by_z [lz*rxy], corr_z [rz]. */
void or_correlation_rmax(const or_lattice* lat, int rmax, int rmax_xy, int rmax_z,
                         int* rxy_eff, int* rz_eff);
void or_measure_correlations(const or_lattice* lat, int rmax, int rmax_xy, int rmax_z,
                             double* corr_x, double* corr_y, double* corr_xy,
                             double* corr_xy_by_z, double* corr_z);

/* ---- plain driver (src/simulation.rs:13-230) --------------------------- */
typedef struct {
  int thermal_sweeps, measure_sweeps, measure_interval;
  double proposal_width;
  int wolff_steps, correlation_interval;
  int correlation_rmax, correlation_rmax_xy, correlation_rmax_z; /* <0 = None */
  int cold_init;
  int use_batch8;
} or_sim_options;

typedef struct {
  double energy, rho_sx, rho_sy, rho_sz, std_rho_sx, std_rho_sy, std_rho_sz;
  double magnetization_squared, chi, acceptance;
  int num_measurements, num_correlation_measurements;
  int rxy, rz;
  /* caller-provided buffers of capacity >= lx/2, lz/2, lz*lx/2: */
  double* corr_x; double* corr_y; double* corr_xy; double* corr_xy_by_z; double* corr_z;
} or_sim_result;

int or_run_theta_simulation(or_lattice* lat, double j_xy, double j_z_mean,
                            double delta_j_z, double temperature,
                            const or_sim_options* opt, or_fastrng* rng,
                            or_sim_result* res);

/* Replay-mode driver: the same flow and estimators as or_run_theta_simulation,
 * with disorder from explicit arrays (when non-NULL) or FastRng(seed), angles
 * and updates from keyed uniforms of `seed` (the GPU driver's contract). */
int or_run_theta_simulation_replay(or_lattice* lat, double j_xy, double j_z_mean,
                                   double delta_j_z, double temperature,
                                   const double* j_xy_array, const double* j_z_array,
                                   const or_sim_options* opt, uint64_t seed,
                                   or_sim_result* res);

/* ---- estimators -------------------------------------------------------- */
double or_mean(const double* x, int n);
double or_sample_std(const double* x, int n);
double or_blocking_stderr(const double* x, int n);
/* carlo-mc BinnedEstimate::from_samples semantics restated from the reference
 * test (src/tests.rs:436-446): bins = block means, stderr = std(bins)/sqrt(nb) */
int or_binned_estimate(const double* samples, int n, int binsize, double* mean,
                       double* stderr_out, double* bins, int* nbins);

#ifdef __cplusplus
}
#endif
#endif
