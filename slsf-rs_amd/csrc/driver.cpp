// driver.cpp -- the plain driver run_theta_simulation (src/simulation.rs:13-230)
// on top of the batch entry points: disorder, initial angles, thermalisation,
// measurement loop with throttled correlations, and the reference's estimators
// (per-direction helicity moduli, blocking errors of the effective series,
// averaged C(r)).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/slsf.h"
#include "common.h"
#include "driver_session.h"

using namespace slsf;

namespace {

struct SplitMix64 {  // FastRng (src/types.rs:9-82)
  uint64_t state;
  uint64_t next() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  bool next_bool() { return (int32_t)(uint32_t)next() < 0; }
};

double mean(const std::vector<double>& x) {
  double s = 0.0;
  for (double v : x) s += v;
  return s / (double)x.size();
}
double mean_square(const std::vector<double>& x) {
  double s = 0.0;
  for (double v : x) s += v * v;
  return s / (double)x.size();
}

struct BatchGuard {
  slsf_batch* b = nullptr;
  DriverSession* s = nullptr;
  ~BatchGuard() {
    driver_free(s);
    slsf_batch_destroy(b);
  }
};

}  // namespace

extern "C" int slsf_run_theta_simulation(int32_t l_x, int32_t l_y, int32_t l_z, double j_xy,
                                         double j_z_mean, double delta_j_z, double temperature,
                                         const double* j_xy_array, const double* j_z_array,
                                         const slsf_simulation_options* o, uint64_t seed,
                                         int32_t precision, slsf_simulation_result* res) {
  clear_error();
  if (!o || !res) return fail(SLSF_E_INVALID, "null argument");
  // validation order of src/simulation.rs:19-40
  if (o->measure_interval <= 0) return fail(SLSF_E_INVALID, "measure_interval must be positive");
  if (o->correlation_interval <= 0)
    return fail(SLSF_E_INVALID, "correlation_interval must be positive");
  const int nmeas = o->measure_sweeps / o->measure_interval;
  if (nmeas == 0)
    return fail(SLSF_E_INVALID,
                "measure_sweeps must include at least one measurement; require measure_sweeps "
                ">= measure_interval");
  if (!temperature_ok(temperature))
    return fail(SLSF_E_INVALID, "temperature T must be positive and finite");
  if (!width_ok(o->proposal_width))
    return fail(SLSF_E_INVALID, "proposal_width must be nonnegative and finite");
  if (l_x <= 0) return fail(SLSF_E_INVALID, "L_x must be positive");
  if (l_y <= 0) return fail(SLSF_E_INVALID, "L_y must be positive");
  if (l_z <= 0) return fail(SLSF_E_INVALID, "L_z must be positive");
  // disorder (src/initialization.rs:33-46 via src/simulation.rs:42)
  std::vector<double> jxy(l_z), jz(l_z);
  if (j_xy_array && j_z_array) {
    jxy.assign(j_xy_array, j_xy_array + l_z);
    jz.assign(j_z_array, j_z_array + l_z);
  } else {
    for (auto& v : jxy) v = j_xy;
    if (delta_j_z < 0.0) return fail(SLSF_E_INVALID, "\xce\xb4J_z must be nonnegative");
    if (j_z_mean - delta_j_z < 0.0)
      return fail(SLSF_E_INVALID, "two-point layer disorder requires J_z_mean - \xce\xb4J_z >= 0");
    SplitMix64 rng{seed};
    for (auto& v : jz)
      v = delta_j_z == 0.0 ? j_z_mean : (rng.next_bool() ? j_z_mean + delta_j_z : j_z_mean - delta_j_z);
  }
  slsf_replica_desc d;
  d.l_x = l_x; d.l_y = l_y; d.l_z = l_z;
  d.wolff_steps = o->wolff_steps;
  d.temperature = temperature;
  d.proposal_width = o->proposal_width;
  d.j_xy = jxy.data();
  d.j_z = jz.data();
  d.seed = seed;
  BatchGuard g;
  if (int rc = slsf_batch_create(&d, 1, precision, &g.b)) return rc;
  if (int rc = slsf_init_angles(g.b, o->cold_init)) return rc;
  int32_t rxy = 0, rz = 0;
  if (int rc = slsf_correlation_rmax(g.b, 0, o->correlation_rmax, o->correlation_rmax_xy,
                                     o->correlation_rmax_z, &rxy, &rz))
    return rc;
  // The whole run is queued on the batch's stream: measurements go to device
  // buffers (the host folds' order, on the device) and are read back once.
  if (int rc = driver_open(g.b, nmeas, rxy, rz, &g.s)) return rc;
  DriverSchedule plan;
  plan.thermal = o->thermal_sweeps > 0 ? o->thermal_sweeps : 0;
  plan.measure_sweeps = o->measure_sweeps;
  plan.measure_interval = o->measure_interval;
  plan.correlation_interval = o->correlation_interval;
  plan.corr = rxy > 0 || rz > 0;
  int64_t ncorr = 0;
  if (int rc = driver_run(g.s, plan, &ncorr)) return rc;
  std::vector<double> series((size_t)nmeas * 8), sums((size_t)driver_sums_size(g.s));
  double acc_sum = 0.0;
  if (int rc = driver_close(g.s, (int64_t)o->thermal_sweeps + o->measure_sweeps, &acc_sum,
                            series.data(), sums.data()))
    return rc;
  std::vector<double> en(nmeas), cx(nmeas), cy(nmeas), cz(nmeas), sx(nmeas), sy(nmeas),
      sz(nmeas), m2(nmeas);
  for (int i = 0; i < nmeas; ++i) {
    const double* v = &series[(size_t)i * 8];
    en[i] = v[0];
    m2[i] = v[1];
    cx[i] = v[2]; cy[i] = v[3]; cz[i] = v[4];
    sx[i] = v[5]; sy[i] = v[6]; sz[i] = v[7];
  }
  const size_t nz = (size_t)l_z;
  const double* cx_sum = sums.data();
  const double* cy_sum = cx_sum + rxy;
  const double* cxy_sum = cy_sum + rxy;
  const double* cbz_sum = cxy_sum + rxy;
  const double* cz_sum = cbz_sum + nz * rxy;
  // estimators (src/simulation.rs:136-194)
  const double vol = (double)l_x * l_y * l_z, beta = 1.0 / temperature;
  const double msx = mean(sx), msy = mean(sy), msz = mean(sz);
  res->rho_sx = mean(cx) / vol - beta * (mean_square(sx) - msx * msx) / vol;
  res->rho_sy = mean(cy) / vol - beta * (mean_square(sy) - msy * msy) / vol;
  res->rho_sz = mean(cz) / vol - beta * (mean_square(sz) - msz * msz) / vol;
  const std::vector<double>* cs3[3] = {&cx, &cy, &cz};
  const std::vector<double>* ss3[3] = {&sx, &sy, &sz};
  const double ms3[3] = {msx, msy, msz};
  double* std3[3] = {&res->std_rho_sx, &res->std_rho_sy, &res->std_rho_sz};
  std::vector<double> eff(en.size());
  for (int a = 0; a < 3; ++a) {
    for (size_t i = 0; i < en.size(); ++i) {
      const double s = (*ss3[a])[i];
      eff[i] = (*cs3[a])[i] / vol - beta * (s * s - ms3[a] * s) / vol;
    }
    *std3[a] = slsf_blocking_stderr(eff.data(), (int64_t)eff.size());
  }
  res->energy = mean(en);
  res->magnetization_squared = mean(m2);
  res->chi = beta * vol * res->magnetization_squared;
  const double norm = ncorr > 0 ? (double)ncorr : 1.0;
  res->rxy = rxy;
  res->rz = rz;
  for (int i = 0; i < rxy; ++i) {
    if (res->corr_x) res->corr_x[i] = cx_sum[i] / norm;
    if (res->corr_y) res->corr_y[i] = cy_sum[i] / norm;
    if (res->corr_xy) res->corr_xy[i] = cxy_sum[i] / norm;
  }
  if (res->corr_xy_by_z)
    for (size_t i = 0; i < nz * rxy; ++i) res->corr_xy_by_z[i] = cbz_sum[i] / norm;
  if (res->corr_z)
    for (int i = 0; i < rz; ++i) res->corr_z[i] = cz_sum[i] / norm;
  const int count = o->thermal_sweeps + o->measure_sweeps;
  res->acceptance = count == 0 ? 0.0 : acc_sum / (double)count;
  res->num_measurements = nmeas;
  res->num_correlation_measurements = (int32_t)ncorr;
  return SLSF_OK;
}
