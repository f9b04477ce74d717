// driver_session.h -- the plain driver's sweep/measure loop kept on the
// batch's stream: observables and correlation sums accumulate in device
// buffers and are read back once, at the end (host-only declarations).
#pragma once
#include <stdint.h>

#include "../../include/slsf.h"

namespace slsf {

struct DriverSession;

// opens a session for replica 0 of `batch`: room for `nmeas` measurements and
// correlations up to rxy / rz; clears the tallies and opens the sweep
int driver_open(slsf_batch* batch, int64_t nmeas, int rxy, int rz, DriverSession** out);
int driver_sweeps(DriverSession* s, int nsweeps);
// measurement `index` of the observables {energy, m^2, cos_x, cos_y, cos_z,
// sin_x, sin_y, sin_z}
int driver_measure(DriverSession* s, int64_t index);
// adds one correlation measurement to the sums {corr_x[rxy], corr_y[rxy],
// corr_xy[rxy], corr_xy_by_z[lz * rxy], corr_z[rz]}
int driver_correlations(DriverSession* s);
// The measurement phase of run_theta_simulation (src/simulation.rs:88-135):
// `thermal` sweeps, then `measure_sweeps` sweeps, measured after every
// `measure_interval`-th; with corr set, measurement m (from 0) also takes
// correlations when m % correlation_interval == 0.  A small FP64 replica runs
// in one workgroup, a chunk of sweeps per launch; any other replica runs the
// HBM kernels sweep by sweep.  *ncorr receives the correlation measurements.
struct DriverSchedule {
  int64_t thermal = 0, measure_sweeps = 0;
  int32_t measure_interval = 1, correlation_interval = 1;
  bool corr = false;
};
int driver_run(DriverSession* s, const DriverSchedule& plan, int64_t* ncorr);
// waits for the stream and reads back: the acceptance summed over `nsweeps`
// sweeps (accepted / V), `nmeas` x 8 observables and the correlation sums
int driver_close(DriverSession* s, int64_t nsweeps, double* acceptance_sum, double* series,
                 double* sums);
void driver_free(DriverSession* s);
int64_t driver_sums_size(const DriverSession* s);

}  // namespace slsf
