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
// waits for the stream and reads back: the acceptance summed over `nsweeps`
// sweeps (accepted / V), `nmeas` x 8 observables and the correlation sums
int driver_close(DriverSession* s, int64_t nsweeps, double* acceptance_sum, double* series,
                 double* sums);
void driver_free(DriverSession* s);
int64_t driver_sums_size(const DriverSession* s);

}  // namespace slsf
