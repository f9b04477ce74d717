// FP32 plain-path kernels (state in HBM, one thread per site).
#include "launch.h"
#include "global_kernels.inc"
#define SLSF_REAL float
#include "kernels_body.inc"
#undef SLSF_REAL

namespace slsf {
void launch_clear_counters(const GlobalState& g, hipStream_t s) {
  k_fill_u64<<<(g.nrep + 255) / 256, 256, 0, s>>>(g.acc, g.nrep);
}
}  // namespace slsf

// ---------------------------------------------------------------------------
// per-measurement accumulation of the observable series into FP64 bins, on the
// device: one thread per replica folds the block partials of k_observables in
// block order, forms the 13 series values and appends a bin mean every
// `binsize` measurements.
// ---------------------------------------------------------------------------
namespace slsf {

__global__ void k_accumulate(const RepDev* __restrict__ reps, int nrep,
                             const double* __restrict__ partials, int obs_blocks, double* accum,
                             double* bins, int32_t* bin_state, int32_t bin_cap) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  const RepDev d = reps[r];
  // measured iff this sweep took part and its index metro_ctr - 1 >= measure_from
  if (!d.active || d.metro_ctr == 0u || d.metro_ctr - 1u < d.measure_from) return;
  const int32_t binsize = d.binsize;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < obs_blocks; ++b)
    for (int k = 0; k < 9; ++k) v[k] += partials[((int64_t)r * obs_blocks + b) * 9 + k];
  const double vol = (double)d.volume;
  const double e = v[0] / vol;
  const double m2 = (v[1] * v[1] + v[2] * v[2]) / (vol * vol);
  const double vals[13] = {e,    e * e, m2,   m2 * m2,     v[3],        v[4],       v[5],
                           v[6], v[7],  v[8], v[6] * v[6], v[7] * v[7], v[8] * v[8]};
  double* acc = accum + (int64_t)r * 13;
  for (int k = 0; k < 13; ++k) acc[k] += vals[k];
  int32_t* st = bin_state + 2 * r;
  st[0] += 1;
  if (st[0] == binsize) {
    if (st[1] < bin_cap)
      for (int k = 0; k < 13; ++k)
        bins[((int64_t)r * 13 + k) * bin_cap + st[1]] = acc[k] / (double)binsize;
    for (int k = 0; k < 13; ++k) acc[k] = 0.0;
    st[0] = 0;
    st[1] += 1;
  }
}

void launch_accumulate(const GlobalState& g, double* accum, double* bins, int32_t* bin_state,
                       int32_t bin_cap, hipStream_t s) {
  k_accumulate<<<(g.nrep + 63) / 64, 64, 0, s>>>(g.reps, g.nrep, g.partials, g.obs_blocks,
                                                  accum, bins, bin_state, bin_cap);
}

}  // namespace slsf
