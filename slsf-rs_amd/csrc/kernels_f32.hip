// FP32 plain-path kernels (state in HBM, one thread per site), plus the
// type-independent launchers of the HBM path.
#include <algorithm>

#include "launch.h"
#include "global_kernels.inc"
#include "slab.inc"
#define SLSF_REAL float
#include "kernels_body.inc"
#undef SLSF_REAL

namespace slsf {

// sweep bookkeeping of the HBM path: a replica that took part adds `metro` to
// its Metropolis counter and its own number of cluster rounds (at most
// `rounds`) to its cluster counter; then every replica is marked for the next
// sweep iff its counter is below its window end.  (metro, rounds) = (0, 0)
// only sets the marks, which is how a run starts.
__global__ void k_sweep_end(RepDev* reps, int nrep, int metro, int rounds) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  RepDev& d = reps[r];
  if (d.active) {
    d.metro_ctr += (uint32_t)metro;
    d.wolff_ctr += (uint32_t)(rounds < d.wolff_steps ? rounds : d.wolff_steps);
  }
  d.active = d.metro_ctr < d.stop_at ? 1u : 0u;
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

void set_extents(GlobalState& g, const RepDev* reps, int n) {
  g.max_volume = 0;
  g.max_plane = 0;
  g.max_lz = 0;
  g.obs_blocks = 0;
  g.max_tiles = 0;
  g.max_tiles8 = 0;
  g.max_tiles4 = 0;
  for (int r = 0; r < n; ++r) {
    g.max_volume = std::max<int64_t>(g.max_volume, reps[r].volume);
    g.max_plane = std::max<int64_t>(g.max_plane, (int64_t)reps[r].lx * reps[r].ly);
    g.max_lz = std::max(g.max_lz, reps[r].lz);
    g.obs_blocks = std::max(g.obs_blocks, measure_blocks_of(reps[r].lx, reps[r].ly, reps[r].lz));
    g.max_tiles = std::max(g.max_tiles, wolff_tiles_of<16>(reps[r].lx, reps[r].ly, reps[r].lz));
    g.max_tiles8 = std::max(g.max_tiles8, wolff_tiles_of<8>(reps[r].lx, reps[r].ly, reps[r].lz));
    g.max_tiles4 = std::max(g.max_tiles4, wolff_tiles_of<4>(reps[r].lx, reps[r].ly, reps[r].lz));
  }
}

// run window of every replica of a production run: measured from its counter
// + thermalization, no sweep once the counter reaches + total; open = 0 closes
// it again (nothing measured, nothing stopped).  Set on the device, so a run
// needs no round trip of the descriptors.
__global__ void k_window(RepDev* reps, int nrep, uint32_t therm, uint32_t total, int32_t binsize,
                         int open) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  RepDev& d = reps[r];
  if (open) {
    d.measure_from = d.metro_ctr + therm;
    d.stop_at = d.metro_ctr + total;
    d.binsize = binsize;
  } else {
    d.measure_from = UINT32_MAX;
    d.stop_at = UINT32_MAX;
  }
}

// per-replica sums of the TALLY_SLOTS slots of a tally
__global__ void k_sum_tallies(const unsigned long long* __restrict__ slots, int nrep,
                              unsigned long long* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  unsigned long long s = 0;
  for (int k = 0; k < TALLY_SLOTS; ++k) s += slots[(int64_t)r * TALLY_SLOTS + k];
  out[r] = s;
}

void launch_window(const GlobalState& g, uint32_t therm, uint32_t total, int32_t binsize,
                   int open, hipStream_t s) {
  k_window<<<(g.nrep + 63) / 64, 64, 0, s>>>(g.reps, g.nrep, therm, total, binsize, open);
}

void launch_sum_tallies(const GlobalState& g, const unsigned long long* slots,
                        unsigned long long* out, hipStream_t s) {
  k_sum_tallies<<<(g.nrep + 63) / 64, 64, 0, s>>>(slots, g.nrep, out);
}

// The slab form wins while a launch has at most one 16-deep tile per two
// compute units (measured per sweep, FP32, slab against per-phase: one L=64
// replica 34.7 against 39.7 us, four L=48 41.5 against 42.6, two L=64 42.6
// against 43.2); beyond that the per-phase kernels' small independent blocks
// fill the device better (eight L=48 55.7 against 50.7, one L=96 59.7 against
// 54.2, four L=64 59.0 against 53.2, one L=128 93.2 against 83.3).
bool hbm_uses_slab(const GlobalState& g, int form) {
  if (form == 1) return true;
  if (form == 2) return false;
  return (int64_t)g.max_tiles * g.nrep <= WOLFF_SHALLOW_BELOW / 2;
}

void launch_clear_counters(const GlobalState& g, hipStream_t s) {
  const int64_t n = (int64_t)g.nrep * TALLY_SLOTS;
  k_fill_u64<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g.acc, n);
  k_fill_u64<<<(unsigned)((n + 255) / 256), 256, 0, s>>>((unsigned long long*)g.csize, n);
}

void launch_sweep_end(const GlobalState& g, int metro, int rounds, hipStream_t s) {
  k_sweep_end<<<(g.nrep + 63) / 64, 64, 0, s>>>(g.reps, g.nrep, metro, rounds);
}

// The end of a measured per-phase sweep without cluster rounds (one workgroup
// per replica; with rounds, the last round's reflection kernel does it): the
// measurement the colour-0 pass took of a pending replica is folded (its
// metro_blocks_of partials, in block order), then the counters advance as in
// k_sweep_end, and a replica measured in this sweep is marked pending for the
// next colour-0 pass (src/model.rs:379-442, as the slab form's bookkeeping).
__global__ void __launch_bounds__(FOLD_THREADS)
k_fold_pending(RepDev* reps, const double* __restrict__ partials, int stride, double* accum,
               double* bins, int32_t* bin_state, int32_t bin_cap, int rounds, int ms) {
  const int r = blockIdx.x;
  RepDev& d = reps[r];
  sweep_bookkeeping(d, r, partials + (int64_t)r * stride * 9,
                    metro_blocks_of(d.lx, d.ly, d.lz, ms), accum, bins, bin_state, bin_cap,
                    rounds);
}

// The closing measurement of a measured per-phase run: the replicas still
// pending fold the partials k_observables took of them.
__global__ void __launch_bounds__(FOLD_THREADS)
k_close_observables(RepDev* reps, const double* __restrict__ partials, int stride, double* accum,
                    double* bins, int32_t* bin_state, int32_t bin_cap) {
  const int r = blockIdx.x;
  RepDev& d = reps[r];
  const bool pend = d.pending;
  __syncthreads();  // every thread has read the flag before thread 0 clears it
  if (!pend) return;
  if (accum)
    fold_measurement(d, r, partials + (int64_t)r * stride * 9, obs_blocks_of(d.lx, d.ly, d.lz),
                     accum, bins, bin_state, bin_cap);
  if (threadIdx.x == 0) d.pending = 0;
}

void launch_fold_pending(const GlobalState& g, double* accum, double* bins, int32_t* bin_state,
                         int32_t bin_cap, int rounds, int measure_sites, hipStream_t s) {
  k_fold_pending<<<g.nrep, FOLD_THREADS, 0, s>>>(g.reps, g.partials, g.obs_blocks, accum, bins,
                                                 bin_state, bin_cap, rounds, measure_sites);
}

void launch_close_observables(const GlobalState& g, double* accum, double* bins,
                              int32_t* bin_state, int32_t bin_cap, hipStream_t s) {
  k_close_observables<<<g.nrep, FOLD_THREADS, 0, s>>>(g.reps, g.partials, g.obs_blocks, accum,
                                                      bins, bin_state, bin_cap);
}

// ---------------------------------------------------------------------------
// device folds of the plain driver: the FP64 sums of the host folds of
// slsf_measure_observables / slsf_measure_correlations in the same order
// (the z-correlation sums plane by plane), so a run that never reads back
// until its end gets the same numbers
// ---------------------------------------------------------------------------
__global__ void k_fold_observables(const RepDev* __restrict__ rep, const double* __restrict__ part,
                                   double* __restrict__ out) {
  const RepDev d = *rep;
  const int nb = obs_blocks_of(d.lx, d.ly, d.lz);
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 9; ++k) v[k] += part[(int64_t)b * 9 + k];
  const double vol = (double)d.volume;
  out[0] = v[0] / vol;
  out[1] = (v[1] * v[1] + v[2] * v[2]) / (vol * vol);
  for (int k = 3; k < 9; ++k) out[k - 1] = v[k];
}

// sums[] = {corr_x[rxy], corr_y[rxy], corr_xy[rxy], corr_xy_by_z[lz * rxy], corr_z[rz]}
// += this measurement's values, in two steps of one block: per (shift, plane)
// the sums over the plane's rows (the x/y sums exactly as the host fold of
// slsf_measure_correlations nests them, src/observables.rs:189-252), then per
// shift the sums over the planes in order
__global__ void __launch_bounds__(256)
k_fold_correlations(const RepDev* __restrict__ rep, const double* __restrict__ hxy,
                    const double* __restrict__ hz, int rxy, int rz, double* __restrict__ planes,
                    double* __restrict__ sums) {
  const RepDev d = *rep;
  const int lz = d.lz, ly = d.ly;
  const int64_t rows = (int64_t)ly * lz;
  const double vol = (double)d.volume, area = (double)d.lx * ly;
  double* px = planes;                     // [rxy][lz]
  double* py = px + (int64_t)rxy * lz;     // [rxy][lz]
  double* pz = py + (int64_t)rxy * lz;     // [rz][lz]
  const int tasks = (rxy + rz) * lz;
  for (int t = threadIdx.x; t < tasks; t += blockDim.x) {
    const int rr = t / lz, z = t - rr * lz;
    if (rr < rxy) {
      double lsx = 0, lsy = 0;
      for (int y = 0; y < ly; ++y) {
        const double* o = hxy + ((int64_t)rr * rows + (int64_t)z * ly + y) * 3;
        lsx += o[0];
        lsx += o[1];
        lsy += o[2];
      }
      px[(int64_t)rr * lz + z] = lsx;
      py[(int64_t)rr * lz + z] = lsy;
    } else {
      const int rzi = rr - rxy;
      double s = 0;
      for (int y = 0; y < ly; ++y) s += hz[(int64_t)rzi * rows + (int64_t)z * ly + y];
      pz[(int64_t)rzi * lz + z] = s;
    }
  }
  __syncthreads();
  double* cx = sums;
  double* cy = cx + rxy;
  double* cxy = cy + rxy;
  double* cbz = cxy + rxy;
  double* cz = cbz + (int64_t)lz * rxy;
  for (int t = threadIdx.x; t < rxy + rz; t += blockDim.x) {
    if (t < rxy) {
      double sx = 0, sy = 0;
      for (int z = 0; z < lz; ++z) {
        const double lsx = px[(int64_t)t * lz + z], lsy = py[(int64_t)t * lz + z];
        sx += lsx;
        sy += lsy;
        cbz[(int64_t)z * rxy + t] += (lsx + lsy) / (2.0 * area);
      }
      cx[t] += sx / vol;
      cy[t] += sy / vol;
      cxy[t] += (sx / vol + sy / vol) / 2.0;
    } else {
      const int rzi = t - rxy;
      double sz = 0;
      for (int z = 0; z < lz; ++z) sz += pz[(int64_t)rzi * lz + z];
      cz[rzi] += sz / vol;
    }
  }
}

void launch_fold_observables(const GlobalState& g, int replica, double* out, hipStream_t s) {
  k_fold_observables<<<1, 1, 0, s>>>(g.reps + replica,
                                     g.partials + (int64_t)replica * g.obs_blocks * 9, out);
}

void launch_fold_correlations(const GlobalState& g, int replica, const double* hxy,
                              const double* hz, int rxy, int rz, double* planes, double* sums,
                              hipStream_t s) {
  if (rxy + rz > 0)
    k_fold_correlations<<<1, 256, 0, s>>>(g.reps + replica, hxy, hz, rxy, rz, planes, sums);
}

}  // namespace slsf
