// launch.h -- host-side launchers of the global-memory kernels.  One set per
// arithmetic type: double (FP64 replay, kernels_f64.hip) and float (FP32 plain
// path, kernels_f32.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "slsf_device.h"

namespace slsf {

struct GlobalState {
  RepDev* reps = nullptr;  // device array [nrep]
  int nrep = 0;
  int64_t max_volume = 0;   // largest replica volume
  int64_t total_sites = 0;  // sum of padded volumes
  void* theta = nullptr;    // R[total_sites]
  void* cs = nullptr;       // (cos, sin) pairs of theta: R2[total_sites]
  void* jxy = nullptr;      // R[total_layers]
  void* jz = nullptr;
  int max_lz = 0;           // largest replica extents (grid sizes)
  int64_t max_plane = 0;
  int max_tiles = 0;        // largest wolff_tiles_of<16>
  int max_tiles8 = 0;       // largest wolff_tiles_of<8> (the shallow tiles of small launches)
  int max_tiles4 = 0;       // largest wolff_tiles_of<4> (slab measurement quarters, smallest launches)
  uint8_t* bondmask = nullptr;        // [total_sites] active-bond bits of the base site
  uint32_t* visited = nullptr;        // [total_sites / 32] cluster membership bitmap
  uint32_t* label = nullptr;          // [total_sites] union-find forest of the cluster step
  WolffDev* wd = nullptr;             // [nrep]
  long long* csize = nullptr;         // [nrep * TALLY_SLOTS] cluster sites, summed until cleared
  unsigned long long* acc = nullptr;  // [nrep * TALLY_SLOTS] accepted moves, summed until cleared
  double* partials = nullptr;         // [nrep * obs_blocks * 9]
  int obs_blocks = 0;                 // max over replicas of measure_blocks_of
  // slab sweeps (enqueue_hbm_sweeps): the second angle buffer, R[total_sites],
  // and the measurement partials of the quarter tiles, [nrep * max_tiles4 * 9]
  void* theta2 = nullptr;
  double* slab_partials = nullptr;
};

// grid extents of a GlobalState from its replica descriptors (host side)
void set_extents(GlobalState& g, const RepDev* reps_host, int n);

template <typename R> void launch_init(const GlobalState& g, int cold, hipStream_t s);
template <typename R> void launch_refresh(const GlobalState& g, hipStream_t s);
// A sweep of the HBM path: launch_sweep_end(g, 0, 0) marks the replicas whose
// window is open, then per sweep launch_metro_phases, launch_wolff for rounds
// 0..wmax-1, then launch_sweep_end(g, 1, wmax) (or, in a measured run,
// launch_fold_pending), which advance the counters and mark the next sweep
// (enqueue_hbm_sweeps composes the sequences).
// With measure (a measured run of per-phase sweeps) the colour-0 phase also
// measures the replicas whose last sweep is pending, into g.partials.
template <typename R> void launch_metro_phases(const GlobalState& g, uint8_t* decisions,
                                               hipStream_t s, bool measure = false);
// The bookkeeping at the end of a per-phase sweep: with accum set, the
// colour-0 pass's partials of the replicas that were pending are folded into
// their bins and the replicas measured in this sweep marked pending
// (launch_fold_pending); either way the counters advance as
// launch_sweep_end(g, 1, rounds) does.
struct SweepEnd {
  double* accum = nullptr;
  double* bins = nullptr;
  int32_t* bin_state = nullptr;
  int32_t bin_cap = 0;
  int rounds = 0;
};
// cluster round `round` (replicas with wolff_steps <= round sit it out): seed,
// bond activations, union-find, reflection; the membership bitmap stays valid
// until the next round.  With `end` (the sweep's last round) the reflection
// kernel also does the sweep's bookkeeping, one workgroup per replica.
template <typename R> void launch_wolff(const GlobalState& g, int round, hipStream_t s,
                                        const SweepEnd* end = nullptr);
void launch_sweep_end(const GlobalState& g, int metro, int rounds, hipStream_t s);
// windowed = 1: only replicas measured in this sweep (active, counter at least
// measure_from) compute their partials; 2: only replicas whose last sweep is
// pending; 0: every replica (a direct measurement)
template <typename R> void launch_observables(const GlobalState& g, hipStream_t s,
                                              int windowed = 0);
// rows = ly * lz of that replica (sizes the grid)
template <typename R> void launch_corr(const GlobalState& g, int replica, int64_t rows, int rxy, int rz,
                                       double* out_xy, double* out_z, hipStream_t s);
// k_corr_planes: out[b] (b < rxy lz, shift b / lz + 1, layer b % lz) = that
// layer's x sum, out[rxy lz + b] its y sum, out[2 rxy lz + k] the z sum of
// shift k + 1; the sums of launch_corr's rows in the host fold's order
// (2 rxy lz + rz doubles; ly, lz of that replica)
template <typename R> void launch_corr_planes(const GlobalState& g, int replica, int ly, int lz,
                                              int rxy, int rz, double* out, hipStream_t s);
// zeroes the accepted-move and cluster-site tallies
void launch_clear_counters(const GlobalState& g, hipStream_t s);
// opens (open = 1) or closes the run window of every replica (RepDev
// measure_from / stop_at / binsize) on the device
void launch_window(const GlobalState& g, uint32_t therm, uint32_t total, int32_t binsize,
                   int open, hipStream_t s);
// out[r] = sum of replica r's TALLY_SLOTS slots
void launch_sum_tallies(const GlobalState& g, const unsigned long long* slots,
                        unsigned long long* out, hipStream_t s);
// the end of a measured per-phase sweep: folds the colour-0 pass's partials of
// the replicas that were pending, advances the counters as
// launch_sweep_end(g, 1, rounds), and marks the replicas measured in this
// sweep as pending
// (measure_sites: sites a thread of the measuring colour-0 pass, which sets
// the number of partials)
void launch_fold_pending(const GlobalState& g, double* accum, double* bins, int32_t* bin_state,
                         int32_t bin_cap, int rounds, int measure_sites, hipStream_t s);
// the end of a measured per-phase run: folds the partials of
// launch_observables(g, s, 2) of the pending replicas and clears the flag
void launch_close_observables(const GlobalState& g, double* accum, double* bins,
                              int32_t* bin_state, int32_t bin_cap, hipStream_t s);

// The FP64 replay of one small replica inside one workgroup
// (k_replay_resident): a chunk of `nsweeps` sweeps of the plain driver, with
// its measurements and correlation measurements.  Sweep j of the chunk is
// step step0 + j of the measurement phase (steps <= 0 thermalise); a step >= 1
// with step % measure_interval == 0 is measurement step / measure_interval - 1,
// written to series[8 m ..] as launch_fold_observables writes it, and with
// corr set, measurement m with m % correlation_interval == 0 adds its
// correlations to sums as launch_fold_correlations does (rows_xy, rows_z and
// planes are that fold's scratch).  Tallies and counters advance as on the
// HBM path.
struct ReplayRun {
  int64_t nsweeps = 0, step0 = 0;
  int32_t measure_interval = 1, correlation_interval = 1;
  int32_t rxy = 0, rz = 0, corr = 0, pad = 0;
  double* series = nullptr;
  double* rows_xy = nullptr;
  double* rows_z = nullptr;
  double* planes = nullptr;
  double* sums = nullptr;
};
constexpr int RR_THREADS = 512;
constexpr int RR_MAX_SITES = 4096;  // angles + (cos, sin) pairs: 96 KiB of LDS
constexpr int RR_MAX_ROWS = 256;    // x-rows of the cluster's bitmaps
constexpr int RR_MAX_CORR_SUMS = 1136;  // 3 rxy + Lz rxy + rz with rxy <= Lx/2, Lz Lx <= 2048
inline bool replay_resident_fits(int lx, int ly, int lz) {
  const int64_t v = (int64_t)lx * ly * lz, rows = (int64_t)ly * lz;
  return lx >= 2 && lx <= 32 && lx % 2 == 0 && ly % 2 == 0 && lz % 2 == 0 &&
         v <= RR_MAX_SITES && rows <= RR_MAX_ROWS && (int64_t)lx * lz <= RR_THREADS;
}
inline bool replay_resident_fits_corr(int lz, int rxy, int rz) {
  return 3 * rxy + lz * rxy + rz <= RR_MAX_CORR_SUMS;
}
void launch_replay_resident(const GlobalState& g, int replica, const ReplayRun& run,
                            hipStream_t s);

// folds of one replica's observables (after launch_observables: out[8] =
// {energy, m^2, cos_x, cos_y, cos_z, sin_x, sin_y, sin_z}) and correlation
// rows (after launch_corr: added into sums) in the host folds' order
void launch_fold_observables(const GlobalState& g, int replica, double* out, hipStream_t s);
// planes: (2 rxy + rz) * lz doubles of scratch for the per-plane sums
void launch_fold_correlations(const GlobalState& g, int replica, const double* hxy,
                              const double* hz, int rxy, int rz, double* planes, double* sums,
                              hipStream_t s);

// device bins of a measured run; accum == nullptr runs without measurement
struct BinTarget {
  double* accum = nullptr;
  double* bins = nullptr;
  int32_t* bin_state = nullptr;
  int32_t bin_cap = 0;
};
// Enqueues `nsweeps` sweeps of the HBM path on s in one of two forms that give
// the same chains (bit for bit) and the same bins up to the order of their FP64
// sums:
//  * slab sweeps (slab.inc): per sweep k_slab (measurement of the incoming
//    state, both Metropolis colours, the first cluster round's bonds and tile
//    components), k_wolff_hook and k_slab_flip (reflection, bins, counters),
//    then the extra cluster rounds; the run ends with the closing measurement.
//    Three launches a sweep, which suits launches too small to fill the device
//    (their time is the latency of the dependent launches).  Leaves the (cos,
//    sin) cache stale (launch_refresh).
//  * per-phase sweeps: two Metropolis colours (the first also measures the
//    state the sweep starts from) and three cluster launches per round, the
//    last reflection also folding the measurement and advancing the counters
//    (five launches a sweep), whose small independent blocks fill the device
//    better when the launch is large.
// form 1 and 2 force the slab and the per-phase form; 0 chooses by the launch
// size (hbm_uses_slab).  Both need g.theta2 and g.slab_partials.  Every sweep
// issues the same launches (the replicas' windows decide who takes part), so the
// sequence can be captured into a graph and replayed.  Returns the launch count.
template <typename R> int64_t enqueue_hbm_sweeps(const GlobalState& g, int64_t nsweeps, int wmax,
                                                 const BinTarget& bt, hipStream_t s,
                                                 int form = 0);
// whether enqueue_hbm_sweeps takes the slab form: forced (form 1 / 2) or, for
// form 0, when the launch has at most one 16-deep tile per two compute units
bool hbm_uses_slab(const GlobalState& g, int form);

}  // namespace slsf
