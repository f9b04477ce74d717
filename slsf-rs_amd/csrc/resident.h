// resident.h -- the FP32 batched kernel in which each replica lives in the
// registers and LDS of one workgroup for a whole launch of many sweeps.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "launch.h"

namespace slsf {

struct ResidentGroup {
  int L = 0;                    // cubic edge of every replica in the group
  int cpt = 1;                  // z-columns per thread (waves per replica = L*RW/64/cpt)
  int sweeps_per_launch = 0;    // 0: default
  std::vector<int32_t> replicas;  // batch indices
  int32_t* d_index = nullptr;   // device copy of `replicas`
};

struct ResidentPlan {
  bool covers_all = false;      // every replica is cubic with a resident L (groups
                                // hold the ones that are; the rest take the HBM path)
  std::vector<ResidentGroup> groups;
  std::vector<hipStream_t> streams;  // one per group after the first
  std::vector<hipEvent_t> events;    // join events of those streams
  hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
};

struct TuneRecord {
  int L, cpt, spl, waves;
  double rate;  // site updates per second of the trial
};

struct ResidentRunArgs {
  int64_t thermalization = 0, sweeps = 0;
  int32_t binsize = 1, sweeps_per_launch = 0, use_graph = 0;
  int32_t replay_stream = 0;  // 1: per-site keyed uniforms of the replay path
  double* accum = nullptr;
  double* bins = nullptr;
  int32_t* bin_state = nullptr;
  int32_t bin_cap = 0;
  unsigned long long* acc_out = nullptr;  // accepted proposals, [replica * TALLY_SLOTS] (summed)
  long long* csize_out = nullptr;         // cluster sites, [replica * TALLY_SLOTS] (summed)
};

struct ResidentRunResult {
  int64_t launches = 0;
  double kernel_ms = 0.0;
};

// L values the resident kernel is built for.
bool resident_supports(int L);
int resident_plan(const std::vector<RepDev>& reps, ResidentPlan& plan);
int resident_run(ResidentPlan& plan, const GlobalState& g, const ResidentRunArgs& a,
                 hipStream_t s, ResidentRunResult* out);
void resident_release(ResidentPlan& plan);
// Enqueues (without waiting) the launches of one group for ra.thermalization +
// ra.sweeps sweeps; returns the number of launches.
int64_t resident_enqueue(const ResidentGroup& grp, const GlobalState& g,
                         const ResidentRunArgs& ra, hipStream_t s);
int resident_autotune(ResidentPlan& plan, const GlobalState& g, hipStream_t s,
                      int trial_sweeps, std::vector<TuneRecord>& records);
std::vector<int> resident_cpt_candidates(int L);
int resident_waves_per_replica(int L, int cpt);

}  // namespace slsf
