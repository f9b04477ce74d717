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
  void* cth = nullptr;
  void* sth = nullptr;
  void* jxy = nullptr;      // R[total_layers]
  void* jz = nullptr;
  uint8_t* bondmask = nullptr;        // [total_sites]
  uint32_t* visited = nullptr;        // [total_sites / 32]
  int32_t* qa = nullptr;              // [total_sites]
  int32_t* qb = nullptr;
  WolffDev* wd = nullptr;             // [nrep]
  long long* csize = nullptr;         // [nrep]
  unsigned long long* acc = nullptr;  // [nrep]
  double* partials = nullptr;         // [nrep * obs_blocks * 9]
  int obs_blocks = 0;
};

template <typename R> void launch_init(const GlobalState& g, int cold, hipStream_t s);
template <typename R> void launch_refresh(const GlobalState& g, hipStream_t s);
// one red/black sweep of every replica, then the sweep counters advance
template <typename R> void launch_metro_sweep(const GlobalState& g, uint8_t* decisions,
                                              hipStream_t s);
// cluster round `round` (replicas with wolff_steps <= round sit it out):
// seed + bond activations + BFS; flip/clear/advance follow separately so that
// a caller can read the membership bitmap in between.
template <typename R> void launch_wolff(const GlobalState& g, int round, hipStream_t s);
template <typename R> void launch_wolff_flip_clear(const GlobalState& g, int round,
                                                   hipStream_t s);
template <typename R> void launch_observables(const GlobalState& g, hipStream_t s);
template <typename R> void launch_corr(const GlobalState& g, int replica, int rxy, int rz,
                                       double* out_xy, double* out_z, hipStream_t s);
void launch_clear_counters(const GlobalState& g, hipStream_t s);

}  // namespace slsf
