// batch_state.h -- the device state a replica batch owns (the opaque
// slsf_batch of the C ABI), shared by the translation units that implement its
// entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/slsf.h"
#include "common.h"
#include "launch.h"
#include "resident.h"

namespace slsf {

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(SLSF_E_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + \
                                     " at " #expr);                                   \
  } while (0)

struct HostReplica {
  int lx, ly, lz, wolff_steps;
  double temperature, width;
  std::vector<double> j_xy, j_z;
  uint64_t seed;
  int64_t volume, padded, site_off;
  int jz_off;
};

}  // namespace slsf

using namespace slsf;

struct slsf_batch {
  int precision = SLSF_FP64_REPLAY;
  int device = 0;                 // the HIP device the batch's memory lives on
  hipStream_t stream = nullptr;
  std::vector<HostReplica> reps;
  std::vector<RepDev> reps_host;  // mirror of the device descriptors
  GlobalState g;
  int64_t total_layers = 0;
  int64_t device_bytes = 0;
  uint8_t* d_decisions = nullptr;
  unsigned long long* d_tally_sum = nullptr;  // [nrep] per-replica sums of a tally
  double* d_accum = nullptr;     // [nrep][NUM_SERIES]
  double* d_bins = nullptr;      // [nrep][NUM_SERIES][bin_cap]
  int32_t* d_bin_state = nullptr;  // [nrep][2]: measurements in open bin, completed bins
  int32_t bin_cap = 0;
  int32_t binsize = 0;
  std::vector<double> acc_sum;   // per replica acceptance sums
  std::vector<int64_t> acc_count;
  std::vector<double> cluster_sum;
  std::vector<int64_t> cluster_count;
  std::vector<void*> allocations;
  ResidentPlan resident;         // FP32 register/LDS-resident groups
  bool initialised = false;
  bool cs_stale = false;         // FP32 sweeps left the (cos, sin) cache behind theta

  ~slsf_batch() {
    resident_release(resident);
    for (void* p : allocations) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
  // frees one allocation made by alloc()
  void release(void* p) {
    if (!p) return;
    auto it = std::find(allocations.begin(), allocations.end(), p);
    if (it == allocations.end()) return;
    allocations.erase(it);
    (void)hipFree(p);
  }
  template <typename T>
  int alloc(T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    HIP_TRY(hipMalloc(&q, bytes));
    allocations.push_back(q);
    device_bytes += (int64_t)bytes;
    *p = (T*)q;
    return SLSF_OK;
  }
};

inline bool is_fp64_batch(const slsf_batch* b) { return b->precision == SLSF_FP64_REPLAY; }

// Brings the (cos, sin) cache up to date after slab sweeps (recomputed from the
// angles with the kernels' own sincos, which gives the cache's values).
int ensure_trig_cache(slsf_batch* b);
