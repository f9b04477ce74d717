// noalloc.hip -- the correlation measurement without device allocation.
//
// slsf_measure_correlations (batch.cpp) allocates its row-sum scratch with
// hipMallocAsync on every call.  slsf_measure_correlations_noalloc computes the
// same values with that scratch, the fold's outputs included, in one
// file-scope device array, `scratch`, of NOALLOC_SCRATCH_DOUBLES doubles
// (16 MiB) that every device the library runs on holds until the process ends.
// It is for callers that cannot allocate device memory.
//
// Host code never names the array in device work: at a device's first call its
// address is taken once with hipGetSymbolAddress (that device current) and kept
// per device, and the kernels receive that pointer.  Every call on a device
// shares the array, which keeps the last call's contents; nothing here needs
// it cleared, because each kernel writes every element it later reads.  Calls
// on one device are serialised by a lock that is held until the call's device
// work has finished, so calls from several host threads never overlap on the
// array (overlapping users of it would race).
#include <hip/hip_runtime.h>

#include <mutex>

#include "batch_state.h"

namespace slsf {
namespace {

constexpr int64_t NOALLOC_SCRATCH_DOUBLES = (int64_t)1 << 21;  // 16 MiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_device_lock[NOALLOC_MAX_DEVICES];
std::mutex g_ptr_lock;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// the address of `scratch` on device dev, which must be current
int scratch_of(int dev, double** out) {
  std::lock_guard<std::mutex> lock(g_ptr_lock);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SLSF_E_DEVICE, std::string("hipGetSymbolAddress of the correlation scratch "
                                             "failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return SLSF_OK;
}

// The host fold of slsf_measure_correlations (src/observables.rs:189-252) on
// the device, term for term in the same order, so both entry points give the
// same bits.  One thread per in-plane shift, then one per z shift.
// out = {corr_x[rxy], corr_y[rxy], corr_xy[rxy], corr_xy_by_z[lz][rxy], corr_z[rz]}
__global__ void k_fold_corr_host_order(const double* __restrict__ hxy,
                                       const double* __restrict__ hz, int rxy, int rz, int ly,
                                       int lz, double vol, double area, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t rows = (int64_t)ly * lz;
  double* cx = out;
  double* cy = cx + rxy;
  double* cxy = cy + rxy;
  double* cbz = cxy + rxy;
  double* cz = cbz + (int64_t)lz * rxy;
  if (t < rxy) {
    double sx = 0, sy = 0;
    for (int z = 0; z < lz; ++z) {
      double lsx = 0, lsy = 0;
      for (int y = 0; y < ly; ++y) {
        const double* o = hxy + ((int64_t)t * rows + (int64_t)z * ly + y) * 3;
        lsx += o[0];
        lsx += o[1];
        lsy += o[2];
      }
      sx += lsx;
      sy += lsy;
      cbz[(int64_t)z * rxy + t] = (lsx + lsy) / (2.0 * area);
    }
    cx[t] = sx / vol;
    cy[t] = sy / vol;
    cxy[t] = (sx / vol + sy / vol) / 2.0;
  } else if (t < rxy + rz) {
    const int r = t - rxy;
    double sz = 0;
    for (int z = 0; z < lz; ++z)
      for (int y = 0; y < ly; ++y) sz += hz[(int64_t)r * rows + (int64_t)z * ly + y];
    cz[r] = sz / vol;
  }
}

// restores the caller's current device when the call returns
struct DeviceGuard {
  int prev = -1;
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

}  // namespace
}  // namespace slsf

using namespace slsf;

extern "C" int slsf_measure_correlations_noalloc(slsf_batch* b, int32_t replica, int32_t rmax,
                                                 int32_t rmax_xy, int32_t rmax_z,
                                                 double* corr_x, double* corr_y,
                                                 double* corr_xy, double* corr_xy_by_z,
                                                 double* corr_z) {
  int32_t rxy = 0, rz = 0;
  if (int rc = slsf_correlation_rmax(b, replica, rmax, rmax_xy, rmax_z, &rxy, &rz)) return rc;
  if (int rc = ensure_trig_cache(b)) return rc;
  const auto& r = b->reps[replica];
  const int64_t rows = (int64_t)r.ly * r.lz;
  const int64_t n_xy = (int64_t)rxy * rows * 3, n_z = (int64_t)rz * rows;
  const int64_t n_out = 3 * (int64_t)rxy + (int64_t)r.lz * rxy + rz;
  const int64_t need = n_xy + n_z + n_out;
  if (need > NOALLOC_SCRATCH_DOUBLES)
    return fail(SLSF_E_INVALID,
                "correlation scratch of " + std::to_string(need) + " doubles exceeds the "
                "no-allocation limit of " + std::to_string(NOALLOC_SCRATCH_DOUBLES) +
                " doubles (16 MiB); use slsf_measure_correlations or smaller radii");
  if (b->device < 0 || b->device >= NOALLOC_MAX_DEVICES)
    return fail(SLSF_E_DEVICE, "device index beyond the no-allocation table");
  DeviceGuard guard;
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != b->device) {
    HIP_TRY(hipSetDevice(b->device));
    guard.prev = cur;
  }
  std::lock_guard<std::mutex> lock(g_device_lock[b->device]);
  double* ptr = nullptr;
  if (int rc = scratch_of(b->device, &ptr)) return rc;
  double* d_xy = ptr;
  double* d_z = ptr + n_xy;
  double* d_out = d_z + n_z;
  if (rxy + rz > 0) {
    if (is_fp64_batch(b)) launch_corr<double>(b->g, replica, rows, rxy, rz, d_xy, d_z, b->stream);
    else launch_corr<float>(b->g, replica, rows, rxy, rz, d_xy, d_z, b->stream);
    k_fold_corr_host_order<<<(rxy + rz + 63) / 64, 64, 0, b->stream>>>(
        d_xy, d_z, rxy, rz, r.ly, r.lz, (double)r.volume, (double)r.lx * r.ly, d_out);
    HIP_TRY(hipGetLastError());
  }
  // the results straight into the caller's arrays
  const double* src[5] = {d_out, d_out + rxy, d_out + 2 * rxy, d_out + 3 * rxy,
                          d_out + 3 * rxy + (int64_t)r.lz * rxy};
  double* dst[5] = {corr_x, corr_y, corr_xy, corr_xy_by_z, corr_z};
  const int64_t len[5] = {rxy, rxy, rxy, (int64_t)r.lz * rxy, rz};
  for (int k = 0; k < 5; ++k)
    if (dst[k] && len[k] > 0)
      HIP_TRY(hipMemcpyAsync(dst[k], src[k], sizeof(double) * (size_t)len[k],
                             hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return SLSF_OK;
}
