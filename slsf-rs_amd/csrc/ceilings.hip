// ceilings.hip -- microbenchmarks of the gfx950 ceilings the Monte Carlo
// kernels run against: LDS read bandwidth (the LDS form of the resident kernel
// moves 32 B of LDS traffic per site update), FP32 FMA issue rate, and the
// issue rates of the hardware v_sin and v_exp (the transcendental ceiling).
// Each kernel keeps every access in bounds of its own buffers and has a fixed
// trip count; the host times it with HIP events.
#include <hip/hip_runtime.h>

#include "../../include/slsf.h"
#include "common.h"

namespace {

constexpr int CT = 256;               // threads per block
constexpr int LDS_VEC = 4096;         // float4 entries: 64 KiB of LDS per block
constexpr int CHAINS = 8;             // independent chains per thread

__global__ void __launch_bounds__(CT) k_lds_read(float* out, int iters) {
  __shared__ float4 buf[LDS_VEC];
  for (int i = threadIdx.x; i < LDS_VEC; i += CT)
    buf[i] = make_float4((float)i, (float)(i + 1), (float)(i + 2), (float)(i + 3));
  __syncthreads();
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  int idx = threadIdx.x;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < CHAINS; ++k) {
      const float4 v = buf[(idx + k * CT) & (LDS_VEC - 1)];  // consecutive lanes: 16 B apart
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
    idx = (idx + 37) & (LDS_VEC - 1);
  }
  out[blockIdx.x * CT + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

template <int OP>
__global__ void __launch_bounds__(CT) k_issue(float* out, float seed, int iters) {
  float a[CHAINS];
#pragma unroll
  for (int i = 0; i < CHAINS; ++i) a[i] = seed + threadIdx.x * 1e-6f + i * 0.01f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < CHAINS; ++i) {
      if (OP == 0) a[i] = __builtin_fmaf(a[i], 0.999f, 0.001f);
      if (OP == 1) a[i] = __builtin_amdgcn_sinf(a[i]);
      if (OP == 2) a[i] = __builtin_amdgcn_exp2f(a[i]);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < CHAINS; ++i) s += a[i];
  out[blockIdx.x * CT + threadIdx.x] = s;
}

// v_exp_f32 of every float x in {-0} U [+0, +inf]: counts the results below 1.
// The resident kernel's Metropolis test u < exp(-beta dE) (u < 1) also decides
// the dE <= 0 case only if there are none.
__global__ void k_exp2_floor(unsigned long long* below) {
  const uint64_t n = 0x7F800001ull;  // bit patterns +0 .. +inf, then -0
  unsigned long long local = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t bits = i == n ? 0x80000000u : (uint32_t)i;
    if (!(__builtin_amdgcn_exp2f(__uint_as_float(bits)) >= 1.0f)) ++local;
  }
  if (local) atomicAdd(below, local);
}

}  // namespace

using slsf::clear_error;
using slsf::fail;

extern "C" int slsf_measure_ceilings(slsf_ceilings* out) {
  clear_error();
  if (!out) return fail(SLSF_E_INVALID, "null argument");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(SLSF_E_DEVICE, "no usable GPU (hipGetDeviceCount found none)");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess)
    return fail(SLSF_E_DEVICE, "hipGetDeviceProperties failed");
  const int blocks = prop.multiProcessorCount * 16;
  float* buf = nullptr;
  if (hipMalloc(&buf, sizeof(float) * (size_t)blocks * CT) != hipSuccess)
    return fail(SLSF_E_DEVICE, "hipMalloc failed");
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipStream_t s = nullptr;
  (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  auto timed = [&](auto launch) -> double {
    launch();  // warm
    (void)hipEventRecord(e0, s);
    launch();
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return ms * 1e-3;
  };
  const int lds_iters = 4096, issue_iters = 4096;
  const double t_lds = timed([&] { k_lds_read<<<blocks, CT, 0, s>>>(buf, lds_iters); });
  const double t_fma = timed([&] { k_issue<0><<<blocks, CT, 0, s>>>(buf, 0.5f, issue_iters); });
  const double t_sin = timed([&] { k_issue<1><<<blocks, CT, 0, s>>>(buf, 0.5f, issue_iters); });
  const double t_exp = timed([&] { k_issue<2><<<blocks, CT, 0, s>>>(buf, 0.5f, issue_iters); });
  const hipError_t err = hipGetLastError();
  (void)hipStreamDestroy(s);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  if (err != hipSuccess) return fail(SLSF_E_DEVICE, hipGetErrorString(err));
  const double lanes = (double)blocks * CT;
  out->lds_bytes_per_s = lanes * lds_iters * CHAINS * 16.0 / t_lds;
  out->fma_wave_instr_per_s = lanes * issue_iters * CHAINS / 64.0 / t_fma;
  out->sin_wave_instr_per_s = lanes * issue_iters * CHAINS / 64.0 / t_sin;
  out->exp_wave_instr_per_s = lanes * issue_iters * CHAINS / 64.0 / t_exp;
  out->compute_units = prop.multiProcessorCount;
  out->reserved = 0;
  return SLSF_OK;
}

extern "C" int slsf_check_exp2_floor(uint64_t* below) {
  clear_error();
  if (!below) return fail(SLSF_E_INVALID, "null argument");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(SLSF_E_DEVICE, "no usable GPU (hipGetDeviceCount found none)");
  unsigned long long* d = nullptr;
  if (hipMalloc(&d, sizeof *d) != hipSuccess) return fail(SLSF_E_DEVICE, "hipMalloc failed");
  unsigned long long h = 0;
  hipError_t err = hipMemset(d, 0, sizeof *d);
  if (err == hipSuccess) {
    k_exp2_floor<<<4096, 256>>>(d);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (err != hipSuccess) return fail(SLSF_E_DEVICE, hipGetErrorString(err));
  *below = h;
  return SLSF_OK;
}
