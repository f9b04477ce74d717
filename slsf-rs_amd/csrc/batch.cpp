// batch.cpp -- C-ABI entry points that own device state: replica batches, the
// reference-mirroring update/measure calls, and production runs with on-device
// binning.  Compiled by hipcc as host code (no kernels here).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/slsf.h"
#include "common.h"
#include "driver_session.h"
#include "batch_state.h"
#include "launch.h"
#include "resident.h"

namespace slsf {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
void clear_error() { g_error.clear(); }
int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}
bool temperature_ok(double t) { return t > 0.0 && std::isfinite(t); }
bool width_ok(double w) { return w >= 0.0 && std::isfinite(w); }

}  // namespace slsf

using namespace slsf;

namespace {

int validate_replica(const slsf_replica_desc& d) {
  if (d.l_x <= 0) return fail(SLSF_E_INVALID, "L_x must be positive");
  if (d.l_y <= 0) return fail(SLSF_E_INVALID, "L_y must be positive");
  if (d.l_z <= 0) return fail(SLSF_E_INVALID, "L_z must be positive");
  if (!temperature_ok(d.temperature))
    return fail(SLSF_E_INVALID, "temperature T must be positive and finite");
  if (!width_ok(d.proposal_width))
    return fail(SLSF_E_INVALID, "proposal_width must be nonnegative and finite");
  if (d.wolff_steps < 0) return fail(SLSF_E_INVALID, "wolff_steps must be nonnegative");
  if (!d.j_xy || !d.j_z) return fail(SLSF_E_INVALID, "layer coupling arrays are required");
  for (int z = 0; z < d.l_z; ++z)
    if (!(d.j_xy[z] >= 0.0) || !(d.j_z[z] >= 0.0))
      return fail(SLSF_E_INVALID, "theta simulation requires nonnegative layer couplings");
  if (d.l_x % 2 || d.l_y % 2 || d.l_z % 2)
    return fail(SLSF_E_INVALID, "red-black Metropolis sweep requires even lattice dimensions");
  const int64_t v = (int64_t)d.l_x * d.l_y * d.l_z;
  if (v >= (int64_t)1 << 31) return fail(SLSF_E_INVALID, "lattice volume must be below 2^31");
  return SLSF_OK;
}

int device_ready() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(SLSF_E_DEVICE, "no usable GPU (hipGetDeviceCount found none)");
  return SLSF_OK;
}

template <typename R>
int upload_couplings(slsf_batch* b) {
  std::vector<R> jxy(b->total_layers), jz(b->total_layers);
  for (const auto& r : b->reps)
    for (int z = 0; z < r.lz; ++z) {
      jxy[r.jz_off + z] = (R)r.j_xy[z];
      jz[r.jz_off + z] = (R)r.j_z[z];
    }
  HIP_TRY(hipMemcpyAsync(b->g.jxy, jxy.data(), sizeof(R) * jxy.size(), hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipMemcpyAsync(b->g.jz, jz.data(), sizeof(R) * jz.size(), hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return SLSF_OK;
}

int sync_reps_from_device(slsf_batch* b) {
  HIP_TRY(hipMemcpyAsync(b->reps_host.data(), b->g.reps, sizeof(RepDev) * b->reps_host.size(),
                         hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return SLSF_OK;
}

int push_reps_to_device(slsf_batch* b) {
  HIP_TRY(hipMemcpyAsync(b->g.reps, b->reps_host.data(), sizeof(RepDev) * b->reps_host.size(),
                         hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return SLSF_OK;
}

bool is_fp64(const slsf_batch* b) { return b->precision == SLSF_FP64_REPLAY; }

// the second angle buffer and tile partials of the slab sweeps, on first use
int ensure_slab_buffers(slsf_batch* b) {
  if (b->g.theta2) return SLSF_OK;
  const size_t rs = is_fp64(b) ? sizeof(double) : sizeof(float);
  if (int rc = b->alloc((char**)&b->g.theta2, rs * (size_t)b->g.total_sites)) return rc;
  return b->alloc(&b->g.slab_partials, (size_t)b->g.nrep * b->g.max_tiles4 * 9);
}

int metro_phases(slsf_batch* b, uint8_t* d_decisions) {
  if (is_fp64(b)) launch_metro_phases<double>(b->g, d_decisions, b->stream);
  else launch_metro_phases<float>(b->g, d_decisions, b->stream);
  HIP_TRY(hipGetLastError());
  return SLSF_OK;
}

// the membership bitmap of this round is copied out when members_host is set
int wolff_round(slsf_batch* b, int round, uint8_t* members_host) {
  if (is_fp64(b)) launch_wolff<double>(b->g, round, b->stream);
  else launch_wolff<float>(b->g, round, b->stream);
  HIP_TRY(hipGetLastError());
  if (members_host) {
    std::vector<uint32_t> words(b->g.total_sites / 32);
    HIP_TRY(hipMemcpyAsync(words.data(), b->g.visited, sizeof(uint32_t) * words.size(),
                           hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    int64_t out = 0;
    for (const auto& r : b->reps)
      for (int64_t i = 0; i < r.volume; ++i, ++out) {
        const int64_t gi = r.site_off + i;
        members_host[out] = (uint8_t)((words[gi >> 5] >> (gi & 31)) & 1u);
      }
  }
  return SLSF_OK;
}

// one full sweep of the HBM path (Metropolis + the replica's cluster rounds)
int hbm_sweep(slsf_batch* b, int wmax) {
  if (int rc = metro_phases(b, nullptr)) return rc;
  for (int k = 0; k < wmax; ++k)
    if (int rc = wolff_round(b, k, nullptr)) return rc;
  launch_sweep_end(b->g, 1, wmax, b->stream);
  HIP_TRY(hipGetLastError());
  return SLSF_OK;
}

int max_wolff_steps(const slsf_batch* b) {
  int m = 0;
  for (const auto& r : b->reps) m = std::max(m, r.wolff_steps);
  return m;
}

// per-replica sums of the TALLY_SLOTS slots of a device tally (summed on the
// device; one small copy back)
int read_tallies(slsf_batch* b, const unsigned long long* slots,
                 std::vector<unsigned long long>& out) {
  out.assign(b->reps.size(), 0ull);
  launch_sum_tallies(b->g, slots, b->d_tally_sum, b->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out.data(), b->d_tally_sum, sizeof(unsigned long long) * out.size(),
                         hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return SLSF_OK;
}

// Reads the device tallies summed since launch_clear_counters over `nsweeps`
// sweeps: the sum of per-sweep acceptances is accepted / V (every sweep of a
// replica visits V sites), added to `acceptance` when given.
int gather_acceptance(slsf_batch* b, int64_t nsweeps, double* acceptance) {
  std::vector<unsigned long long> acc;
  if (int rc = read_tallies(b, b->g.acc, acc)) return rc;
  for (size_t r = 0; r < acc.size(); ++r) {
    const double a = (double)acc[r] / (double)b->reps[r].volume;
    b->acc_sum[r] += a;
    b->acc_count[r] += nsweeps;
    if (acceptance) acceptance[r] += a;
  }
  return SLSF_OK;
}

// cluster sites summed since launch_clear_counters; `clusters[r]` is how many
// clusters replica r grew meanwhile
int gather_clusters(slsf_batch* b, const std::vector<int64_t>& clusters, int64_t* cluster_size) {
  std::vector<unsigned long long> cs;
  if (int rc = read_tallies(b, (unsigned long long*)b->g.csize, cs)) return rc;
  for (size_t r = 0; r < cs.size(); ++r) {
    b->cluster_sum[r] += (double)cs[r];
    b->cluster_count[r] += clusters[r];
    if (cluster_size) cluster_size[r] = cs[r];
  }
  return SLSF_OK;
}

std::vector<int64_t> clusters_in(const slsf_batch* b, int64_t nsweeps) {
  std::vector<int64_t> n(b->reps.size());
  for (size_t r = 0; r < n.size(); ++r) n[r] = nsweeps * b->reps[r].wolff_steps;
  return n;
}

}  // namespace

int ensure_trig_cache(slsf_batch* b) {
  if (!b->cs_stale) return SLSF_OK;
  if (is_fp64(b)) launch_refresh<double>(b->g, b->stream);
  else launch_refresh<float>(b->g, b->stream);
  HIP_TRY(hipGetLastError());
  b->cs_stale = false;
  return SLSF_OK;
}

// The per-phase kernels read and keep the (cos, sin) cache in FP64; in FP32
// they form the pairs from the angles and leave the cache behind them, for
// the readers of it (correlations, checkpoints) to refresh.
int before_per_phase(slsf_batch* b) { return is_fp64(b) ? ensure_trig_cache(b) : SLSF_OK; }
void after_per_phase(slsf_batch* b) {
  if (!is_fp64(b)) b->cs_stale = true;
}

extern "C" {

const char* slsf_last_error(void) { return g_error.c_str(); }
int32_t slsf_abi_version(void) { return 1; }

int slsf_batch_create(const slsf_replica_desc* replicas, int32_t n, int32_t precision,
                      slsf_batch** out) {
  clear_error();
  if (!out) return fail(SLSF_E_INVALID, "output handle pointer is null");
  *out = nullptr;
  if (n <= 0 || !replicas) return fail(SLSF_E_INVALID, "a batch needs at least one replica");
  if (precision != SLSF_FP64_REPLAY && precision != SLSF_FP32)
    return fail(SLSF_E_INVALID, "precision must be SLSF_FP64_REPLAY or SLSF_FP32");
  for (int i = 0; i < n; ++i)
    if (int rc = validate_replica(replicas[i])) return rc;
  if (int rc = device_ready()) return rc;

  auto b = std::make_unique<slsf_batch>();
  b->precision = precision;
  HIP_TRY(hipGetDevice(&b->device));
  HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  int64_t site_off = 0;
  int jz_off = 0;
  for (int i = 0; i < n; ++i) {
    const auto& d = replicas[i];
    HostReplica r;
    r.lx = d.l_x; r.ly = d.l_y; r.lz = d.l_z;
    r.wolff_steps = d.wolff_steps;
    r.temperature = d.temperature;
    r.width = d.proposal_width;
    r.j_xy.assign(d.j_xy, d.j_xy + d.l_z);
    r.j_z.assign(d.j_z, d.j_z + d.l_z);
    r.seed = d.seed;
    r.volume = (int64_t)d.l_x * d.l_y * d.l_z;
    r.padded = (r.volume + 63) / 64 * 64;
    r.site_off = site_off;
    r.jz_off = jz_off;
    site_off += r.padded;
    jz_off += d.l_z;
    b->reps.push_back(std::move(r));
  }
  b->total_layers = jz_off;
  b->reps_host.resize(n);
  for (int i = 0; i < n; ++i) {
    const auto& r = b->reps[i];
    RepDev& d = b->reps_host[i];
    std::memset(&d, 0, sizeof d);
    d.lx = r.lx; d.ly = r.ly; d.lz = r.lz; d.wolff_steps = r.wolff_steps;
    d.site_off = r.site_off;
    d.volume = r.volume;
    d.jz_off = r.jz_off;
    d.key0 = (uint32_t)r.seed;
    d.key1 = (uint32_t)(r.seed >> 32);
    d.beta = 1.0 / r.temperature;  // src/updates.rs:453
    d.width = r.width;
    d.measure_from = UINT32_MAX;  // outside slsf_run nothing is binned
    d.stop_at = UINT32_MAX;       // and nothing stops a replica
    d.active = 1;
    d.binsize = 1;
    set_divisors(d);
  }
  GlobalState& g = b->g;
  g.nrep = n;
  g.total_sites = site_off;
  set_extents(g, b->reps_host.data(), n);
  const size_t rs = precision == SLSF_FP64_REPLAY ? sizeof(double) : sizeof(float);
  slsf_batch* bp = b.get();
  int rc = SLSF_OK;
  rc = rc ? rc : bp->alloc((RepDev**)&g.reps, (size_t)n);
  rc = rc ? rc : bp->alloc((char**)&g.theta, rs * (size_t)site_off);
  rc = rc ? rc : bp->alloc((char**)&g.cs, 2 * rs * (size_t)site_off);
  rc = rc ? rc : bp->alloc((char**)&g.jxy, rs * (size_t)jz_off);
  rc = rc ? rc : bp->alloc((char**)&g.jz, rs * (size_t)jz_off);
  rc = rc ? rc : bp->alloc(&g.bondmask, (size_t)site_off);
  rc = rc ? rc : bp->alloc(&g.visited, (size_t)site_off / 32);
  rc = rc ? rc : bp->alloc(&g.label, (size_t)site_off);
  rc = rc ? rc : bp->alloc(&g.wd, (size_t)n);
  rc = rc ? rc : bp->alloc(&g.csize, (size_t)n * TALLY_SLOTS);
  rc = rc ? rc : bp->alloc(&g.acc, (size_t)n * TALLY_SLOTS);
  rc = rc ? rc : bp->alloc(&g.partials, (size_t)n * g.obs_blocks * 9);
  rc = rc ? rc : bp->alloc(&bp->d_tally_sum, (size_t)n);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(g.visited, 0, sizeof(uint32_t) * (size_t)(site_off / 32), b->stream));
  HIP_TRY(hipMemsetAsync(g.theta, 0, rs * (size_t)site_off, b->stream));
  if ((rc = push_reps_to_device(bp))) return rc;
  rc = precision == SLSF_FP64_REPLAY ? upload_couplings<double>(bp) : upload_couplings<float>(bp);
  if (rc) return rc;
  b->acc_sum.assign(n, 0.0);
  b->acc_count.assign(n, 0);
  b->cluster_sum.assign(n, 0.0);
  b->cluster_count.assign(n, 0);
  if (precision == SLSF_FP32) {
    if ((rc = resident_plan(bp->reps_host, bp->resident))) return rc;
  }
  *out = b.release();
  return SLSF_OK;
}

void slsf_batch_destroy(slsf_batch* batch) { delete batch; }
int32_t slsf_batch_size(const slsf_batch* b) { return b ? (int32_t)b->reps.size() : 0; }
int32_t slsf_batch_precision(const slsf_batch* b) { return b ? b->precision : -1; }
int64_t slsf_batch_device_bytes(const slsf_batch* b) { return b ? b->device_bytes : 0; }

int slsf_init_angles(slsf_batch* b, int32_t cold) {
  clear_error();
  if (!b) return fail(SLSF_E_INVALID, "batch handle is null");
  for (auto& d : b->reps_host) { d.metro_ctr = 0; d.wolff_ctr = 0; }
  if (int rc = push_reps_to_device(b)) return rc;
  if (is_fp64(b)) launch_init<double>(b->g, cold, b->stream);
  else launch_init<float>(b->g, cold, b->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(b->stream));
  std::fill(b->acc_sum.begin(), b->acc_sum.end(), 0.0);
  std::fill(b->acc_count.begin(), b->acc_count.end(), 0);
  std::fill(b->cluster_sum.begin(), b->cluster_sum.end(), 0.0);
  std::fill(b->cluster_count.begin(), b->cluster_count.end(), 0);
  b->initialised = true;
  b->cs_stale = false;
  return SLSF_OK;
}

int slsf_get_theta(slsf_batch* b, int32_t replica, double* theta_out) {
  clear_error();
  if (!b || !theta_out) return fail(SLSF_E_INVALID, "null argument");
  if (replica < 0 || replica >= (int)b->reps.size())
    return fail(SLSF_E_INVALID, "replica index out of range");
  const auto& r = b->reps[replica];
  if (is_fp64(b)) {
    HIP_TRY(hipMemcpyAsync(theta_out, (double*)b->g.theta + r.site_off,
                           sizeof(double) * r.volume, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
  } else {
    // FP32 state is kept in turns; report radians like the FP64 state
    std::vector<float> tmp(r.volume);
    HIP_TRY(hipMemcpyAsync(tmp.data(), (float*)b->g.theta + r.site_off,
                           sizeof(float) * r.volume, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (int64_t i = 0; i < r.volume; ++i) theta_out[i] = (double)tmp[i] * (2.0 * M_PI);
  }
  return SLSF_OK;
}

int slsf_set_theta(slsf_batch* b, int32_t replica, const double* theta) {
  clear_error();
  if (!b || !theta) return fail(SLSF_E_INVALID, "null argument");
  if (replica < 0 || replica >= (int)b->reps.size())
    return fail(SLSF_E_INVALID, "replica index out of range");
  const auto& r = b->reps[replica];
  if (is_fp64(b)) {
    HIP_TRY(hipMemcpyAsync((double*)b->g.theta + r.site_off, theta, sizeof(double) * r.volume,
                           hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    launch_refresh<double>(b->g, b->stream);
  } else {
    std::vector<float> tmp(r.volume);
    for (int64_t i = 0; i < r.volume; ++i) {
      double t = theta[i] / (2.0 * M_PI);
      t -= std::floor(t);
      float f = (float)t;
      tmp[i] = f >= 1.0f ? 0.0f : f;
    }
    HIP_TRY(hipMemcpyAsync((float*)b->g.theta + r.site_off, tmp.data(), sizeof(float) * r.volume,
                           hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    launch_refresh<float>(b->g, b->stream);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->initialised = true;
  b->cs_stale = false;
  return SLSF_OK;
}

int slsf_get_counters(const slsf_batch* bc, uint32_t* metro, uint32_t* wolff) {
  clear_error();
  slsf_batch* b = const_cast<slsf_batch*>(bc);
  if (!b) return fail(SLSF_E_INVALID, "batch handle is null");
  if (int rc = sync_reps_from_device(b)) return rc;
  for (size_t r = 0; r < b->reps.size(); ++r) {
    if (metro) metro[r] = b->reps_host[r].metro_ctr;
    if (wolff) wolff[r] = b->reps_host[r].wolff_ctr;
  }
  return SLSF_OK;
}

int slsf_set_counters(slsf_batch* b, uint32_t metro, uint32_t wolff) {
  clear_error();
  if (!b) return fail(SLSF_E_INVALID, "batch handle is null");
  if (int rc = sync_reps_from_device(b)) return rc;
  for (auto& d : b->reps_host) { d.metro_ctr = metro; d.wolff_ctr = wolff; }
  return push_reps_to_device(b);
}

int slsf_metropolis_sweep(slsf_batch* b, double* acceptance, uint8_t* decisions) {
  clear_error();
  if (!b) return fail(SLSF_E_INVALID, "batch handle is null");
  if (!b->initialised) return fail(SLSF_E_STATE, "angles are not initialised");
  if (int rc = before_per_phase(b)) return rc;
  if (decisions && !b->d_decisions)
    if (int rc = b->alloc(&b->d_decisions, (size_t)b->g.total_sites)) return rc;
  if (acceptance) std::fill(acceptance, acceptance + b->reps.size(), 0.0);
  launch_clear_counters(b->g, b->stream);
  launch_sweep_end(b->g, 0, 0, b->stream);
  if (int rc = metro_phases(b, decisions ? b->d_decisions : nullptr)) return rc;
  after_per_phase(b);
  launch_sweep_end(b->g, 1, 0, b->stream);
  if (int rc = gather_acceptance(b, 1, acceptance)) return rc;
  if (decisions) {
    std::vector<uint8_t> all(b->g.total_sites);
    HIP_TRY(hipMemcpyAsync(all.data(), b->d_decisions, all.size(), hipMemcpyDeviceToHost,
                           b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    int64_t o = 0;
    for (const auto& r : b->reps) {
      std::memcpy(decisions + o, all.data() + r.site_off, (size_t)r.volume);
      o += r.volume;
    }
  }
  return SLSF_OK;
}

int slsf_wolff_cluster_step(slsf_batch* b, int64_t* cluster_size, uint8_t* members) {
  clear_error();
  if (!b) return fail(SLSF_E_INVALID, "batch handle is null");
  if (!b->initialised) return fail(SLSF_E_STATE, "angles are not initialised");
  if (int rc = before_per_phase(b)) return rc;
  // a standalone cluster step runs one cluster for every replica (round 0 with
  // wolff_steps forced >= 1 would change the descriptor; use a dedicated path)
  if (int rc = sync_reps_from_device(b)) return rc;
  std::vector<int> saved(b->reps_host.size());
  bool patched = false;
  for (size_t r = 0; r < saved.size(); ++r) {
    saved[r] = b->reps_host[r].wolff_steps;
    if (saved[r] < 1) { b->reps_host[r].wolff_steps = 1; patched = true; }
  }
  if (patched) if (int rc = push_reps_to_device(b)) return rc;
  launch_clear_counters(b->g, b->stream);
  launch_sweep_end(b->g, 0, 0, b->stream);
  int rc = wolff_round(b, 0, members);
  after_per_phase(b);
  if (!rc) {
    launch_sweep_end(b->g, 0, 1, b->stream);
    rc = gather_clusters(b, std::vector<int64_t>(b->reps.size(), 1), cluster_size);
  }
  if (patched) {
    if (int rc2 = sync_reps_from_device(b)) return rc2;
    for (size_t r = 0; r < saved.size(); ++r) b->reps_host[r].wolff_steps = saved[r];
    if (int rc2 = push_reps_to_device(b)) return rc2;
  }
  return rc;
}

int slsf_sweep(slsf_batch* b, int32_t nsweeps, double* acceptance_sum) {
  clear_error();
  if (!b) return fail(SLSF_E_INVALID, "batch handle is null");
  if (!b->initialised) return fail(SLSF_E_STATE, "angles are not initialised");
  if (int rc = before_per_phase(b)) return rc;
  if (nsweeps < 0) return fail(SLSF_E_INVALID, "nsweeps must be nonnegative");
  const int wmax = max_wolff_steps(b);
  launch_clear_counters(b->g, b->stream);
  launch_sweep_end(b->g, 0, 0, b->stream);
  after_per_phase(b);
  for (int s = 0; s < nsweeps; ++s)
    if (int rc = hbm_sweep(b, wmax)) return rc;
  if (int rc = gather_acceptance(b, nsweeps, acceptance_sum)) return rc;
  return gather_clusters(b, clusters_in(b, nsweeps), nullptr);
}

int slsf_measure_observables(slsf_batch* b, slsf_observables* out) {
  clear_error();
  if (!b || !out) return fail(SLSF_E_INVALID, "null argument");
  if (int rc = before_per_phase(b)) return rc;
  if (is_fp64(b)) launch_observables<double>(b->g, b->stream);
  else launch_observables<float>(b->g, b->stream);
  HIP_TRY(hipGetLastError());
  const int stride = b->g.obs_blocks;
  std::vector<double> part((size_t)b->reps.size() * stride * 9);
  HIP_TRY(hipMemcpyAsync(part.data(), b->g.partials, sizeof(double) * part.size(),
                         hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (size_t r = 0; r < b->reps.size(); ++r) {
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const auto& rr = b->reps[r];
    const int nb = obs_blocks_of(rr.lx, rr.ly, rr.lz);
    for (int blk = 0; blk < nb; ++blk)
      for (int k = 0; k < 9; ++k) v[k] += part[((size_t)r * stride + blk) * 9 + k];
    const double vol = (double)b->reps[r].volume;
    out[r].energy = v[0] / vol;
    out[r].magnetization_squared = (v[1] * v[1] + v[2] * v[2]) / (vol * vol);
    out[r].cos_x = v[3]; out[r].cos_y = v[4]; out[r].cos_z = v[5];
    out[r].sin_x = v[6]; out[r].sin_y = v[7]; out[r].sin_z = v[8];
  }
  return SLSF_OK;
}

int slsf_correlation_rmax(const slsf_batch* b, int32_t replica, int32_t rmax, int32_t rmax_xy,
                          int32_t rmax_z, int32_t* rxy, int32_t* rz) {
  clear_error();
  if (!b || !rxy || !rz) return fail(SLSF_E_INVALID, "null argument");
  if (replica < 0 || replica >= (int)b->reps.size())
    return fail(SLSF_E_INVALID, "replica index out of range");
  const auto& r = b->reps[replica];
  // src/observables.rs:177-181
  const int req_xy = rmax_xy >= 0 ? rmax_xy : (rmax >= 0 ? rmax : std::min(r.lx / 2, r.ly / 2));
  const int req_z = rmax_z >= 0 ? rmax_z : (rmax >= 0 ? rmax : r.lz / 2);
  *rxy = std::min(std::min(req_xy, r.lx / 2), r.ly / 2);
  *rz = std::min(req_z, r.lz / 2);
  return SLSF_OK;
}

int slsf_measure_correlations(slsf_batch* b, int32_t replica, int32_t rmax, int32_t rmax_xy,
                              int32_t rmax_z, double* corr_x, double* corr_y, double* corr_xy,
                              double* corr_xy_by_z, double* corr_z) {
  int32_t rxy = 0, rz = 0;
  if (int rc = slsf_correlation_rmax(b, replica, rmax, rmax_xy, rmax_z, &rxy, &rz)) return rc;
  if (int rc = ensure_trig_cache(b)) return rc;
  const auto& r = b->reps[replica];
  const int64_t rows = (int64_t)r.ly * r.lz;
  double* d_xy = nullptr;
  double* d_z = nullptr;
  HIP_TRY(hipMallocAsync((void**)&d_xy, sizeof(double) * (size_t)std::max<int64_t>(1, rxy * rows * 3), b->stream));
  HIP_TRY(hipMallocAsync((void**)&d_z, sizeof(double) * (size_t)std::max<int64_t>(1, rz * rows), b->stream));
  if (is_fp64(b)) launch_corr<double>(b->g, replica, rows, rxy, rz, d_xy, d_z, b->stream);
  else launch_corr<float>(b->g, replica, rows, rxy, rz, d_xy, d_z, b->stream);
  HIP_TRY(hipGetLastError());
  std::vector<double> hxy((size_t)rxy * rows * 3), hz((size_t)rz * rows);
  if (!hxy.empty())
    HIP_TRY(hipMemcpyAsync(hxy.data(), d_xy, sizeof(double) * hxy.size(), hipMemcpyDeviceToHost,
                           b->stream));
  if (!hz.empty())
    HIP_TRY(hipMemcpyAsync(hz.data(), d_z, sizeof(double) * hz.size(), hipMemcpyDeviceToHost,
                           b->stream));
  HIP_TRY(hipFreeAsync(d_xy, b->stream));
  HIP_TRY(hipFreeAsync(d_z, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  // host reduction in the reference's order (src/observables.rs:189-252)
  const double vol = (double)r.volume, area = (double)r.lx * r.ly;
  for (int rr = 1; rr <= rxy; ++rr) {
    double sx = 0, sy = 0;
    for (int z = 0; z < r.lz; ++z) {
      double lsx = 0, lsy = 0;
      for (int y = 0; y < r.ly; ++y) {
        const double* o = &hxy[(((size_t)(rr - 1) * rows) + (size_t)z * r.ly + y) * 3];
        lsx += o[0];
        lsx += o[1];
        lsy += o[2];
      }
      sx += lsx;
      sy += lsy;
      if (corr_xy_by_z) corr_xy_by_z[(size_t)z * rxy + (rr - 1)] = (lsx + lsy) / (2.0 * area);
    }
    if (corr_x) corr_x[rr - 1] = sx / vol;
    if (corr_y) corr_y[rr - 1] = sy / vol;
    if (corr_xy) corr_xy[rr - 1] = (sx / vol + sy / vol) / 2.0;
  }
  for (int rr = 1; rr <= rz; ++rr) {
    double sz = 0;
    for (int z = 0; z < r.lz; ++z)
      for (int y = 0; y < r.ly; ++y) sz += hz[((size_t)(rr - 1) * rows) + (size_t)z * r.ly + y];
    if (corr_z) corr_z[rr - 1] = sz / vol;
  }
  return SLSF_OK;
}

int slsf_get_acceptance(slsf_batch* b, double* acceptance) {
  clear_error();
  if (!b || !acceptance) return fail(SLSF_E_INVALID, "null argument");
  for (size_t r = 0; r < b->reps.size(); ++r)
    acceptance[r] = b->acc_count[r] ? b->acc_sum[r] / (double)b->acc_count[r] : 0.0;
  return SLSF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// production runs
// ---------------------------------------------------------------------------
namespace {

int ensure_accumulators(slsf_batch* b) {
  const size_t nrep = b->reps.size();
  if (b->d_accum) return SLSF_OK;
  if (int rc = b->alloc(&b->d_accum, nrep * SLSF_NUM_SERIES)) return rc;
  if (int rc = b->alloc(&b->d_bin_state, nrep * 2)) return rc;
  HIP_TRY(hipMemsetAsync(b->d_accum, 0, sizeof(double) * nrep * SLSF_NUM_SERIES, b->stream));
  HIP_TRY(hipMemsetAsync(b->d_bin_state, 0, sizeof(int32_t) * nrep * 2, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return SLSF_OK;
}

// room for `need` completed bins per (replica, series); a larger buffer takes
// at least twice the old capacity, and the completed bins move over in one
// pitched copy
int grow_bins(slsf_batch* b, int64_t need) {
  if (need <= b->bin_cap) return SLSF_OK;
  if (need > INT32_MAX) return fail(SLSF_E_INVALID, "too many bins for one batch");
  need = std::max<int64_t>(need, 2 * (int64_t)b->bin_cap);
  need = std::min<int64_t>(need, INT32_MAX);
  const size_t rows = b->reps.size() * SLSF_NUM_SERIES;
  double* nb = nullptr;
  if (int rc = b->alloc(&nb, rows * (size_t)need)) return rc;
  if (b->d_bins) {  // keep completed bins
    HIP_TRY(hipMemcpy2DAsync(nb, sizeof(double) * (size_t)need, b->d_bins,
                             sizeof(double) * (size_t)b->bin_cap,
                             sizeof(double) * (size_t)b->bin_cap, rows,
                             hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->release(b->d_bins);
    b->device_bytes -= (int64_t)(sizeof(double) * rows * (size_t)b->bin_cap);
  }
  b->d_bins = nb;
  b->bin_cap = (int32_t)need;
  return SLSF_OK;
}

int ensure_bins(slsf_batch* b, const slsf_run_options* opt) {
  const size_t nrep = b->reps.size();
  if (int rc = ensure_accumulators(b)) return rc;
  if (opt->sweeps == 0) return SLSF_OK;
  std::vector<int32_t> st(nrep * 2);
  HIP_TRY(hipMemcpyAsync(st.data(), b->d_bin_state, sizeof(int32_t) * st.size(),
                         hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  int32_t have = 0;
  bool measured = false;
  for (size_t r = 0; r < nrep; ++r) {
    have = std::max(have, st[2 * r + 1]);
    measured = measured || st[2 * r] != 0 || st[2 * r + 1] != 0;
  }
  if (b->binsize != 0 && b->binsize != opt->binsize && measured)
    return fail(SLSF_E_STATE, "binsize must not change between measured runs of one batch");
  b->binsize = opt->binsize;
  return grow_bins(b, (int64_t)have + opt->sweeps / opt->binsize + 1);
}

}  // namespace

extern "C" {

const char* slsf_series_name(int32_t s) {
  static const char* names[SLSF_NUM_SERIES] = {
      "Energy",          "EnergySquared",   "MagnetizationSquared", "MagnetizationFourth",
      "_helicity_cos_x", "_helicity_cos_y", "_helicity_cos_z",      "_helicity_sin_x",
      "_helicity_sin_y", "_helicity_sin_z", "_helicity_sin2_x",     "_helicity_sin2_y",
      "_helicity_sin2_z"};
  return (s >= 0 && s < SLSF_NUM_SERIES) ? names[s] : "";
}

int slsf_run(slsf_batch* b, const slsf_run_options* opt, slsf_run_stats* stats) {
  clear_error();
  if (!b || !opt) return fail(SLSF_E_INVALID, "null argument");
  if (!b->initialised) return fail(SLSF_E_STATE, "angles are not initialised");
  if (opt->thermalization < 0 || opt->sweeps < 0)
    return fail(SLSF_E_INVALID, "sweep counts must be nonnegative");
  if (opt->sweeps > 0 && opt->binsize <= 0) return fail(SLSF_E_INVALID, "binsize must be positive");
  if (opt->stream != SLSF_STREAM_PRODUCTION && opt->stream != SLSF_STREAM_REPLAY)
    return fail(SLSF_E_INVALID, "stream must be SLSF_STREAM_PRODUCTION or SLSF_STREAM_REPLAY");
  if (int rc = ensure_bins(b, opt)) return rc;
  // run window of every replica: measured from counter + thermalization, stop
  // after thermalization + sweeps (kernels enforce both per replica); set and
  // closed again on the device, in stream order
  launch_window(b->g, (uint32_t)opt->thermalization,
                (uint32_t)(opt->thermalization + opt->sweeps),
                opt->binsize > 0 ? opt->binsize : 1, 1, b->stream);
  HIP_TRY(hipGetLastError());
  struct RestoreWindow {
    slsf_batch* b;
    ~RestoreWindow() {
      launch_window(b->g, 0, 0, 1, 0, b->stream);
      (void)hipStreamSynchronize(b->stream);
    }
  } restore{b};
  const auto t0 = std::chrono::steady_clock::now();
  int64_t launches = 0;
  double kernel_ms = 0.0;
  // FP32: every cubic replica of a resident size runs the resident kernel and
  // the others the HBM kernels, so a replica's path (and stream) depends on its
  // own size only, never on what it is batched with.  FP64 replay: HBM kernels.
  const int64_t total = opt->thermalization + opt->sweeps;
  const bool resident_part = b->precision == SLSF_FP32 && !b->resident.groups.empty();
  const bool hbm_part = !resident_part || !b->resident.covers_all;
  if (resident_part) {
    ResidentRunArgs a;
    a.thermalization = opt->thermalization;
    a.sweeps = opt->sweeps;
    a.binsize = opt->binsize;
    a.sweeps_per_launch = opt->sweeps_per_launch;
    a.use_graph = opt->use_graph;
    a.replay_stream = opt->stream == SLSF_STREAM_REPLAY;
    a.accum = b->d_accum;
    a.bins = b->d_bins;
    a.bin_state = b->d_bin_state;
    a.bin_cap = b->bin_cap;
    a.acc_out = b->g.acc;
    a.csize_out = b->g.csize;
    ResidentRunResult res;  // zeroes the tallies of every replica first
    b->cs_stale = true;     // the resident kernel writes angles only
    if (int rc = resident_run(b->resident, b->g, a, b->stream, &res)) return rc;
    launches += res.launches;
    kernel_ms += res.kernel_ms;
  }
  if (hbm_part) {
    // HBM path: tallies stay on the device until the end of the run; the
    // replicas' windows cut thermalization (no measurement) and the tail, and
    // replicas the resident kernel already ran to their window end sit out
    const int wmax = max_wolff_steps(b);
    if (opt->hbm_form < 0 || opt->hbm_form > 2)
      return fail(SLSF_E_INVALID, "hbm_form must be 0 (by launch size), 1 (slab) or 2 (per-phase)");
    if (int rc = ensure_slab_buffers(b)) return rc;
    const bool slab = hbm_uses_slab(b->g, opt->hbm_form);
    if (slab) b->cs_stale = true;  // slab sweeps leave the trig cache behind
    else if (int rc = before_per_phase(b)) return rc;
    after_per_phase(b);
    if (!resident_part) {
      launch_clear_counters(b->g, b->stream);
      launches += 2;
    }
    launch_sweep_end(b->g, 0, 0, b->stream);
    launches += 1;
    BinTarget bt;
    if (opt->sweeps > 0) {
      bt.accum = b->d_accum;
      bt.bins = b->d_bins;
      bt.bin_state = b->d_bin_state;
      bt.bin_cap = b->bin_cap;
    }
    auto enqueue = [&](int64_t n) {
      return is_fp64(b) ? enqueue_hbm_sweeps<double>(b->g, n, wmax, bt, b->stream, opt->hbm_form)
                        : enqueue_hbm_sweeps<float>(b->g, n, wmax, bt, b->stream, opt->hbm_form);
    };
    if (opt->use_graph && total > 0) {
      // one graph of `chunk` sweeps, replayed; sweeps past the window are no-ops
      const int64_t chunk = std::min<int64_t>(
          total, opt->sweeps_per_launch > 0 ? opt->sweeps_per_launch : 20);
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int64_t per = enqueue(chunk);
      hipError_t e = hipStreamEndCapture(b->stream, &graph);
      if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      for (int64_t done = 0; e == hipSuccess && done < total; done += chunk) {
        e = hipGraphLaunch(exec, b->stream);
        launches += per;
      }
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      HIP_TRY(e);
    } else {
      launches += enqueue(total);
    }
    HIP_TRY(hipGetLastError());
  }
  if (int rc = gather_acceptance(b, total, nullptr)) return rc;
  if (int rc = gather_clusters(b, clusters_in(b, total), nullptr)) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  const auto t1 = std::chrono::steady_clock::now();
  if (stats) {
    stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    stats->kernel_ms = kernel_ms;
    int64_t su = 0;
    double acc = 0, cl = 0;
    int64_t ncl = 0;
    for (size_t r = 0; r < b->reps.size(); ++r) {
      su += b->reps[r].volume * (opt->thermalization + opt->sweeps);
      acc += b->acc_count[r] ? b->acc_sum[r] / (double)b->acc_count[r] : 0.0;
      cl += b->cluster_sum[r];
      ncl += b->cluster_count[r];
    }
    stats->site_updates = su;
    stats->launches = launches;
    stats->acceptance_mean = acc / (double)b->reps.size();
    stats->mean_cluster_size = ncl ? cl / (double)ncl : 0.0;
  }
  return SLSF_OK;
}

int slsf_autotune(slsf_batch* b, int32_t trial_sweeps, slsf_tune_record* records, int32_t cap,
                  int32_t* count) {
  clear_error();
  if (!b) return fail(SLSF_E_INVALID, "batch handle is null");
  if (count) *count = 0;
  if (b->precision != SLSF_FP32 || b->resident.groups.empty())
    return SLSF_OK;  // nothing to tune: the HBM kernels have no candidates
  if (!b->initialised) return fail(SLSF_E_STATE, "angles are not initialised");
  std::vector<TuneRecord> recs;
  if (int rc = resident_autotune(b->resident, b->g, b->stream, std::max(trial_sweeps, 1), recs))
    return rc;
  int32_t n = 0;
  for (const auto& r : recs) {
    if (records && n < cap) {
      slsf_tune_record& o = records[n];
      o.l = r.L;
      o.columns_per_thread = r.cpt;
      o.waves_per_replica = r.waves;
      o.sweeps_per_launch = r.spl;
      o.site_updates_per_s = r.rate;
      o.chosen = 0;
      o.reserved = 0;
      for (const auto& grp : b->resident.groups)
        if (grp.L == r.L && grp.cpt == r.cpt) o.chosen = 1;
    }
    ++n;
  }
  if (count) *count = n;
  return SLSF_OK;
}

int slsf_reserve_bins(slsf_batch* b, int32_t nbins) {
  clear_error();
  if (!b) return fail(SLSF_E_INVALID, "batch handle is null");
  if (nbins < 0) return fail(SLSF_E_INVALID, "nbins must be nonnegative");
  if (int rc = ensure_accumulators(b)) return rc;
  return grow_bins(b, nbins);
}

int slsf_get_bins(slsf_batch* b, int32_t replica, int32_t series, double* out, int32_t cap,
                  int32_t* nbins) {
  clear_error();
  if (!b || !nbins) return fail(SLSF_E_INVALID, "null argument");
  if (replica < 0 || replica >= (int)b->reps.size())
    return fail(SLSF_E_INVALID, "replica index out of range");
  if (series < 0 || series >= SLSF_NUM_SERIES) return fail(SLSF_E_INVALID, "unknown series");
  if (!b->d_bins) { *nbins = 0; return SLSF_OK; }
  int32_t st[2];
  HIP_TRY(hipMemcpy(st, b->d_bin_state + 2 * replica, sizeof st, hipMemcpyDeviceToHost));
  *nbins = st[1];
  const int32_t n = std::min(cap, st[1]);
  if (n > 0 && out)
    HIP_TRY(hipMemcpy(out, b->d_bins + ((size_t)replica * SLSF_NUM_SERIES + series) * b->bin_cap,
                      sizeof(double) * n, hipMemcpyDeviceToHost));
  return SLSF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// checkpoint image: the whole run state of a batch as bytes
// ---------------------------------------------------------------------------
namespace {

constexpr char CKPT_MAGIC[8] = {'S', 'L', 'S', 'F', 'C', 'K', 'P', 'T'};
constexpr uint32_t CKPT_VERSION = 1;

struct CkptHeader {
  char magic[8];
  uint32_t version;
  int32_t precision;
  int32_t nrep;
  int32_t binsize;
  int32_t nbins;       // completed bins stored per replica and series
  int32_t measured;    // 1: the accumulators and bin counters follow
  int64_t total_sites;
};
struct CkptReplica {   // identity of a replica: a checkpoint only loads into the same one
  int32_t lx, ly, lz, wolff_steps;
  uint64_t seed;
  double temperature, width;
};
struct CkptChain {     // chain position and run tallies
  uint32_t metro_ctr, wolff_ctr;
  double acc_sum;
  int64_t acc_count;
  double cluster_sum;
  int64_t cluster_count;
};
static_assert(sizeof(CkptHeader) == 40 && sizeof(CkptReplica) == 40 && sizeof(CkptChain) == 40,
              "checkpoint records are packed");

size_t real_bytes(const slsf_batch* b) { return is_fp64(b) ? 8 : 4; }

int ckpt_layout(slsf_batch* b, CkptHeader& h, size_t& bytes) {
  const size_t nrep = b->reps.size();
  std::memcpy(h.magic, CKPT_MAGIC, 8);
  h.version = CKPT_VERSION;
  h.precision = b->precision;
  h.nrep = (int32_t)nrep;
  h.binsize = b->binsize;
  h.nbins = 0;
  h.measured = b->d_accum ? 1 : 0;
  h.total_sites = b->g.total_sites;
  if (b->d_bin_state) {
    std::vector<int32_t> st(nrep * 2);
    HIP_TRY(hipMemcpy(st.data(), b->d_bin_state, sizeof(int32_t) * st.size(),
                      hipMemcpyDeviceToHost));
    for (size_t r = 0; r < nrep; ++r) h.nbins = std::max(h.nbins, std::min(st[2 * r + 1], b->bin_cap));
  }
  bytes = sizeof(CkptHeader) + nrep * (sizeof(CkptReplica) + sizeof(CkptChain)) +
          (size_t)h.total_sites * 3 * real_bytes(b);
  if (h.measured)
    bytes += nrep * SLSF_NUM_SERIES * sizeof(double) + nrep * 2 * sizeof(int32_t) +
             nrep * SLSF_NUM_SERIES * (size_t)h.nbins * sizeof(double);
  return SLSF_OK;
}

CkptReplica identity(const HostReplica& r) {
  CkptReplica c;
  c.lx = r.lx;
  c.ly = r.ly;
  c.lz = r.lz;
  c.wolff_steps = r.wolff_steps;
  c.seed = r.seed;
  c.temperature = r.temperature;
  c.width = r.width;
  return c;
}

}  // namespace

extern "C" {

int slsf_checkpoint_size(slsf_batch* b, int64_t* bytes) {
  clear_error();
  if (!b || !bytes) return fail(SLSF_E_INVALID, "null argument");
  CkptHeader h;
  size_t n = 0;
  if (int rc = ckpt_layout(b, h, n)) return rc;
  *bytes = (int64_t)n;
  return SLSF_OK;
}

int slsf_checkpoint_save(slsf_batch* b, void* buffer, int64_t bytes) {
  clear_error();
  if (!b || !buffer) return fail(SLSF_E_INVALID, "null argument");
  if (!b->initialised) return fail(SLSF_E_STATE, "angles are not initialised");
  CkptHeader h;
  size_t need = 0;
  if (int rc = ckpt_layout(b, h, need)) return rc;
  if (bytes < (int64_t)need) return fail(SLSF_E_INVALID, "checkpoint buffer is too small");
  if (int rc = ensure_trig_cache(b)) return rc;
  if (int rc = sync_reps_from_device(b)) return rc;
  const size_t nrep = b->reps.size();
  char* p = static_cast<char*>(buffer);
  std::memcpy(p, &h, sizeof h);
  p += sizeof h;
  for (size_t r = 0; r < nrep; ++r) {
    const CkptReplica c = identity(b->reps[r]);
    std::memcpy(p, &c, sizeof c);
    p += sizeof c;
  }
  for (size_t r = 0; r < nrep; ++r) {
    CkptChain c;
    c.metro_ctr = b->reps_host[r].metro_ctr;
    c.wolff_ctr = b->reps_host[r].wolff_ctr;
    c.acc_sum = b->acc_sum[r];
    c.acc_count = b->acc_count[r];
    c.cluster_sum = b->cluster_sum[r];
    c.cluster_count = b->cluster_count[r];
    std::memcpy(p, &c, sizeof c);
    p += sizeof c;
  }
  // angles and their trig cache (refreshed above) exactly as the kernels hold them
  const size_t tb = (size_t)h.total_sites * real_bytes(b);
  HIP_TRY(hipMemcpy(p, b->g.theta, tb, hipMemcpyDeviceToHost));
  p += tb;
  HIP_TRY(hipMemcpy(p, b->g.cs, 2 * tb, hipMemcpyDeviceToHost));
  p += 2 * tb;
  if (h.measured) {
    HIP_TRY(hipMemcpy(p, b->d_accum, sizeof(double) * nrep * SLSF_NUM_SERIES,
                      hipMemcpyDeviceToHost));
    p += sizeof(double) * nrep * SLSF_NUM_SERIES;
    HIP_TRY(hipMemcpy(p, b->d_bin_state, sizeof(int32_t) * nrep * 2, hipMemcpyDeviceToHost));
    p += sizeof(int32_t) * nrep * 2;
    if (h.nbins > 0)
      HIP_TRY(hipMemcpy2D(p, sizeof(double) * (size_t)h.nbins, b->d_bins,
                          sizeof(double) * (size_t)b->bin_cap, sizeof(double) * (size_t)h.nbins,
                          nrep * SLSF_NUM_SERIES, hipMemcpyDeviceToHost));
  }
  return SLSF_OK;
}

int slsf_checkpoint_load(slsf_batch* b, const void* buffer, int64_t bytes) {
  clear_error();
  if (!b || !buffer) return fail(SLSF_E_INVALID, "null argument");
  const size_t nrep = b->reps.size();
  const char* p = static_cast<const char*>(buffer);
  CkptHeader h;
  if (bytes < (int64_t)sizeof h) return fail(SLSF_E_INVALID, "checkpoint image is truncated");
  std::memcpy(&h, p, sizeof h);
  p += sizeof h;
  if (std::memcmp(h.magic, CKPT_MAGIC, 8) != 0 || h.version != CKPT_VERSION)
    return fail(SLSF_E_INVALID, "not a checkpoint image of this library version");
  if (h.precision != b->precision)
    return fail(SLSF_E_INVALID, "checkpoint precision does not match the batch");
  if (h.nrep != (int32_t)nrep || h.total_sites != b->g.total_sites)
    return fail(SLSF_E_INVALID, "checkpoint replica count does not match the batch");
  if (h.nbins < 0 || h.binsize < 0 || (h.measured != 0 && h.measured != 1))
    return fail(SLSF_E_INVALID, "checkpoint header is corrupt");
  const size_t tb = (size_t)h.total_sites * real_bytes(b);
  size_t need = sizeof h + nrep * (sizeof(CkptReplica) + sizeof(CkptChain)) + 3 * tb;
  if (h.measured)
    need += nrep * SLSF_NUM_SERIES * sizeof(double) + nrep * 2 * sizeof(int32_t) +
            nrep * SLSF_NUM_SERIES * (size_t)h.nbins * sizeof(double);
  if (bytes != (int64_t)need) return fail(SLSF_E_INVALID, "checkpoint image has the wrong size");
  for (size_t r = 0; r < nrep; ++r) {
    CkptReplica c;
    std::memcpy(&c, p, sizeof c);
    p += sizeof c;
    const CkptReplica mine = identity(b->reps[r]);
    if (std::memcmp(&c, &mine, sizeof c) != 0)
      return fail(SLSF_E_INVALID, "checkpoint replica " + std::to_string(r) +
                                      " differs from the batch's (size, seed, temperature, "
                                      "proposal width or wolff steps)");
  }
  std::vector<CkptChain> chain(nrep);
  std::memcpy(chain.data(), p, sizeof(CkptChain) * nrep);
  p += sizeof(CkptChain) * nrep;
  HIP_TRY(hipMemcpy(b->g.theta, p, tb, hipMemcpyHostToDevice));
  p += tb;
  HIP_TRY(hipMemcpy(b->g.cs, p, 2 * tb, hipMemcpyHostToDevice));
  p += 2 * tb;
  for (size_t r = 0; r < nrep; ++r) {
    b->reps_host[r].metro_ctr = chain[r].metro_ctr;
    b->reps_host[r].wolff_ctr = chain[r].wolff_ctr;
    b->acc_sum[r] = chain[r].acc_sum;
    b->acc_count[r] = chain[r].acc_count;
    b->cluster_sum[r] = chain[r].cluster_sum;
    b->cluster_count[r] = chain[r].cluster_count;
  }
  if (int rc = push_reps_to_device(b)) return rc;
  if (h.measured) {
    slsf_run_options none{};  // allocates the accumulators and bin counters
    if (int rc = ensure_bins(b, &none)) return rc;
    HIP_TRY(hipMemcpy(b->d_accum, p, sizeof(double) * nrep * SLSF_NUM_SERIES,
                      hipMemcpyHostToDevice));
    p += sizeof(double) * nrep * SLSF_NUM_SERIES;
    HIP_TRY(hipMemcpy(b->d_bin_state, p, sizeof(int32_t) * nrep * 2, hipMemcpyHostToDevice));
    p += sizeof(int32_t) * nrep * 2;
    if (int rc = grow_bins(b, h.nbins)) return rc;
    if (h.nbins > 0) {
      HIP_TRY(hipMemcpy2D(b->d_bins, sizeof(double) * (size_t)b->bin_cap, p,
                          sizeof(double) * (size_t)h.nbins, sizeof(double) * (size_t)h.nbins,
                          nrep * SLSF_NUM_SERIES, hipMemcpyHostToDevice));
      p += sizeof(double) * (size_t)h.nbins * nrep * SLSF_NUM_SERIES;
    }
  } else if (b->d_accum) {  // the image has no measurement: none is open afterwards
    HIP_TRY(hipMemset(b->d_accum, 0, sizeof(double) * nrep * SLSF_NUM_SERIES));
    HIP_TRY(hipMemset(b->d_bin_state, 0, sizeof(int32_t) * nrep * 2));
  }
  b->binsize = h.binsize;
  b->initialised = true;
  b->cs_stale = false;  // the image holds the cache of its angles
  return SLSF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the plain driver's session: sweeps, measurements and correlation sums stay
// on the stream until driver_close
// ---------------------------------------------------------------------------
namespace slsf {

struct DriverSession {
  slsf_batch* b = nullptr;
  double* series = nullptr;   // [nmeas][8]
  double* rows_xy = nullptr;  // [rxy][ly * lz][3]
  double* rows_z = nullptr;   // [rz][ly * lz]
  double* sums = nullptr;
  double* planes = nullptr;   // [2 rxy + rz][lz] per-plane sums of one measurement
  int64_t nmeas = 0, nsums = 0;
  int rxy = 0, rz = 0, wmax = 0;
};

int driver_open(slsf_batch* b, int64_t nmeas, int rxy, int rz, DriverSession** out) {
  if (!b || !out) return fail(SLSF_E_INVALID, "null argument");
  if (!b->initialised) return fail(SLSF_E_STATE, "angles are not initialised");
  if (int rc = ensure_trig_cache(b)) return rc;
  auto s = std::make_unique<DriverSession>();
  s->b = b;
  s->nmeas = nmeas;
  s->rxy = rxy;
  s->rz = rz;
  s->wmax = max_wolff_steps(b);
  const auto& r = b->reps[0];
  const int64_t rows = (int64_t)r.ly * r.lz;
  s->nsums = 3 * (int64_t)rxy + (int64_t)r.lz * rxy + rz;
  if (int rc = b->alloc(&s->series, (size_t)std::max<int64_t>(1, nmeas) * 8)) return rc;
  if (int rc = b->alloc(&s->rows_xy, (size_t)std::max<int64_t>(1, rxy * rows * 3))) return rc;
  if (int rc = b->alloc(&s->rows_z, (size_t)std::max<int64_t>(1, rz * rows))) return rc;
  if (int rc = b->alloc(&s->sums, (size_t)std::max<int64_t>(1, s->nsums))) return rc;
  if (int rc = b->alloc(&s->planes, (size_t)std::max<int64_t>(1, (2 * rxy + rz) * (int64_t)r.lz)))
    return rc;
  HIP_TRY(hipMemsetAsync(s->sums, 0, sizeof(double) * (size_t)std::max<int64_t>(1, s->nsums),
                         b->stream));
  launch_clear_counters(b->g, b->stream);
  launch_sweep_end(b->g, 0, 0, b->stream);
  HIP_TRY(hipGetLastError());
  *out = s.release();
  return SLSF_OK;
}

int driver_sweeps(DriverSession* s, int nsweeps) {
  after_per_phase(s->b);
  for (int k = 0; k < nsweeps; ++k)
    if (int rc = hbm_sweep(s->b, s->wmax)) return rc;
  return SLSF_OK;
}

int driver_measure(DriverSession* s, int64_t index) {
  if (index < 0 || index >= s->nmeas) return fail(SLSF_E_INVALID, "measurement index out of range");
  slsf_batch* b = s->b;
  if (is_fp64(b)) launch_observables<double>(b->g, b->stream);
  else launch_observables<float>(b->g, b->stream);
  launch_fold_observables(b->g, 0, s->series + index * 8, b->stream);
  HIP_TRY(hipGetLastError());
  return SLSF_OK;
}

int driver_correlations(DriverSession* s) {
  if (s->rxy + s->rz == 0) return SLSF_OK;
  slsf_batch* b = s->b;
  if (int rc = ensure_trig_cache(b)) return rc;  // the row dot products read the cache
  const int64_t rows = (int64_t)b->reps[0].ly * b->reps[0].lz;
  if (is_fp64(b)) launch_corr<double>(b->g, 0, rows, s->rxy, s->rz, s->rows_xy, s->rows_z, b->stream);
  else launch_corr<float>(b->g, 0, rows, s->rxy, s->rz, s->rows_xy, s->rows_z, b->stream);
  launch_fold_correlations(b->g, 0, s->rows_xy, s->rows_z, s->rxy, s->rz, s->planes, s->sums,
                           b->stream);
  HIP_TRY(hipGetLastError());
  return SLSF_OK;
}

// sweeps per launch of the one-workgroup replay: about 10-20 ms of work at
// L = 16, so a long run is still cut into launches the host can interleave
constexpr int64_t REPLAY_CHUNK = 1024;

int driver_run(DriverSession* s, const DriverSchedule& plan, int64_t* ncorr) {
  if (plan.thermal < 0 || plan.measure_sweeps < 0 || plan.measure_interval <= 0 ||
      plan.correlation_interval <= 0)
    return fail(SLSF_E_INVALID, "invalid driver schedule");
  slsf_batch* b = s->b;
  const int64_t nmeas = plan.measure_sweeps / plan.measure_interval;
  if (nmeas > s->nmeas) return fail(SLSF_E_INVALID, "more measurements than the session holds");
  const bool corr = plan.corr && s->rxy + s->rz > 0;
  const auto& r = b->reps_host[0];
  if (is_fp64(b) && b->reps.size() == 1 && replay_resident_fits(r.lx, r.ly, r.lz) &&
      replay_resident_fits_corr(r.lz, s->rxy, s->rz)) {
    ReplayRun run;
    run.measure_interval = plan.measure_interval;
    run.correlation_interval = plan.correlation_interval;
    run.rxy = s->rxy;
    run.rz = s->rz;
    run.corr = corr ? 1 : 0;
    run.series = s->series;
    run.rows_xy = s->rows_xy;
    run.rows_z = s->rows_z;
    run.planes = s->planes;
    run.sums = s->sums;
    const int64_t total = plan.thermal + plan.measure_sweeps;
    for (int64_t done = 0; done < total; done += REPLAY_CHUNK) {
      run.nsweeps = std::min(REPLAY_CHUNK, total - done);
      run.step0 = done - plan.thermal + 1;
      launch_replay_resident(b->g, 0, run, b->stream);
      HIP_TRY(hipGetLastError());
    }
    *ncorr = corr ? (nmeas + plan.correlation_interval - 1) / plan.correlation_interval : 0;
    return SLSF_OK;
  }
  if (int rc = driver_sweeps(s, (int)plan.thermal)) return rc;
  int64_t idx = 0, nc = 0;
  for (int64_t step = 1; step <= plan.measure_sweeps; ++step) {
    if (int rc = driver_sweeps(s, 1)) return rc;
    if (step % plan.measure_interval != 0) continue;
    if (int rc = driver_measure(s, idx)) return rc;
    if (corr && idx % plan.correlation_interval == 0) {
      if (int rc = driver_correlations(s)) return rc;
      ++nc;
    }
    ++idx;
  }
  *ncorr = nc;
  return SLSF_OK;
}

int driver_close(DriverSession* s, int64_t nsweeps, double* acceptance_sum, double* series,
                 double* sums) {
  slsf_batch* b = s->b;
  if (series && s->nmeas > 0)
    HIP_TRY(hipMemcpyAsync(series, s->series, sizeof(double) * (size_t)s->nmeas * 8,
                           hipMemcpyDeviceToHost, b->stream));
  if (sums && s->nsums > 0)
    HIP_TRY(hipMemcpyAsync(sums, s->sums, sizeof(double) * (size_t)s->nsums,
                           hipMemcpyDeviceToHost, b->stream));
  if (int rc = gather_acceptance(b, nsweeps, acceptance_sum)) return rc;  // synchronises
  return gather_clusters(b, clusters_in(b, nsweeps), nullptr);
}

void driver_free(DriverSession* s) { delete s; }
int64_t driver_sums_size(const DriverSession* s) { return s->nsums; }

}  // namespace slsf
