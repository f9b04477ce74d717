// server.cpp -- continuous batching of independent Monte Carlo jobs.
//
// One pool per lattice size L holds `cap` slots of FP32 state in HBM (angles in
// turns, trig cache, couplings, descriptors, FP64 accumulators and bins).  A job
// is admitted into a free slot of its size (the pool is created on first use if
// the device-memory budget allows), runs in chunks of sweeps together with every
// other active slot of the pool -- the register/LDS-resident kernel for
// L <= 32, the HBM kernels above -- and retires when its counter reaches its
// window end.  Pools are driven on separate streams, so jobs of different sizes
// overlap on the device.  Per-replica keyed uniforms and per-replica reductions
// make a job's bins independent of the mix it ran in.
//
// A job that asks for correlation series is sampled on the reference's
// cadence (after the sweeps of the measurements m with m % interval == 0):
// a pool's chunk ends at the earliest such sweep of its active slots, and
// after it k_corr_planes forms the sampled slots' plane sums, which the host
// folds and bins per job.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/slsf.h"
#include "common.h"
#include "launch.h"
#include "resident.h"

using namespace slsf;

namespace {

constexpr int32_t kBinCap = 256;  // bins per slot

#define SRV_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(SLSF_E_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + \
                                     " at " #expr);                                   \
  } while (0)

int64_t padded_volume(int L) { return ((int64_t)L * L * L + 63) / 64 * 64; }

// device bytes of one slot: angles + trig cache (12 B/site), couplings,
// descriptor, accumulators and bins; the HBM path adds its Wolff scratch
// (bond masks 1 B, union-find labels 4 B, membership bits), block partials,
// and the slab sweeps' second angle buffer (4 B) and quarter-tile partials
int64_t slot_bytes(int L) {
  const int64_t vp = padded_volume(L);
  int64_t b = vp * 12 + 2 * L * 4 + (int64_t)sizeof(RepDev) + 13 * 8 + 2 * 4 +
              13 * 8 * (int64_t)kBinCap + 2 * 8 * TALLY_SLOTS + 4;
  if (!resident_supports(L))
    b += vp * 5 + vp / 8 + (int64_t)measure_blocks_of(L, L, L) * 9 * 8 + (int64_t)sizeof(WolffDev) +
         vp * 4 + (int64_t)((L + 15) / 16) * ((L + 15) / 16) * ((L + 3) / 4) * 9 * 8;
  return b;
}

struct SplitMix64 {
  uint64_t state;
  uint64_t next() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double f64() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// plane sums of one correlation sample: 2 rxy l + rz doubles with rxy, rz <= l/2
int64_t corr_plane_doubles(int L) { return (int64_t)L * L + L; }
// device bytes a pool of size L holds besides its slots (the plane sums)
int64_t pool_fixed_bytes(int L) {
  return std::max<int64_t>(corr_plane_doubles(L) * (int64_t)sizeof(double), 256);
}

struct Job {
  int64_t id;
  slsf_job_request req;
  int32_t binsize;
  double submit_s, start_s = -1.0;
  int slot = -1;
  // correlation series: radii, cadence, the sweep count of the next sample
  // (0: none left), the open bin's sums and the complete bins' means
  int32_t rxy = 0, rz = 0, interval = 1;
  int64_t next_sample = 0;
  int32_t in_bin = 0;
  std::vector<double> bin_sum, bin_means;
  bool corr() const { return rxy > 0 || rz > 0; }
  int nseries() const { return 3 * rxy + req.l * rxy + rz; }
};

// Sweep count after which measurement m = k * interval is taken (the first
// such count above `done`), or 0 when the window has no further sample.
int64_t next_sample_after(const Job& j, int64_t done) {
  const int64_t therm = j.req.thermalization;
  const int64_t m = std::max<int64_t>(0, done - therm);  // measurements taken
  const int64_t k = (m + j.interval - 1) / j.interval * j.interval;
  return k < j.req.sweeps ? therm + k + 1 : 0;
}

struct Pool {
  int L = 0;
  bool resident = false;
  int cap = 0;
  int64_t vp = 0;
  GlobalState g;
  double* accum = nullptr;
  double* bins = nullptr;
  int32_t* bin_state = nullptr;
  std::vector<void*> allocs;
  int64_t bytes = 0;
  hipStream_t stream = nullptr;
  std::vector<int64_t> slot_job;  // job id or -1
  std::vector<RepDev> host;       // descriptor mirror
  std::vector<int32_t> active_idx;
  ResidentGroup group;            // active slots of a resident pool
  bool running = false;
  bool cs_stale = false;          // chunks left the trig cache behind the angles
  std::vector<int64_t> done;      // each slot's sweep count after the last chunk
  double* corr_planes = nullptr;  // one correlation sample's plane sums
  std::vector<double> corr_host;

  ~Pool() {
    for (void* p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
  template <typename T>
  int alloc(T** p, size_t count) {
    void* q = nullptr;
    const size_t b = std::max<size_t>(count * sizeof(T), 256);
    SRV_TRY(hipMalloc(&q, b));
    SRV_TRY(hipMemset(q, 0, b));
    allocs.push_back(q);
    bytes += (int64_t)b;
    *p = (T*)q;
    return SLSF_OK;
  }
};

double now_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct slsf_server {
  int64_t budget = 0, used = 0;
  int32_t max_slots = 0, chunk = 0;
  std::chrono::steady_clock::time_point t0;
  std::map<int, std::unique_ptr<Pool>> pools;
  std::deque<Job> queue;
  std::map<int64_t, Job> running;
  std::deque<slsf_job_result> done;
  std::map<int64_t, std::vector<double>> corr_done;  // finished jobs' series, [series][bin]
  int64_t next_id = 1, finished = 0, site_updates = 0;
};

namespace {

int create_pool(slsf_server* s, int L, std::unique_ptr<Pool>& out) {
  const int64_t per = slot_bytes(L);
  const int64_t room = s->budget - s->used - pool_fixed_bytes(L);
  const int cap = (int)std::min<int64_t>(s->max_slots, std::max<int64_t>(room, 0) / per);
  if (cap <= 0) return SLSF_OK;  // not now: the job waits
  auto p = std::make_unique<Pool>();
  p->L = L;
  p->resident = resident_supports(L);
  p->cap = cap;
  p->vp = padded_volume(L);
  SRV_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  GlobalState& g = p->g;
  g.nrep = cap;
  g.total_sites = p->vp * cap;
  RepDev shape;
  std::memset(&shape, 0, sizeof shape);
  shape.lx = shape.ly = shape.lz = L;
  shape.volume = (int64_t)L * L * L;
  set_extents(g, &shape, 1);
  const size_t n = (size_t)g.total_sites;
  int rc = SLSF_OK;
  rc = rc ? rc : p->alloc((RepDev**)&g.reps, (size_t)cap);
  rc = rc ? rc : p->alloc((float**)&g.theta, n);
  rc = rc ? rc : p->alloc((float2**)&g.cs, n);
  rc = rc ? rc : p->alloc((float**)&g.jxy, (size_t)cap * L);
  rc = rc ? rc : p->alloc((float**)&g.jz, (size_t)cap * L);
  rc = rc ? rc : p->alloc(&g.acc, (size_t)cap * TALLY_SLOTS);
  rc = rc ? rc : p->alloc(&g.csize, (size_t)cap * TALLY_SLOTS);
  rc = rc ? rc : p->alloc(&p->accum, (size_t)cap * 13);
  rc = rc ? rc : p->alloc(&p->bin_state, (size_t)cap * 2);
  rc = rc ? rc : p->alloc(&p->bins, (size_t)cap * 13 * kBinCap);
  rc = rc ? rc : p->alloc(&p->group.d_index, (size_t)cap);
  rc = rc ? rc : p->alloc(&p->corr_planes, (size_t)corr_plane_doubles(L));
  if (!p->resident) {
    rc = rc ? rc : p->alloc(&g.bondmask, n);
    rc = rc ? rc : p->alloc(&g.visited, n / 32);
    rc = rc ? rc : p->alloc(&g.label, n);
    rc = rc ? rc : p->alloc(&g.wd, (size_t)cap);
    rc = rc ? rc : p->alloc(&g.partials, (size_t)cap * g.obs_blocks * 9);
    rc = rc ? rc : p->alloc((float**)&g.theta2, n);
    rc = rc ? rc : p->alloc(&g.slab_partials, (size_t)cap * g.max_tiles4 * 9);
  }
  if (rc) return rc;
  p->group.L = L;
  p->group.cpt = 1;
  p->slot_job.assign(cap, -1);
  p->done.assign(cap, 0);
  p->host.assign(cap, RepDev{});
  s->used += p->bytes;
  out = std::move(p);
  return SLSF_OK;
}

// Frees every pool other than size L that has no active slot and returns its
// bytes to the budget, so that a job of a size without a pool is not kept out
// by idle pools of other sizes.  A pool with no active slot has no work in
// flight: its last chunk was waited for when its slots retired.
void release_idle_pools(slsf_server* s, int L) {
  for (auto it = s->pools.begin(); it != s->pools.end();) {
    Pool& p = *it->second;
    const bool idle = std::none_of(p.slot_job.begin(), p.slot_job.end(),
                                   [](int64_t j) { return j >= 0; });
    if (it->first != L && idle) {
      s->used -= p.bytes;
      it = s->pools.erase(it);
    } else {
      ++it;
    }
  }
}

int admit(slsf_server* s, Pool& p, int slot, Job& j) {
  const int L = p.L;
  const slsf_job_request& q = j.req;
  RepDev d;
  std::memset(&d, 0, sizeof d);
  d.lx = d.ly = d.lz = L;
  d.wolff_steps = q.wolff_steps;
  d.site_off = p.vp * slot;
  d.volume = (int64_t)L * L * L;
  d.jz_off = slot * L;
  d.key0 = (uint32_t)q.seed;
  d.key1 = (uint32_t)(q.seed >> 32);
  d.measure_from = (uint32_t)q.thermalization;
  d.stop_at = (uint32_t)(q.thermalization + q.sweeps);
  d.active = 1;
  d.binsize = j.binsize;
  d.beta = 1.0 / q.temperature;
  d.width = q.proposal_width;
  set_divisors(d);
  p.host[slot] = d;
  // couplings: J_xy[z] ~ U[min, max] from SplitMix64(disorder_seed), J_z = j_perp
  std::vector<float> jxy(L), jz(L);
  SplitMix64 rng{q.disorder_seed};
  for (int z = 0; z < L; ++z) {
    jxy[z] = (float)(q.j_xy_min + (q.j_xy_max - q.j_xy_min) * rng.f64());
    jz[z] = (float)q.j_perp;
  }
  SRV_TRY(hipMemcpyAsync(p.g.reps + slot, &p.host[slot], sizeof(RepDev), hipMemcpyHostToDevice,
                         p.stream));
  SRV_TRY(hipMemcpyAsync((float*)p.g.jxy + (size_t)slot * L, jxy.data(), sizeof(float) * L,
                         hipMemcpyHostToDevice, p.stream));
  SRV_TRY(hipMemcpyAsync((float*)p.g.jz + (size_t)slot * L, jz.data(), sizeof(float) * L,
                         hipMemcpyHostToDevice, p.stream));
  SRV_TRY(hipMemsetAsync(p.accum + (size_t)slot * 13, 0, 13 * sizeof(double), p.stream));
  SRV_TRY(hipMemsetAsync(p.bin_state + (size_t)slot * 2, 0, 2 * sizeof(int32_t), p.stream));
  SRV_TRY(hipMemsetAsync(p.g.acc + (size_t)slot * TALLY_SLOTS, 0,
                         sizeof(unsigned long long) * TALLY_SLOTS, p.stream));
  SRV_TRY(hipMemsetAsync(p.g.csize + (size_t)slot * TALLY_SLOTS, 0,
                         sizeof(long long) * TALLY_SLOTS, p.stream));
  GlobalState one = p.g;  // keyed initial angles of this slot only
  one.reps = p.g.reps + slot;
  one.nrep = 1;
  launch_init<float>(one, 0, p.stream);
  SRV_TRY(hipGetLastError());
  // the host vectors above must outlive the async copies
  SRV_TRY(hipStreamSynchronize(p.stream));
  p.slot_job[slot] = j.id;
  p.done[slot] = 0;
  j.slot = slot;
  j.start_s = now_since(s->t0);
  if (j.corr()) {
    j.next_sample = next_sample_after(j, 0);
    j.bin_sum.assign(j.nseries(), 0.0);
  }
  return SLSF_OK;
}

// Sweeps of the pool's next chunk: the server's chunk, ended early at the
// next correlation sample of any active slot.
int64_t chunk_sweeps(slsf_server* s, const Pool& p) {
  int64_t n = s->chunk;
  for (int slot : p.active_idx) {
    auto it = s->running.find(p.slot_job[slot]);
    if (it == s->running.end()) continue;
    const Job& j = it->second;
    if (j.corr() && j.next_sample > 0) n = std::min(n, j.next_sample - p.done[slot]);
  }
  return std::max<int64_t>(n, 1);
}

// One correlation sample of the job in `slot` (its angles are at the sample's
// sweep): plane sums on the device, the fold of slsf_measure_correlations on
// the host, added to the job's open bin.
int sample_correlations(Pool& p, int slot, Job& j) {
  const int L = p.L, rxy = j.rxy, rz = j.rz;
  if (p.cs_stale) {  // chunks left the trig cache behind the angles: this slot's pairs
    GlobalState one = p.g;
    one.reps = p.g.reps + slot;
    one.nrep = 1;
    launch_refresh<float>(one, p.stream);
  }
  launch_corr_planes<float>(p.g, slot, L, L, rxy, rz, p.corr_planes, p.stream);
  SRV_TRY(hipGetLastError());
  const size_t n = (size_t)2 * rxy * L + rz;
  p.corr_host.resize(std::max<size_t>(n, 1));
  SRV_TRY(hipMemcpyAsync(p.corr_host.data(), p.corr_planes, sizeof(double) * n,
                         hipMemcpyDeviceToHost, p.stream));
  SRV_TRY(hipStreamSynchronize(p.stream));
  const double* px = p.corr_host.data();
  const double* py = px + (size_t)rxy * L;
  const double* pz = py + (size_t)rxy * L;
  const double vol = (double)L * L * L, area = (double)L * L;
  double* sum = j.bin_sum.data();
  // series order: CorrX, CorrY, CorrXY, CorrXY_z (layer-major), CorrZ
  for (int rr = 0; rr < rxy; ++rr) {
    double sx = 0.0, sy = 0.0;
    for (int z = 0; z < L; ++z) {
      const double lsx = px[(size_t)rr * L + z], lsy = py[(size_t)rr * L + z];
      sx += lsx;
      sy += lsy;
      sum[3 * rxy + (size_t)z * rxy + rr] += (lsx + lsy) / (2.0 * area);
    }
    sum[rr] += sx / vol;
    sum[rxy + rr] += sy / vol;
    sum[2 * rxy + rr] += (sx / vol + sy / vol) / 2.0;
  }
  for (int rr = 0; rr < rz; ++rr) sum[3 * rxy + (size_t)L * rxy + rr] += pz[rr] / vol;
  if (++j.in_bin == j.binsize) {
    for (double& v : j.bin_sum) {
      j.bin_means.push_back(v / j.binsize);
      v = 0.0;
    }
    j.in_bin = 0;
  }
  return SLSF_OK;
}

int run_chunk(slsf_server* s, Pool& p) {
  p.active_idx.clear();
  for (int i = 0; i < p.cap; ++i)
    if (p.slot_job[i] >= 0) p.active_idx.push_back(i);
  p.running = !p.active_idx.empty();
  if (!p.running) return SLSF_OK;
  const int64_t n = chunk_sweeps(s, p);
  if (p.resident) {
    p.group.replicas = p.active_idx;
    SRV_TRY(hipMemcpyAsync(p.group.d_index, p.active_idx.data(),
                           sizeof(int32_t) * p.active_idx.size(), hipMemcpyHostToDevice,
                           p.stream));
    ResidentRunArgs ra;
    ra.thermalization = n;  // each slot's window cuts its own tail
    ra.sweeps = 0;
    ra.sweeps_per_launch = n;
    ra.accum = p.accum;
    ra.bins = p.bins;
    ra.bin_state = p.bin_state;
    ra.bin_cap = kBinCap;
    ra.acc_out = p.g.acc;
    ra.csize_out = p.g.csize;
    resident_enqueue(p.group, p.g, ra, p.stream);
    p.cs_stale = true;  // the resident kernel writes angles only
  } else {
    int wmax = 0;
    for (int i : p.active_idx) wmax = std::max(wmax, p.host[i].wolff_steps);
    GlobalState g = p.g;  // launch only up to the highest active slot
    g.nrep = p.active_idx.back() + 1;
    // neither FP32 form keeps the trig cache: the correlation measurement
    // refreshes it when it reads it
    p.cs_stale = true;
    launch_sweep_end(g, 0, 0, p.stream);  // marks the slots whose window is open
    BinTarget bt;
    bt.accum = p.accum;
    bt.bins = p.bins;
    bt.bin_state = p.bin_state;
    bt.bin_cap = kBinCap;
    (void)enqueue_hbm_sweeps<float>(g, n, wmax, bt, p.stream);
  }
  SRV_TRY(hipGetLastError());
  return SLSF_OK;
}

int retire(slsf_server* s, Pool& p, int32_t* nfinished) {
  if (!p.running) return SLSF_OK;
  SRV_TRY(hipStreamSynchronize(p.stream));
  std::vector<RepDev> devs(p.cap);
  std::vector<unsigned long long> slots((size_t)p.cap * TALLY_SLOTS), acc(p.cap, 0ull);
  SRV_TRY(hipMemcpy(devs.data(), p.g.reps, sizeof(RepDev) * p.cap, hipMemcpyDeviceToHost));
  SRV_TRY(hipMemcpy(slots.data(), p.g.acc, sizeof(unsigned long long) * slots.size(),
                    hipMemcpyDeviceToHost));
  for (int i = 0; i < p.cap; ++i)
    for (int k = 0; k < TALLY_SLOTS; ++k) acc[i] += slots[(size_t)i * TALLY_SLOTS + k];
  for (int slot : p.active_idx) {
    p.done[slot] = devs[slot].metro_ctr;
    auto it = s->running.find(p.slot_job[slot]);
    if (it == s->running.end() || !it->second.corr()) continue;
    Job& j = it->second;
    if (j.next_sample > 0 && p.done[slot] == j.next_sample) {
      if (int rc = sample_correlations(p, slot, j)) return rc;
      j.next_sample = next_sample_after(j, p.done[slot]);
    }
  }
  const double t = now_since(s->t0);
  for (int slot : p.active_idx) {
    const RepDev& d = devs[slot];
    if (d.metro_ctr < d.stop_at) continue;
    auto it = s->running.find(p.slot_job[slot]);
    if (it == s->running.end()) continue;
    Job& j = it->second;
    int32_t st[2];
    SRV_TRY(hipMemcpy(st, p.bin_state + 2 * slot, sizeof st, hipMemcpyDeviceToHost));
    const int nb = std::min(st[1], kBinCap);
    std::vector<double> bins((size_t)13 * std::max(nb, 1));
    for (int k = 0; k < 13 && nb > 0; ++k)
      SRV_TRY(hipMemcpy(bins.data() + (size_t)k * nb, p.bins + ((size_t)slot * 13 + k) * kBinCap,
                        sizeof(double) * nb, hipMemcpyDeviceToHost));
    slsf_job_result r;
    std::memset(&r, 0, sizeof r);
    r.job_id = j.id;
    r.l = p.L;
    r.nbins = nb;
    r.submit_s = j.submit_s;
    r.start_s = j.start_s;
    r.finish_s = t;
    const int64_t total = j.req.thermalization + j.req.sweeps;
    const double vol = (double)d.volume;
    r.acceptance = total > 0 ? (double)acc[slot] / (vol * (double)total) : 0.0;
    r.site_updates = (int64_t)vol * total;
    if (nb > 0) {
      (void)slsf_evaluate(bins.data(), nb, vol, 1.0 / j.req.temperature, r.value, r.error);
      for (int k = 0; k < 13; ++k) {
        double m = 0.0;
        for (int b = 0; b < nb; ++b) m += bins[(size_t)k * nb + b];
        r.series_mean[k] = m / nb;
      }
    }
    if (j.corr()) {
      const int ns = j.nseries(), cb = (int)(j.bin_means.size() / (size_t)ns);
      r.corr_rxy = j.rxy;
      r.corr_rz = j.rz;
      r.corr_nbins = cb;
      r.corr_nseries = ns;
      std::vector<double> series((size_t)ns * cb);
      for (int b = 0; b < cb; ++b)
        for (int k = 0; k < ns; ++k)
          series[(size_t)k * cb + b] = j.bin_means[(size_t)b * ns + k];
      s->corr_done[j.id] = std::move(series);
    }
    s->site_updates += r.site_updates;
    s->done.push_back(r);
    s->running.erase(it);
    p.slot_job[slot] = -1;
    ++s->finished;
    if (nfinished) ++*nfinished;
  }
  return SLSF_OK;
}

}  // namespace

extern "C" {

int64_t slsf_server_slot_bytes(int32_t l) { return l > 0 ? slot_bytes(l) : 0; }

int slsf_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
  clear_error();
  if (!free_bytes || !total_bytes) return fail(SLSF_E_INVALID, "null argument");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(SLSF_E_DEVICE, "no usable GPU (hipGetDeviceCount found none)");
  size_t f = 0, t = 0;
  SRV_TRY(hipMemGetInfo(&f, &t));
  *free_bytes = (int64_t)f;
  *total_bytes = (int64_t)t;
  return SLSF_OK;
}

int slsf_server_create(int64_t budget, int32_t max_slots, int32_t chunk, slsf_server** out) {
  clear_error();
  if (!out) return fail(SLSF_E_INVALID, "output handle pointer is null");
  *out = nullptr;
  if (budget <= 0 || max_slots <= 0 || chunk <= 0)
    return fail(SLSF_E_INVALID, "budget, slots and chunk must be positive");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(SLSF_E_DEVICE, "no usable GPU (hipGetDeviceCount found none)");
  auto s = std::make_unique<slsf_server>();
  s->budget = budget;
  s->max_slots = max_slots;
  s->chunk = chunk;
  s->t0 = std::chrono::steady_clock::now();
  *out = s.release();
  return SLSF_OK;
}

void slsf_server_destroy(slsf_server* s) { delete s; }

int slsf_server_submit(slsf_server* s, const slsf_job_request* q, int64_t* job_id) {
  clear_error();
  if (!s || !q) return fail(SLSF_E_INVALID, "null argument");
  if (q->l <= 0) return fail(SLSF_E_INVALID, "L_x must be positive");
  if (q->l % 2) return fail(SLSF_E_INVALID, "red-black Metropolis sweep requires even lattice dimensions");
  if (!temperature_ok(q->temperature))
    return fail(SLSF_E_INVALID, "temperature T must be positive and finite");
  if (!width_ok(q->proposal_width))
    return fail(SLSF_E_INVALID, "proposal_width must be nonnegative and finite");
  if (!(q->j_xy_min >= 0.0) || !(q->j_xy_max >= q->j_xy_min) || !(q->j_perp >= 0.0))
    return fail(SLSF_E_INVALID, "theta simulation requires nonnegative layer couplings");
  if (q->wolff_steps < 0 || q->thermalization < 0 || q->sweeps <= 0 || q->binsize <= 0)
    return fail(SLSF_E_INVALID, "sweeps and binsize must be positive");
  if (q->thermalization + q->sweeps >= (int64_t)UINT32_MAX)
    return fail(SLSF_E_INVALID, "a job must have fewer than 2^32 sweeps");
  if ((int64_t)q->l * q->l * q->l >= (int64_t)1 << 31)
    return fail(SLSF_E_INVALID, "lattice volume must be below 2^31");
  if (q->correlation_rmax_xy < 0 || q->correlation_rmax_z < 0)
    return fail(SLSF_E_INVALID, "correlation radii must be nonnegative");
  if (slot_bytes(q->l) + pool_fixed_bytes(q->l) > s->budget)
    return fail(SLSF_E_ADMISSION, "a pool of one L=" + std::to_string(q->l) + " slot needs " +
                                      std::to_string(slot_bytes(q->l) + pool_fixed_bytes(q->l)) +
                                      " bytes, more than the device budget of " +
                                      std::to_string(s->budget));
  Job j;
  j.id = s->next_id++;
  j.req = *q;
  j.binsize = std::max<int32_t>(q->binsize, (int32_t)((q->sweeps + kBinCap - 1) / kBinCap));
  // src/model.rs:360-365
  j.rxy = std::min(q->correlation_rmax_xy, q->l / 2);
  j.rz = std::min(q->correlation_rmax_z, q->l / 2);
  j.interval = std::max(1, q->correlation_interval);
  j.submit_s = now_since(s->t0);
  s->queue.push_back(j);
  if (job_id) *job_id = j.id;
  return SLSF_OK;
}

int slsf_server_step(slsf_server* s, int32_t* finished) {
  clear_error();
  if (!s) return fail(SLSF_E_INVALID, "null argument");
  if (finished) *finished = 0;
  // admission: oldest first; a job whose size has no free slot waits without
  // blocking jobs of other sizes
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    auto pit = s->pools.find(it->req.l);
    if (pit == s->pools.end()) {
      std::unique_ptr<Pool> p;
      if (int rc = create_pool(s, it->req.l, p)) return rc;
      if (!p) {  // no room: give back idle pools of other sizes and try again
        release_idle_pools(s, it->req.l);
        if (int rc = create_pool(s, it->req.l, p)) return rc;
      }
      if (!p) { ++it; continue; }
      pit = s->pools.emplace(it->req.l, std::move(p)).first;
    }
    Pool& pool = *pit->second;
    int slot = -1;
    for (int i = 0; i < pool.cap; ++i)
      if (pool.slot_job[i] < 0) { slot = i; break; }
    if (slot < 0) { ++it; continue; }
    if (int rc = admit(s, pool, slot, *it)) return rc;
    s->running.emplace(it->id, *it);
    it = s->queue.erase(it);
  }
  // one chunk for every pool with active slots, all enqueued before any wait
  for (auto& kv : s->pools)
    if (int rc = run_chunk(s, *kv.second)) return rc;
  for (auto& kv : s->pools)
    if (int rc = retire(s, *kv.second, finished)) return rc;
  return SLSF_OK;
}

int slsf_server_poll(slsf_server* s, slsf_job_result* out, int32_t cap, int32_t* n) {
  clear_error();
  if (!s || !n) return fail(SLSF_E_INVALID, "null argument");
  int32_t k = 0;
  while (k < cap && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *n = k;
  return SLSF_OK;
}

int slsf_server_correlations(slsf_server* s, int64_t job_id, double* out, int64_t cap,
                             int64_t* count) {
  clear_error();
  if (!s || !count) return fail(SLSF_E_INVALID, "null argument");
  auto it = s->corr_done.find(job_id);
  if (it == s->corr_done.end())
    return fail(SLSF_E_INVALID, "job " + std::to_string(job_id) +
                                    " has no correlation series to take (not finished, "
                                    "not measuring correlations, or taken already)");
  *count = (int64_t)it->second.size();
  if (!out) return SLSF_OK;
  if (cap < *count) return fail(SLSF_E_INVALID, "output buffer too small");
  std::copy(it->second.begin(), it->second.end(), out);
  s->corr_done.erase(it);
  return SLSF_OK;
}

int slsf_server_get_stats(const slsf_server* s, slsf_server_stats* out) {
  clear_error();
  if (!s || !out) return fail(SLSF_E_INVALID, "null argument");
  out->queued = (int64_t)s->queue.size();
  out->active = (int64_t)s->running.size();
  out->finished = s->finished;
  out->device_bytes = s->used;
  out->site_updates = s->site_updates;
  out->clock_s = now_since(s->t0);
  return SLSF_OK;
}

}  // extern "C"
