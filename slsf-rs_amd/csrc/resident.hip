// resident.hip -- the FP32 batched kernel: each replica lives in one workgroup
// for a whole launch of many sweeps (Metropolis + Wolff + measurement).
//
// Mapping (cubic L in {8, 12, 16, 24, 32}).  Thread t owns the z-column
// (x, y) = (t % RW, t / RW), RW = next power of two >= L (lanes x >= L idle).
// A site's state is its angle in turns (revolutions, [0, 1)) in one FP32
// register; the hardware v_sin/v_cos take turns, so a Metropolis update costs
// the sin/cos of the old and the proposed angle and one exp2 for the
// Boltzmann factor, with the neighbours' (cos, sin) read from LDS.  Couplings
// J_xy[z], J_z[z], J_z[z-1] sit in an LDS table.
//
// Metropolis, colour p (src/updates.rs:440-486): all threads write the
// (cos, sin) of their opposite-colour sites into LDS slot idx/2 (every pair of
// x-neighbours holds one site of each colour), barrier, then every thread
// updates its colour-p sites reading all six neighbours' (cos, sin) from LDS
// (the two z-neighbours are its own column's sites, exchanged with the rest).
//
// Wolff (src/updates.rs:551-620): r_i = cos(theta_i - phi) for every site goes
// to LDS, then every site decides its three canonical bonds (+x, +y, +z) with
// its keyed uniforms; the decisions become LDS bitmaps, one 32-bit word per
// (y, z) row (wave ballots).  The cluster is the seed's component of that bond
// graph -- the same set the reference's DFS builds from the same uniforms --
// found by a monotone flood fill over rows (x-runs closed inside the word by
// carry arithmetic, y and z through the words 1, 2 and 4 rows away along
// straight bonded lines) until a round changes nothing.  Members are reflected
// in registers.
//
// Measurement (src/observables.rs:80-154): taken in the colour-0 Metropolis
// pass of the next sweep, which already reads every colour-0 site's six
// neighbours and computes its (cos, sin) before moving it: each colour-0 site
// sums its six bonds (each bond has exactly one colour-0 end), all sites sum
// (cos, sin); FP32 per thread, FP64 across the workgroup, accumulated into
// FP64 bins on the device.
//
// Uniforms are the library's keyed Philox stream (same counters as the FP64
// replay path; FP32 uses the top 24 bits of the words), so a replica's chain
// does not depend on the batch it runs in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>

#include "../../include/slsf.h"
#include "common.h"
#include "resident.h"

namespace slsf {

template <int L, int CPT>
struct RShape {
  static constexpr int RW = L <= 8 ? 8 : (L <= 16 ? 16 : 32);
  static constexpr int YS = L / CPT;  // y stride between a thread's columns
  static constexpr int NT = RW * YS;
  static constexpr int V = L * L * L;
  static constexpr int ROWS = L * L;
  static constexpr int NW = NT / 64;
  static constexpr int RPW = 64 / RW;  // y values per wave
  static constexpr uint32_t MASK = (L == 32) ? 0xffffffffu : ((1u << L) - 1u);
  static_assert(NT % 64 == 0, "workgroup must be whole waves");
  static_assert((L / 2) % 2 == 0, "pairs of same-colour sites per column");
  static_assert(YS % 2 == 0, "a thread's columns share their colour parity");
  static_assert(4 * ROWS <= V, "the fill's four path-mask arrays fit in the exchange buffer");
};

struct ResArgs {
  RepDev* reps;
  const int32_t* index;  // batch replica index per workgroup
  float* theta;
  const float* jxy;
  const float* jz;
  double* accum;
  double* bins;
  int32_t* bin_state;
  unsigned long long* acc_out;
  long long* csize_out;
  int32_t bin_cap, nsweeps;
};


// Rotation (right by s < L) and reversal of an L-bit periodic row held in the
// low bits of a word (L = 32: v_alignbit_b32 and v_bfrev_b32).
template <int L>
SLSF_D uint32_t rotr_row(uint32_t w, uint32_t s) {
  if (L == 32) return __builtin_amdgcn_alignbit(w, w, s);
  return ((w >> s) | (w << (L - s))) & RShape<L, 1>::MASK;
}
template <int L>
SLSF_D uint32_t rev_row(uint32_t w) {
  return __builtin_bitreverse32(w) >> (32 - L);
}

// Runs of a periodic row under its active x-bonds b (bit i: sites i and i+1
// mod L bonded), prepared once per cluster for run_closure.  The row is
// rotated right by s so that an inactive bond sits at position L-1 (with
// every bond active, bond L-1 is cut: the chain 0..L-1 is still one run);
// br are the rotated bonds, eu their run tops (~br), and bd/ed the same for
// the bit-reversed row, which turns downward propagation into upward.
struct RowRuns {
  uint32_t s, back, br, eu, bd, ed;  // back = (L - s) % L undoes the rotation
};
template <int L>
SLSF_D RowRuns row_runs(uint32_t b) {
  constexpr uint32_t M = RShape<L, 1>::MASK;
  const uint32_t open = ~b & M;
  const uint32_t j = open ? (uint32_t)__builtin_ctz(open) : (uint32_t)(L - 1);
  RowRuns r;
  r.s = (j + 1 == (uint32_t)L) ? 0u : j + 1;
  r.back = r.s ? (uint32_t)L - r.s : 0u;
  r.br = rotr_row<L>(b, r.s) & (M >> 1);
  r.eu = ~r.br & M;
  r.bd = rev_row<L>(r.br) >> 1;
  r.ed = ~r.bd & M;
  return r;
}
// Closure of the marked sites n through the row's runs.  Upward, with the
// seeds q & br (seeds at run tops move nowhere up): eu - seeds borrows from
// each seed up to the top of its run and stops there (every run, the last
// included, ends at a bit of eu), so (eu - seeds) ^ eu is every site from a
// run's lowest seed to its top; the reversed row does the same downward.
template <int L>
SLSF_D uint32_t run_closure(uint32_t n, const RowRuns& r) {
  uint32_t q = rotr_row<L>(n, r.s);
  q |= (r.eu - (q & r.br)) ^ r.eu;
  uint32_t v = rev_row<L>(q);
  v |= (r.ed - (v & r.bd)) ^ r.ed;
  return rotr_row<L>(rev_row<L>(v), r.back);
}


// An empty asm makes a value opaque to the optimizer.  The exchange and the
// Metropolis step read their angles through opq(): otherwise the compiler
// rewrites cos(acc ? t1 : t0) as acc ? cos(t1) : cos(t0) and keeps the
// Metropolis step's sines and cosines alive from one phase to the next, which
// the register budget of the 1024-thread L=32 workgroup cannot hold.  (The
// Wolff projection cos(theta - phi) shares nothing with them and reads the
// angles directly, saving the copy an opq() of a live value costs.)
SLSF_D float opq(float v) {
  asm("" : "+v"(v));
  return v;
}

// An LDS pointer made opaque, kept in the LDS address space: a loop over
// compile-time offsets from it compiles to ds_* instructions with immediate
// offsets (up to 64 KiB from the base) and no per-access address arithmetic.
typedef __attribute__((address_space(3))) float lds_f32;
SLSF_D lds_f32* lds_base(float* p) {
  lds_f32* q = (lds_f32*)p;
  asm("" : "+v"(q));
  return q;
}

// The 24-bit integer of a uniform word as a float (exact): u24(w) = w24(w) 2^-24.
// Comparisons and affine maps of u are rewritten on w24 where the scaling by
// a power of two stays exact, which saves the multiply and keeps every value.
SLSF_D float w24(uint32_t w) { return (float)(w >> 8); }

// Bond uniforms of the production stream: G sites of a column take their 3G
// uniforms as consecutive 24-bit fields (most significant bits first) of the
// words of NC keyed calls, w0..w3 of call 0, then call 1, ...: 16 sites from
// 9 calls where L allows, 8 from 5, 4 from 3.
template <int L>
struct BondGroup {
  static constexpr int G = (L % 16 == 0) ? 16 : ((L % 8 == 0) ? 8 : 4);
  static constexpr int NC = (G * 72 + 127) / 128;
  // the first site whose uniforms reach into call c's first word (bit 128 c):
  // field f ends at bit 24 f + 23, so f >= floor(16 c / 3)
  static constexpr int first_site(int c) { return (16 * c / 3) / 3; }
};
// 24-bit field f of the words of NC calls (f compile-time after unrolling)
template <int NC>
SLSF_D uint32_t field24(const Words (&w)[NC], int f) {
  const int bit = 24 * f, wi = bit >> 5, off = bit & 31;
  const uint32_t a = word(w[wi >> 2], wi & 3);
  if (off == 0) return a >> 8;
  if (off == 8) return a & 0xFFFFFFu;
  const uint32_t b = word(w[(wi + 1) >> 2], (wi + 1) & 3);
  // a field across two words is three whole bytes of them: one v_perm_b32
  // (selector bytes 0-3 pick bytes of b, 4-7 bytes of a, 0x0C a zero byte).
  // off 16: (a.byte1, a.byte0, b.byte3); off 24: (a.byte0, b.byte3, b.byte2)
  return __builtin_amdgcn_perm(a, b, off == 16 ? 0x0C050403u : 0x0C040302u);
}

// hardware transcendentals: v_sin/v_cos take revolutions, v_exp is 2^x
SLSF_D float sin_t(float t) { return __builtin_amdgcn_sinf(t); }
SLSF_D float cos_t(float t) { return __builtin_amdgcn_cosf(t); }
SLSF_D float exp2_hw(float x) { return __builtin_amdgcn_exp2f(x); }
SLSF_D float fract_hw(float x) { return __builtin_amdgcn_fractf(x); }

// sum of N doubles over the workgroup; result valid in thread 0
template <int N, int NW>
SLSF_D void wg_sum(double (&v)[N], double* red) {
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) red[wave * N + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) {
      double s = 0.0;
      for (int w = 0; w < NW; ++w) s += red[w * N + k];
      v[k] = s;
    }
}

// FP32 partial sums of one measurement over a thread's colour-0 sites: every
// bond has exactly one colour-0 end, so each colour-0 site (cos, sin) = (c, s)
// adds its six bonds: cos(a - b) = ca cb + sa sb summed per direction, and
// sin along +e, sin(t_i - t_{i+e}) + sin(t_{i-e} - t_i) = s (c+ - c-) -
// c (s+ - s-); the magnetization takes the site and its colour-1 z+ neighbour.
struct MeasPart {
  float ec[3], es[3], mx, my;
};
// Sums of a site's neighbour pairs shared by its local field and its
// measurement: (cos, sin) of the two x neighbours and of the two y
// neighbours, and the coupled z field, each with explicit fmas so that every
// caller rounds them alike.
struct PairSums {
  float sxc, sxs, syc, sys, hzc, hzs;
};
SLSF_D PairSums pair_sums(float2 nxp, float2 nxm, float2 nyp, float2 nym, float2 nzp,
                          float2 nzm, float4 jj) {
  return {nxp.x + nxm.x,
          nxp.y + nxm.y,
          nyp.x + nym.x,
          nyp.y + nym.y,
          __builtin_fmaf(jj.y, nzp.x, jj.z * nzm.x),
          __builtin_fmaf(jj.y, nzp.y, jj.z * nzm.y)};
}
SLSF_D void measure_site(MeasPart& m, float c, float s, const PairSums& q, float2 nxp,
                         float2 nxm, float2 nyp, float2 nym, float2 nzp, float2 nzm,
                         float4 jj) {
  m.ec[0] = __builtin_fmaf(jj.x, __builtin_fmaf(c, q.sxc, s * q.sxs), m.ec[0]);
  m.ec[1] = __builtin_fmaf(jj.x, __builtin_fmaf(c, q.syc, s * q.sys), m.ec[1]);
  m.ec[2] = __builtin_fmaf(s, q.hzs, __builtin_fmaf(c, q.hzc, m.ec[2]));
  m.es[0] = __builtin_fmaf(jj.x, __builtin_fmaf(s, nxp.x - nxm.x, -(c * (nxp.y - nxm.y))),
                           m.es[0]);
  m.es[1] = __builtin_fmaf(jj.x, __builtin_fmaf(s, nyp.x - nym.x, -(c * (nyp.y - nym.y))),
                           m.es[1]);
  const float dzc = __builtin_fmaf(jj.y, nzp.x, -(jj.z * nzm.x));
  const float dzs = __builtin_fmaf(jj.y, nzp.y, -(jj.z * nzm.y));
  m.es[2] = __builtin_fmaf(s, dzc, __builtin_fmaf(-c, dzs, m.es[2]));
  m.mx += c + nzp.x;
  m.my += s + nzp.y;
}

// PAIRED (production stream): the (proposal, accept) uniforms of up to eight
// same-colour sites of a column are consecutive 24-bit fields of the words of
// up to three keyed calls (tags METRO_GROUP + c, counter index = the group's
// first site); otherwise one call per site with the replay path's counters
// (tag METRO), which lets an FP32 run consume exactly the uniforms the FP64
// replay path uses.
// The workgroup's LDS.  The row bitmaps and the coupling table come first, so
// their offsets stay below 64 KiB and fold into the immediate offsets of the
// LDS instructions; the exchange buffer is addressed from per-thread bases.
template <int L, int CPT>
struct ResLds {
  using S = RShape<L, CPT>;
  uint32_t bxm[S::ROWS], bym[S::ROWS], bzm[S::ROWS], vis[S::ROWS];
  float4 jt[L];  // {J_xy[z], J_z[z], J_z[z-1], 0}
  double red[S::NW * 9];
  double acc[13];   // the replica's open-bin accumulators for the launch
  int32_t bst[2];   // and its bin counters (sweeps in the open bin, bins done)
  int flags[3];
  float2 xbuf[S::V / 2];  // opposite-colour (cos, sin); Wolff: r[V], then path masks
};

template <int L, int CPT, bool PAIRED>
__global__ void __launch_bounds__((RShape<L, CPT>::NT)) k_resident(ResArgs a) {
  using S = RShape<L, CPT>;
  __shared__ ResLds<L, CPT> sm;
  float2* const xbuf = sm.xbuf;
  uint32_t* const bxm = sm.bxm;
  uint32_t* const bym = sm.bym;
  uint32_t* const bzm = sm.bzm;
  uint32_t* const vis = sm.vis;
  float4* const jt = sm.jt;
  double* const red = sm.red;
  int* const flags = sm.flags;
  float* rbuf = reinterpret_cast<float*>(xbuf);

  const int rep = a.index[blockIdx.x];
  const RepDev d = a.reps[rep];
  const int t = threadIdx.x;
  int x = t % S::RW, y = t / S::RW;  // column j of this thread is (x, y + j * YS)
  const bool own = x < L;
  const int64_t base = d.site_off;
  const float beta_l2e = (float)(d.beta * 1.4426950408889634);
  const float wturn = (float)(d.width * 0.15915494309189535);  // width in turns
  const uint32_t k0 = d.key0, k1 = d.key1;
  uint32_t mctr = d.metro_ctr, wctr = d.wolff_ctr;

  // Angles of this thread's columns in turns, [0, 1) (the FP32 state
  // format), by colour: th[j][c * L/2 + k] is the colour-c site of column j,
  // at z = 2k + (c ^ zpar) with zpar = (x + y) & 1 (one parity for all of a
  // thread's columns, YS being even), so every colour phase names its sites
  // at compile time.  Coordinates are made opaque (empty asm) at the start of
  // every phase, which keeps per-site address arithmetic from being hoisted
  // out of the sweep loop into hundreds of live registers.
  constexpr int H = L / 2;
  float th[CPT][L];
  {
    const int zpar = (x + y) & 1;
#pragma unroll
    for (int j = 0; j < CPT; ++j)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < H; ++k) {
          const int z = 2 * k + (c ^ zpar);
          th[j][c * H + k] = own ? a.theta[base + x + L * (y + j * S::YS + L * z)] : 0.0f;
        }
  }
  if (t < L) {
    const int zm = (t == 0) ? L - 1 : t - 1;
    jt[t] = make_float4(a.jxy[d.jz_off + t], a.jz[d.jz_off + t], a.jz[d.jz_off + zm], 0.0f);
  }
  if (t < 3) flags[t] = 0;
  // the replica's measurement state comes in for the launch; a launch with
  // no measurement state (the autotuner's trials, which only thermalize and
  // pass no arrays) neither reads nor writes it
  const bool has_bins = a.accum != nullptr && a.bin_state != nullptr;
  if (has_bins && t < 13) sm.acc[t] = a.accum[(int64_t)rep * 13 + t];
  if (has_bins && t < 2) sm.bst[t] = a.bin_state[2 * rep + t];
  __syncthreads();

  uint32_t nacc_wave = 0;  // wave-uniform: accepted proposals of this wave
  long long ncluster = 0;  // thread 0: cluster sites

  // LDS slot of site (x, y, z) is (x + L y + L^2 z) / 2 = (x + L y) / 2 + HL z
  // (L^2 z is even): a per-thread base plus a compile-time plane offset
  constexpr int HL = L * L / 2;

  // write (cos, sin) of this thread's sites of colour `col` to LDS slot idx/2
  auto exchange = [&](int col) __attribute__((always_inline)) {
    asm volatile("" : "+v"(x), "+v"(y));
    const int z1 = (col + x + y) & 1;
    if (!own) return;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int b = ((x + L * (y + j * S::YS)) >> 1) + HL * z1;
#pragma unroll
      for (int k = 0; k < H; ++k) {
        const float tv = opq(th[j][col * H + k]);
        xbuf[b + 2 * HL * k] = make_float2(cos_t(tv), sin_t(tv));
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  };

  // Metropolis update of this thread's colour-p sites; with `meas` (colour 0
  // only) the same pass also takes the measurement of the state it starts
  // from, sharing the neighbour reads and the sites' (cos, sin)
  MeasPart part;
  auto update = [&](int p, bool meas) __attribute__((always_inline)) {
    asm volatile("" : "+v"(x), "+v"(y));
    const int y0 = y;
    const int z0 = (p + x + y) & 1;
    if (!own) return;
    const int xp = (x + 1 == L) ? 0 : x + 1, xm = (x == 0) ? L - 1 : x - 1;
    const float4* jz0 = jt + z0;
    Words w{0u, 0u, 0u, 0u};
    Words wm[3];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
    const int y = y0 + j * S::YS;
    const int yp = (y + 1 == L) ? 0 : y + 1, ym = (y == 0) ? L - 1 : y - 1;
    // slots of the first site (z = z0) of the column and of its neighbours
    const int s0 = HL * z0;
    const int bxp = ((xp + L * y) >> 1) + s0, bxm = ((xm + L * y) >> 1) + s0;
    const int byp = ((x + L * yp) >> 1) + s0, bym = ((x + L * ym) >> 1) + s0;
    const int bz = ((x + L * y) >> 1) + s0;
#pragma unroll
    for (int k = 0; k < L / 2; ++k) {
      const float t0 = opq(th[j][p * H + k]);
      const int zo = 2 * HL * k;  // plane offset of z = 2k + z0
      // z neighbours, periodic: z = L - 1 (last k, z0 = 1) wraps up to 0,
      // z = 0 (first k, z0 = 0) wraps down to L - 1
      const int dzp = (k == L / 2 - 1) ? (z0 ? -HL * (L - 1) : HL) : HL;
      const int dzm = (k == 0) ? (z0 ? -HL : HL * (L - 1)) : -HL;
      const float2 nxp = xbuf[bxp + zo];
      const float2 nxm = xbuf[bxm + zo];
      const float2 nyp = xbuf[byp + zo];
      const float2 nym = xbuf[bym + zo];
      const float2 nzp = xbuf[bz + zo + dzp];
      const float2 nzm = xbuf[bz + zo + dzm];
      const float4 jj = jz0[2 * k];
      const int row = L * (y + L * (2 * k + z0));
      const PairSums q = pair_sums(nxp, nxm, nyp, nym, nzp, nzm, jj);
      const float hx = __builtin_fmaf(jj.x, q.sxc + q.syc, q.hzc);
      const float hy = __builtin_fmaf(jj.x, q.sxs + q.sys, q.hzs);
      float wp, ua;  // wp: the proposal's 24-bit integer as a float
      if (PAIRED) {
        // group of sites k0g .. k0g + 7 (fewer at the column's end)
        const int m = k % 8, k0g = k - m;  // compile-time once the loop is unrolled
        const int n = (L / 2 - k0g) < 8 ? (L / 2 - k0g) : 8;
        const int nc = (48 * n + 127) / 128;
        const int grow = L * (y + L * (2 * k0g + z0));
#pragma unroll
        for (int c = 0; c < nc; ++c)  // each call just before the first site it serves
          if ((16 * c / 3) / 2 == m) wm[c] = keyed(k0, k1, (uint32_t)(x + grow), mctr,
                                                   TAG_METRO_GROUP + c);
        wp = (float)field24<3>(wm, 2 * m);
        ua = (float)field24<3>(wm, 2 * m + 1) * 0x1.0p-24f;
      } else {
        w = keyed(k0, k1, (uint32_t)(x + row), mctr, TAG_METRO);
        wp = w24(w.w0);
        ua = u24(w.w2);
      }
      // 2u - 1 with u = wp 2^-24, in one exact fma
      // explicit fmas here and below: an expression left for the compiler to
      // contract may be fused differently in different instances of this
      // kernel (columns per thread), and a replica's chain must not depend on
      // which instance runs it
      const float t1 = fract_hw(__builtin_fmaf(wturn, __builtin_fmaf(wp, 0x1.0p-23f, -1.0f), t0));
      const float c0 = cos_t(t0), s0 = sin_t(t0), c1 = cos_t(t1), s1 = sin_t(t1);
      if (meas) measure_site(part, c0, s0, q, nxp, nxm, nyp, nym, nzp, nzm, jj);
      const float de = __builtin_fmaf(c0 - c1, hx, (s0 - s1) * hy);
      const float pacc = exp2_hw(-de * beta_l2e);  // evaluated always: no branch
      // accept iff de <= 0 or u < exp(-beta de) (src/updates.rs:255-264).  With
      // u < 1 the second test alone decides both: v_exp_f32 returns at least 1
      // for every argument >= 0 (checked over all such floats on the device).
      // The ballot of the one compare is its lane mask: a scalar count per wave.
      const bool acc = ua < pacc;
      nacc_wave += __popcll(__builtin_amdgcn_ballot_w64(acc));
      th[j][p * H + k] = acc ? t1 : t0;
      // one site in flight per thread: keeps the register budget of the
      // 1024-thread L=32 workgroup (occupancy supplies the latency hiding)
      __builtin_amdgcn_sched_barrier(0);
    }
    }
  };

  // folds the partial sums of one measurement into the replica's FP64 bins:
  // wave sums (DPP lane moves and row swaps, no LDS round trips), then wave 0
  // finishes with one lane per quantity -- lane k < 8 adds quantity k over
  // the waves in wave order, lane k < 13 accumulates series k -- against the
  // accumulators held in LDS for the launch, so the other waves never wait on
  // a global memory round trip
  auto measure_finish = [&]() __attribute__((always_inline)) {
    // quantities: mx, my, cos_x..z, sin_x..z (the energy is -(cos_x + cos_y +
    // cos_z), formed from their FP64 totals)
    double v[8] = {(double)part.mx,    (double)part.my,    (double)part.ec[0], (double)part.ec[1],
                   (double)part.ec[2], (double)part.es[0], (double)part.es[1], (double)part.es[2]};
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = wave_total_f64(v[k]);
    if ((t & 63) == 0)
#pragma unroll
      for (int k = 0; k < 8; ++k) red[(t >> 6) * 8 + k] = v[k];
    __syncthreads();
    if (t < 64) {
      double s = 0.0;
      if (t < 8)
        for (int w = 0; w < S::NW; ++w) s += red[w * 8 + t];
      const double s0 = __shfl(s, 0, 64), s1 = __shfl(s, 1, 64);
      const double c0 = __shfl(s, 2, 64), c1 = __shfl(s, 3, 64), c2 = __shfl(s, 4, 64);
      // series: e, e^2, m^2, m^4, cos_x..z, sin_x..z, sin_x^2..sin_z^2
      const double sk = __shfl(s, t < 10 ? t - 2 : t - 5, 64);
      const double vol = (double)S::V;
      const double e = -(c0 + c1 + c2) / vol;
      const double m2 = __builtin_fma(s0, s0, s1 * s1) / (vol * vol);
      const double val = t == 0   ? e
                         : t == 1 ? e * e
                         : t == 2 ? m2
                         : t == 3 ? m2 * m2
                         : t < 10 ? sk
                                  : sk * sk;
      const int cnt = sm.bst[0] + 1, done = sm.bst[1];
      const bool full = cnt == d.binsize;
      if (t < 13) {
        const double acc = sm.acc[t] + val;
        if (full && done < a.bin_cap)
          a.bins[((int64_t)rep * 13 + t) * a.bin_cap + done] = acc / (double)d.binsize;
        sm.acc[t] = full ? 0.0 : acc;
      }
      if (t == 0) {
        sm.bst[0] = full ? 0 : cnt;
        sm.bst[1] = full ? done + 1 : done;
      }
    }
  };

  // a measurement on its own (the last sweep of a launch); needs the colour-1
  // exchange in LDS and sums exactly what the fused pass sums
  auto measure = [&]() __attribute__((always_inline)) {
    asm volatile("" : "+v"(x), "+v"(y));
    part = MeasPart{};
    if (own) {
      const int z0 = (x + y) & 1;  // colour-0 sites of the column: z = 2k + z0
      const int xp = (x + 1 == L) ? 0 : x + 1, xm = (x == 0) ? L - 1 : x - 1;
      const float4* jz0 = jt + z0;
      const int y0 = y;
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
      const int y = y0 + j * S::YS;
      const int yp = (y + 1 == L) ? 0 : y + 1, ym = (y == 0) ? L - 1 : y - 1;
      const int s0 = HL * z0;
      const int bxp = ((xp + L * y) >> 1) + s0, bxm = ((xm + L * y) >> 1) + s0;
      const int byp = ((x + L * yp) >> 1) + s0, bym = ((x + L * ym) >> 1) + s0;
      const int bz = ((x + L * y) >> 1) + s0;
#pragma unroll
      for (int k = 0; k < L / 2; ++k) {
        const float t0 = opq(th[j][k]);
        const int zo = 2 * HL * k;
        const float2 nxp = xbuf[bxp + zo];
        const float2 nxm = xbuf[bxm + zo];
        const float2 nyp = xbuf[byp + zo];
        const float2 nym = xbuf[bym + zo];
        // colour-1 neighbours in z, read back from the exchange (periodic)
        const int dzp = (k == L / 2 - 1) ? (z0 ? -HL * (L - 1) : HL) : HL;
        const int dzm = (k == 0) ? (z0 ? -HL : HL * (L - 1)) : -HL;
        const float2 nzp = xbuf[bz + zo + dzp];
        const float2 nzm = xbuf[bz + zo + dzm];
        const float4 jj = jz0[2 * k];
        measure_site(part, cos_t(t0), sin_t(t0), pair_sums(nxp, nxm, nyp, nym, nzp, nzm, jj),
                     nxp, nxm, nyp, nym, nzp, nzm, jj);
        __builtin_amdgcn_sched_barrier(0);
      }
      }
    }
    measure_finish();
  };

  auto wolff = [&]() __attribute__((always_inline)) {
    asm volatile("" : "+v"(x), "+v"(y));
    // the thread index made opaque too: the row loops below then compute their
    // addresses here instead of keeping (or spilling) them across the sweep loop
    int tw = t;
    asm volatile("" : "+v"(tw));
    const Words ws = keyed(k0, k1, 0u, wctr, TAG_WSEED);
    const float phit = u24(ws.w0);  // phi in turns
    const int seed = (int)(((uint64_t)ws.w2 * (uint64_t)S::V) >> 32);
    // r_i = cos(theta_i - phi) (src/updates.rs:579-582), stored scaled by
    // sqrt(2 beta log2 e), so that a bond's exponent -2 beta J r_i r_j (in
    // base 2) is the product of J and the two stored values
    const float rsc = __builtin_sqrtf(2.0f * beta_l2e);
    if (own) {
      const int zpar = (x + y) & 1;
#pragma unroll
      for (int j = 0; j < CPT; ++j)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          // the column's colour-c sites z = 2k + (c ^ zpar) lie two planes
          // apart: from two bases (k < H/2, k >= H/2) every offset fits the
          // 16-bit immediate of ds_write
          float* col = rbuf + (x + L * (y + j * S::YS + L * (c ^ zpar)));
          lds_f32* r0 = lds_base(col);
          lds_f32* r1 = lds_base(col + (H / 2) * 2 * L * L);
#pragma unroll
          for (int k = 0; k < H; ++k) {
            const float v = cos_t(th[j][c * H + k] - phit) * rsc;
            if (k < H / 2) r0[2 * k * L * L] = v;
            else r1[2 * (k - H / 2) * L * L] = v;
            __builtin_amdgcn_sched_barrier(0);
          }
        }
    }
    __syncthreads();
    const int wave = tw >> 6, lane = tw & 63;
    const int xp = (x + 1 == L) ? 0 : x + 1;
    // Bond decisions read r from LDS only, so this loop stays rolled; it takes
    // G sites (z .. z+G-1) of a column per step.  Production stream: the 3G
    // bond uniforms of the G sites are consecutive 24-bit fields of the words
    // of BondGroup<G>::NC keyed calls (tags WBOND_GROUP + c, index = the first
    // site), each call made just before the first site that reaches into it;
    // replay stream: one call per site (tag WBOND, words 0..2), as in the
    // FP64 path.
    using BG = BondGroup<L>;
#pragma unroll 1
    for (int jq = 0; jq < CPT * L; jq += BG::G) {
      const int j = jq / L, zq = jq % L;
      const int yj = y + j * S::YS;
      const int yp = (yj + 1 == L) ? 0 : yj + 1;
      constexpr int PL = L * L;  // index step of one plane
      // Per-thread bases of the group, made opaque once: every site of the
      // group reads its r values, couplings and row words at compile-time
      // offsets from them.  (Left to itself the compiler merges zq + q into one
      // index and rebuilds every address of every site.)  bzl is the +z
      // neighbour of the group's last site, which wraps at the lattice's top.
      int b0 = x + L * (yj + L * zq), bx = xp + L * (yj + L * zq), by = x + L * (yp + L * zq);
      int bzl = (zq + BG::G == L) ? x + L * yj : b0 + BG::G * PL;
      int bj = zq, brow = wave * S::RPW + lane + j * S::YS + L * zq;
      asm volatile("" : "+v"(b0), "+v"(bx), "+v"(by), "+v"(bzl), "+v"(bj), "+v"(brow));
      const uint32_t i0 = (uint32_t)b0;
      Words wq[BG::NC];
#pragma unroll
      for (int q = 0; q < BG::G; ++q) {
        if (PAIRED && own) {
#pragma unroll
          for (int c = 0; c < BG::NC; ++c)
            if (BG::first_site(c) == q) wq[c] = keyed(k0, k1, i0, wctr, TAG_WBOND_GROUP + c);
        }
        unsigned long long gx = 0, gy = 0, gz = 0;  // bond ballots of the wave
        if (own) {
          const int i = b0 + q * PL;  // site (x, yj, zq + q)
          const float ri = rbuf[i];
          const float rxp = rbuf[bx + q * PL];
          const float ryp = rbuf[by + q * PL];
          const float rzp = (q + 1 < BG::G) ? rbuf[i + PL] : rbuf[bzl];
          const float4 jj = jt[bj + q];
          float wv[3];  // the bond uniforms' 24-bit integers as floats
          if (PAIRED) {
#pragma unroll
            for (int a = 0; a < 3; ++a) wv[a] = (float)field24<BG::NC>(wq, 3 * q + a);
          } else {
            const Words w = keyed(k0, k1, (uint32_t)i, wctr, TAG_WBOND);
            wv[0] = w24(w.w0);
            wv[1] = w24(w.w1);
            wv[2] = w24(w.w2);
          }
          // 2 beta log2(e) J r_i r_j of the three bonds (r scaled as stored)
          const float qxy = jj.x * ri, qz = jj.y * ri;
          const float px = qxy * rxp, py = qxy * ryp, pz = qz * rzp;
          // active iff p > 0 and u < 1 - exp(-2 beta p), as u 2^24 < (1 - e) 2^24
          // with the right side in one fma.  One compare decides both: for
          // p <= 0, e >= 1 makes the right side <= 0 <= u 2^24.  Its ballot is
          // the compare's lane mask (lanes outside the lattice add no bits).
          auto active = [&](float wu, float pb) {
            const float e = exp2_hw(-pb);
            return __builtin_amdgcn_ballot_w64(wu < __builtin_fmaf(-e, 0x1.0p24f, 0x1.0p24f));
          };
          gx = active(wv[0], px);
          gy = active(wv[1], py);
          gz = active(wv[2], pz);
        }
        if (lane < S::RPW) {
          const int row = brow + L * q;
          const int sh = lane * S::RW;
          bxm[row] = (uint32_t)(gx >> sh) & S::MASK;
          bym[row] = (uint32_t)(gy >> sh) & S::MASK;
          bzm[row] = (uint32_t)(gz >> sh) & S::MASK;
        }
      }
    }
    for (int r = tw; r < S::ROWS; r += S::NT) vis[r] = (r == seed / L) ? (1u << (seed % L)) : 0u;
    __syncthreads();
    // Straight-line path masks, in the exchange buffer (free during the fill):
    // bit x of py2[r] says the y-bonds of column x from row r to row r + 2 (same
    // z) are both active, py4 the same over four rows; pz2/pz4 along z.  A round
    // then moves a mark up to four rows in a straight line, which cuts the
    // number of rounds without changing the fixed point.
    uint32_t* py2 = reinterpret_cast<uint32_t*>(xbuf);
    uint32_t* pz2 = py2 + S::ROWS;
    uint32_t* py4 = pz2 + S::ROWS;
    uint32_t* pz4 = py4 + S::ROWS;
    for (int r = tw; r < S::ROWS; r += S::NT) {
      const int ry = r % L, rz = r / L;
      py2[r] = bym[r] & bym[((ry + 1) % L) + L * rz];
      pz2[r] = bzm[r] & bzm[ry + L * ((rz + 1) % L)];
    }
    __syncthreads();
    for (int r = tw; r < S::ROWS; r += S::NT) {
      const int ry = r % L, rz = r / L;
      py4[r] = py2[r] & py2[((ry + 2) % L) + L * rz];
      pz4[r] = pz2[r] & pz2[ry + L * ((rz + 2) % L)];
    }
    __syncthreads();
    // Everything about a row that the rounds leave unchanged stays in
    // registers for the whole fill: its own marks (only this thread writes
    // them), its runs, and for the 12 rows a mark can come from (1, 2 and 4
    // rows away in y and in z, both ways) that row's index and the bond or
    // path mask that lets the mark through.
    constexpr int RPT = (S::ROWS + S::NT - 1) / S::NT;
    constexpr int NB = 12;
    RowRuns runs[RPT];
    const uint32_t* nb_row[RPT][NB];
    uint32_t nb_mask[RPT][NB], mine[RPT];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = tw + i * S::NT;
      const bool valid = r < S::ROWS;
      const int rr = valid ? r : 0;
      const int ry = rr % L, rz = rr / L;
      const int ym = ((ry + L - 1) % L) + L * rz, yp = ((ry + 1) % L) + L * rz;
      const int zm = ry + L * ((rz + L - 1) % L), zp = ry + L * ((rz + 1) % L);
      const int y2m = ((ry + L - 2) % L) + L * rz, y2p = ((ry + 2) % L) + L * rz;
      const int z2m = ry + L * ((rz + L - 2) % L), z2p = ry + L * ((rz + 2) % L);
      const int y4m = ((ry + L - 4) % L) + L * rz, y4p = ((ry + 4) % L) + L * rz;
      const int z4m = ry + L * ((rz + L - 4) % L), z4p = ry + L * ((rz + 4) % L);
      const int rows[NB] = {ym, yp, zm, zp, y2m, y2p, z2m, z2p, y4m, y4p, z4m, z4p};
      const uint32_t masks[NB] = {bym[ym], bym[rr], bzm[zm], bzm[rr], py2[y2m], py2[rr],
                                  pz2[z2m], pz2[rr], py4[y4m], py4[rr], pz4[z4m], pz4[rr]};
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        nb_row[i][b] = vis + rows[b];
        nb_mask[i][b] = valid ? masks[b] : 0u;
      }
      runs[i] = row_runs<L>(valid ? bxm[rr] : 0u);
      mine[i] = valid ? vis[rr] : 0u;
    }
    // flood fill to the fixed point; the flag of round r lives in slot r % 3,
    // which thread 0 clears only in round r + 2, after every thread read it
    for (int round = 0;; ++round) {
      bool changed = false;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int r = tw + i * S::NT;
        if (RPT * S::NT > S::ROWS && r >= S::ROWS) break;
        uint32_t n = mine[i];
        for (int b = 0; b < NB; ++b) n |= *nb_row[i][b] & nb_mask[i][b];
        n = run_closure<L>(n, runs[i]);
        if (n != mine[i]) {
          vis[r] = n;
          mine[i] = n;
          changed = true;
        }
      }
      if (changed) flags[round % 3] = 1;
      if (t == 0) flags[(round + 1) % 3] = 0;
      __syncthreads();
      if (flags[round % 3] == 0) break;
    }
    // reflection theta -> 2 phi + pi - theta (src/updates.rs:503-506), in turns
    if (own) {
      const float refl = 2.0f * phit + 0.5f;
      const int zpar = (x + y) & 1;
#pragma unroll
      for (int j = 0; j < CPT; ++j)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int k = 0; k < H; ++k) {
            const int z = 2 * k + (c ^ zpar);
            const bool in = (vis[y + j * S::YS + L * z] >> x) & 1u;
            float& v = th[j][c * H + k];
            v = in ? fract_hw(refl - v) : v;
            __builtin_amdgcn_sched_barrier(0);
          }
    }
    int members = 0;
    for (int r = tw; r < S::ROWS; r += S::NT) members += __popc(vis[r]);
    ncluster += members;
    wctr += 1u;
  };

  // the replica's own window decides how many sweeps it does and which are
  // measured (uniform per workgroup); a measurement is taken at the next
  // sweep's first exchange, or after the loop for the last sweep
  bool pending = false;
  for (int sw = 0; sw < a.nsweeps && mctr < d.stop_at; ++sw) {
    exchange(1);
    __syncthreads();
    if (pending) {
      part = MeasPart{};
      update(0, true);
      measure_finish();
    } else {
      update(0, false);
    }
    __syncthreads();
    exchange(0);
    __syncthreads();
    update(1, false);
    __syncthreads();
    pending = mctr >= d.measure_from;
    mctr += 1u;
    for (int k = 0; k < d.wolff_steps; ++k) {
      wolff();
      __syncthreads();
    }
  }
  if (pending) {
    exchange(1);
    __syncthreads();
    measure();
  }

  // write back the angles (turns), counters, tallies.  The (cos, sin) cache in
  // HBM is left behind the angles: the library refreshes it (with the same
  // v_sin/v_cos) where a caller reads it, which a launch of sweeps does not
  if (own) {
    const int zpar = (x + y) & 1;
#pragma unroll
    for (int j = 0; j < CPT; ++j)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < H; ++k) {
          const int z = 2 * k + (c ^ zpar);
          const int64_t g = base + x + L * (y + j * S::YS + L * z);
          a.theta[g] = th[j][c * H + k];
        }
  }
  if (has_bins && t < 13) a.accum[(int64_t)rep * 13 + t] = sm.acc[t];
  if (has_bins && t < 2) a.bin_state[2 * rep + t] = sm.bst[t];
  double tallies[2] = {(t & 63) == 0 ? (double)nacc_wave : 0.0, (double)ncluster};
  __syncthreads();
  wg_sum<2, S::NW>(tallies, red);
  if (t == 0) {
    a.acc_out[(int64_t)rep * TALLY_SLOTS] += (unsigned long long)tallies[0];
    a.csize_out[(int64_t)rep * TALLY_SLOTS] += (long long)tallies[1];
    a.reps[rep].metro_ctr = mctr;
    a.reps[rep].wolff_ctr = wctr;
  }
}

__global__ void k_zero_tallies(unsigned long long* acc, long long* cs, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    acc[i] = 0ull;
    cs[i] = 0ll;
  }
}

bool resident_supports(int L) { return L == 8 || L == 12 || L == 16 || L == 24 || L == 32; }
static int default_cpt(int L);

int resident_plan(const std::vector<RepDev>& reps, ResidentPlan& plan) {
  plan.groups.clear();
  plan.covers_all = true;
  std::map<int, std::vector<int32_t>> by_l;
  for (size_t i = 0; i < reps.size(); ++i) {
    const RepDev& r = reps[i];
    if (r.lx == r.ly && r.ly == r.lz && resident_supports(r.lx)) by_l[r.lx].push_back((int32_t)i);
    else plan.covers_all = false;
  }
  for (auto& kv : by_l) {
    ResidentGroup g;
    g.L = kv.first;
    g.replicas = kv.second;
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(int32_t) * g.replicas.size()) != hipSuccess)
      return fail(SLSF_E_DEVICE, "hipMalloc failed for the resident plan");
    if (hipMemcpy(p, g.replicas.data(), sizeof(int32_t) * g.replicas.size(),
                  hipMemcpyHostToDevice) != hipSuccess)
      return fail(SLSF_E_DEVICE, "hipMemcpy failed for the resident plan");
    g.d_index = (int32_t*)p;
    g.cpt = default_cpt(g.L);
    plan.groups.push_back(std::move(g));
  }
  return SLSF_OK;
}

// A launch loads and stores its replicas, ends with a measurement of its own,
// and leaves compute units idle while its last workgroups finish: 1000 sweeps
// a launch make all of that small (+2% at L=32 over 50; the chains and bins do
// not depend on the launch length).
static int default_sweeps_per_launch(int) { return 1000; }
static int default_cpt(int L) { (void)L; return 1; }

template <int L, int CPT>
static void launch_resident(const ResArgs& a, int nblocks, bool paired, hipStream_t s) {
  if (paired) k_resident<L, CPT, true><<<nblocks, RShape<L, CPT>::NT, 0, s>>>(a);
  else k_resident<L, CPT, false><<<nblocks, RShape<L, CPT>::NT, 0, s>>>(a);
}

template <int L>
static void launch_cpt(int cpt, const ResArgs& a, int nblocks, bool paired, hipStream_t s) {
  if (cpt == 2) launch_resident<L, 2>(a, nblocks, paired, s);
  else launch_resident<L, 1>(a, nblocks, paired, s);
}

static void launch_any(int L, int cpt, const ResArgs& a, int nblocks, bool paired,
                       hipStream_t s) {
  switch (L) {
    case 8: launch_resident<8, 1>(a, nblocks, paired, s); break;  // L/2 = 4: one column
    case 12: launch_resident<12, 1>(a, nblocks, paired, s); break;
    case 16: launch_cpt<16>(cpt, a, nblocks, paired, s); break;
    case 24: launch_cpt<24>(cpt, a, nblocks, paired, s); break;
    default: launch_cpt<32>(cpt, a, nblocks, paired, s); break;
  }
}

std::vector<int> resident_cpt_candidates(int L) {
  if (L == 16 || L == 24 || L == 32) return {1, 2};
  return {1};
}

int resident_waves_per_replica(int L, int cpt) {
  const int rw = L <= 8 ? 8 : (L <= 16 ? 16 : 32);
  return rw * (L / cpt) / 64;
}

static ResArgs make_args(const ResidentGroup& grp, const GlobalState& g,
                         const ResidentRunArgs& ra) {
  ResArgs a;
  a.reps = g.reps;
  a.index = grp.d_index;
  a.theta = (float*)g.theta;
  a.jxy = (const float*)g.jxy;
  a.jz = (const float*)g.jz;
  a.accum = ra.accum;
  a.bins = ra.bins;
  a.bin_state = ra.bin_state;
  a.acc_out = ra.acc_out;
  a.csize_out = ra.csize_out;
  a.bin_cap = ra.bin_cap;
  a.nsweeps = 0;
  return a;
}

// Enqueues every launch of one group for the run on stream s; returns the count.
static int64_t enqueue_group(const ResidentGroup& grp, const GlobalState& g,
                             const ResidentRunArgs& ra, hipStream_t s) {
  ResArgs a = make_args(grp, g, ra);
  const int spl = ra.sweeps_per_launch > 0   ? ra.sweeps_per_launch
                  : grp.sweeps_per_launch > 0 ? grp.sweeps_per_launch
                                              : default_sweeps_per_launch(grp.L);
  const int nb = (int)grp.replicas.size();
  int64_t launches = 0;
  int64_t left = ra.thermalization + ra.sweeps;  // the replicas' windows cut the tail
  while (left > 0) {
    a.nsweeps = (int)std::min<int64_t>(left, spl);
    launch_any(grp.L, grp.cpt, a, nb, !ra.replay_stream, s);
    ++launches;
    left -= a.nsweeps;
  }
  return launches;
}

int64_t resident_enqueue(const ResidentGroup& grp, const GlobalState& g,
                         const ResidentRunArgs& ra, hipStream_t s) {
  return enqueue_group(grp, g, ra, s);
}

#define RES_TRY(expr, msg)                                    \
  do {                                                        \
    if ((expr) != hipSuccess) return fail(SLSF_E_DEVICE, msg); \
  } while (0)

int resident_run(ResidentPlan& plan, const GlobalState& g, const ResidentRunArgs& ra,
                 hipStream_t s, ResidentRunResult* out) {
  const size_t ng = plan.groups.size();
  // one stream per extra group so uneven sizes share the GPU (default priority)
  while (plan.streams.size() + 1 < ng) {
    hipStream_t t;
    RES_TRY(hipStreamCreateWithFlags(&t, hipStreamNonBlocking), "hipStreamCreate failed");
    plan.streams.push_back(t);
    hipEvent_t e;
    RES_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate failed");
    plan.events.push_back(e);
  }
  if (!plan.fork) {
    RES_TRY(hipEventCreateWithFlags(&plan.fork, hipEventDisableTiming), "hipEventCreate failed");
    RES_TRY(hipEventCreate(&plan.t0), "hipEventCreate failed");
    RES_TRY(hipEventCreate(&plan.t1), "hipEventCreate failed");
  }
  k_zero_tallies<<<(g.nrep * TALLY_SLOTS + 255) / 256, 256, 0, s>>>(ra.acc_out, ra.csize_out,
                                                                    g.nrep * TALLY_SLOTS);
  RES_TRY(hipEventRecord(plan.t0, s), "hipEventRecord failed");
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (ra.use_graph)
    RES_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "graph capture failed");
  int64_t launches = 0;
  if (ng > 1) RES_TRY(hipEventRecord(plan.fork, s), "hipEventRecord failed");
  for (size_t i = 0; i < ng; ++i) {
    hipStream_t gs = i == 0 ? s : plan.streams[i - 1];
    if (i > 0) RES_TRY(hipStreamWaitEvent(gs, plan.fork, 0), "hipStreamWaitEvent failed");
    launches += enqueue_group(plan.groups[i], g, ra, gs);
  }
  for (size_t i = 1; i < ng; ++i) {
    RES_TRY(hipEventRecord(plan.events[i - 1], plan.streams[i - 1]), "hipEventRecord failed");
    RES_TRY(hipStreamWaitEvent(s, plan.events[i - 1], 0), "hipStreamWaitEvent failed");
  }
  if (ra.use_graph) {
    RES_TRY(hipStreamEndCapture(s, &graph), "graph capture failed");
    RES_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate failed");
    RES_TRY(hipGraphLaunch(exec, s), "graph launch failed");
  }
  RES_TRY(hipEventRecord(plan.t1, s), "hipEventRecord failed");
  if (hipGetLastError() != hipSuccess) return fail(SLSF_E_DEVICE, "resident launch failed");
  const hipError_t e = hipEventSynchronize(plan.t1);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(SLSF_E_DEVICE, std::string("resident kernel failed: ") +
                                                      hipGetErrorString(e));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, plan.t0, plan.t1);
  out->kernel_ms = ms;
  out->launches = launches;
  return SLSF_OK;
}

void resident_release(ResidentPlan& plan) {
  for (auto& grp : plan.groups)
    if (grp.d_index) (void)hipFree(grp.d_index);
  for (hipStream_t t : plan.streams) (void)hipStreamDestroy(t);
  for (hipEvent_t e : plan.events) (void)hipEventDestroy(e);
  if (plan.fork) (void)hipEventDestroy(plan.fork);
  if (plan.t0) (void)hipEventDestroy(plan.t0);
  if (plan.t1) (void)hipEventDestroy(plan.t1);
  plan = ResidentPlan();
}

// Times every columns-per-thread candidate of each group (waves per replica
// L/4 or L/8 at L = 16, 24, 32) on the batch's own replicas without
// measurement, then restores the replicas' angles, trig caches and counters, so
// tuning leaves no trace in the chains.  Every candidate runs the same sweeps in
// launches of the same length, twice in alternation, and keeps its faster time,
// so neither launch overhead nor the order of the trials favours one of them.
// The launch length stays the library's default: longer launches only gain
// (fewer loads and stores of the replicas, fewer tails of idle compute units),
// and the default is already the longest.
int resident_autotune(ResidentPlan& plan, const GlobalState& g, hipStream_t s,
                      int trial_sweeps, std::vector<TuneRecord>& records) {
  records.clear();
  const size_t n = (size_t)g.total_sites;
  float *th = nullptr, *c = nullptr;
  RepDev* rd = nullptr;
  RES_TRY(hipMalloc((void**)&th, n * sizeof(float)), "hipMalloc failed");
  RES_TRY(hipMalloc((void**)&c, n * 2 * sizeof(float)), "hipMalloc failed");
  RES_TRY(hipMalloc((void**)&rd, (size_t)g.nrep * sizeof(RepDev)), "hipMalloc failed");
  auto copy_all = [&](bool save) -> hipError_t {
    hipError_t e = hipSuccess;
    auto cp = [&](void* dst, const void* src, size_t b) {
      if (e == hipSuccess) e = hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToDevice, s);
    };
    if (save) {
      cp(th, g.theta, n * 4); cp(c, g.cs, n * 8);
      cp(rd, g.reps, g.nrep * sizeof(RepDev));
    } else {
      cp(g.theta, th, n * 4); cp(g.cs, c, n * 8);
      cp(g.reps, rd, g.nrep * sizeof(RepDev));
    }
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  };
  int rc = copy_all(true) == hipSuccess ? SLSF_OK : fail(SLSF_E_DEVICE, "autotune save failed");
  unsigned long long* acc = nullptr;
  long long* cs = nullptr;
  if (!rc && (hipMalloc((void**)&acc, (size_t)g.nrep * TALLY_SLOTS * 8) != hipSuccess ||
              hipMalloc((void**)&cs, (size_t)g.nrep * TALLY_SLOTS * 8) != hipSuccess))
    rc = fail(SLSF_E_DEVICE, "hipMalloc failed");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
  }
  const int sweeps = std::max(trial_sweeps, 1);
  for (size_t gi = 0; !rc && gi < plan.groups.size(); ++gi) {
    ResidentGroup& grp = plan.groups[gi];
    const std::vector<int> cands = resident_cpt_candidates(grp.L);
    std::vector<float> best_ms(cands.size(), 0.f);
    for (int pass = 0; !rc && pass < 2; ++pass)
      for (size_t k = 0; !rc && k < cands.size(); ++k) {
        ResidentGroup trial = grp;
        trial.cpt = cands[k];
        trial.sweeps_per_launch = sweeps;  // the trial is one launch
        ResidentRunArgs ra;
        ra.thermalization = sweeps;
        ra.acc_out = acc;
        ra.csize_out = cs;
        ResidentRunArgs warm = ra;  // loads the code object and warms the caches
        warm.thermalization = 1;
        enqueue_group(trial, g, warm, s);
        (void)hipEventRecord(e0, s);
        enqueue_group(trial, g, ra, s);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) {
          rc = fail(SLSF_E_DEVICE, "autotune trial failed");
          break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (pass == 0 || ms < best_ms[k]) best_ms[k] = ms;
      }
    if (rc) break;
    size_t best = 0;
    for (size_t k = 0; k < cands.size(); ++k) {
      const double rate = (double)grp.replicas.size() * grp.L * grp.L * grp.L * (double)sweeps /
                          (best_ms[k] * 1e-3);
      records.push_back(TuneRecord{grp.L, cands[k], default_sweeps_per_launch(grp.L),
                                   resident_waves_per_replica(grp.L, cands[k]), rate});
      if (best_ms[k] < best_ms[best]) best = k;
    }
    grp.cpt = cands[best];
    grp.sweeps_per_launch = 0;  // the default launch length
  }
  if (!rc && copy_all(false) != hipSuccess) rc = fail(SLSF_E_DEVICE, "autotune restore failed");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(th); (void)hipFree(c); (void)hipFree(rd);
  if (acc) (void)hipFree(acc);
  if (cs) (void)hipFree(cs);
  return rc;
}

}  // namespace slsf
