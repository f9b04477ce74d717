// slsf_device.h -- primitives shared by every kernel of the library.
//
//  * Philox4x32-10 counter-based RNG (Salmon et al., SC'11).  Every uniform of
//    the Monte Carlo is keyed by (replica seed, counter): no stream state lives
//    on the device, results do not depend on how replicas are batched, and a
//    checkpoint is just theta plus the sweep counters.
//  * FP64 sincos/exp from Cody-Waite range reduction and Taylor polynomials,
//    constants derived from exact rational arithmetic (pi by Machin's formula,
//    ln 2 by its binary series, 1/n! correctly rounded).  The FP64 kernels are
//    compiled with -ffp-contract=off so these routines give the same bits on the
//    GPU as the same expression tree gives on an x86 host.
//  * Counter layout of the keyed uniforms:
//        ctr = {index, counter, tag, 0}, key = {seed lo32, seed hi32}
//      tag INIT : index = site, counter = 0          -> theta = 2 pi u53(w0,w1)
//      tag METRO: index = site, counter = sweep      -> proposal u(w0,w1), accept u(w2,w3)
//      tag WSEED: index = 0,    counter = cluster    -> phi u(w0,w1), seed site (w2*V)>>32
//      tag WBOND: index = base site of bond (base, base+e_a), counter = cluster
//                                                    -> bond uniform from word a
//      tags METRO_GROUP + c, c < 3 (FP32 resident kernel, production stream):
//        index = the first of up to eight same-colour sites z, z+2, ... of a
//        column (site k of the colour's L/2 with k % 8 == 0), counter = sweep;
//        the words of the calls (three for eight sites, two for four), read as
//        consecutive 24-bit fields, are (proposal, accept) of each site in turn
//      tags WBOND_GROUP + c, c < NC (FP32 resident kernel, production stream):
//        index = site (x, y, z) with z % G == 0, counter = cluster, G = 16, 8
//        or 4 sites of a column (the largest that divides L) and NC = 9, 5 or 3
//        calls; the words of the calls, in order, read as consecutive 24-bit
//        fields (most significant bits first), are the +x, +y, +z bond
//        uniforms of the sites z .. z+G-1
//    FP64 uniforms use 53 bits (two words), FP32 uniforms the top 24 bits of the
//    first word of the pair; bond uniforms are w*2^-32 (FP64) or (w>>8)*2^-24
//    (a packed field f: f*2^-24).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SLSF_HD __host__ __device__ __forceinline__
#define SLSF_D __device__ __forceinline__

namespace slsf {

constexpr uint32_t TAG_INIT = 1;
constexpr uint32_t TAG_METRO = 2;
constexpr uint32_t TAG_WSEED = 3;
constexpr uint32_t TAG_WBOND = 4;
constexpr uint32_t TAG_WBOND_GROUP = 16;  // .. 24: up to nine calls per sixteen sites
constexpr uint32_t TAG_METRO_GROUP = 32;  // .. 34: three calls per eight sites

constexpr double TWO_PI = 3.141592653589793 * 2.0;  // src/types.rs:6
constexpr double PI = 3.141592653589793;
constexpr float TWO_PI_F = 6.28318530717958647692f;
constexpr float PI_F = 3.14159265358979323846f;
constexpr float INV_TWO_PI_F = 0.15915494309189533577f;
constexpr float LOG2E_F = 1.44269504088896340736f;

struct Words {
  uint32_t w0, w1, w2, w3;
};

SLSF_HD Words philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                            uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
#if defined(__HIP_DEVICE_COMPILE__)
    // one three-input XOR (v_bitop3_b32, truth table 0x96) per output word
    c0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
    c2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
#else
    c0 = hi1 ^ c1 ^ k0;
    c2 = hi0 ^ c3 ^ k1;
#endif
    c1 = lo1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return Words{c0, c1, c2, c3};
}

SLSF_HD uint32_t word(const Words& w, int k) {
  return k == 0 ? w.w0 : k == 1 ? w.w1 : k == 2 ? w.w2 : w.w3;
}

SLSF_HD Words keyed(uint32_t k0, uint32_t k1, uint32_t index, uint32_t counter,
                    uint32_t tag) {
  return philox4x32_10(index, counter, tag, 0u, k0, k1);
}

SLSF_HD double u53(uint32_t hi, uint32_t lo) {
  const uint64_t bits = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
  return (double)bits * 0x1.0p-53;
}
SLSF_HD float u24(uint32_t w) { return (float)(w >> 8) * 0x1.0p-24f; }

// ---- FP64 math ------------------------------------------------------------
namespace fp64 {
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double PIO2_1 = 0x1.921fb54400000p+0;  // 33 significant bits
constexpr double PIO2_2 = 0x1.0b4611a600000p-34; // 33 significant bits
constexpr double PIO2_3 = 0x1.3198a2e037073p-69;
constexpr double LOG2E = 0x1.71547652b82fep+0;
constexpr double LN2_HI = 0x1.62e42ff000000p-1;  // 32 significant bits
constexpr double LN2_LO = -0x1.718432a1b0e26p-35;
// 1/n! rounded to nearest double, n = 0..18
constexpr double F0 = 1.0, F1 = 0x1.0000000000000p+0, F2 = 0x1.0000000000000p-1,
                 F3 = 0x1.5555555555555p-3, F4 = 0x1.5555555555555p-5,
                 F5 = 0x1.1111111111111p-7, F6 = 0x1.6c16c16c16c17p-10,
                 F7 = 0x1.a01a01a01a01ap-13, F8 = 0x1.a01a01a01a01ap-16,
                 F9 = 0x1.71de3a556c734p-19, F10 = 0x1.27e4fb7789f5cp-22,
                 F11 = 0x1.ae64567f544e4p-26, F12 = 0x1.1eed8eff8d898p-29,
                 F13 = 0x1.6124613a86d09p-33, F14 = 0x1.93974a8c07c9dp-37,
                 F15 = 0x1.ae7f3e733b81fp-41, F16 = 0x1.ae7f3e733b81fp-45,
                 F17 = 0x1.952c77030ad4ap-49, F18 = 0x1.6827863b97d97p-53;
}  // namespace fp64

// sin and cos of x: k = nearest multiple of pi/2, r = x - k pi/2 in three
// parts, Taylor series of sin to r^17 and cos to r^18 on |r| <= pi/4.
SLSF_HD void dsincos(double x, double* s_out, double* c_out) {
  using namespace fp64;
  const double kf = floor(x * TWO_OVER_PI + 0.5);
  const double r = ((x - kf * PIO2_1) - kf * PIO2_2) - kf * PIO2_3;
  const double r2 = r * r;
  double ps = F17;
  ps = -F15 + r2 * ps;
  ps = F13 + r2 * ps;
  ps = -F11 + r2 * ps;
  ps = F9 + r2 * ps;
  ps = -F7 + r2 * ps;
  ps = F5 + r2 * ps;
  ps = -F3 + r2 * ps;
  const double s = r + (r * r2) * ps;
  double pc = -F18;
  pc = F16 + r2 * pc;
  pc = -F14 + r2 * pc;
  pc = F12 + r2 * pc;
  pc = -F10 + r2 * pc;
  pc = F8 + r2 * pc;
  pc = -F6 + r2 * pc;
  pc = F4 + r2 * pc;
  pc = -F2 + r2 * pc;
  const double c = 1.0 + r2 * pc;
  const int q = (int)((int64_t)kf & 3);
  if (q == 0) { *s_out = s; *c_out = c; }
  else if (q == 1) { *s_out = c; *c_out = -s; }
  else if (q == 2) { *s_out = -s; *c_out = -c; }
  else { *s_out = -c; *c_out = s; }
}

// exp(x): k = nearest integer to x/ln2, r = x - k ln2 in two parts, Taylor to
// r^13, scaled by 2^k built from its exponent bits.  x < -708 returns 0.
SLSF_HD double dexp(double x) {
  using namespace fp64;
  if (x > 709.0) return __builtin_inf();
  if (x < -708.0) return 0.0;
  const double kf = floor(x * LOG2E + 0.5);
  const double r = (x - kf * LN2_HI) - kf * LN2_LO;
  double p = F13;
  p = F12 + r * p;
  p = F11 + r * p;
  p = F10 + r * p;
  p = F9 + r * p;
  p = F8 + r * p;
  p = F7 + r * p;
  p = F6 + r * p;
  p = F5 + r * p;
  p = F4 + r * p;
  p = F3 + r * p;
  p = F2 + r * p;
  p = F1 + r * p;
  p = F0 + r * p;
  const uint64_t bits = (uint64_t)((int64_t)kf + 1023) << 52;
  return p * __builtin_bit_cast(double, bits);
}

SLSF_HD double wrap_angle(double t) {  // src/types.rs:335-343
  if (t >= TWO_PI) return t - TWO_PI;
  if (t < 0.0) return t + TWO_PI;
  return t;
}
SLSF_HD float wrap_angle(float t) {
  if (t >= TWO_PI_F) return t - TWO_PI_F;
  if (t < 0.0f) return t + TWO_PI_F;
  return t;
}

// ---- exact division by a lattice extent -----------------------------------
// floor(n / d) for 32-bit n and d as a multiply-high by m = ceil(2^64 / d)
// (m = 0 encodes d = 1).  With n = q d + r, n m / 2^64 = n / d + n e / (d 2^64)
// for some 0 <= e < d, and the excess stays below 1 - r/d because n < 2^32 and
// d < 2^32, so the floor is q exactly.
SLSF_HD uint64_t divisor_magic(uint32_t d) {
  return d <= 1u ? 0ull : (~0ull) / d + 1ull;  // floor((2^64 - 1) / d) + 1 = ceil(2^64 / d)
}
SLSF_HD uint32_t fast_div(uint32_t n, uint64_t m) {
  if (m == 0ull) return n;
  const uint64_t lo = (uint64_t)n * (uint32_t)m;
  const uint64_t hi = (uint64_t)n * (uint32_t)(m >> 32);
  return (uint32_t)((hi + (lo >> 32)) >> 32);
}

// ---- per-replica descriptor shared by the global-memory kernels ------------
struct RepDev {
  int32_t lx, ly, lz, wolff_steps;
  int64_t site_off;    // first site of this replica in the concatenated arrays
  int64_t volume;
  int32_t jz_off;      // first layer of this replica in the coupling arrays
  uint32_t key0, key1;
  uint32_t metro_ctr;  // Metropolis sweeps done (counter of TAG_METRO)
  uint32_t wolff_ctr;  // Wolff clusters done (counter of TAG_WSEED/TAG_WBOND)
  uint32_t measure_from;  // sweeps with counter >= this are measured
  uint32_t stop_at;       // the replica does no sweep once metro_ctr reaches this
  uint32_t active;        // HBM path: this replica takes part in the current sweep
  int32_t binsize;        // measurements per bin
  int32_t pending;        // slab path: the state after the last sweep is still to be measured
  double beta, width;
  uint64_t div_lx, div_hx, div_plane;  // divisor_magic of lx, lx/2, lx*ly
  uint64_t pad1[2];
};
static_assert(sizeof(RepDev) == 128, "RepDev layout");

// fills the divisor fields from lx, ly (host side, after the sizes are set)
inline void set_divisors(RepDev& d) {
  d.div_lx = divisor_magic((uint32_t)d.lx);
  d.div_hx = divisor_magic((uint32_t)(d.lx / 2));
  d.div_plane = divisor_magic((uint32_t)d.lx * (uint32_t)d.ly);
}

// Per-replica tallies (accepted moves, cluster sites) are spread over
// TALLY_SLOTS counters so that the blocks of one launch do not all update one
// address; a tally is the sum of its replica's slots.
constexpr int TALLY_SLOTS = 64;

// block partition of the fused observables: blocks never straddle planes; a
// function of the replica's own sizes only, so its FP64 sums do not depend on
// the batch.
// A block of OBS_THREADS threads sums one chunk of a plane, obs_per_thread
// sites per thread: 4 for planes of 1025 to 8192 sites (more blocks for the
// small single-replica lattices, which are latency-bound: 4 instead of 8 gave
// +2-3% on one L=64 replica) and 8 otherwise (with block sums from wave
// totals, 8 instead of 32 on the larger planes: four L=96 replicas 128 -> 122
// us a sweep, four L=128 unchanged).  The partition depends on the replica's
// extents only.
constexpr int OBS_THREADS = 256;
SLSF_HD int obs_per_thread(int lx, int ly) {
  return lx * ly > 8192 ? 8 : (lx * ly > 1024 ? 4 : 8);
}
SLSF_HD int obs_chunks_per_plane(int lx, int ly) {
  const int chunk = OBS_THREADS * obs_per_thread(lx, ly);
  return (lx * ly + chunk - 1) / chunk;
}
SLSF_HD int obs_blocks_of(int lx, int ly, int lz) { return lz * obs_chunks_per_plane(lx, ly); }

// k_metro: METRO_SITES sites of one colour a thread, 256 threads a block, the
// blocks of a plane's colour half numbered from 0.  The per-phase sweeps'
// measurement rides on the colour-0 pass, whose FP32 threads take
// 2 METRO_SITES sites each (in groups of METRO_SITES), so that a block
// reduction of the measurement serves twice the sites (FP64 threads, whose
// registers are full already, keep METRO_SITES); one partial per block of
// that pass (its count, like the partition, depends on the replica's extents
// and the precision only)
constexpr int METRO_SITES = 2;
template <typename R>
constexpr int metro_measure_sites() { return sizeof(R) == 4 ? 2 * METRO_SITES : METRO_SITES; }
SLSF_HD int metro_blocks_per_plane(int lx, int ly, int ms) {
  return ((lx / 2) * ly + 256 * ms - 1) / (256 * ms);
}
SLSF_HD int metro_blocks_of(int lx, int ly, int lz, int ms) {
  return lz * metro_blocks_per_plane(lx, ly, ms);
}
// partials a replica needs for either measurement of the per-phase sweeps
SLSF_HD int measure_blocks_of(int lx, int ly, int lz) {
  const int a = obs_blocks_of(lx, ly, lz), b = metro_blocks_of(lx, ly, lz, METRO_SITES);
  return a > b ? a : b;
}

struct WolffDev {  // per replica, per cluster step
  double phi, sphi, cphi;
  int64_t seed_site;
  uint32_t step;   // cluster counter used for the keyed uniforms
  int32_t active;  // 0 when this replica does no cluster in this round
  int32_t swept;   // slab path: the replica's Metropolis sweep ran in this sweep's k_slab
  int32_t pad;
};


// ---- wave sums of doubles in a fixed order ---------------------------------
// A DPP lane move of a double (both 32-bit halves) under control CTRL.
template <int CTRL>
SLSF_D double dpp_f64(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, 0xF, 0xF, true);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// v + v[l ^ 16] and v + v[l ^ 32] in every lane l, from gfx950's row swaps:
// v_permlane16_swap(a, b) leaves rows (a0, b0, a2, b2) in its first result
// and (a1, b1, a3, b3) in its second (measured), v_permlane32_swap the two
// halves likewise, so with a = b = v the two results hold a lane's own row and
// its partner row, and their sum is the pair sum.  VALU moves instead of two
// LDS round trips (ds_bpermute) per 32-bit half.
SLSF_D double swap_sum16_f64(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
  const auto pl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto ph = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  return __builtin_bit_cast(double, ((uint64_t)ph[0] << 32) | pl[0]) +
         __builtin_bit_cast(double, ((uint64_t)ph[1] << 32) | pl[1]);
}
SLSF_D double swap_sum32_f64(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
  const auto pl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto ph = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __builtin_bit_cast(double, ((uint64_t)ph[0] << 32) | pl[0]) +
         __builtin_bit_cast(double, ((uint64_t)ph[1] << 32) | pl[1]);
}
// Sum over the 64 lanes of a wave, every lane receiving the total: pairs,
// quads, half rows and rows of 16 through DPP lane moves (each partial formed
// from the same two operands on both lanes that hold it), then the row sums
// r0 + r1 and (r0 + r1) + (r2 + r3) through row swaps (IEEE addition commutes,
// so every lane gets the one value).  The order of wave_sum_f64, bit for bit,
// without its scalar reads: no scalar registers, for kernels whose occupancy
// their SGPR count sets.
SLSF_D double wave_total_f64(double v) {
  v = v + dpp_f64<0xB1>(v);
  v = v + dpp_f64<0x4E>(v);
  v = v + dpp_f64<0x141>(v);
  v = v + dpp_f64<0x140>(v);
  return swap_sum32_f64(swap_sum16_f64(v));
}

}  // namespace slsf
