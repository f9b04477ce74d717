/*
 * slsf_oracle.c -- CPU restatement of the SLSF-rs hot path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * file's library, and only as the checker.  Every function cites the reference
 * file:line it restates.  Build: make -C oracle (gcc -O2 -ffp-contract=off).
 */
#include "slsf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_TWO_PI (3.141592653589793 * 2.0) /* src/types.rs:6 */
#define OR_PI 3.141592653589793

static char g_err[256];
const char* or_last_error(void) { return g_err; }
static int fail(const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return 1;
}

/* ===================================================================== */
/* SplitMix64 FastRng -- src/types.rs:7-82                               */
/* ===================================================================== */
#define SPLITMIX64_GAMMA 0x9E3779B97F4A7C15ull

void or_rng_seed(or_fastrng* r, uint64_t seed) { /* seed_from_u64, :48-54 */
  r->seed = seed;
  r->state = seed;
  r->draws = 0;
}

uint64_t or_rng_next_u64(or_fastrng* r) { /* splitmix64_next, :18-24 */
  r->state += SPLITMIX64_GAMMA;
  uint64_t z = r->state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  r->draws += 1;
  return z ^ (z >> 31);
}

uint32_t or_rng_next_u32(or_fastrng* r) { return (uint32_t)or_rng_next_u64(r); } /* :60-62 */

void or_rng_set_position(or_fastrng* r, uint64_t position) { /* :31-36 */
  r->state = r->seed + SPLITMIX64_GAMMA * position;
  r->draws = position;
}

/* rand 0.10 Standard f64: 53 high bits scaled by 2^-53 (third-party; unpinned) */
double or_rng_f64(or_fastrng* r) {
  return (double)(or_rng_next_u64(r) >> 11) * 0x1.0p-53;
}

/* rand Standard bool: sign bit of next_u32 (third-party; unpinned) */
int or_rng_bool(or_fastrng* r) { return (int32_t)or_rng_next_u32(r) < 0; }

/* rand random_range(0..n) for u64/usize: Canon's widening-multiply method with
 * one optional extra word (third-party; unpinned). */
uint64_t or_rng_range(or_fastrng* r, uint64_t n) {
  unsigned __int128 m = (unsigned __int128)or_rng_next_u64(r) * n;
  uint64_t result = (uint64_t)(m >> 64);
  uint64_t lo = (uint64_t)m;
  if (lo > (uint64_t)(0 - n)) {
    unsigned __int128 m2 = (unsigned __int128)or_rng_next_u64(r) * n;
    uint64_t hi2 = (uint64_t)(m2 >> 64);
    uint64_t sum = lo + hi2;
    if (sum < lo) result += 1;
  }
  return result;
}

/* ===================================================================== */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                     */
/* ===================================================================== */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Replay-mode counter layout: ctr = {index, counter, tag, 0}, key = seed. */
enum { TAG_INIT = 1, TAG_METRO = 2, TAG_WSEED = 3, TAG_WBOND = 4 };

static void keyed(uint64_t seed, uint32_t index, uint32_t counter, uint32_t tag,
                  uint32_t w[4]) {
  const uint32_t ctr[4] = {index, counter, tag, 0u};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  or_philox4x32_10(ctr, key, w);
}

static double u53(uint32_t hi, uint32_t lo) {
  uint64_t bits = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
  return (double)bits * 0x1.0p-53;
}

/* ===================================================================== */
/* FP64 sincos / exp: Cody-Waite reduction + Taylor polynomials.          */
/* Constants from oracle/gen_math_constants.py (exact rational arithmetic).*/
/* Evaluated with contraction off so the device library can match bits.   */
/* ===================================================================== */
#define OC_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define OC_PIO2_1 0x1.921fb54400000p+0
#define OC_PIO2_2 0x1.0b4611a600000p-34
#define OC_PIO2_3 0x1.3198a2e037073p-69
#define OC_LOG2E 0x1.71547652b82fep+0
#define OC_LN2_HI 0x1.62e42ff000000p-1
#define OC_LN2_LO -0x1.718432a1b0e26p-35

static const double INV_FACT[19] = {
    1.0, 0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3,
    0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10,
    0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19,
    0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29,
    0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41,
    0x1.ae7f3e733b81fp-45, 0x1.952c77030ad4ap-49, 0x1.6827863b97d97p-53};

void or_sincos(double x, double* s_out, double* c_out) {
  double kf = floor(x * OC_TWO_OVER_PI + 0.5);
  double r = ((x - kf * OC_PIO2_1) - kf * OC_PIO2_2) - kf * OC_PIO2_3;
  double r2 = r * r;
  /* sin r = r + r^3 * (-1/3! + r^2 (1/5! - ...)) up to r^17 */
  double ps = INV_FACT[17];
  ps = -INV_FACT[15] + r2 * ps;
  ps = INV_FACT[13] + r2 * ps;
  ps = -INV_FACT[11] + r2 * ps;
  ps = INV_FACT[9] + r2 * ps;
  ps = -INV_FACT[7] + r2 * ps;
  ps = INV_FACT[5] + r2 * ps;
  ps = -INV_FACT[3] + r2 * ps;
  double s = r + (r * r2) * ps;
  /* cos r = 1 + r^2 * (-1/2! + r^2 (1/4! - ...)) up to r^18 */
  double pc = -INV_FACT[18];
  pc = INV_FACT[16] + r2 * pc;
  pc = -INV_FACT[14] + r2 * pc;
  pc = INV_FACT[12] + r2 * pc;
  pc = -INV_FACT[10] + r2 * pc;
  pc = INV_FACT[8] + r2 * pc;
  pc = -INV_FACT[6] + r2 * pc;
  pc = INV_FACT[4] + r2 * pc;
  pc = -INV_FACT[2] + r2 * pc;
  double c = 1.0 + r2 * pc;
  int64_t k = (int64_t)kf;
  switch ((int)(k & 3)) {
    case 0: *s_out = s; *c_out = c; break;
    case 1: *s_out = c; *c_out = -s; break;
    case 2: *s_out = -s; *c_out = -c; break;
    default: *s_out = -c; *c_out = s; break;
  }
}

double or_exp(double x) {
  if (x > 709.0) return INFINITY;
  if (x < -708.0) return 0.0;
  double kf = floor(x * OC_LOG2E + 0.5);
  double r = (x - kf * OC_LN2_HI) - kf * OC_LN2_LO;
  double p = INV_FACT[13];
  for (int n = 12; n >= 0; --n) p = INV_FACT[n] + r * p;
  int64_t k = (int64_t)kf;
  uint64_t bits = (uint64_t)(k + 1023) << 52;
  double scale;
  memcpy(&scale, &bits, sizeof scale);
  return p * scale;
}

/* reference-mode math: libm stands in for sleef u10 (src/fast_math.rs:3) */
static void ref_sincos(double x, double* s, double* c) { *s = sin(x); *c = cos(x); }

/* Vector reference-mode math for the batch steps (src/fast_math.rs:5-25: sleef's
 * sincos_u10/exp_u10 on 4 or 8 lanes).  Only the timed CPU-baseline builds define
 * OR_VECTOR_MATH (8: AVX-512, 4: AVX2); they call glibc's libmvec, the system's
 * vector libm (1-4 ulp, as sleef u10), through its x86_64 vector ABI names.  The
 * strict checker build keeps the scalar libm calls above. */
#if defined(OR_VECTOR_MATH)
#include <immintrin.h>
__m256d _ZGVdN4v_sin(__m256d);
__m256d _ZGVdN4v_cos(__m256d);
__m256d _ZGVdN4v_exp(__m256d);
static void ref_sincos_n(int lanes, const double* x, double* s, double* c) {
#if OR_VECTOR_MATH == 8
  if (lanes == 8) {
    extern __m512d _ZGVeN8v_sin(__m512d);
    extern __m512d _ZGVeN8v_cos(__m512d);
    const __m512d v = _mm512_loadu_pd(x);
    _mm512_storeu_pd(s, _ZGVeN8v_sin(v));
    _mm512_storeu_pd(c, _ZGVeN8v_cos(v));
    return;
  }
#endif
  for (int k = 0; k < lanes; k += 4) {
    const __m256d v = _mm256_loadu_pd(x + k);
    _mm256_storeu_pd(s + k, _ZGVdN4v_sin(v));
    _mm256_storeu_pd(c + k, _ZGVdN4v_cos(v));
  }
}
static void ref_exp_n(int lanes, const double* x, double* out) {
#if OR_VECTOR_MATH == 8
  if (lanes == 8) {
    extern __m512d _ZGVeN8v_exp(__m512d);
    _mm512_storeu_pd(out, _ZGVeN8v_exp(_mm512_loadu_pd(x)));
    return;
  }
#endif
  for (int k = 0; k < lanes; k += 4) _mm256_storeu_pd(out + k, _ZGVdN4v_exp(_mm256_loadu_pd(x + k)));
}
#endif

/* ===================================================================== */
/* Lattice, scratch -- src/types.rs:109-251,335-366                      */
/* ===================================================================== */
int or_lattice_new(or_lattice* lat, int lx, int ly, int lz) {
  memset(lat, 0, sizeof *lat);
  if (lx <= 0) return fail("L_x must be positive"); /* :122-130 */
  if (ly <= 0) return fail("L_y must be positive");
  if (lz <= 0) return fail("L_z must be positive");
  size_t v = (size_t)lx * ly * lz;
  lat->lx = lx; lat->ly = ly; lat->lz = lz;
  lat->theta = calloc(v, sizeof(double));
  lat->cos_t = calloc(v, sizeof(double));
  lat->sin_t = calloc(v, sizeof(double));
  lat->j_xy = calloc((size_t)lz, sizeof(double));
  lat->j_z = calloc((size_t)lz, sizeof(double));
  lat->in_cluster = calloc(v, 1);
  lat->stack = calloc(v, sizeof(int));
  lat->cluster = calloc(v, sizeof(int));
  for (size_t i = 0; i < v; ++i) lat->cos_t[i] = 1.0; /* scratch of theta = 0 */
  return 0;
}

void or_lattice_free(or_lattice* lat) {
  free(lat->theta); free(lat->cos_t); free(lat->sin_t); free(lat->j_xy);
  free(lat->j_z); free(lat->in_cluster); free(lat->stack); free(lat->cluster);
  memset(lat, 0, sizeof *lat);
}

static inline size_t vol(const or_lattice* l) { return (size_t)l->lx * l->ly * l->lz; }
static inline int plus1(int i, int l) { return i + 1 < l ? i + 1 : 0; }   /* :354-360 */
static inline int minus1(int i, int l) { return i > 0 ? i - 1 : l - 1; } /* :362-366 */
static inline size_t idx3(const or_lattice* l, int x, int y, int z) {
  return (size_t)x + (size_t)l->lx * ((size_t)y + (size_t)l->ly * z); /* :146-149 */
}

static inline double wrap_angle(double t) { /* :335-343 */
  if (t >= OR_TWO_PI) return t - OR_TWO_PI;
  if (t < 0.0) return t + OR_TWO_PI;
  return t;
}

void or_scratch_refresh(or_lattice* lat, int replay_math) { /* :183-191 */
  size_t v = vol(lat);
  for (size_t i = 0; i < v; ++i) {
    if (replay_math) or_sincos(lat->theta[i], &lat->sin_t[i], &lat->cos_t[i]);
    else ref_sincos(lat->theta[i], &lat->sin_t[i], &lat->cos_t[i]);
  }
}

static int validate_temperature(double t) { /* src/types.rs:373-378 */
  if (!(t > 0.0 && isfinite(t))) return fail("temperature T must be positive and finite");
  return 0;
}
static int validate_width(double w) { /* src/types.rs:380-385 */
  if (!(w >= 0.0 && isfinite(w))) return fail("proposal_width must be nonnegative and finite");
  return 0;
}
static int validate_even(const or_lattice* l) { /* src/updates.rs:429-438 */
  if (l->lx % 2 || l->ly % 2 || l->lz % 2)
    return fail("red-black Metropolis sweep requires even lattice dimensions");
  return 0;
}

/* ===================================================================== */
/* Initialisation -- src/initialization.rs                               */
/* ===================================================================== */
int or_two_point_disorder(double* values, int n, double mean, double delta,
                          or_fastrng* rng, const char* name) { /* :5-31 */
  if (delta < 0.0) {
    snprintf(g_err, sizeof g_err, "\xce\xb4%s must be nonnegative", name);
    return 1;
  }
  double lower = mean - delta, upper = mean + delta;
  if (lower < 0.0) {
    snprintf(g_err, sizeof g_err,
             "two-point layer disorder requires %s_mean - \xce\xb4%s >= 0", name, name);
    return 1;
  }
  if (delta == 0.0) {
    for (int i = 0; i < n; ++i) values[i] = mean;
  } else {
    for (int i = 0; i < n; ++i) values[i] = or_rng_bool(rng) ? upper : lower;
  }
  return 0;
}

/* Extension (no reference counterpart): iid uniform layer couplings on [lo, hi],
 * one f64 draw per layer from the same FastRng. */
int or_uniform_disorder(double* values, int n, double lo, double hi, or_fastrng* rng) {
  if (!(hi >= lo) || lo < 0.0) return fail("uniform layer disorder requires 0 <= lo <= hi");
  for (int i = 0; i < n; ++i) values[i] = lo + (hi - lo) * or_rng_f64(rng);
  return 0;
}

int or_initialize_disorder(or_lattice* lat, double j_xy, double j_z_mean,
                           double delta_j_z, or_fastrng* rng) { /* :33-46 */
  for (int z = 0; z < lat->lz; ++z) lat->j_xy[z] = j_xy;
  return or_two_point_disorder(lat->j_z, lat->lz, j_z_mean, delta_j_z, rng, "J_z");
}

void or_initialize_angles(or_lattice* lat, int cold, or_fastrng* rng) { /* :48-62 */
  size_t v = vol(lat);
  for (size_t i = 0; i < v; ++i) lat->theta[i] = cold ? 0.0 : OR_TWO_PI * or_rng_f64(rng);
}

void or_initialize_angles_keyed(or_lattice* lat, uint64_t seed) {
  size_t v = vol(lat);
  uint32_t w[4];
  for (size_t i = 0; i < v; ++i) {
    keyed(seed, (uint32_t)i, 0u, TAG_INIT, w);
    lat->theta[i] = OR_TWO_PI * u53(w[0], w[1]);
  }
}

/* ===================================================================== */
/* Metropolis -- src/updates.rs:12-53,169-486                            */
/* ===================================================================== */
int or_vector_exp_min_uphill(int lanes) { return lanes == 4 ? 2 : lanes == 8 ? 3 : lanes; }

void or_metropolis_accept_probabilities(const double* delta, int lanes, double beta,
                                        double* prob) { /* :21-53 */
  double arg[8];
  int uphill_lane[8], uphill = 0;
  for (int lane = 0; lane < lanes; ++lane) {
    if (delta[lane] > 0.0) {
      arg[uphill] = -delta[lane] * beta;
      uphill_lane[uphill++] = lane;
    }
  }
  for (int lane = 0; lane < lanes; ++lane) prob[lane] = 1.0;
#if defined(OR_VECTOR_MATH)
  /* vector exp of all lanes once enough of them are uphill (:42-51) */
  if (lanes >= 4 && uphill >= or_vector_exp_min_uphill(lanes)) {
    double e[8];
    for (int k = uphill; k < lanes; ++k) arg[k] = 0.0;
    ref_exp_n(lanes, arg, e);
    for (int k = 0; k < uphill; ++k) prob[uphill_lane[k]] = e[k];
    return;
  }
#endif
  /* scalar exp (the checker build takes libm for both, value-neutral there) */
  for (int k = 0; k < uphill; ++k) prob[uphill_lane[k]] = exp(arg[k]);
}

/* local_field with the reference's summation order (:169-231) */
static void local_field(const or_lattice* l, size_t idx, int x, int y, int z,
                        double* hx, double* hy) {
  size_t plane = (size_t)l->lx * l->ly, v = vol(l);
  size_t xp = (x + 1 == l->lx) ? idx + 1 - l->lx : idx + 1;
  size_t xm = (x == 0) ? idx + l->lx - 1 : idx - 1;
  size_t yp = (y + 1 == l->ly) ? idx + l->lx - plane : idx + l->lx;
  size_t ym = (y == 0) ? idx + plane - l->lx : idx - l->lx;
  size_t zp = (z + 1 == l->lz) ? idx + plane - v : idx + plane;
  int zm_l = (z == 0) ? l->lz - 1 : z - 1;
  size_t zm = (z == 0) ? idx + v - plane : idx - plane;
  double jxy = l->j_xy[z], jzp = l->j_z[z], jzm = l->j_z[zm_l];
  const double* c = l->cos_t;
  const double* s = l->sin_t;
  *hx = jxy * (c[xp] + c[xm] + c[yp] + c[ym]) + jzp * c[zp] + jzm * c[zm];
  *hy = jxy * (s[xp] + s[xm] + s[yp] + s[ym]) + jzp * s[zp] + jzm * s[zm];
}

static inline double local_energy(double s, double c, double hx, double hy) { /* :233-236 */
  return -(c * hx + s * hy);
}

/* scalar reference step (:238-271) */
static int step_at_ref(or_lattice* l, double width, double beta, size_t idx, int x,
                       int y, int z, or_fastrng* rng) {
  double t_old = l->theta[idx];
  double t_new = wrap_angle(t_old + width * (2.0 * or_rng_f64(rng) - 1.0));
  double ns, nc, hx, hy;
  ref_sincos(t_new, &ns, &nc);
  local_field(l, idx, x, y, z, &hx, &hy);
  double old_e = local_energy(l->sin_t[idx], l->cos_t[idx], hx, hy);
  double de = local_energy(ns, nc, hx, hy) - old_e;
  if (de <= 0.0 || or_rng_f64(rng) < exp(-de * beta)) {
    l->theta[idx] = t_new;
    l->sin_t[idx] = ns;
    l->cos_t[idx] = nc;
    return 1;
  }
  return 0;
}

/* LANES same-colour sites x0, x0+2, ... (:300-356) */
static int step_batch_ref(or_lattice* l, int lanes, double width, double beta, int x0,
                          int y, int z, or_fastrng* rng) {
  size_t idxs[8];
  double tn[8], hx[8], hy[8], oe[8], ns[8], nc[8], de[8], prob[8];
  for (int k = 0; k < lanes; ++k) {
    int x = x0 + 2 * k;
    size_t idx = idx3(l, x, y, z);
    idxs[k] = idx;
    tn[k] = wrap_angle(l->theta[idx] + width * (2.0 * or_rng_f64(rng) - 1.0));
    local_field(l, idx, x, y, z, &hx[k], &hy[k]);
    oe[k] = local_energy(l->sin_t[idx], l->cos_t[idx], hx[k], hy[k]);
  }
#if defined(OR_VECTOR_MATH)
  ref_sincos_n(lanes, tn, ns, nc); /* vector sincos of the lanes (:338) */
#else
  for (int k = 0; k < lanes; ++k) ref_sincos(tn[k], &ns[k], &nc[k]);
#endif
  for (int k = 0; k < lanes; ++k) de[k] = local_energy(ns[k], nc[k], hx[k], hy[k]) - oe[k];
  or_metropolis_accept_probabilities(de, lanes, beta, prob);
  int accepted = 0;
  for (int k = 0; k < lanes; ++k) {
    if (de[k] <= 0.0 || or_rng_f64(rng) < prob[k]) {
      l->theta[idxs[k]] = tn[k];
      l->sin_t[idxs[k]] = ns[k];
      l->cos_t[idxs[k]] = nc[k];
      accepted++;
    }
  }
  return accepted;
}

int or_metropolis_sweep_ref(or_lattice* l, double temperature, double width,
                            int use_batch8, or_fastrng* rng, double* acceptance) {
  /* :440-486 */
  if (validate_temperature(temperature) || validate_width(width) || validate_even(l)) return 1;
  double beta = 1.0 / temperature;
  size_t accepted = 0;
  for (int parity = 0; parity < 2; ++parity)
    for (int z = 0; z < l->lz; ++z)
      for (int y = 0; y < l->ly; ++y) {
        int x = (parity + y + z) & 1;
        if (use_batch8)
          while (x + 14 < l->lx) { accepted += step_batch_ref(l, 8, width, beta, x, y, z, rng); x += 16; }
        while (x + 6 < l->lx) { accepted += step_batch_ref(l, 4, width, beta, x, y, z, rng); x += 8; }
        while (x < l->lx) {
          accepted += step_at_ref(l, width, beta, idx3(l, x, y, z), x, y, z, rng);
          x += 2;
        }
      }
  *acceptance = (double)accepted / (double)vol(l);
  return 0;
}

int or_local_metropolis_step_ref(or_lattice* l, double temperature, double width,
                                 or_fastrng* rng, int* accepted) { /* :410-427,273-298 */
  if (validate_temperature(temperature) || validate_width(width)) return 1;
  or_scratch_refresh(l, 0); /* fresh ThetaScratch per call, as the reference */
  size_t v = vol(l), plane = (size_t)l->lx * l->ly;
  size_t idx = (size_t)or_rng_range(rng, v);
  int z = (int)(idx / plane);
  size_t xy = idx - (size_t)z * plane;
  int y = (int)(xy / l->lx);
  int x = (int)(xy - (size_t)y * l->lx);
  *accepted = step_at_ref(l, width, 1.0 / temperature, idx, x, y, z, rng);
  return 0;
}

/* Replay-mode sweep: identical rule, uniforms keyed by (seed, site, sweep). */
int or_metropolis_sweep_replay(or_lattice* l, double temperature, double width,
                               uint64_t seed, uint32_t sweep, uint8_t* decisions,
                               double* acceptance) {
  if (validate_temperature(temperature) || validate_width(width) || validate_even(l)) return 1;
  double beta = 1.0 / temperature;
  size_t accepted = 0;
  uint32_t w[4];
  for (int parity = 0; parity < 2; ++parity)
    for (int z = 0; z < l->lz; ++z)
      for (int y = 0; y < l->ly; ++y)
        for (int x = (parity + y + z) & 1; x < l->lx; x += 2) {
          size_t idx = idx3(l, x, y, z);
          keyed(seed, (uint32_t)idx, sweep, TAG_METRO, w);
          double u_prop = u53(w[0], w[1]), u_acc = u53(w[2], w[3]);
          double t_new = wrap_angle(l->theta[idx] + width * (2.0 * u_prop - 1.0));
          double ns, nc, hx, hy;
          or_sincos(t_new, &ns, &nc);
          local_field(l, idx, x, y, z, &hx, &hy);
          double old_e = local_energy(l->sin_t[idx], l->cos_t[idx], hx, hy);
          double de = local_energy(ns, nc, hx, hy) - old_e;
          int acc = de <= 0.0 || u_acc < or_exp(-de * beta);
          if (acc) {
            l->theta[idx] = t_new;
            l->sin_t[idx] = ns;
            l->cos_t[idx] = nc;
            accepted++;
          }
          if (decisions) decisions[idx] = (uint8_t)acc;
        }
  *acceptance = (double)accepted / (double)vol(l);
  return 0;
}

/* ===================================================================== */
/* Wolff -- src/updates.rs:131-167,498-629                               */
/* ===================================================================== */
double or_wolff_add_probability(double beta, double coupling, double ri, double rj) {
  double a = -2.0 * beta * coupling * ri * rj; /* :499-501 */
  return 1.0 - exp(a < 0.0 ? a : 0.0);
}

static double add_probability_replay(double beta, double coupling, double ri, double rj) {
  double a = -2.0 * beta * coupling * ri * rj;
  return 1.0 - or_exp(a < 0.0 ? a : 0.0);
}

double or_wolff_reflect_angle(double theta, double phi) { /* :503-506 */
  return wrap_angle(2.0 * phi + OR_PI - theta);
}

/* neighbour list in the reference order x+, x-, y+, y-, z+, z- (:131-167).
 * For replay: canonical bond = (base site, axis) with base the minus end. */
typedef struct { size_t site; double coupling; size_t base; int axis; } nbr_t;

static void neighbours(const or_lattice* l, int x, int y, int z, nbr_t nb[6]) {
  int xp = plus1(x, l->lx), xm = minus1(x, l->lx);
  int yp = plus1(y, l->ly), ym = minus1(y, l->ly);
  int zp = plus1(z, l->lz), zm = minus1(z, l->lz);
  size_t me = idx3(l, x, y, z);
  nb[0] = (nbr_t){idx3(l, xp, y, z), l->j_xy[z], me, 0};
  nb[1] = (nbr_t){idx3(l, xm, y, z), l->j_xy[z], idx3(l, xm, y, z), 0};
  nb[2] = (nbr_t){idx3(l, x, yp, z), l->j_xy[z], me, 1};
  nb[3] = (nbr_t){idx3(l, x, ym, z), l->j_xy[z], idx3(l, x, ym, z), 1};
  nb[4] = (nbr_t){idx3(l, x, y, zp), l->j_z[z], me, 2};
  nb[5] = (nbr_t){idx3(l, x, y, zm), l->j_z[zm], idx3(l, x, y, zm), 2};
}

static void site_xyz(const or_lattice* l, size_t idx, int* x, int* y, int* z) {
  size_t plane = (size_t)l->lx * l->ly;
  *z = (int)(idx / plane);
  size_t r = idx - (size_t)(*z) * plane;
  *y = (int)(r / l->lx);
  *x = (int)(r - (size_t)(*y) * l->lx);
}

int or_wolff_step_ref(or_lattice* l, double temperature, int use_scratch,
                      or_fastrng* rng, int64_t* cluster_size) { /* :551-620 */
  if (validate_temperature(temperature)) return 1;
  double beta = 1.0 / temperature;
  double phi = OR_TWO_PI * or_rng_f64(rng);
  double sphi, cphi;
  ref_sincos(phi, &sphi, &cphi);
  int x0 = (int)or_rng_range(rng, (uint64_t)l->lx);
  int y0 = (int)or_rng_range(rng, (uint64_t)l->ly);
  int z0 = (int)or_rng_range(rng, (uint64_t)l->lz);
  size_t sp = 0, nc = 0;
  size_t seed_idx = idx3(l, x0, y0, z0);
  l->in_cluster[seed_idx] = 1;
  l->stack[sp++] = (int)seed_idx;
  l->cluster[nc++] = (int)seed_idx;
  while (sp > 0) {
    size_t idx = (size_t)l->stack[--sp];
    int x, y, z;
    site_xyz(l, idx, &x, &y, &z);
    double ri = use_scratch ? l->cos_t[idx] * cphi + l->sin_t[idx] * sphi
                            : cos(l->theta[idx] - phi);
    nbr_t nb[6];
    neighbours(l, x, y, z, nb);
    for (int k = 0; k < 6; ++k) { /* try_add_wolff_neighbor :508-549 */
      size_t n = nb[k].site;
      if (l->in_cluster[n]) continue;
      double rj = use_scratch ? l->cos_t[n] * cphi + l->sin_t[n] * sphi
                              : cos(l->theta[n] - phi);
      if (nb[k].coupling * ri * rj <= 0.0) continue;
      if (or_rng_f64(rng) < or_wolff_add_probability(beta, nb[k].coupling, ri, rj)) {
        l->in_cluster[n] = 1;
        l->stack[sp++] = (int)n;
        l->cluster[nc++] = (int)n;
      }
    }
  }
  for (size_t k = 0; k < nc; ++k) { /* :602-618 */
    size_t idx = (size_t)l->cluster[k];
    double t = or_wolff_reflect_angle(l->theta[idx], phi);
    l->theta[idx] = t;
    if (use_scratch) ref_sincos(t, &l->sin_t[idx], &l->cos_t[idx]);
    l->in_cluster[idx] = 0;
  }
  *cluster_size = (int64_t)nc;
  return 0;
}

/* Replay-mode Wolff.  Same DFS and tests as the reference; the bond (base, axis)
 * draws its uniform from the key (seed, base, step) word[axis] * 2^-32, and its
 * probability is evaluated with the base endpoint first.  The cluster is then the
 * seed's component in the graph of active bonds, independent of traversal order. */
int or_wolff_step_replay(or_lattice* l, double temperature, uint64_t seed,
                         uint32_t step, uint8_t* members, int64_t* cluster_size) {
  if (validate_temperature(temperature)) return 1;
  double beta = 1.0 / temperature;
  size_t v = vol(l);
  uint32_t w[4];
  keyed(seed, 0u, step, TAG_WSEED, w);
  double phi = OR_TWO_PI * u53(w[0], w[1]);
  size_t seed_idx = (size_t)(((uint64_t)w[2] * (uint64_t)v) >> 32);
  double sphi, cphi;
  or_sincos(phi, &sphi, &cphi);
  size_t sp = 0, nc = 0;
  l->in_cluster[seed_idx] = 1;
  l->stack[sp++] = (int)seed_idx;
  l->cluster[nc++] = (int)seed_idx;
  while (sp > 0) {
    size_t idx = (size_t)l->stack[--sp];
    int x, y, z;
    site_xyz(l, idx, &x, &y, &z);
    double ri = l->cos_t[idx] * cphi + l->sin_t[idx] * sphi;
    nbr_t nb[6];
    neighbours(l, x, y, z, nb);
    for (int k = 0; k < 6; ++k) {
      size_t n = nb[k].site;
      if (l->in_cluster[n]) continue;
      double rj = l->cos_t[n] * cphi + l->sin_t[n] * sphi;
      double r_base = (nb[k].base == idx) ? ri : rj;
      double r_other = (nb[k].base == idx) ? rj : ri;
      if (nb[k].coupling * r_base * r_other <= 0.0) continue;
      keyed(seed, (uint32_t)nb[k].base, step, TAG_WBOND, w);
      double u = (double)w[nb[k].axis] * 0x1.0p-32;
      if (u < add_probability_replay(beta, nb[k].coupling, r_base, r_other)) {
        l->in_cluster[n] = 1;
        l->stack[sp++] = (int)n;
        l->cluster[nc++] = (int)n;
      }
    }
  }
  if (members) memset(members, 0, v);
  for (size_t k = 0; k < nc; ++k) {
    size_t idx = (size_t)l->cluster[k];
    double t = or_wolff_reflect_angle(l->theta[idx], phi);
    l->theta[idx] = t;
    or_sincos(t, &l->sin_t[idx], &l->cos_t[idx]);
    l->in_cluster[idx] = 0;
    if (members) members[idx] = 1;
  }
  *cluster_size = (int64_t)nc;
  return 0;
}

/* ===================================================================== */
/* Observables -- src/observables.rs                                     */
/* ===================================================================== */
void or_measure_observables(const or_lattice* l, or_observables* o) { /* :80-154 */
  double e = 0, mx = 0, my = 0, cx = 0, cy = 0, cz = 0, sx = 0, sy = 0, sz = 0;
  const double* C = l->cos_t;
  const double* S = l->sin_t;
  for (int z = 0; z < l->lz; ++z)
    for (int y = 0; y < l->ly; ++y)
      for (int x = 0; x < l->lx; ++x) {
        size_t i = idx3(l, x, y, z);
        double c = C[i], s = S[i];
        mx += c;
        my += s;
        size_t ix = idx3(l, plus1(x, l->lx), y, z);
        size_t iy = idx3(l, x, plus1(y, l->ly), z);
        size_t iz = idx3(l, x, y, plus1(z, l->lz));
        double dxc = c * C[ix] + s * S[ix], dxs = s * C[ix] - c * S[ix];
        double dyc = c * C[iy] + s * S[iy], dys = s * C[iy] - c * S[iy];
        double dzc = c * C[iz] + s * S[iz], dzs = s * C[iz] - c * S[iz];
        double jcx = l->j_xy[z] * dxc, jcy = l->j_xy[z] * dyc, jcz = l->j_z[z] * dzc;
        cx += jcx; cy += jcy; cz += jcz;
        sx += l->j_xy[z] * dxs;
        sy += l->j_xy[z] * dys;
        sz += l->j_z[z] * dzs;
        e -= jcx + jcy + jcz;
      }
  double v = (double)vol(l);
  o->energy = e / v;
  o->magnetization_squared = (mx * mx + my * my) / (v * v);
  o->cos_x = cx; o->cos_y = cy; o->cos_z = cz;
  o->sin_x = sx; o->sin_y = sy; o->sin_z = sz;
}

double or_measure_energy_direct(const or_lattice* l) { /* :45-61 */
  double e = 0.0;
  for (int z = 0; z < l->lz; ++z)
    for (int y = 0; y < l->ly; ++y)
      for (int x = 0; x < l->lx; ++x) {
        double t = l->theta[idx3(l, x, y, z)];
        e -= l->j_xy[z] * cos(t - l->theta[idx3(l, plus1(x, l->lx), y, z)]);
        e -= l->j_xy[z] * cos(t - l->theta[idx3(l, x, plus1(y, l->ly), z)]);
        e -= l->j_z[z] * cos(t - l->theta[idx3(l, x, y, plus1(z, l->lz))]);
      }
  return e / (double)vol(l);
}

double or_measure_magnetization_squared_direct(const or_lattice* l) { /* :63-69 */
  double mx = 0, my = 0;
  size_t v = vol(l);
  for (size_t i = 0; i < v; ++i) { mx += cos(l->theta[i]); my += sin(l->theta[i]); }
  double dv = (double)v;
  return (mx * mx + my * my) / (dv * dv);
}

void or_helicity_sums_direct(const or_lattice* l, double out[6]) { /* :267-295 */
  double cx = 0, cy = 0, cz = 0, sx = 0, sy = 0, sz = 0;
  for (int z = 0; z < l->lz; ++z)
    for (int y = 0; y < l->ly; ++y)
      for (int x = 0; x < l->lx; ++x) {
        double t = l->theta[idx3(l, x, y, z)];
        double dx = t - l->theta[idx3(l, plus1(x, l->lx), y, z)];
        double dy = t - l->theta[idx3(l, x, plus1(y, l->ly), z)];
        double dz = t - l->theta[idx3(l, x, y, plus1(z, l->lz))];
        cx += l->j_xy[z] * cos(dx); cy += l->j_xy[z] * cos(dy); cz += l->j_z[z] * cos(dz);
        sx += l->j_xy[z] * sin(dx); sy += l->j_xy[z] * sin(dy); sz += l->j_z[z] * sin(dz);
      }
  out[0] = cx; out[1] = cy; out[2] = cz; out[3] = sx; out[4] = sy; out[5] = sz;
}

int or_measure_helicity_modulus(const or_lattice* l, double temperature, double out[3]) {
  /* :297-310 */
  if (validate_temperature(temperature)) return 1;
  double v = (double)vol(l), beta = 1.0 / temperature, h[6];
  or_helicity_sums_direct(l, h);
  for (int a = 0; a < 3; ++a) out[a] = h[a] / v - beta * (h[3 + a] * h[3 + a]) / v;
  return 0;
}

/* cached_correlation_dot_x8 (:13-43): 8 lane partial sums, ordered lane
 * reduction, then a scalar tail. */
static double dot_x8(const double* c, const double* s, size_t a0, size_t b0, size_t len) {
  double lane[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t i = 0;
  while (i + 8 <= len) {
    for (int k = 0; k < 8; ++k)
      lane[k] += c[a0 + i + k] * c[b0 + i + k] + s[a0 + i + k] * s[b0 + i + k];
    i += 8;
  }
  double sum = 0.0;
  for (int k = 0; k < 8; ++k) sum += lane[k];
  while (i < len) { sum += c[a0 + i] * c[b0 + i] + s[a0 + i] * s[b0 + i]; ++i; }
  return sum;
}

void or_correlation_rmax(const or_lattice* l, int rmax, int rmax_xy, int rmax_z,
                         int* rxy_eff, int* rz_eff) { /* :177-181 */
  int mxy = (l->lx / 2) < (l->ly / 2) ? l->lx / 2 : l->ly / 2;
  int req_xy = rmax_xy >= 0 ? rmax_xy : (rmax >= 0 ? rmax : mxy);
  int req_z = rmax_z >= 0 ? rmax_z : (rmax >= 0 ? rmax : l->lz / 2);
  int e = req_xy < l->lx / 2 ? req_xy : l->lx / 2;
  *rxy_eff = e < l->ly / 2 ? e : l->ly / 2;
  *rz_eff = req_z < l->lz / 2 ? req_z : l->lz / 2;
}

void or_measure_correlations(const or_lattice* l, int rmax, int rmax_xy, int rmax_z,
                             double* corr_x, double* corr_y, double* corr_xy,
                             double* corr_xy_by_z, double* corr_z) { /* :166-265 */
  int rxy, rz;
  or_correlation_rmax(l, rmax, rmax_xy, rmax_z, &rxy, &rz);
  const double* C = l->cos_t;
  const double* S = l->sin_t;
  double v = (double)vol(l), area = (double)l->lx * l->ly;
  size_t lx = (size_t)l->lx, plane = (size_t)l->lx * l->ly;
  for (int r = 1; r <= rxy; ++r) {
    double sx = 0, sy = 0;
    for (int z = 0; z < l->lz; ++z) {
      size_t zb = plane * z;
      double lsx = 0, lsy = 0;
      for (int y = 0; y < l->ly; ++y) {
        size_t row = zb + lx * y;
        lsx += dot_x8(C, S, row, row + r, lx - r);
        lsx += dot_x8(C, S, row + lx - r, row, (size_t)r);
        size_t ysh = (size_t)((y + r) % l->ly);
        lsy += dot_x8(C, S, row, zb + lx * ysh, lx);
      }
      sx += lsx;
      sy += lsy;
      corr_xy_by_z[(size_t)z * rxy + (r - 1)] = (lsx + lsy) / (2.0 * area);
    }
    corr_x[r - 1] = sx / v;
    corr_y[r - 1] = sy / v;
  }
  for (int r = 1; r <= rz; ++r) {
    double sz = 0;
    for (int z = 0; z < l->lz; ++z) {
      size_t zb = plane * z, zs = plane * ((z + r) % l->lz);
      for (int y = 0; y < l->ly; ++y) sz += dot_x8(C, S, zb + lx * y, zs + lx * y, lx);
    }
    corr_z[r - 1] = sz / v;
  }
  for (int r = 0; r < rxy; ++r) corr_xy[r] = (corr_x[r] + corr_y[r]) / 2.0;
}

/* ===================================================================== */
/* Estimators -- src/simulation.rs:197-230                               */
/* ===================================================================== */
double or_mean(const double* x, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += x[i];
  return s / (double)n;
}

double or_sample_std(const double* x, int n) {
  if (n <= 1) return 0.0;
  double m = or_mean(x, n), v = 0.0;
  for (int i = 0; i < n; ++i) v += (x[i] - m) * (x[i] - m);
  return sqrt(v / (double)(n - 1));
}

double or_blocking_stderr(const double* x, int n) {
  if (n <= 1) return 0.0;
  double* d = malloc(sizeof(double) * (size_t)n);
  memcpy(d, x, sizeof(double) * (size_t)n);
  double se = or_sample_std(d, n) / sqrt((double)n);
  while (n >= 16) {
    int even = n - (n % 2);
    for (int i = 0; i < even; i += 2) d[i / 2] = (d[i] + d[i + 1]) / 2.0;
    n = even / 2;
    double cand = or_sample_std(d, n) / sqrt((double)n);
    if (cand > se) se = cand;
  }
  free(d);
  return se;
}

int or_binned_estimate(const double* samples, int n, int binsize, double* mean,
                       double* stderr_out, double* bins, int* nbins) {
  if (binsize <= 0 || binsize > n) return fail("binsize is larger than the sample series");
  int nb = n / binsize;
  for (int b = 0; b < nb; ++b) {
    double s = 0.0;
    for (int k = 0; k < binsize; ++k) s += samples[b * binsize + k];
    bins[b] = s / (double)binsize;
  }
  *nbins = nb;
  *mean = or_mean(bins, nb);
  *stderr_out = nb > 1 ? or_sample_std(bins, nb) / sqrt((double)nb) : 0.0;
  return 0;
}

/* ===================================================================== */
/* Plain driver -- src/simulation.rs:13-195                              */
/* ===================================================================== */
static double mean_square(const double* x, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += x[i] * x[i];
  return s / (double)n;
}

static int run_sim(or_lattice* l, double j_xy, double j_z_mean, double delta_j_z,
                   double temperature, const double* jxy_arr, const double* jz_arr,
                   const or_sim_options* o, or_fastrng* rng, int replay, uint64_t seed,
                   or_sim_result* res);

int or_run_theta_simulation(or_lattice* l, double j_xy, double j_z_mean, double delta_j_z,
                            double temperature, const or_sim_options* o, or_fastrng* rng,
                            or_sim_result* res) {
  return run_sim(l, j_xy, j_z_mean, delta_j_z, temperature, NULL, NULL, o, rng, 0, 0, res);
}

int or_run_theta_simulation_replay(or_lattice* l, double j_xy, double j_z_mean,
                                   double delta_j_z, double temperature,
                                   const double* jxy_arr, const double* jz_arr,
                                   const or_sim_options* o, uint64_t seed, or_sim_result* res) {
  or_fastrng rng;
  or_rng_seed(&rng, seed);
  return run_sim(l, j_xy, j_z_mean, delta_j_z, temperature, jxy_arr, jz_arr, o, &rng, 1, seed,
                 res);
}

/* one sweep of the driver: Metropolis + wolff_steps clusters, either mode */
static int driver_sweep(or_lattice* l, double temperature, const or_sim_options* o,
                        or_fastrng* rng, int replay, uint64_t seed, uint32_t* metro_ctr,
                        uint32_t* wolff_ctr, double* a) {
  int64_t cs;
  if (replay) {
    if (or_metropolis_sweep_replay(l, temperature, o->proposal_width, seed, (*metro_ctr)++,
                                   NULL, a)) return 1;
    for (int k = 0; k < o->wolff_steps; ++k)
      if (or_wolff_step_replay(l, temperature, seed, (*wolff_ctr)++, NULL, &cs)) return 1;
  } else {
    if (or_metropolis_sweep_ref(l, temperature, o->proposal_width, o->use_batch8, rng, a))
      return 1;
    for (int k = 0; k < o->wolff_steps; ++k)
      if (or_wolff_step_ref(l, temperature, 1, rng, &cs)) return 1;
  }
  return 0;
}

static int run_sim(or_lattice* l, double j_xy, double j_z_mean, double delta_j_z,
                   double temperature, const double* jxy_arr, const double* jz_arr,
                   const or_sim_options* o, or_fastrng* rng, int replay, uint64_t seed,
                   or_sim_result* res) {
  if (o->measure_interval <= 0) return fail("measure_interval must be positive");
  if (o->correlation_interval <= 0) return fail("correlation_interval must be positive");
  int nmeas = o->measure_sweeps / o->measure_interval;
  if (nmeas == 0)
    return fail("measure_sweeps must include at least one measurement; require "
                "measure_sweeps >= measure_interval");
  if (validate_temperature(temperature) || validate_width(o->proposal_width)) return 1;
  int rxy, rz;
  or_correlation_rmax(l, o->correlation_rmax, o->correlation_rmax_xy, o->correlation_rmax_z,
                      &rxy, &rz);
  if (jxy_arr && jz_arr) {
    for (int z = 0; z < l->lz; ++z) { l->j_xy[z] = jxy_arr[z]; l->j_z[z] = jz_arr[z]; }
  } else if (or_initialize_disorder(l, j_xy, j_z_mean, delta_j_z, rng)) {
    return 1;
  }
  for (int z = 0; z < l->lz; ++z)
    if (l->j_xy[z] < 0.0 || l->j_z[z] < 0.0)
      return fail("theta simulation requires nonnegative layer couplings");
  if (replay) {
    if (o->cold_init) { for (size_t i = 0; i < vol(l); ++i) l->theta[i] = 0.0; }
    else or_initialize_angles_keyed(l, seed);
  } else {
    or_initialize_angles(l, o->cold_init, rng);
  }
  or_scratch_refresh(l, replay);
  double acc_sum = 0.0, a;
  int acc_count = 0;
  uint32_t metro_ctr = 0, wolff_ctr = 0;
  for (int t = 0; t < o->thermal_sweeps; ++t) {
    if (driver_sweep(l, temperature, o, rng, replay, seed, &metro_ctr, &wolff_ctr, &a)) return 1;
    acc_sum += a; acc_count++;
  }
  double* series = calloc((size_t)nmeas * 8, sizeof(double));
  double *en = series, *cxs = series + nmeas, *cys = series + 2 * nmeas,
         *czs = series + 3 * nmeas, *sxs = series + 4 * nmeas, *sys = series + 5 * nmeas,
         *szs = series + 6 * nmeas, *m2 = series + 7 * nmeas;
  size_t lz = (size_t)l->lz;
  double* cx_sum = calloc((size_t)rxy + 1, sizeof(double));
  double* cy_sum = calloc((size_t)rxy + 1, sizeof(double));
  double* cxy_sum = calloc((size_t)rxy + 1, sizeof(double));
  double* cbz_sum = calloc(lz * (size_t)rxy + 1, sizeof(double));
  double* cz_sum = calloc((size_t)rz + 1, sizeof(double));
  double* tcx = calloc((size_t)rxy + 1, sizeof(double));
  double* tcy = calloc((size_t)rxy + 1, sizeof(double));
  double* tcxy = calloc((size_t)rxy + 1, sizeof(double));
  double* tcbz = calloc(lz * (size_t)rxy + 1, sizeof(double));
  double* tcz = calloc((size_t)rz + 1, sizeof(double));
  int nm = 0, ncorr = 0;
  for (int step = 1; step <= o->measure_sweeps; ++step) {
    if (driver_sweep(l, temperature, o, rng, replay, seed, &metro_ctr, &wolff_ctr, &a)) return 1;
    acc_sum += a; acc_count++;
    if (step % o->measure_interval == 0) {
      or_observables ob;
      or_measure_observables(l, &ob);
      en[nm] = ob.energy; cxs[nm] = ob.cos_x; cys[nm] = ob.cos_y; czs[nm] = ob.cos_z;
      sxs[nm] = ob.sin_x; sys[nm] = ob.sin_y; szs[nm] = ob.sin_z;
      m2[nm] = ob.magnetization_squared;
      nm++;
      if ((rxy > 0 || rz > 0) && (nm - 1) % o->correlation_interval == 0) {
        or_measure_correlations(l, -1, rxy, rz, tcx, tcy, tcxy, tcbz, tcz);
        for (int i = 0; i < rxy; ++i) { cx_sum[i] += tcx[i]; cy_sum[i] += tcy[i]; cxy_sum[i] += tcxy[i]; }
        for (size_t i = 0; i < lz * (size_t)rxy; ++i) cbz_sum[i] += tcbz[i];
        for (int i = 0; i < rz; ++i) cz_sum[i] += tcz[i];
        ncorr++;
      }
    }
  }
  double v = (double)vol(l), beta = 1.0 / temperature;
  double msx = or_mean(sxs, nm), msy = or_mean(sys, nm), msz = or_mean(szs, nm);
  res->rho_sx = or_mean(cxs, nm) / v - beta * (mean_square(sxs, nm) - msx * msx) / v;
  res->rho_sy = or_mean(cys, nm) / v - beta * (mean_square(sys, nm) - msy * msy) / v;
  res->rho_sz = or_mean(czs, nm) / v - beta * (mean_square(szs, nm) - msz * msz) / v;
  double* eff = malloc(sizeof(double) * (size_t)nm);
  const double* cs3[3] = {cxs, cys, czs};
  const double* ss3[3] = {sxs, sys, szs};
  double ms3[3] = {msx, msy, msz};
  double* std3[3] = {&res->std_rho_sx, &res->std_rho_sy, &res->std_rho_sz};
  for (int a3 = 0; a3 < 3; ++a3) { /* :145-159 */
    for (int i = 0; i < nm; ++i) {
      double s = ss3[a3][i];
      eff[i] = cs3[a3][i] / v - beta * (s * s - ms3[a3] * s) / v;
    }
    *std3[a3] = or_blocking_stderr(eff, nm);
  }
  res->energy = or_mean(en, nm);
  res->magnetization_squared = or_mean(m2, nm);
  res->chi = beta * v * res->magnetization_squared;
  double norm = ncorr > 0 ? (double)ncorr : 1.0;
  res->rxy = rxy;
  res->rz = rz;
  for (int i = 0; i < rxy; ++i) {
    res->corr_x[i] = cx_sum[i] / norm;
    res->corr_y[i] = cy_sum[i] / norm;
    res->corr_xy[i] = cxy_sum[i] / norm;
  }
  for (size_t i = 0; i < lz * (size_t)rxy; ++i) res->corr_xy_by_z[i] = cbz_sum[i] / norm;
  for (int i = 0; i < rz; ++i) res->corr_z[i] = cz_sum[i] / norm;
  res->acceptance = acc_count == 0 ? 0.0 : acc_sum / (double)acc_count;
  res->num_measurements = nmeas;
  res->num_correlation_measurements = ncorr;
  free(series); free(eff); free(cx_sum); free(cy_sum); free(cxy_sum); free(cbz_sum);
  free(cz_sum); free(tcx); free(tcy); free(tcxy); free(tcbz); free(tcz);
  return 0;
}
