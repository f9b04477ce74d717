// estimators.cpp -- FP64 statistics on the host: binning, blocking error,
// jackknife, and the reference's derived observables.
#include <cmath>
#include <vector>

#include "../../include/slsf.h"
#include "common.h"

using namespace slsf;

namespace {

double mean(const double* x, int64_t n) {  // src/simulation.rs:197-199
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += x[i];
  return s / (double)n;
}

double sample_std(const double* x, int64_t n) {  // src/simulation.rs:201-208
  if (n <= 1) return 0.0;
  const double m = mean(x, n);
  double v = 0.0;
  for (int64_t i = 0; i < n; ++i) v += (x[i] - m) * (x[i] - m);
  return std::sqrt(v / (double)(n - 1));
}

// derived observable k of the reference's Evaluator (src/model.rs:857-888) and
// the Binder ratio, from the means of the 13 series
double eval_one(int k, const double* m, double volume, double beta) {
  switch (k) {
    case SLSF_EVAL_MAGNETIZATION: return std::sqrt(m[SLSF_SERIES_M2]);
    case SLSF_EVAL_CHI: return beta * volume * m[SLSF_SERIES_M2];
    case SLSF_EVAL_SPECIFIC_HEAT:
      return beta * beta * volume *
             (m[SLSF_SERIES_ENERGY2] - m[SLSF_SERIES_ENERGY] * m[SLSF_SERIES_ENERGY]);
    case SLSF_EVAL_RHO_XY: {
      const double rx = m[SLSF_SERIES_COS_X] / volume -
                        beta * (m[SLSF_SERIES_SIN2_X] - m[SLSF_SERIES_SIN_X] * m[SLSF_SERIES_SIN_X]) / volume;
      const double ry = m[SLSF_SERIES_COS_Y] / volume -
                        beta * (m[SLSF_SERIES_SIN2_Y] - m[SLSF_SERIES_SIN_Y] * m[SLSF_SERIES_SIN_Y]) / volume;
      return (rx + ry) / 2.0;
    }
    case SLSF_EVAL_RHO_Z:
      return m[SLSF_SERIES_COS_Z] / volume -
             beta * (m[SLSF_SERIES_SIN2_Z] - m[SLSF_SERIES_SIN_Z] * m[SLSF_SERIES_SIN_Z]) / volume;
    case SLSF_EVAL_BINDER: return m[SLSF_SERIES_M4] / (m[SLSF_SERIES_M2] * m[SLSF_SERIES_M2]);
    case SLSF_EVAL_ENERGY: return m[SLSF_SERIES_ENERGY];
    default: return NAN;
  }
}

// Disorder-averaged evaluable k over the realizations r with keep[r] (all
// when keep is null): the mean of the per-realization evaluables, except the
// Binder ratio, the ratio of the disorder averages of <m^4> and <m^2>^2.
double disorder_eval(int k, const double* means, int n_real, const std::vector<char>* keep,
                     double volume, double beta) {
  double sum = 0.0, m2 = 0.0, m4 = 0.0;
  int n = 0;
  for (int r = 0; r < n_real; ++r) {
    if (keep && !(*keep)[r]) continue;
    const double* m = means + (size_t)r * SLSF_NUM_SERIES;
    if (k == SLSF_EVAL_BINDER) {
      m2 += m[SLSF_SERIES_M2];
      m4 += m[SLSF_SERIES_M4];
    } else {
      sum += eval_one(k, m, volume, beta);
    }
    ++n;
  }
  if (k == SLSF_EVAL_BINDER) {
    m2 /= (double)n;
    m4 /= (double)n;
    return m4 / (m2 * m2);
  }
  return sum / (double)n;
}

// jackknife standard error of the leave-one-out values loo[0..n)
double jackknife_error(const std::vector<double>& loo) {
  const int64_t n = (int64_t)loo.size();
  const double m = mean(loo.data(), n);
  double v = 0.0;
  for (double x : loo) v += (x - m) * (x - m);
  return std::sqrt((double)(n - 1) / (double)n * v);
}

// first sign change of a - b over the temperature grid, linear interpolation
bool first_crossing(const double* t, const std::vector<double>& a, const std::vector<double>& b,
                    double* tc) {
  const size_t n = a.size();
  for (size_t k = 0; k + 1 < n; ++k) {
    const double d0 = a[k] - b[k], d1 = a[k + 1] - b[k + 1];
    if (d0 == 0.0) {
      *tc = t[k];
      return true;
    }
    if ((d0 < 0.0) != (d1 < 0.0) && d1 != d0) {
      *tc = t[k] + (t[k + 1] - t[k]) * d0 / (d0 - d1);
      return true;
    }
  }
  if (n > 0 && a[n - 1] == b[n - 1]) {
    *tc = t[n - 1];
    return true;
  }
  return false;
}

}  // namespace

extern "C" {

int slsf_binned_estimate(const double* samples, int64_t n, int32_t binsize, double* mean_out,
                         double* stderr_out, double* bins, int64_t* nbins) {
  clear_error();
  if (binsize <= 0 || binsize > n)
    return fail(SLSF_E_INVALID, "binsize is larger than the sample series");
  const int64_t nb = n / binsize;
  std::vector<double> local;
  double* out = bins;
  if (!out) { local.resize(nb); out = local.data(); }
  for (int64_t b = 0; b < nb; ++b) {
    double s = 0.0;
    for (int32_t k = 0; k < binsize; ++k) s += samples[b * binsize + k];
    out[b] = s / (double)binsize;
  }
  if (nbins) *nbins = nb;
  if (mean_out) *mean_out = mean(out, nb);
  if (stderr_out) *stderr_out = nb > 1 ? sample_std(out, nb) / std::sqrt((double)nb) : 0.0;
  return SLSF_OK;
}

double slsf_blocking_stderr(const double* x, int64_t n) {  // src/simulation.rs:210-226
  if (n <= 1) return 0.0;
  std::vector<double> d(x, x + n);
  double se = sample_std(d.data(), n) / std::sqrt((double)n);
  while ((int64_t)d.size() >= 16) {
    const size_t even = d.size() - (d.size() % 2);
    std::vector<double> h(even / 2);
    for (size_t i = 0; i < even; i += 2) h[i / 2] = (d[i] + d[i + 1]) / 2.0;
    d.swap(h);
    const double cand = sample_std(d.data(), (int64_t)d.size()) / std::sqrt((double)d.size());
    if (cand > se) se = cand;
  }
  return se;
}

const char* slsf_eval_name(int32_t k) {
  static const char* names[SLSF_NUM_EVALS] = {"Magnetization", "Chi",    "SpecificHeat", "RhoXY",
                                              "RhoZ",          "Binder", "Energy"};
  return (k >= 0 && k < SLSF_NUM_EVALS) ? names[k] : "";
}

int slsf_jackknife(const double* loo, int64_t n, double f_all, double* estimate,
                   double* stderr_out) {
  clear_error();
  if (!loo || n < 2) return fail(SLSF_E_INVALID, "jackknife needs at least two bins");
  const double m = mean(loo, n);
  double v = 0.0;
  for (int64_t i = 0; i < n; ++i) v += (loo[i] - m) * (loo[i] - m);
  if (estimate) *estimate = (double)n * f_all - (double)(n - 1) * m;
  if (stderr_out) *stderr_out = std::sqrt((double)(n - 1) / (double)n * v);
  return SLSF_OK;
}

int slsf_evaluate(const double* bins, int64_t nbins, double volume, double beta, double* value,
                  double* error) {
  clear_error();
  if (!bins || !value || nbins < 1) return fail(SLSF_E_INVALID, "no bins to evaluate");
  double full[SLSF_NUM_SERIES];
  for (int s = 0; s < SLSF_NUM_SERIES; ++s) full[s] = mean(bins + (size_t)s * nbins, nbins);
  for (int k = 0; k < SLSF_NUM_EVALS; ++k) value[k] = eval_one(k, full, volume, beta);
  if (!error) return SLSF_OK;
  if (nbins < 2) {
    for (int k = 0; k < SLSF_NUM_EVALS; ++k) error[k] = 0.0;
    return SLSF_OK;
  }
  std::vector<double> loo((size_t)SLSF_NUM_EVALS * nbins);
  double m[SLSF_NUM_SERIES];
  for (int64_t i = 0; i < nbins; ++i) {
    for (int s = 0; s < SLSF_NUM_SERIES; ++s)
      m[s] = (full[s] * (double)nbins - bins[(size_t)s * nbins + i]) / (double)(nbins - 1);
    for (int k = 0; k < SLSF_NUM_EVALS; ++k) loo[(size_t)k * nbins + i] = eval_one(k, m, volume, beta);
  }
  for (int k = 0; k < SLSF_NUM_EVALS; ++k) {
    const double* l = loo.data() + (size_t)k * nbins;
    const double lm = mean(l, nbins);
    double v = 0.0;
    for (int64_t i = 0; i < nbins; ++i) v += (l[i] - lm) * (l[i] - lm);
    error[k] = std::sqrt((double)(nbins - 1) / (double)nbins * v);
  }
  return SLSF_OK;
}

int slsf_disorder_average(const double* means, int32_t n_real, double volume, double beta,
                          double* value, double* error) {
  clear_error();
  if (!means || !value) return fail(SLSF_E_INVALID, "null argument");
  if (n_real < 2) return fail(SLSF_E_INVALID, "a disorder average needs at least two realizations");
  std::vector<char> keep((size_t)n_real, 1);
  std::vector<double> loo((size_t)n_real);
  for (int k = 0; k < SLSF_NUM_EVALS; ++k) {
    value[k] = disorder_eval(k, means, n_real, nullptr, volume, beta);
    if (!error) continue;
    for (int r = 0; r < n_real; ++r) {
      keep[r] = 0;
      loo[r] = disorder_eval(k, means, n_real, &keep, volume, beta);
      keep[r] = 1;
    }
    error[k] = jackknife_error(loo);
  }
  return SLSF_OK;
}

int slsf_crossing(const double* temps, int32_t n_temps, int32_t n_real, int32_t eval,
                  const double* means_a, double volume_a, double scale_a,
                  const double* means_b, double volume_b, double scale_b, int32_t* found,
                  double* t_cross, double* t_error) {
  clear_error();
  if (!temps || !means_a || !means_b || !found || !t_cross)
    return fail(SLSF_E_INVALID, "null argument");
  if (n_temps < 2) return fail(SLSF_E_INVALID, "a crossing needs at least two temperatures");
  if (n_real < 2) return fail(SLSF_E_INVALID, "a crossing needs at least two realizations");
  if (eval < 0 || eval >= SLSF_NUM_EVALS) return fail(SLSF_E_INVALID, "unknown evaluable");
  const size_t per_t = (size_t)n_real * SLSF_NUM_SERIES;
  std::vector<double> a((size_t)n_temps), b((size_t)n_temps);
  auto curves = [&](const std::vector<char>* keep) {
    for (int t = 0; t < n_temps; ++t) {
      const double beta = 1.0 / temps[t];
      a[t] = scale_a * disorder_eval(eval, means_a + t * per_t, n_real, keep, volume_a, beta);
      b[t] = scale_b * disorder_eval(eval, means_b + t * per_t, n_real, keep, volume_b, beta);
    }
  };
  curves(nullptr);
  double tc = 0.0;
  *found = first_crossing(temps, a, b, &tc) ? 1 : 0;
  *t_cross = *found ? tc : NAN;
  if (t_error) *t_error = NAN;
  if (!*found || !t_error) return SLSF_OK;
  std::vector<char> keep((size_t)n_real, 1);
  std::vector<double> loo((size_t)n_real);
  for (int r = 0; r < n_real; ++r) {
    keep[r] = 0;
    curves(&keep);
    keep[r] = 1;
    if (!first_crossing(temps, a, b, &loo[r])) return SLSF_OK;  // error stays NaN
  }
  *t_error = jackknife_error(loo);
  return SLSF_OK;
}

}  // extern "C"
