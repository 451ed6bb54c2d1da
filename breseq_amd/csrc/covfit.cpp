// covfit.cpp -- coverage-distribution fit on the host (a few thousand histogram bins; not worth a
// kernel): censored negative-binomial fit by multi-restart Nelder-Mead and the deletion
// propagation cutoff from its quantile.
//   fit_censored_negative_binomial  coverage_distribution.cpp:115-326
//   fit_coverage_distribution       coverage_distribution.cpp:345-399
//   nelder_mead_minimize            stats.cpp:2289-2392
//   qnbinom_mu / dnbinom_mu         stats.cpp:2047-2071, stats.h:92-100
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "host_util.h"

namespace bsq {

// Regularized incomplete beta I_x(a, b) by Lentz's continued fraction.
double reg_inc_beta(double a, double b, double x) {
  if (x <= 0.0) return 0.0;
  if (x >= 1.0) return 1.0;
  const double lbt = std::lgamma(a + b) - std::lgamma(a) - std::lgamma(b) + a * std::log(x) + b * std::log1p(-x);
  auto cf = [](double aa, double bb, double xx) {
    const double tiny = 1e-300;
    const double qab = aa + bb, qap = aa + 1.0, qam = aa - 1.0;
    double c = 1.0, d = 1.0 - qab * xx / qap;
    if (std::fabs(d) < tiny) d = tiny;
    d = 1.0 / d;
    double h = d;
    for (int m = 1; m < 1000; m++) {
      const int m2 = 2 * m;
      double an = m * (bb - m) * xx / ((qam + m2) * (aa + m2));
      d = 1.0 + an * d;
      if (std::fabs(d) < tiny) d = tiny;
      c = 1.0 + an / c;
      if (std::fabs(c) < tiny) c = tiny;
      d = 1.0 / d;
      h *= d * c;
      an = -(aa + m) * (qab + m) * xx / ((aa + m2) * (qap + m2));
      d = 1.0 + an * d;
      if (std::fabs(d) < tiny) d = tiny;
      c = 1.0 + an / c;
      if (std::fabs(c) < tiny) c = tiny;
      d = 1.0 / d;
      const double del = d * c;
      h *= del;
      if (std::fabs(del - 1.0) < 1e-16) break;
    }
    return h;
  };
  if (x < (a + 1.0) / (a + b + 2.0)) return std::exp(lbt) * cf(a, b, x) / a;
  return 1.0 - std::exp(lbt) * cf(b, a, 1.0 - x) / b;
}

// negative binomial CDF, Cephes nbdtr(k, n, p) = I_p(n, k + 1)
static double nbdtr(double k, double n, double p) { return reg_inc_beta(n, k + 1.0, p); }

static double dnbinom_mu(double k, double size, double mu) {
  const double p = size / (size + mu);
  const double lp = std::lgamma(k + size) - std::lgamma(size) - std::lgamma(k + 1) + size * std::log(p) + k * std::log(1.0 - p);
  return std::exp(lp);
}

static uint32_t qnbinom_mu(double target, double size, double mu) {
  const double p = size / (size + mu);
  if (target <= 0.0) return 0;
  uint64_t hi = 1;
  while (nbdtr((double)hi, size, p) < target) {
    hi *= 2;
    if (hi > ((uint64_t)1 << 40)) break;
  }
  uint64_t lo = hi / 2;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (nbdtr((double)mid, size, p) < target) lo = mid + 1;
    else hi = mid;
  }
  return (uint32_t)lo;
}

struct NmResult {
  std::vector<double> estimate;
  bool converged;
};

static NmResult nelder_mead(const std::function<double(const std::vector<double>&)>& f, const std::vector<double>& x0,
                            uint32_t max_iter = 1000, double tol = 1e-8) {
  const size_t n = x0.size();
  std::vector<std::vector<double>> sx(n + 1, x0);
  for (size_t i = 0; i < n; i++) sx[i + 1][i] += (x0[i] != 0.0) ? 0.05 * std::fabs(x0[i]) : 0.00025;
  std::vector<double> v(n + 1);
  for (size_t i = 0; i <= n; i++) v[i] = f(sx[i]);
  bool conv = false;
  for (uint32_t it = 0; it < max_iter; it++) {
    std::vector<size_t> ord(n + 1);
    for (size_t i = 0; i <= n; i++) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&v](size_t a, size_t b) { return v[a] < v[b]; });
    std::vector<std::vector<double>> s2(n + 1);
    std::vector<double> v2(n + 1);
    for (size_t i = 0; i <= n; i++) {
      s2[i] = sx[ord[i]];
      v2[i] = v[ord[i]];
    }
    sx = s2;
    v = v2;
    if ((v[n] - v[0]) < tol) {
      conv = true;
      break;
    }
    std::vector<double> c(n, 0.0);
    for (size_t i = 0; i < n; i++)
      for (size_t j = 0; j < n; j++) c[j] += sx[i][j];
    for (size_t j = 0; j < n; j++) c[j] /= (double)n;
    std::vector<double> xr(n);
    for (size_t j = 0; j < n; j++) xr[j] = c[j] + 1.0 * (c[j] - sx[n][j]);
    const double fr = f(xr);
    if (fr < v[0]) {
      std::vector<double> xe(n);
      for (size_t j = 0; j < n; j++) xe[j] = c[j] + 2.0 * (xr[j] - c[j]);
      const double fe = f(xe);
      if (fe < fr) {
        sx[n] = xe;
        v[n] = fe;
      } else {
        sx[n] = xr;
        v[n] = fr;
      }
    } else if (fr < v[n - 1]) {
      sx[n] = xr;
      v[n] = fr;
    } else {
      std::vector<double> xc(n);
      for (size_t j = 0; j < n; j++) xc[j] = c[j] + 0.5 * (sx[n][j] - c[j]);
      const double fc = f(xc);
      if (fc < v[n]) {
        sx[n] = xc;
        v[n] = fc;
      } else {
        for (size_t i = 1; i <= n; i++) {
          for (size_t j = 0; j < n; j++) sx[i][j] = sx[0][j] + 0.5 * (sx[i][j] - sx[0][j]);
          v[i] = f(sx[i]);
        }
      }
    }
  }
  size_t best = 0;
  for (size_t i = 1; i <= n; i++)
    if (v[i] < v[best]) best = i;
  return NmResult{sx[best], conv && std::isfinite(v[best])};
}

static void fit_censored_nb(const std::vector<double>& n, uint32_t N, bsq_coverage_fit& res, double& nb_mu, double& nb_size) {
  nb_mu = nb_size = 0;
  double total = 0;
  for (uint32_t i = 1; i <= N; i++) total += n[i];
  if (total == 0) return;
  double m = 0;
  for (uint32_t i = 1; i <= N; i++) m += i * n[i];
  m /= total;
  double v = 0;
  if (total > 1) {
    for (uint32_t i = 1; i <= N; i++) v += n[i] * (i - m) * (i - m);
    v /= (total - 1);
  }
  res.average = m;
  res.variance = v;
  res.relative_variance = (m > 0) ? v / m : 0.0;
  std::vector<double> ma(N + 1, 0.0);
  std::vector<bool> ok(N + 1, false);
  if (N >= 5) {
    for (uint32_t i = 3; i <= N - 2; i++) {
      ma[i] = (n[i - 2] + n[i - 1] + n[i] + n[i + 1] + n[i + 2]) / 5.0;
      ok[i] = true;
    }
  } else {
    for (uint32_t i = 1; i <= N; i++) {
      ma[i] = n[i];
      ok[i] = true;
    }
  }
  const uint32_t min_i = std::max((uint32_t)(m / 4.0), (uint32_t)1);
  double max_n = 0;
  uint32_t max_i = 0;
  for (uint32_t i = min_i; i <= N; i++)
    if (ok[i] && ma[i] > max_n) {
      max_n = ma[i];
      max_i = i;
    }
  const uint32_t start_i = std::max((uint32_t)std::floor(max_i * 0.5), (uint32_t)1);
  const uint32_t end_i = std::min((uint32_t)std::ceil(max_i * 1.5), N);
  if (start_i == end_i) return;
  res.censor_start = start_i;
  res.censor_end = end_i;
  uint32_t per_bin = (end_i - start_i) / 1000;
  std::vector<double> x;
  uint32_t sf, ef;
  if (per_bin > 1) {
    sf = start_i / per_bin;
    ef = (uint32_t)std::ceil((double)end_i / per_bin);
    x.assign(ef + 1, 0.0);
    for (uint32_t i = sf; i <= ef; i++)
      for (uint32_t j = 1; j <= per_bin; j++) {
        const uint32_t idx = i * per_bin + j;
        if (idx <= N) x[i] += n[idx];
      }
  } else {
    per_bin = 1;
    sf = start_i;
    ef = end_i;
    x.assign(ef + 1, 0.0);
    for (uint32_t i = 1; i <= ef; i++) x[i] = n[i];
  }
  double inner = 0;
  for (uint32_t i = sf; i <= ef; i++) inner += x[i];
  double num = 0, den = 0;
  for (uint32_t i = 1; i <= ef; i++) {
    num += i * x[i];
    den += x[i];
  }
  const double mean_est = num / den;
  auto f_nb = [&](const std::vector<double>& par) -> double {
    const double mu = std::exp(par[0]), size = std::exp(par[1]);
    if (!std::isfinite(mu) || !std::isfinite(size)) return 1e10;
    std::vector<double> dist(ef + 1, 0.0);
    double tot = 0;
    for (uint32_t i = sf; i <= ef; i++) {
      dist[i] = dnbinom_mu((double)i, size, mu);
      tot += dist[i];
    }
    if (!(tot > 0) || !std::isfinite(tot)) return 1e10;
    double l = 0;
    for (uint32_t i = sf; i <= ef; i++) {
      const double d = (x[i] / inner) - (dist[i] / tot);
      l += d * d;
    }
    return std::isfinite(l) ? l : 1e10;
  };
  const double tries[6] = {mean_est, (double)ef, (double)sf, 1.0 * (ef + sf) / 4.0, 2.0 * (ef + sf) / 4.0,
                           3.0 * (ef + sf) / 4.0};
  double best = HUGE_VAL;
  for (double tm : tries) {
    double ts = 100000;
    for (int k = 0; k < 8; k++) {
      ts /= 10.0;
      NmResult r = nelder_mead(f_nb, {std::log(tm), std::log(ts)});
      if (!r.converged) continue;
      const double obj = f_nb(r.estimate);
      if (obj < best) {
        best = obj;
        nb_mu = std::exp(r.estimate[0]);
        nb_size = std::exp(r.estimate[1]);
      }
    }
  }
  if (!(best < HUGE_VAL)) nb_mu = nb_size = 0;
  double incl = 0;
  if (nb_mu > 0) {
    const double p = nb_size / (nb_size + nb_mu);
    incl = nbdtr((double)ef, nb_size, p) - nbdtr((double)sf, nb_size, p);
    if (incl >= 0.01 && per_bin > 1) nb_mu *= per_bin;
  }
  if (incl < 0.01) nb_mu = nb_size = 0;
}

}  // namespace bsq

using namespace bsq;

extern "C" int bsq_fit_coverage_distribution(const uint32_t* hist, int64_t len, double sequence_length, bsq_coverage_fit* out,
                                             bsq_status* status) {
  try {
    if (!out || (len > 0 && !hist) || !(sequence_length > 0)) throw BsqError(BSQ_ERR_ARG, "bad arguments");
    *out = bsq_coverage_fit{};
    // read_coverage_histogram (coverage_distribution.cpp:34-65): rows j >= 1 as written
    uint32_t N = 0;
    for (int64_t j = 1; j < len; j++) N = (uint32_t)j;
    std::vector<double> n(N + 1, 0.0);
    for (int64_t j = 1; j < len; j++) n[j] = hist[j];
    if (len <= 1) n.assign(1, 0.0);
    double mu = 0, size = 0;
    fit_censored_nb(n, N, *out, mu, size);
    out->nb_fit_mu = mu;
    out->nb_fit_size = size;
    const double m = out->average, v = out->variance;
    const double pr_cutoff = 0.05 / std::sqrt(sequence_length);
    double dp;
    if (mu > 0) {
      dp = (double)qnbinom_mu(pr_cutoff, size, mu);
    } else {
      const double se = (1.0 / (v - m)) * m * m;
      if (se > 0 && std::isfinite(se)) dp = (double)qnbinom_mu(pr_cutoff, se, m);
      else dp = -1;
      if (!(dp >= 1)) dp = m * 0.1;
    }
    if (dp < 1) dp = 1;
    if (mu <= 3 && m <= 3) dp = -1;
    out->deletion_coverage_propagation_cutoff = dp;
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    set_status(status, e.code, e.what());
    return e.code;
  }
}
