// covfit.cpp -- the coverage-distribution stage on the host: the censored negative-binomial fit of
// each coverage group's unique-only coverage histogram and the deletion propagation cutoff drawn
// from its quantile (coverage_distribution.cpp:115-399, 510-633).
//
// The fit is a few hundred histogram bins and a few thousand sequential objective evaluations:
// latency work with no data parallelism worth a kernel, whose printed results depend on every
// rounding of glibc's log/exp/pow.  It therefore runs on host cores, laid out for them:
//   * the 48 Nelder-Mead restarts (6 starting means x 8 starting sizes) are independent, so they run
//     concurrently on a pool of host threads; the best converged restart is then chosen in the
//     reference's restart order with a strict '<', which gives the sequential loop's choice;
//   * an objective evaluation hoists everything that does not depend on the bin (log Gamma of the
//     size, log p, log(1-p)) and the per-bin constants of the window (log Gamma(k+1), the observed
//     shares) out of the bin loop.  Hoisting a pure function's call does not change its value, and
//     the remaining operations are the reference's, in its order, so every objective value, every
//     simplex decision and the final estimate are the reference's bit for bit;
//   * the simplex is three fixed vertices in registers and stack arrays: no allocation per step.
// Inside a pipelined session the stage runs on its own host thread while the device runs the
// identify column pass (the fit decides only the MC runs, which come after it).
//
// The special functions are the ones breseq computes with: log Gamma and the incomplete beta
// integral of the Cephes library (S. Moshier, Cephes Math Library 2.8, 2000), whose coefficient
// tables and evaluation order fix the last bits of the NB pmf and CDF.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "covfit.h"
#include "host_util.h"

namespace bsq {
namespace nb {

// ---- special functions (Cephes method) --------------------------------------------------------

// Horner evaluation c[0] x^n + ... + c[n], or with an implied leading 1 (monic).
static inline double horner(double x, const double* c, int n) {
  double a = c[0];
  for (int i = 1; i <= n; i++) a = a * x + c[i];
  return a;
}
static inline double horner_monic(double x, const double* c, int n) {
  double a = x + c[0];
  for (int i = 1; i < n; i++) a = a * x + c[i];
  return a;
}

// log Gamma(x), x > 0, in the form breseq's NB pmf uses (stats.cpp:534-650): upward/downward
// recurrence into [2, 3) and a rational approximation there below 13; Stirling's series above.
// The rational and series coefficients are evaluated as Horner chains in breseq's order.
namespace lg {
constexpr double kNum[6] = {-1378.25152569120859100, -38801.6315134637840924, -331612.992738871184744,
                            -1162370.97492762307383, -1721737.00820839662146, -853555.664245765465627};
constexpr double kDen[7] = {1.0, -351.815701436523470549, -17064.2106651881159223, -220528.590553854454839,
                            -1139334.44367982507207, -2532523.07177582951285, -2018891.41433532773231};
// Stirling series coefficients as breseq states them: mantissa times a power of ten, folded once
constexpr double kSt[5] = {8.11614167470508450300 * 0.0001, -5.95061904284301438324 * 0.0001,
                           7.93650340457716943945 * 0.0001, -2.77777777730099687205 * 0.001,
                           8.33333333333331927722 * 0.01};
constexpr double kStLarge0 = 7.9365079365079365079365 * 0.0001, kStLarge1 = 2.7777777777777777777778 * 0.001,
                 kStLarge2 = 0.0833333333333333333333;
constexpr double kHalfLog2Pi = 0.91893853320467274178;
}  // namespace lg

double log_gamma(double x) {
  if (x < 13) {
    double z = 1, s = 0, u = x;
    while (u >= 3) {
      s = s - 1;
      u = x + s;
      z = z * u;
    }
    while (u < 2) {
      z = z / u;
      s = s + 1;
      u = x + s;
    }
    if (z < 0) z = -z;
    if (u == 2) return std::log(z);
    s = s - 2;
    const double y = x + s;
    return std::log(z) + y * horner(y, lg::kNum, 5) / horner(y, lg::kDen, 6);
  }
  double q = (x - 0.5) * std::log(x) - x + lg::kHalfLog2Pi;
  if (x > 100000000) return q;
  const double p = 1 / (x * x);
  if (x >= 1000.0) return q + ((lg::kStLarge0 * p - lg::kStLarge1) * p + lg::kStLarge2) / x;
  return q + horner(p, lg::kSt, 4) / x;
}

// Cephes gamma (positive arguments) and its log, used by the incomplete beta integral.
namespace cg {
constexpr double kP[7] = {1.60119522476751861407E-4, 1.19135147006586384913E-3, 1.04213797561761569935E-2,
                          4.76367800457137231464E-2, 2.07448227648435975150E-1, 4.94214826801497100753E-1,
                          9.99999999999999996796E-1};
constexpr double kQ[8] = {-2.31581873324120129819E-5, 5.39605580493303397842E-4, -4.45641913851797240494E-3,
                          1.18139785222060435552E-2,  3.58236398605498653373E-2, -2.34591795718243348568E-1,
                          7.14304917030273074085E-2,  1.00000000000000000320E0};
constexpr double kStir[5] = {7.87311395793093628397E-4, -2.29549961613378126380E-4, -2.68132617805781232825E-3,
                             3.47222221605458667310E-3, 8.33333333333482257126E-2};
constexpr double kA[5] = {8.11614167470508450300E-4, -5.95061904284301438324E-4, 7.93650340457716943945E-4,
                          -2.77777777730099687205E-3, 8.33333333333331927722E-2};
constexpr double kB[6] = {-1.37825152569120859100E3, -3.88016315134637840924E4, -3.31612992738871184744E5,
                          -1.16237097492762307383E6, -1.72173700820839662146E6, -8.53555664245765465627E5};
constexpr double kC[6] = {-3.51815701436523470549E2, -1.70642106651881159223E4, -2.20528590553854454839E5,
                          -1.13933444367982507207E6, -2.53252307177582951285E6, -2.01889141433532773231E6};
constexpr double kMaxGam = 171.624376956302725, kMaxStir = 143.01608, kSqrt2Pi = 2.50662827463100050242E0;
constexpr double kHalfLog2Pi = 0.91893853320467274178, kMaxLgm = 2.556348e305;
constexpr double kMaxLog = 7.09782712893383996732E2, kMinLog = -7.451332191019412076235E2;
constexpr double kEps = 1.11022302462515654042E-16, kBig = 4.503599627370496e15, kBigInv = 2.22044604925031308085e-16;

// Gamma(x) for x > 33 by Stirling's formula
static double stirling(double x) {
  const double iw = 1.0 / x;
  const double w = 1.0 + iw * horner(iw, kStir, 4);
  const double e = std::exp(x);
  double y;
  if (x > kMaxStir) {
    const double v = std::pow(x, 0.5 * x - 0.25);
    y = v * (v / e);
  } else {
    y = std::pow(x, x - 0.5) / e;
  }
  return kSqrt2Pi * y * w;
}

static double gamma_pos(double x) {
  if (x > 33.0) return stirling(x);
  double z = 1.0;
  while (x >= 3.0) {
    x -= 1.0;
    z *= x;
  }
  while (x < 2.0) {
    if (x < 1.e-9) return (x == 0.0) ? INFINITY : z / ((1.0 + 0.5772156649015329 * x) * x);
    z /= x;
    x += 1.0;
  }
  if (x == 2.0) return z;
  x -= 2.0;
  return z * horner(x, kP, 6) / horner(x, kQ, 7);
}

static double lgam_pos(double x) {
  if (x < 13.0) {
    double z = 1.0, s = 0.0, u = x;
    while (u >= 3.0) {
      s -= 1.0;
      u = x + s;
      z *= u;
    }
    while (u < 2.0) {
      if (u == 0.0) return INFINITY;
      z /= u;
      s += 1.0;
      u = x + s;
    }
    if (z < 0.0) z = -z;
    if (u == 2.0) return std::log(z);
    s -= 2.0;
    const double y = x + s;
    return std::log(z) + y * horner(y, kB, 5) / horner_monic(y, kC, 6);
  }
  if (x > kMaxLgm) return INFINITY;
  double q = (x - 0.5) * std::log(x) - x + kHalfLog2Pi;
  if (x > 1.0e8) return q;
  const double p = 1.0 / (x * x);
  if (x >= 1000.0)
    return q + ((7.9365079365079365079365e-4 * p - 2.7777777777777777777778e-3) * p + 0.0833333333333333333333) / x;
  return q + horner(p, kA, 4) / x;
}

// Gamma(a+b) / (Gamma(a) Gamma(b)) x^a ... prefactor helpers of the incomplete beta integral
static double power_series(double a, double b, double x) {
  const double ai = 1.0 / a;
  double u = (1.0 - b) * x;
  double v = u / (a + 1.0);
  const double first = v;
  double t = u, n = 2.0, s = 0.0;
  const double stop = kEps * ai;
  while (std::fabs(v) > stop) {
    u = (n - b) * x / n;
    t *= u;
    v = t / (a + n);
    s += v;
    n += 1.0;
  }
  s += first;
  s += ai;
  u = a * std::log(x);
  if ((a + b) < kMaxGam && std::fabs(u) < kMaxLog) return s * (gamma_pos(a + b) / (gamma_pos(a) * gamma_pos(b))) * std::pow(x, a);
  t = lgam_pos(a + b) - lgam_pos(a) - lgam_pos(b) + u + std::log(s);
  return (t < kMinLog) ? 0.0 : std::exp(t);
}

// Continued-fraction expansions of the incomplete beta integral.  Both are the same three-term
// recurrence p_k = p_{k-1} + d_k p_{k-2} (and q likewise) with two partial numerators per step;
// they differ in the argument (x, or x / (1-x)) and in how the numerators' factors move.
struct CfSpec {
  double z;
  double f1, f2, f3, f4, f5, f6, f7, f8;       // starting factors
  double df2, df6;                           // per-step change of f2 and f6 (the others are fixed)
};

static double continued_fraction(const CfSpec& c) {
  double f1 = c.f1, f2 = c.f2, f3 = c.f3, f4 = c.f4, f5 = c.f5, f6 = c.f6, f7 = c.f7, f8 = c.f8;
  double pm2 = 0.0, qm2 = 1.0, pm1 = 1.0, qm1 = 1.0, ans = 1.0, r = 1.0;
  const double thresh = 3.0 * kEps;
  for (int n = 0; n < 300; n++) {
    double d = -(c.z * f1 * f2) / (f3 * f4);
    double pk = pm1 + pm2 * d, qk = qm1 + qm2 * d;
    pm2 = pm1;
    pm1 = pk;
    qm2 = qm1;
    qm1 = qk;
    d = (c.z * f5 * f6) / (f7 * f8);
    pk = pm1 + pm2 * d;
    qk = qm1 + qm2 * d;
    pm2 = pm1;
    pm1 = pk;
    qm2 = qm1;
    qm1 = qk;
    if (qk != 0) r = pk / qk;
    double t;
    if (r != 0) {
      t = std::fabs((ans - r) / r);
      ans = r;
    } else {
      t = 1.0;
    }
    if (t < thresh) return ans;
    f1 += 1.0;
    f2 += c.df2;
    f3 += 2.0;
    f4 += 2.0;
    f5 += 1.0;
    f6 += c.df6;
    f7 += 2.0;
    f8 += 2.0;
    if ((std::fabs(qk) + std::fabs(pk)) > kBig) {
      pm2 *= kBigInv;
      pm1 *= kBigInv;
      qm2 *= kBigInv;
      qm1 *= kBigInv;
    }
    if ((std::fabs(qk) < kBigInv) || (std::fabs(pk) < kBigInv)) {
      pm2 *= kBig;
      pm1 *= kBig;
      qm2 *= kBig;
      qm1 *= kBig;
    }
  }
  return ans;
}

// Regularized incomplete beta integral I_x(a, b), a, b > 0.
static double inc_beta(double a0, double b0, double x0) {
  if (a0 <= 0.0 || b0 <= 0.0) return 0.0;
  if (x0 <= 0.0) return 0.0;
  if (x0 >= 1.0) return (x0 == 1.0) ? 1.0 : 0.0;
  if ((b0 * x0) <= 1.0 && x0 <= 0.95) return power_series(a0, b0, x0);
  // above the mean the symmetric integral 1 - I_{1-x}(b, a) converges faster
  const bool swap = x0 > (a0 / (a0 + b0));
  const double a = swap ? b0 : a0, b = swap ? a0 : b0;
  const double x = swap ? 1.0 - x0 : x0, xc = swap ? x0 : 1.0 - x0;
  double t;
  if (swap && (b * x) <= 1.0 && x <= 0.95) {
    t = power_series(a, b, x);
  } else {
    double w;
    if (x * (a + b - 2.0) - (a - 1.0) < 0.0)
      w = continued_fraction(CfSpec{x, a, a + b, a, a + 1.0, 1.0, b - 1.0, a + 1.0, a + 2.0, 1.0, -1.0});
    else
      w = continued_fraction(CfSpec{x / (1.0 - x), a, b - 1.0, a, a + 1.0, 1.0, a + b, a + 1.0, a + 2.0, -1.0, 1.0}) / xc;
    double y = a * std::log(x);
    t = b * std::log(xc);
    if ((a + b) < kMaxGam && std::fabs(y) < kMaxLog && std::fabs(t) < kMaxLog) {
      t = std::pow(xc, b);
      t *= std::pow(x, a);
      t /= a;
      t *= w;
      t *= gamma_pos(a + b) / (gamma_pos(a) * gamma_pos(b));
    } else {
      y += t + lgam_pos(a + b) - lgam_pos(a) - lgam_pos(b);
      y += std::log(w / a);
      t = (y < kMinLog) ? 0.0 : std::exp(y);
    }
  }
  if (swap) t = (t <= kEps) ? 1.0 - kEps : 1.0 - t;
  return t;
}
}  // namespace cg

// P(X <= k) of a negative binomial with `size` successes and success probability p
double nb_cdf(double k, double size, double p) { return cg::inc_beta(size, k + 1, p); }

// smallest k with P(X <= k) >= target, (size, mu) parametrization (stats.cpp:2047-2071)
uint32_t nb_quantile(double target, double size, double mu) {
  const double p = size / (size + mu);
  if (target <= 0.0) return 0;
  uint64_t hi = 1;
  while (nb_cdf((double)hi, size, p) < target) hi *= 2;
  uint64_t lo = hi / 2;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (nb_cdf((double)mid, size, p) < target) lo = mid + 1;
    else hi = mid;
  }
  return (uint32_t)lo;
}

// ---- the censored least-squares objective ---------------------------------------------------

// The fitting window [sf, ef] of the (possibly coarse-grained) histogram with what every
// evaluation reuses: the observed shares x_i / inner and log Gamma(i + 1).
struct Window {
  uint32_t sf = 0, ef = 0, per_bin = 1;
  double inner = 0, mean_estimate = 0;
  std::vector<double> x;        // coarse-grained counts, indexed 0..ef
  std::vector<double> share;    // x_i / inner, i = sf..ef (offset by sf)
  std::vector<double> lg_k1;    // log Gamma(i + 1)
  int bins() const { return (int)(ef - sf + 1); }
};

// sum_i (x_i/inner - d_i/D)^2 over the window, d_i = NB pmf(i; size = e^ls, mu = e^lm),
// D = sum_i d_i; 1e10 when a parameter or a sum is not finite.  `dist` is caller scratch of
// bins() doubles.
static double objective(const Window& W, double lm, double ls, double* dist) {
  const double mu = std::exp(lm), size = std::exp(ls);
  if (!std::isfinite(mu) || !std::isfinite(size)) return 1e10;
  const double p = size / (size + mu);
  const double lg_size = log_gamma(size);
  const double log_p = std::log(p), log_q = std::log(1.0 - p);
  const double size_log_p = size * log_p;
  const int nb = W.bins();
  double total = 0;
  for (int j = 0; j < nb; j++) {
    const double k = (double)(W.sf + (uint32_t)j);
    const double lpmf = log_gamma(k + size) - lg_size - W.lg_k1[j] + size_log_p + k * log_q;
    dist[j] = std::exp(lpmf);
    total += dist[j];
  }
  if (!(total > 0) || !std::isfinite(total)) return 1e10;
  double l = 0;
  for (int j = 0; j < nb; j++) {
    const double d = W.share[j] - dist[j] / total;
    l += d * d;
  }
  return std::isfinite(l) ? l : 1e10;
}

// ---- Nelder-Mead on two parameters ------------------------------------------------------------

struct Restart {
  double lm0, ls0;       // starting point (log mu, log size)
  double lm, ls;         // estimate
  double value;          // objective at the estimate (when converged)
  bool converged;
};

// The reference's simplex method (stats.cpp:2289-2392: reflection 1, expansion 2, contraction
// 0.5, shrink 0.5, at most 1000 steps, converged when the value spread is below 1e-8), on three
// fixed vertices.  The vertices are reordered by value each step as libstdc++'s insertion sort
// orders three elements.
static void simplex_minimize(const Window& W, Restart& R, double* dist) {
  double v[3][2] = {{R.lm0, R.ls0}, {R.lm0, R.ls0}, {R.lm0, R.ls0}};
  v[1][0] += (R.lm0 != 0.0) ? 0.05 * std::fabs(R.lm0) : 0.00025;
  v[2][1] += (R.ls0 != 0.0) ? 0.05 * std::fabs(R.ls0) : 0.00025;
  double f[3];
  for (int i = 0; i < 3; i++) f[i] = objective(W, v[i][0], v[i][1], dist);
  bool conv = false;
  for (int it = 0; it < 1000; it++) {
    int o[3] = {0, 1, 2};
    for (int i = 1; i < 3; i++) {
      const int val = o[i];
      if (f[val] < f[o[0]]) {
        for (int j = i; j > 0; j--) o[j] = o[j - 1];
        o[0] = val;
      } else {
        int j = i;
        while (f[val] < f[o[j - 1]]) {
          o[j] = o[j - 1];
          j--;
        }
        o[j] = val;
      }
    }
    double sv[3][2], sf[3];
    for (int i = 0; i < 3; i++) {
      sv[i][0] = v[o[i]][0];
      sv[i][1] = v[o[i]][1];
      sf[i] = f[o[i]];
    }
    std::memcpy(v, sv, sizeof v);
    std::memcpy(f, sf, sizeof f);
    if ((f[2] - f[0]) < 1e-8) {
      conv = true;
      break;
    }
    double c[2];
    for (int j = 0; j < 2; j++) c[j] = ((0.0 + v[0][j]) + v[1][j]) / 2.0;
    const double xr[2] = {c[0] + 1.0 * (c[0] - v[2][0]), c[1] + 1.0 * (c[1] - v[2][1])};
    const double fr = objective(W, xr[0], xr[1], dist);
    if (fr < f[0]) {
      const double xe[2] = {c[0] + 2.0 * (xr[0] - c[0]), c[1] + 2.0 * (xr[1] - c[1])};
      const double fe = objective(W, xe[0], xe[1], dist);
      const bool take_e = fe < fr;
      v[2][0] = take_e ? xe[0] : xr[0];
      v[2][1] = take_e ? xe[1] : xr[1];
      f[2] = take_e ? fe : fr;
    } else if (fr < f[1]) {
      v[2][0] = xr[0];
      v[2][1] = xr[1];
      f[2] = fr;
    } else {
      const double xc[2] = {c[0] + 0.5 * (v[2][0] - c[0]), c[1] + 0.5 * (v[2][1] - c[1])};
      const double fc = objective(W, xc[0], xc[1], dist);
      if (fc < f[2]) {
        v[2][0] = xc[0];
        v[2][1] = xc[1];
        f[2] = fc;
      } else {
        for (int i = 1; i < 3; i++) {
          for (int j = 0; j < 2; j++) v[i][j] = v[0][j] + 0.5 * (v[i][j] - v[0][j]);
          f[i] = objective(W, v[i][0], v[i][1], dist);
        }
      }
    }
  }
  int best = 0;
  for (int i = 1; i < 3; i++)
    if (f[i] < f[best]) best = i;
  R.lm = v[best][0];
  R.ls = v[best][1];
  R.converged = conv && std::isfinite(f[best]);
  R.value = R.converged ? objective(W, R.lm, R.ls, dist) : HUGE_VAL;
}

// ---- the fit ------------------------------------------------------------------------------------

struct Fit {
  double size = 0, mu = 0, average = 0, variance = 0, relative_variance = 0, scale = 0;
  uint32_t censor_start = 0, censor_end = 0;
};

// Moments, censoring window and coarse-graining of histogram n[0..N] (coverage_distribution.cpp:
// 119-221).  Returns false when there is nothing to fit (empty histogram or an empty window).
static bool prepare_window(const std::vector<double>& n, uint32_t N, Fit& r, Window& W) {
  double total = 0;
  for (uint32_t i = 1; i <= N; i++) total += n[i];
  if (total == 0) return false;
  double m = 0;
  for (uint32_t i = 1; i <= N; i++) m += i * n[i];
  m /= total;
  double v = 0;
  if (total > 1) {
    for (uint32_t i = 1; i <= N; i++) v += n[i] * (i - m) * (i - m);
    v /= (total - 1);
  }
  r.average = m;
  r.variance = v;
  r.relative_variance = (m > 0) ? v / m : 0.0;
  // the peak of the five-point moving average (defined 3..N-2), or of the raw counts for N < 5,
  // at or above a quarter of the mean; first maximum wins
  const uint32_t min_i = std::max((uint32_t)(m / 4.0), (uint32_t)1);
  double peak = 0;
  uint32_t peak_i = 0;
  for (uint32_t i = min_i; i <= N; i++) {
    double a;
    if (N >= 5) {
      if (i < 3 || i > N - 2) continue;
      a = (n[i - 2] + n[i - 1] + n[i] + n[i + 1] + n[i + 2]) / 5.0;
    } else {
      a = n[i];
    }
    if (a > peak) {
      peak = a;
      peak_i = i;
    }
  }
  const uint32_t start_i = std::max((uint32_t)std::floor(peak_i * 0.5), (uint32_t)1);
  const uint32_t end_i = std::min((uint32_t)std::ceil(peak_i * 1.5), N);
  if (start_i == end_i) return false;
  r.censor_start = start_i;
  r.censor_end = end_i;
  // coarse-graining to at most ~1000-2000 bins
  uint32_t per_bin = (end_i - start_i) / 1000;
  if (per_bin > 1) {
    W.sf = start_i / per_bin;
    W.ef = (uint32_t)std::ceil((double)end_i / per_bin);
    W.x.assign(W.ef + 1, 0.0);
    for (uint32_t i = W.sf; i <= W.ef; i++)
      for (uint32_t j = 1; j <= per_bin; j++) {
        const uint32_t idx = i * per_bin + j;
        if (idx <= N) W.x[i] += n[idx];
      }
  } else {
    per_bin = 1;
    W.sf = start_i;
    W.ef = end_i;
    W.x.assign(W.ef + 1, 0.0);
    for (uint32_t i = 1; i <= W.ef; i++) W.x[i] = n[i];
  }
  W.per_bin = per_bin;
  W.inner = 0;
  for (uint32_t i = W.sf; i <= W.ef; i++) W.inner += W.x[i];
  double num = 0, den = 0;
  for (uint32_t i = 1; i <= W.ef; i++) {
    num += i * W.x[i];
    den += W.x[i];
  }
  W.mean_estimate = num / den;
  const int nb = W.bins();
  W.share.resize((size_t)nb);
  W.lg_k1.resize((size_t)nb);
  for (int j = 0; j < nb; j++) {
    const uint32_t i = W.sf + (uint32_t)j;
    W.share[(size_t)j] = W.x[i] / W.inner;
    W.lg_k1[(size_t)j] = log_gamma((double)i + 1);
  }
  return true;
}

static Fit fit_censored(const std::vector<double>& n, uint32_t N, int threads) {
  Fit r;
  Window W;
  if (!prepare_window(n, N, r, W)) return r;
  // restarts in the reference's order: six starting means, each with sizes 10^4 ... 10^-3 formed by
  // repeated division
  const double sum_ends = (double)(W.ef + W.sf);
  const double means[6] = {W.mean_estimate, (double)W.ef, (double)W.sf, 1.0 * sum_ends / 4.0,
                           2.0 * sum_ends / 4.0, 3.0 * sum_ends / 4.0};
  std::vector<Restart> rs;
  for (double tm : means) {
    double ts = 100000;
    for (int k = 0; k < 8; k++) {
      ts /= 10.0;
      Restart R{};
      R.lm0 = std::log(tm);
      R.ls0 = std::log(ts);
      rs.push_back(R);
    }
  }
  auto restart = [&](int i) {
    thread_local std::vector<double> dist;
    if (dist.size() < (size_t)W.bins()) dist.resize((size_t)W.bins());
    simplex_minimize(W, rs[(size_t)i], dist.data());
  };
  if (threads > 1) {
    pool_run((int)rs.size(), restart);
  } else {
    for (int i = 0; i < (int)rs.size(); i++) restart(i);
  }
  double best = HUGE_VAL, mu = 0, size = 0;
  for (const Restart& R : rs) {
    if (!R.converged) continue;
    if (R.value < best) {
      best = R.value;
      mu = std::exp(R.lm);
      size = std::exp(R.ls);
    }
  }
  if (!(best < HUGE_VAL)) mu = size = 0;
  double incl = 0;
  if (mu > 0) {
    const double p = size / (size + mu);
    incl = nb_cdf((double)W.ef, size, p) - nb_cdf((double)W.sf, size, p);
    if (incl >= 0.01 && W.per_bin > 1) mu = mu * W.per_bin;
  }
  if (incl < 0.01) mu = size = 0;
  r.mu = mu;
  r.size = size;
  if (mu > 0) r.scale = W.inner / incl;
  return r;
}

}  // namespace nb

// The deletion propagation cutoff of one coverage group (coverage_distribution.cpp:345-399):
// the (0.05/sqrt(length)) quantile of the fitted NB; fallbacks when the fit failed.
void fit_coverage_group(const std::vector<double>& n, uint32_t N, uint32_t group_length, int threads,
                        bsq_coverage_fit& out, double& scale) {
  const nb::Fit f = nb::fit_censored(n, N, threads);
  const double pr_cutoff = 0.05 / std::sqrt(group_length);
  const double m = f.average, v = f.variance;
  double dp;
  if (f.mu > 0) {
    dp = (double)nb::nb_quantile(pr_cutoff, f.size, f.mu);
  } else {
    // variance = mu + mu^2 / size, solved for size
    const double se = (1.0 / (v - m)) * m * m;
    dp = ((se > 0) && std::isfinite(se)) ? (double)nb::nb_quantile(pr_cutoff, se, m) : -1;
    if (!(dp >= 1)) dp = m * 0.1;
  }
  if (dp < 1) dp = 1;
  if ((f.mu <= 3) && (m <= 3)) dp = -1;   // both fit and mean very low: the sequence is deleted
  out.nb_fit_size = f.size;
  out.nb_fit_mu = f.mu;
  out.average = f.average;
  out.variance = f.variance;
  out.relative_variance = f.relative_variance;
  out.deletion_coverage_propagation_cutoff = dp;
  out.censor_start = f.censor_start;
  out.censor_end = f.censor_end;
  scale = f.scale;
}

}  // namespace bsq

using namespace bsq;

extern "C" int bsq_fit_coverage_distribution(const uint32_t* hist, int64_t len, double sequence_length, bsq_coverage_fit* out,
                                             bsq_status* status) {
  try {
    if (!out || (len > 0 && !hist) || !(sequence_length > 0)) throw BsqError(BSQ_ERR_ARG, "bad arguments");
    *out = bsq_coverage_fit{};
    // the histogram as read back from its file (coverage_distribution.cpp:34-65): rows j >= 1
    const uint32_t N = (len > 1) ? (uint32_t)(len - 1) : 0;
    std::vector<double> n((size_t)N + 1, 0.0);
    for (int64_t j = 1; j < len; j++) n[(size_t)j] = hist[j];
    double scale = 0;
    fit_coverage_group(n, N, (uint32_t)sequence_length, host_threads(), *out, scale);
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    set_status(status, e.code, e.what());
    return e.code;
  }
}


namespace bsq {

// The coverage-distribution stage over all coverage groups (coverage_distribution.cpp:510-633):
// per group, the summed length of its sequences (as 32-bit, like the reference's uint32_t sum),
// the fit of its histogram and the summary of every member.  Targeted sequencing fits nothing.
void coverage_stage(const bsq_coverage_stage_args* a, bsq_coverage_summary* per_target, int threads) {
  if (!a || !per_target || a->n_targets < 0 || (a->n_targets > 0 && !a->target_lengths))
    throw BsqError(BSQ_ERR_ARG, "bad coverage stage arguments");
  for (int t = 0; t < a->n_targets; t++) per_target[t] = bsq_coverage_summary{};
  if (a->targeted_sequencing) return;
  const int G = a->n_coverage_groups;
  if (G < 0 || (G > 0 && (!a->histograms || !a->histogram_lengths))) throw BsqError(BSQ_ERR_ARG, "bad coverage groups");
  for (int t = 0; t < a->n_targets; t++) {
    const int g = a->coverage_group ? a->coverage_group[t] : 0;
    if (g < 0 || g >= G) throw BsqError(BSQ_ERR_ARG, "coverage group of a target out of range");
  }
  for (int g = 0; g < G; g++) {
    uint32_t length = 0;
    for (int t = 0; t < a->n_targets; t++)
      if ((a->coverage_group ? a->coverage_group[t] : 0) == g) length += (uint32_t)a->target_lengths[t];
    const int64_t len = a->histogram_lengths[g];
    const double* h = a->histograms[g];
    if (len > 0 && !h) throw BsqError(BSQ_ERR_ARG, "missing coverage histogram");
    const uint32_t N = (len > 1) ? (uint32_t)(len - 1) : 0;
    std::vector<double> n((size_t)N + 1, 0.0);
    for (int64_t j = 1; j < len; j++) n[(size_t)j] = h[j];
    bsq_coverage_fit f{};
    double scale = 0;
    fit_coverage_group(n, N, length, threads, f, scale);
    bsq_coverage_summary s{};
    s.nbinom_size_parameter = f.nb_fit_size;
    s.nbinom_mean_parameter = f.nb_fit_mu;
    if (s.nbinom_mean_parameter != 0) {
      s.nbinom_prob_parameter = s.nbinom_size_parameter / (s.nbinom_mean_parameter + s.nbinom_size_parameter);
      s.nbinom_variance = s.nbinom_mean_parameter + std::pow(s.nbinom_mean_parameter, 2) / s.nbinom_size_parameter;
      s.nbinom_relative_variance = s.nbinom_variance / s.nbinom_mean_parameter;
    }
    s.average = f.average;
    s.variance = f.variance;
    s.relative_variance = f.relative_variance;
    s.deletion_coverage_propagation_cutoff = f.deletion_coverage_propagation_cutoff;
    for (int t = 0; t < a->n_targets; t++)
      if ((a->coverage_group ? a->coverage_group[t] : 0) == g) per_target[t] = s;
  }
}

// Mean of histogram rows j >= 1 in the fit's own summation order; the stage's cutoff can be
// negative only when this is at most 3.
double coverage_histogram_mean(const double* h, int64_t len) {
  double total = 0, m = 0;
  for (int64_t i = 1; i < len; i++) total += h[i];
  if (total == 0) return 0.0;
  for (int64_t i = 1; i < len; i++) m += (uint32_t)i * h[i];
  return m / total;
}

}  // namespace bsq

extern "C" int bsq_analyze_unique_coverage_distributions(const bsq_coverage_stage_args* args,
                                                         bsq_coverage_summary* per_target, bsq_status* status) {
  try {
    coverage_stage(args, per_target, host_threads());
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    set_status(status, e.code, e.what());
    return e.code;
  }
}

// The histogram file as `in >> coverage >> count` reads it after the header line: whitespace-
// separated (unsigned, double) pairs until the first that does not parse.
extern "C" int bsq_read_coverage_histogram(const char* text, int64_t text_len, double* out, int64_t cap, int64_t* len,
                                           bsq_status* status) {
  try {
    if ((!text && text_len > 0) || text_len < 0 || !len || (cap > 0 && !out)) throw BsqError(BSQ_ERR_ARG, "bad arguments");
    std::string s(text ? text : "", (size_t)text_len);
    size_t i = s.find('\n');
    i = (i == std::string::npos) ? s.size() : i + 1;
    std::vector<std::pair<uint32_t, double>> rows;
    const char* p = s.c_str() + i;
    for (;;) {
      char* e = nullptr;
      while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f') p++;
      if (*p == '-' || *p == 0) break;   // an unsigned extraction fails on a sign or at the end
      errno = 0;
      const unsigned long long c = strtoull(p, &e, 10);
      if (e == p || errno == ERANGE || c > 0xFFFFFFFFull) break;
      p = e;
      const double v = strtod(p, &e);
      if (e == p) break;
      p = e;
      rows.emplace_back((uint32_t)c, v);
    }
    uint32_t mx = 0;
    for (const auto& r : rows) mx = std::max(mx, r.first);
    const int64_t n = rows.empty() ? 1 : (int64_t)mx + 1;
    *len = n;
    if (cap > 0) {
      for (int64_t j = 0; j < std::min(cap, n); j++) out[j] = 0.0;
      for (const auto& r : rows)
        if ((int64_t)r.first < cap) out[r.first] = r.second;
    }
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    set_status(status, e.code, e.what());
    return e.code;
  }
}
