/*
 * covfit_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Single-threaded CPU restatement of breseq's coverage-distribution fit, the stage between
 * error_count and identify_mutations that turns the unique-only coverage histogram into the
 * deletion propagation cutoff.  It follows the reference's arithmetic step for step (built with
 * -ffp-contract=off), including the Cephes special functions breseq carries in stats.cpp, so that
 * its results are what the reference computes on the same histogram.
 *
 * Function                                 reference                          pinned by (tests/...)
 * -------------------------------------------------------------------------------------------------
 * read_coverage_histogram (rows j >= 1)    coverage_distribution.cpp:34-65    test_covfit_oracle.py
 *                                                                             (rows by hand)
 * lngamma (ALGLIB/Cephes log gamma)        stats.cpp:534-650                  scipy.special.gammaln
 * gamma, stirf, lgam, polevl, p1evl        stats.cpp:258-498 (Cephes 2.8)     scipy.special.gamma / gammaln
 * incbet, incbcf, incbd, pseries, nbdtr    stats.cpp:982-991, 1385-1660       scipy.special.betainc,
 *                                                                             scipy.stats.nbinom.cdf
 * dnbinom_mu                               stats.h:92-100                     scipy.stats.nbinom.pmf
 * qnbinom_mu                               stats.cpp:2047-2071                scipy.stats.nbinom.ppf
 * nelder_mead_minimize                     stats.cpp:2289-2392                quadratic bowl minimum;
 *                                                                             the fit's optimum vs scipy
 * fit_censored_negative_binomial           coverage_distribution.cpp:115-326  window and moments by hand;
 *                                                                             objective minimum vs scipy
 * fit_coverage_distribution (cutoff and    coverage_distribution.cpp:345-399  fallbacks by hand (failed fit,
 *   its fallbacks)                                                            low coverage, 10% of mean)
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <vector>

#include "oracle.h"

namespace orc_fit {

using std::vector;

// ---- Cephes constants and tables (stats.cpp:117-255) ----
static const double kPi = 3.14159265358979323846264338327950288419716939937510582097494459230781640628620899862803482534211706;
static const double kGamP[] = {1.60119522476751861407E-4, 1.19135147006586384913E-3, 1.04213797561761569935E-2,
                               4.76367800457137231464E-2, 2.07448227648435975150E-1, 4.94214826801497100753E-1,
                               9.99999999999999996796E-1};
static const double kGamQ[] = {-2.31581873324120129819E-5, 5.39605580493303397842E-4, -4.45641913851797240494E-3,
                               1.18139785222060435552E-2,  3.58236398605498653373E-2, -2.34591795718243348568E-1,
                               7.14304917030273074085E-2,  1.00000000000000000320E0};
static const double kMaxGam = 171.624376956302725;
static const double kStir[5] = {7.87311395793093628397E-4, -2.29549961613378126380E-4, -2.68132617805781232825E-3,
                                3.47222221605458667310E-3, 8.33333333333482257126E-2};
static const double kMaxStir = 143.01608;
static const double kSqtPi = 2.50662827463100050242E0;
static const double kLs2Pi = 0.91893853320467274178;
static const double kMaxLgm = 2.556348e305;
static const double kLgA[] = {8.11614167470508450300E-4, -5.95061904284301438324E-4, 7.93650340457716943945E-4,
                              -2.77777777730099687205E-3, 8.33333333333331927722E-2};
static const double kLgB[] = {-1.37825152569120859100E3, -3.88016315134637840924E4, -3.31612992738871184744E5,
                              -1.16237097492762307383E6, -1.72173700820839662146E6, -8.53555664245765465627E5};
static const double kLgC[] = {-3.51815701436523470549E2, -1.70642106651881159223E4, -2.20528590553854454839E5,
                              -1.13933444367982507207E6, -2.53252307177582951285E6, -2.01889141433532773231E6};
static const double kMaxLog = 7.09782712893383996732E2;
static const double kMinLog = -7.451332191019412076235E2;
static const double kMachEp = 1.11022302462515654042E-16;
static const double kBig = 4.503599627370496e15;
static const double kBigInv = 2.22044604925031308085e-16;

static double polevl(double x, const double* c, int n) {
  double a = c[0];
  for (int i = 1; i <= n; i++) a = a * x + c[i];
  return a;
}

static double p1evl(double x, const double* c, int n) {
  double a = x + c[0];
  for (int i = 1; i < n; i++) a = a * x + c[i];
  return a;
}

static double stirf(double x) {
  double w = 1.0 / x;
  w = 1.0 + w * polevl(w, kStir, 4);
  double y = std::exp(x);
  if (x > kMaxStir) {
    const double v = std::pow(x, 0.5 * x - 0.25);
    y = v * (v / y);
  } else {
    y = std::pow(x, x - 0.5) / y;
  }
  return kSqtPi * y * w;
}

// Cephes gamma (stats.cpp:320-398), positive and negative arguments
static double cgamma(double x) {
  int sg = 1;
  double p, q = std::fabs(x), z;
  if (q > 33.0) {
    if (x < 0.0) {
      p = std::floor(q);
      if (p == q) return NAN;
      const uint32_t i = (uint32_t)p;
      if ((i & 1) == 0) sg = -1;
      z = q - p;
      if (z > 0.5) {
        p += 1.0;
        z = q - p;
      }
      z = q * std::sin(kPi * z);
      if (z == 0.0) return sg * INFINITY;
      z = std::fabs(z);
      z = kPi / (z * stirf(q));
    } else {
      z = stirf(x);
    }
    return sg * z;
  }
  z = 1.0;
  while (x >= 3.0) {
    x -= 1.0;
    z *= x;
  }
  while (x < 0.0) {
    if (x > -1.E-9) goto small;
    z /= x;
    x += 1.0;
  }
  while (x < 2.0) {
    if (x < 1.e-9) goto small;
    z /= x;
    x += 1.0;
  }
  if (x == 2.0) return z;
  x -= 2.0;
  p = polevl(x, kGamP, 6);
  q = polevl(x, kGamQ, 7);
  return z * p / q;
small:
  if (x == 0.0) return INFINITY;
  return z / ((1.0 + 0.5772156649015329 * x) * x);
}

// Cephes lgam (stats.cpp:409-498) for positive arguments (all the fit uses)
static double lgam(double x) {
  double p, q, u, z;
  if (x < 13.0) {
    z = 1.0;
    p = 0.0;
    u = x;
    while (u >= 3.0) {
      p -= 1.0;
      u = x + p;
      z *= u;
    }
    while (u < 2.0) {
      if (u == 0.0) return INFINITY;
      z /= u;
      p += 1.0;
      u = x + p;
    }
    if (z < 0.0) z = -z;
    if (u == 2.0) return std::log(z);
    p -= 2.0;
    x = x + p;
    p = x * polevl(x, kLgB, 5) / p1evl(x, kLgC, 6);
    return std::log(z) + p;
  }
  if (x > kMaxLgm) return INFINITY;
  q = (x - 0.5) * std::log(x) - x + kLs2Pi;
  if (x > 1.0e8) return q;
  p = 1.0 / (x * x);
  if (x >= 1000.0)
    q += ((7.9365079365079365079365e-4 * p - 2.7777777777777777777778e-3) * p + 0.0833333333333333333333) / x;
  else
    q += polevl(p, kLgA, 4) / x;
  return q;
}

// lngamma (stats.cpp:534-650): the log gamma the NB pmf uses, positive arguments.  The reference
// writes its polynomials as Horner chains with literal coefficients, some given as a product with
// a power of ten; those products are folded to the same doubles here.
static double lngamma(double x) {
  const double logpi = 1.14472988584940017414;
  const double ls2pi = 0.91893853320467274178;
  if (x < -34.0) {
    const double q = -x;
    const double w = lngamma(q);
    double p = (int32_t)std::floor(q);
    double z = q - p;
    if (z > 0.5) {
      p = p + 1;
      z = p - q;
    }
    z = q * std::sin(M_PI * z);
    return logpi - std::log(z) - w;
  }
  if (x < 13) {
    double z = 1, p = 0, u = x;
    while (u >= 3) {
      p = p - 1;
      u = x + p;
      z = z * u;
    }
    while (u < 2) {
      z = z / u;
      p = p + 1;
      u = x + p;
    }
    if (z < 0) z = -z;
    if (u == 2) return std::log(z);
    p = p - 2;
    x = x + p;
    double b = -1378.25152569120859100;
    b = -38801.6315134637840924 + x * b;
    b = -331612.992738871184744 + x * b;
    b = -1162370.97492762307383 + x * b;
    b = -1721737.00820839662146 + x * b;
    b = -853555.664245765465627 + x * b;
    double c = 1;
    c = -351.815701436523470549 + x * c;
    c = -17064.2106651881159223 + x * c;
    c = -220528.590553854454839 + x * c;
    c = -1139334.44367982507207 + x * c;
    c = -2532523.07177582951285 + x * c;
    c = -2018891.41433532773231 + x * c;
    p = x * b / c;
    return std::log(z) + p;
  }
  double q = (x - 0.5) * std::log(x) - x + ls2pi;
  if (x > 100000000) return q;
  const double p = 1 / (x * x);
  if (x >= 1000.0) {
    q = q + ((7.9365079365079365079365 * 0.0001 * p - 2.7777777777777777777778 * 0.001) * p + 0.0833333333333333333333) / x;
  } else {
    double a = 8.11614167470508450300 * 0.0001;
    a = -5.95061904284301438324 * 0.0001 + p * a;
    a = 7.93650340457716943945 * 0.0001 + p * a;
    a = -2.77777777730099687205 * 0.001 + p * a;
    a = 8.33333333333331927722 * 0.01 + p * a;
    q = q + a / x;
  }
  return q;
}

static double pseries(double a, double b, double x) {
  const double ai = 1.0 / a;
  double u = (1.0 - b) * x;
  double v = u / (a + 1.0);
  const double t1 = v;
  double t = u, n = 2.0, s = 0.0;
  const double z = kMachEp * ai;
  while (std::fabs(v) > z) {
    u = (n - b) * x / n;
    t *= u;
    v = t / (a + n);
    s += v;
    n += 1.0;
  }
  s += t1;
  s += ai;
  u = a * std::log(x);
  if ((a + b) < kMaxGam && std::fabs(u) < kMaxLog) {
    t = cgamma(a + b) / (cgamma(a) * cgamma(b));
    s = s * t * std::pow(x, a);
  } else {
    t = lgam(a + b) - lgam(a) - lgam(b) + u + std::log(s);
    s = (t < kMinLog) ? 0.0 : std::exp(t);
  }
  return s;
}

// the two continued fractions of incbet (stats.cpp:1475-1630); `second` selects expansion #2
static double incb_cf(double a, double b, double x, bool second) {
  double k1 = a, k2 = second ? b - 1.0 : a + b, k3 = a, k4 = a + 1.0, k5 = 1.0, k6 = second ? a + b : b - 1.0,
         k7 = a + 1.0, k8 = a + 2.0;
  if (!second) k7 = k4;
  double pkm2 = 0.0, qkm2 = 1.0, pkm1 = 1.0, qkm1 = 1.0, ans = 1.0, r = 1.0, t;
  const double z = second ? x / (1.0 - x) : x;
  const double thresh = 3.0 * kMachEp;
  int n = 0;
  do {
    double xk = -(z * k1 * k2) / (k3 * k4);
    double pk = pkm1 + pkm2 * xk;
    double qk = qkm1 + qkm2 * xk;
    pkm2 = pkm1;
    pkm1 = pk;
    qkm2 = qkm1;
    qkm1 = qk;
    xk = (z * k5 * k6) / (k7 * k8);
    pk = pkm1 + pkm2 * xk;
    qk = qkm1 + qkm2 * xk;
    pkm2 = pkm1;
    pkm1 = pk;
    qkm2 = qkm1;
    qkm1 = qk;
    if (qk != 0) r = pk / qk;
    if (r != 0) {
      t = std::fabs((ans - r) / r);
      ans = r;
    } else {
      t = 1.0;
    }
    if (t < thresh) return ans;
    k1 += 1.0;
    if (second) k2 -= 1.0;
    else k2 += 1.0;
    k3 += 2.0;
    k4 += 2.0;
    k5 += 1.0;
    if (second) k6 += 1.0;
    else k6 -= 1.0;
    k7 += 2.0;
    k8 += 2.0;
    if ((std::fabs(qk) + std::fabs(pk)) > kBig) {
      pkm2 *= kBigInv;
      pkm1 *= kBigInv;
      qkm2 *= kBigInv;
      qkm1 *= kBigInv;
    }
    if ((std::fabs(qk) < kBigInv) || (std::fabs(pk) < kBigInv)) {
      pkm2 *= kBig;
      pkm1 *= kBig;
      qkm2 *= kBig;
      qkm1 *= kBig;
    }
  } while (++n < 300);
  return ans;
}

// incbet (stats.cpp:1385-1472)
static double incbet(double aa, double bb, double xx) {
  if (aa <= 0.0 || bb <= 0.0) return 0.0;
  if (xx <= 0.0 || xx >= 1.0) {
    if (xx == 0.0) return 0.0;
    if (xx == 1.0) return 1.0;
    return 0.0;
  }
  double a, b, t, x, xc, w, y;
  bool flag = false;
  if ((bb * xx) <= 1.0 && xx <= 0.95) {
    t = pseries(aa, bb, xx);
    goto done;
  }
  w = 1.0 - xx;
  if (xx > (aa / (aa + bb))) {
    flag = true;
    a = bb;
    b = aa;
    xc = xx;
    x = w;
  } else {
    a = aa;
    b = bb;
    xc = w;
    x = xx;
  }
  if (flag && (b * x) <= 1.0 && x <= 0.95) {
    t = pseries(a, b, x);
    goto done;
  }
  y = x * (a + b - 2.0) - (a - 1.0);
  if (y < 0.0) w = incb_cf(a, b, x, false);
  else w = incb_cf(a, b, x, true) / xc;
  y = a * std::log(x);
  t = b * std::log(xc);
  if ((a + b) < kMaxGam && std::fabs(y) < kMaxLog && std::fabs(t) < kMaxLog) {
    t = std::pow(xc, b);
    t *= std::pow(x, a);
    t /= a;
    t *= w;
    t *= cgamma(a + b) / (cgamma(a) * cgamma(b));
    goto done;
  }
  y += t + lgam(a + b) - lgam(a) - lgam(b);
  y += std::log(w / a);
  t = (y < kMinLog) ? 0.0 : std::exp(y);
done:
  if (flag) t = (t <= kMachEp) ? 1.0 - kMachEp : 1.0 - t;
  return t;
}

static double nbdtr(double k, double n, double p) { return incbet(n, k + 1, p); }

static double dnbinom_mu(double k, double size, double mu) {
  const double p = size / (size + mu);
  const double lp = lngamma(k + size) - lngamma(size) - lngamma(k + 1) + size * std::log(p) + k * std::log(1.0 - p);
  return std::exp(lp);
}

static uint32_t qnbinom_mu(double target, double size, double mu) {
  const double p = size / (size + mu);
  if (target <= 0.0) return 0;
  uint64_t hi = 1;
  while (nbdtr((double)hi, size, p) < target) hi *= 2;
  uint64_t lo = hi / 2;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (nbdtr((double)mid, size, p) < target) lo = mid + 1;
    else hi = mid;
  }
  return (uint32_t)lo;
}

struct NmOut {
  vector<double> est;
  bool converged;
};

// nelder_mead_minimize (stats.cpp:2289-2392); max_iterations 1000, tolerance 1e-8 (stats.h defaults)
static NmOut nelder_mead(const std::function<double(const vector<double>&)>& f, const vector<double>& x0,
                         uint32_t max_iter = 1000, double tol = 1e-8) {
  const size_t n = x0.size();
  vector<vector<double>> sx(n + 1, x0);
  for (size_t i = 0; i < n; i++) sx[i + 1][i] += (x0[i] != 0.0) ? 0.05 * std::fabs(x0[i]) : 0.00025;
  vector<double> v(n + 1);
  for (size_t i = 0; i <= n; i++) v[i] = f(sx[i]);
  bool conv = false;
  for (uint32_t it = 0; it < max_iter; it++) {
    // std::sort on n+1 = 3 indices is libstdc++'s insertion sort: an element smaller than the
    // first goes to the front, any other moves left past larger elements only
    vector<size_t> ord(n + 1);
    for (size_t i = 0; i <= n; i++) ord[i] = i;
    for (size_t i = 1; i <= n; i++) {
      const size_t val = ord[i];
      if (v[val] < v[ord[0]]) {
        for (size_t j = i; j > 0; j--) ord[j] = ord[j - 1];
        ord[0] = val;
      } else {
        size_t j = i;
        while (v[val] < v[ord[j - 1]]) {
          ord[j] = ord[j - 1];
          j--;
        }
        ord[j] = val;
      }
    }
    vector<vector<double>> s2(n + 1);
    vector<double> v2(n + 1);
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
    vector<double> c(n, 0.0);
    for (size_t i = 0; i < n; i++)
      for (size_t j = 0; j < n; j++) c[j] += sx[i][j];
    for (size_t j = 0; j < n; j++) c[j] /= (double)n;
    vector<double> xr(n);
    for (size_t j = 0; j < n; j++) xr[j] = c[j] + 1.0 * (c[j] - sx[n][j]);
    const double fr = f(xr);
    if (fr < v[0]) {
      vector<double> xe(n);
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
      vector<double> xc(n);
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
  return NmOut{sx[best], conv && std::isfinite(v[best])};
}

struct FitOut {
  double size = 0, mu = 0, average = 0, variance = 0, relative_variance = 0, scale = 0;
  uint32_t censor_start = 0, censor_end = 0;
  uint32_t restarts_converged = 0;
  double best_objective = HUGE_VAL;
};

// fit_censored_negative_binomial (coverage_distribution.cpp:115-326)
static FitOut fit_censored_nb(const vector<double>& n, uint32_t N) {
  FitOut r;
  double total = 0;
  for (uint32_t i = 1; i <= N; i++) total += n[i];
  if (total == 0) return r;
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
  vector<double> ma(N + 1, 0.0);
  vector<bool> ok(N + 1, false);
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
  if (start_i == end_i) return r;
  r.censor_start = start_i;
  r.censor_end = end_i;
  uint32_t per_bin = (end_i - start_i) / 1000;
  vector<double> x;
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
  auto f_nb = [&](const vector<double>& par) -> double {
    const double mu = std::exp(par[0]), size = std::exp(par[1]);
    if (!std::isfinite(mu) || !std::isfinite(size)) return 1e10;
    vector<double> dist(ef + 1, 0.0);
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
  vector<double> tries = {mean_est, (double)ef, (double)sf, 1.0 * (ef + sf) / 4.0, 2.0 * (ef + sf) / 4.0,
                          3.0 * (ef + sf) / 4.0};
  double mu = 0, size = 0, best = HUGE_VAL;
  for (double tm : tries) {
    double ts = 100000;
    for (int k = 0; k < 8; k++) {
      ts /= 10.0;
      NmOut o = nelder_mead(f_nb, {std::log(tm), std::log(ts)});
      if (!o.converged) continue;
      r.restarts_converged++;
      const double obj = f_nb(o.est);
      if (obj < best) {
        best = obj;
        mu = std::exp(o.est[0]);
        size = std::exp(o.est[1]);
      }
    }
  }
  r.best_objective = best;
  if (!(best < HUGE_VAL)) mu = size = 0;
  double incl = 0;
  if (mu > 0) {
    const double p = size / (size + mu);
    incl = nbdtr((double)ef, size, p) - nbdtr((double)sf, size, p);
    if (incl >= 0.01 && per_bin > 1) mu = mu * per_bin;
  }
  if (incl < 0.01) mu = size = 0;
  r.mu = mu;
  r.size = size;
  if (mu > 0) r.scale = inner / incl;
  return r;
}

}  // namespace orc_fit

using namespace orc_fit;

extern "C" {

// fit_coverage_distribution (coverage_distribution.cpp:345-399) on a histogram given as the rows
// a coverage distribution file holds: hist[j] = positions with unique coverage j, rows j >= 1
// (read_coverage_histogram, :34-65; index 0 is ignored).  sequence_length is the summed length of
// the coverage group's sequences (analyze_unique_coverage_distribution, :524-530), whose
// 0.05 / sqrt(length) is the deletion propagation probability cutoff.  out[0..9] = nb_fit_size,
// nb_fit_mu, average, variance, relative_variance, deletion_coverage_propagation_cutoff,
// censor_start, censor_end, nb_fit_scale, restarts that converged.
int orc_fit_coverage_distribution(const uint32_t* hist, int64_t len, uint32_t sequence_length, double* out) {
  uint32_t N = 0;
  for (int64_t j = 1; j < len; j++) N = (uint32_t)j;
  vector<double> n(N + 1, 0.0);
  for (int64_t j = 1; j < len; j++) n[(size_t)j] = hist[j];
  const FitOut f = fit_censored_nb(n, N);
  const double pr_cutoff = 0.05 / std::sqrt(sequence_length);
  const double m = f.average, v = f.variance;
  double dp;
  if (f.mu > 0) {
    dp = (double)qnbinom_mu(pr_cutoff, f.size, f.mu);
  } else {
    const double se = (1.0 / (v - m)) * m * m;
    if ((se > 0) && std::isfinite(se)) dp = (double)qnbinom_mu(pr_cutoff, se, m);
    else dp = -1;
    if (!(dp >= 1)) dp = m * 0.1;
  }
  if (dp < 1) dp = 1;
  if ((f.mu <= 3) && (m <= 3)) dp = -1;
  out[0] = f.size;
  out[1] = f.mu;
  out[2] = f.average;
  out[3] = f.variance;
  out[4] = f.relative_variance;
  out[5] = dp;
  out[6] = f.censor_start;
  out[7] = f.censor_end;
  out[8] = f.scale;
  out[9] = f.restarts_converged;
  return 0;
}

// The special functions alone, for their pins: which = 0 lngamma(x), 1 Cephes gamma(x),
// 2 Cephes lgam(x), 3 incbet(a, b, x), 4 nbdtr(k, n, p), 5 dnbinom_mu(k, size, mu),
// 6 qnbinom_mu(target, size, mu).
double orc_special(int32_t which, double a, double b, double x) {
  switch (which) {
    case 0: return lngamma(a);
    case 1: return cgamma(a);
    case 2: return lgam(a);
    case 3: return incbet(a, b, x);
    case 4: return nbdtr(a, b, x);
    case 5: return dnbinom_mu(a, b, x);
    case 6: return (double)qnbinom_mu(a, b, x);
    default: return NAN;
  }
}

// nelder_mead_minimize on the 2-D quadratic (p0 - c0)^2 + w (p1 - c1)^2 from (x0, x1); writes
// est[2] and returns whether it converged.
int orc_nelder_mead_quadratic(double c0, double c1, double w, double x0, double x1, double* est) {
  auto f = [&](const vector<double>& p) { return (p[0] - c0) * (p[0] - c0) + w * (p[1] - c1) * (p[1] - c1); };
  NmOut o = nelder_mead(f, {x0, x1});
  est[0] = o.est[0];
  est[1] = o.est[1];
  return o.converged ? 1 : 0;
}

}  // extern "C"
