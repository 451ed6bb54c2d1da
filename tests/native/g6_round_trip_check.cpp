// g6_round_trip (host_util.cpp) against the text route it replaces: strtod of the %g text.
// Usage: g6_round_trip_check <count> <seed>; exit status 0 when every value agrees bit for bit.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../breseq_amd/csrc/host_util.h"

static long g_bad = 0, g_n = 0;

static void check(double x) {
  const double want = strtod(bsq::fmt_g(x).c_str(), nullptr);
  const double got = bsq::g6_round_trip(x);
  g_n++;
  if (std::memcmp(&want, &got, sizeof(double)) != 0) {
    if (g_bad < 20) printf("mismatch at %.17g: text %.17g, direct %.17g\n", x, want, got);
    g_bad++;
  }
}

static void around(double x) {
  check(x);
  check(-x);
  double u = x, d = x;
  for (int i = 0; i < 3; i++) {
    u = std::nextafter(u, INFINITY);
    d = std::nextafter(d, 0.0);
    check(u);
    check(d);
    check(-u);
  }
}

int main(int argc, char** argv) {
  const long count = argc > 1 ? atol(argv[1]) : 100000;
  std::mt19937_64 rng(argc > 2 ? strtoull(argv[2], nullptr, 10) : 1);
  std::uniform_real_distribution<double> ex(-12.0, 16.0);
  std::uniform_int_distribution<uint64_t> cnt(0, 2000000000ull);
  // outside the covered range, and the special values
  const double special[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 1e-9, 1e15, 5e-324, 1.7976931348623157e308};
  for (double s : special) around(s);
  for (long i = 0; i < count; i++) {
    // magnitudes spread over [1e-12, 1e16]
    check(std::pow(10.0, ex(rng)) * (rng() & 1 ? 1 : -1));
    // the error-rate table's values: log10(c + 1) - log10(total + 5)
    const uint64_t c = cnt(rng) >> (rng() % 40), t = c + (cnt(rng) >> (rng() % 40));
    check(std::log10((double)c + 1.0) - std::log10((double)t + 5.0));
    // bit patterns of doubles in the covered range
    uint64_t b = rng();
    b = (b & 0x800FFFFFFFFFFFFFull) | ((uint64_t)(993 + rng() % 80) << 52);
    double x;
    std::memcpy(&x, &b, sizeof(x));
    check(x);
  }
  // exact ties at the sixth digit: (w + 1/2) * 10^j for j <= 0, and (2w + 1) / (2^(k+1) 5^k) when
  // 5^k divides 2w + 1; then the neighbours of each tie and of each power of ten
  for (uint64_t w = 100000; w < 1000000; w += 1 + rng() % 97) {
    for (int j = 0; j <= 9; j++) around(((double)w + 0.5) * std::pow(10.0, j));
    uint64_t p5 = 1;
    for (int k = 1; k <= 8; k++) {
      p5 *= 5;
      if ((2 * w + 1) % p5 == 0) around(std::ldexp((double)((2 * w + 1) / p5), -(k + 1)));
    }
  }
  for (int j = -9; j <= 15; j++) around(std::pow(10.0, j));
  for (int j = -9; j <= 14; j++) around(9.999995 * std::pow(10.0, j));
  printf("%ld values, %ld mismatches\n", g_n, g_bad);
  return g_bad == 0 ? 0 : 1;
}
