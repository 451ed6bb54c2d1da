// Checks bsq::dec::format_sci / format_fixed (breseq_amd/csrc/dec_format.h) against std::to_chars,
// which is specified to write the correctly rounded text of printf("%.*e") / printf("%.*f").
// Usage: dec_format_check <count> <seed> [max precision]; prints the first mismatch and exits 1, else exits 0.
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../breseq_amd/csrc/dec_format.h"

static int check(double v, int prec, bool sci) {
  char want[512], got[bsq::dec::kMaxText];
  int nw;
  if (std::isnan(v)) {
    std::memcpy(want, "NA", 2);
    nw = 2;
  } else {
    const auto r = std::to_chars(want, want + sizeof(want), v, sci ? std::chars_format::scientific : std::chars_format::fixed, prec);
    nw = (int)(r.ptr - want);
  }
  const int ng = sci ? bsq::dec::format_sci(v, prec, got) : bsq::dec::format_fixed(v, prec, got);
  if (ng != nw || std::memcmp(want, got, (size_t)nw) != 0) {
    std::printf("mismatch v=%a prec=%d sci=%d want=%.*s got=%.*s\n", v, prec, (int)sci, nw, want, ng > 0 ? ng : 0, got);
    return 1;
  }
  return 0;
}

int main(int argc, char** argv) {
  const long count = argc > 1 ? std::atol(argv[1]) : 100000;
  std::mt19937_64 rng(argc > 2 ? std::atol(argv[2]) : 1);
  // special and boundary values
  const double specials[] = {0.0, -0.0, 1.0, -1.0, 0.5, 1.5, 2.5, 0.125, 0.375, 9.5, 99.5, 0.05, 0.15, 0.25, 1e-5,
                             5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 9.999999999999999e22, 1e23,
                             123456789012345678.0, 0.1, 0.3, 2.0 / 3.0, 1.0 / 3.0, 999.9999999, 0.99995, 9.9995e-7,
                             INFINITY, -INFINITY, NAN, 4503599627370496.5, 9007199254740993.0, 1e-300, 1e300};
  for (double v : specials)
    for (int p = 0; p <= 20; p++)
      if (check(v, p, true) || check(v, p, false)) return 1;
  std::uniform_int_distribution<int> prec(0, argc > 3 ? std::atoi(argv[3]) : 20);
  for (long i = 0; i < count; i++) {
    uint64_t b = rng();
    double v;
    const int kind = (int)(i % 4);
    if (kind == 0) {
      std::memcpy(&v, &b, 8);   // any bit pattern
    } else if (kind == 1) {
      v = std::ldexp((double)(b >> 11), -53) * std::pow(10.0, (int)(b % 40) - 20);   // moderate magnitudes
    } else if (kind == 2) {
      // exact ties at the printed precision: k / 2^j with small j
      v = std::ldexp((double)(int64_t)(b % 200001) - 100000.0, -(int)(b % 12));
    } else {
      v = (double)(int64_t)(b % 2000001) / 1000.0;   // fixed-point decimals near ties
    }
    const int p = prec(rng);
    if (check(v, p, true) || check(v, p, false)) return 1;
  }
  std::printf("ok %ld\n", count);
  return 0;
}
