// dec_format.h -- correctly rounded decimal text of an IEEE double, as printf("%.*e") and
// printf("%.*f") write it (round half to even on the exact binary value), for host and device code.
//
// The value m * 2^e (m < 2^53) is expanded exactly: the integer part as decimal digits of a big
// integer (repeated division by 10^9), the fraction as a big integer F over 2^s that yields nine
// digits per multiplication by 10^9.  Digits are generated only as far as the precision needs,
// plus one more and a sticky bit for the rounding decision.  No floating-point arithmetic takes
// part, so the result does not depend on the device's rounding modes or library routines.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define BSQ_HD __host__ __device__
#else
#define BSQ_HD
#endif

namespace bsq {
namespace dec {

constexpr int kMaxPrecision = 60;    // digits after the point
constexpr int kDigitCap = 420;       // integer digits (<= 309) or leading fraction zeros (<= 323) + precision + 2
constexpr int kMaxText = 400;        // longest text written (fixed format of 1e308 with 60 decimals)
constexpr int kWords = 36;           // 32-bit words of the fraction numerator: s <= 1074 bits plus a spare word

// Exact decimal digits of |v| (finite): int_digits[0..n_int) (no leading zeros; none when the
// integer part is 0), then the fraction as F / 2^s in words (little-endian, 32-bit).
struct Expansion {
  char digits[kDigitCap];
  int n_int;           // integer digits in digits[0..n_int)
  int n;               // digits produced so far (integer digits, then fraction digits)
  uint32_t F[kWords];  // fraction numerator below 2^s (s <= 1074), zero above
  int nw;              // words of F that can be non-zero (0: an integer, no fraction)
  int s;               // fraction bits
};

BSQ_HD inline bool frac_nonzero(const Expansion& x) {
  uint32_t any = 0;
#pragma unroll
  for (int i = 0; i < kWords; i++) any |= x.F[i];
  return any != 0;
}

// Nine more fraction digits per step: F *= 10^9, and the part at or above bit s (below 10^9, so
// within the words s/32 and s/32 + 1) is the next nine digits.  Every loop over F has a constant
// trip count, so F can live in registers.
BSQ_HD inline void more_digits(Expansion& x, int want) {
  const int wi = x.s >> 5, sh = x.s & 31;
  const uint32_t keep_mask = sh ? ((1u << sh) - 1u) : 0u;
  while (x.n < want && x.n + 9 <= kDigitCap) {
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < kWords; i++) {
      const uint64_t t = (uint64_t)x.F[i] * 1000000000ull + carry;
      x.F[i] = (uint32_t)t;
      carry = t >> 32;
    }
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < kWords; i++) {
      if (i == wi) lo = x.F[i];
      if (i == wi + 1) hi = x.F[i];
    }
    uint32_t chunk = (uint32_t)((((uint64_t)hi << 32) | lo) >> sh);
#pragma unroll
    for (int i = 0; i < kWords; i++) {
      if (i == wi) x.F[i] &= keep_mask;
      else if (i > wi) x.F[i] = 0;
    }
    for (int k = 8; k >= 0; k--) {
      x.digits[x.n + k] = (char)('0' + chunk % 10u);
      chunk /= 10u;
    }
    x.n += 9;
  }
}

// Set up the expansion of |v| = m * 2^e (v finite).
BSQ_HD inline void expand(Expansion& x, uint64_t bits) {
  const int be = (int)((bits >> 52) & 0x7ff);
  uint64_t m = bits & ((1ull << 52) - 1ull);
  int e;
  if (be == 0) {
    e = -1074;
  } else {
    m |= 1ull << 52;
    e = be - 1075;
  }
  x.n = 0;
  x.n_int = 0;
  x.nw = 0;
  x.s = 0;
#pragma unroll
  for (int i = 0; i < kWords; i++) x.F[i] = 0;
  // integer part
  uint32_t I[34];
  int ni = 0;
  if (e >= 0) {
    // m << e as words
    const int ws = e >> 5, bs = e & 31;
    for (int i = 0; i < ws; i++) I[i] = 0;
    const uint64_t lo = m << bs;
    const uint64_t hi = bs ? (m >> (64 - bs)) : 0ull;
    I[ws] = (uint32_t)lo;
    I[ws + 1] = (uint32_t)(lo >> 32);
    I[ws + 2] = (uint32_t)hi;
    ni = ws + 3;
  } else if (-e < 64) {
    const uint64_t ip = m >> (-e);
    I[0] = (uint32_t)ip;
    I[1] = (uint32_t)(ip >> 32);
    ni = 2;
    x.s = -e;
  } else {
    x.s = -e;
  }
  while (ni > 0 && I[ni - 1] == 0) ni--;
  // integer digits, least significant chunk first
  char rev[kDigitCap];
  int nr = 0;
  while (ni > 0) {
    uint64_t rem = 0;
    for (int i = ni - 1; i >= 0; i--) {
      const uint64_t cur = (rem << 32) | I[i];
      I[i] = (uint32_t)(cur / 1000000000ull);
      rem = cur % 1000000000ull;
    }
    while (ni > 0 && I[ni - 1] == 0) ni--;
    uint32_t c = (uint32_t)rem;
    for (int k = 0; k < 9; k++) {
      rev[nr++] = (char)('0' + c % 10u);
      c /= 10u;
    }
  }
  while (nr > 0 && rev[nr - 1] == '0') nr--;
  for (int k = 0; k < nr; k++) x.digits[k] = rev[nr - 1 - k];
  x.n_int = nr;
  x.n = nr;
  // fraction: the low s bits of m
  if (x.s > 0) {
    const int s = x.s;
    x.nw = (s >> 5) + 2;   // room for the digits above bit s after a multiplication
    const uint64_t f = (s >= 64) ? m : (m & ((1ull << s) - 1ull));
    x.F[0] = (uint32_t)f;
    x.F[1] = (uint32_t)(f >> 32);
  } else {
    x.F[0] = x.F[1] = 0;
  }
}

// Digit k of the expansion, generating fraction digits as needed.
BSQ_HD inline char digit_at(Expansion& x, int k) {
  if (k >= x.n && x.nw == 0) return '0';   // an integer: no fraction digits
  if (k >= x.n) more_digits(x, k + 1);
  return k < x.n ? x.digits[k] : '0';
}

// Rounds the kept digits keep[0..n) up by one unit in the last place when the discarded part
// (first discarded digit `next`, anything non-zero after it `rest`) is above half, or exactly half
// and the last kept digit is odd.  Returns true when the carry runs out of the top digit.
BSQ_HD inline bool round_half_even(char* keep, int n, char next, bool rest) {
  const bool odd = n > 0 && ((keep[n - 1] - '0') & 1);
  const bool up = next > '5' || (next == '5' && (rest || odd));
  if (!up) return false;
  for (int i = n - 1; i >= 0; i--) {
    if (keep[i] == '9') {
      keep[i] = '0';
    } else {
      keep[i]++;
      return false;
    }
  }
  return true;
}

// Whether any digit from position k on (and the unexpanded fraction) is non-zero.
BSQ_HD inline bool any_from(const Expansion& x, int k) {
  for (int i = k; i < x.n; i++)
    if (x.digits[i] != '0') return true;
  return frac_nonzero(x);
}

BSQ_HD inline int put_special(uint64_t bits, char* out) {
  const bool neg = (bits >> 63) != 0;
  int n = 0;
  const bool is_nan = ((bits >> 52) & 0x7ff) == 0x7ff && (bits & ((1ull << 52) - 1ull));
  if (is_nan) {
    out[0] = 'N';
    out[1] = 'A';
    return 2;
  }
  if (neg) out[n++] = '-';
  out[n++] = 'i';
  out[n++] = 'n';
  out[n++] = 'f';
  return n;
}

// printf("%.*e", precision, v) for v not NaN (NaN writes "NA"); returns the length, or -1 when the
// precision is above kMaxPrecision.  out holds at least kMaxText bytes.
BSQ_HD inline int format_sci(double v, int precision, char* out) {
  if (precision < 0 || precision > kMaxPrecision) return -1;
  uint64_t bits;
  __builtin_memcpy(&bits, &v, 8);
  if (((bits >> 52) & 0x7ff) == 0x7ff) return put_special(bits, out);
  const bool neg = (bits >> 63) != 0;
  int n = 0;
  if (neg) out[n++] = '-';
  char keep[kMaxPrecision + 2];
  int exp10 = 0;
  if ((bits & ~(1ull << 63)) == 0) {
    for (int i = 0; i <= precision; i++) keep[i] = '0';
  } else {
    Expansion x;
    expand(x, bits & ~(1ull << 63));
    // first non-zero digit
    int f = 0;
    if (x.n_int > 0) {
      f = 0;
      exp10 = x.n_int - 1;
    } else {
      for (;;) {
        const char c = digit_at(x, f);
        if (c != '0') break;
        f++;
      }
      exp10 = -(f + 1);
    }
    for (int i = 0; i <= precision; i++) keep[i] = digit_at(x, f + i);
    const char next = digit_at(x, f + precision + 1);
    const bool rest = any_from(x, f + precision + 2);
    if (round_half_even(keep, precision + 1, next, rest)) {
      keep[0] = '1';
      for (int i = 1; i <= precision; i++) keep[i] = '0';
      exp10++;
    }
  }
  out[n++] = keep[0];
  if (precision > 0) {
    out[n++] = '.';
    for (int i = 1; i <= precision; i++) out[n++] = keep[i];
  }
  out[n++] = 'e';
  out[n++] = exp10 < 0 ? '-' : '+';
  int a = exp10 < 0 ? -exp10 : exp10;
  if (a >= 100) {
    out[n++] = (char)('0' + a / 100);
    a %= 100;
  }
  out[n++] = (char)('0' + a / 10);
  out[n++] = (char)('0' + a % 10);
  return n;
}

// printf("%.*f", precision, v) for v not NaN (NaN writes "NA"); returns the length, or -1 when the
// precision is above kMaxPrecision.
BSQ_HD inline int format_fixed(double v, int precision, char* out) {
  if (precision < 0 || precision > kMaxPrecision) return -1;
  uint64_t bits;
  __builtin_memcpy(&bits, &v, 8);
  if (((bits >> 52) & 0x7ff) == 0x7ff) return put_special(bits, out);
  const bool neg = (bits >> 63) != 0;
  int n = 0;
  if (neg) out[n++] = '-';
  Expansion x;
  expand(x, bits & ~(1ull << 63));
  // kept digits: the integer digits (a single 0 when none), then `precision` fraction digits
  char keep[kDigitCap + 1];
  int nk = 0;
  if (x.n_int == 0) keep[nk++] = '0';
  for (int i = 0; i < x.n_int; i++) keep[nk++] = x.digits[i];
  for (int i = 0; i < precision; i++) keep[nk++] = digit_at(x, x.n_int + i);
  const char next = digit_at(x, x.n_int + precision);
  const bool rest = any_from(x, x.n_int + precision + 1);
  bool carry = round_half_even(keep, nk, next, rest);
  if (carry) out[n++] = '1';
  const int int_len = nk - precision;
  for (int i = 0; i < int_len; i++) out[n++] = keep[i];
  if (precision > 0) {
    out[n++] = '.';
    for (int i = int_len; i < nk; i++) out[n++] = keep[i];
  }
  return n;
}

// %.*e when scientific, else %.*f: one entry for callers that pick the form per value, so that
// threads formatting different fields run the same code.
BSQ_HD inline int format_number(double v, int precision, bool scientific, char* out) {
  return scientific ? format_sci(v, precision, out) : format_fixed(v, precision, out);
}

}  // namespace dec
}  // namespace bsq
