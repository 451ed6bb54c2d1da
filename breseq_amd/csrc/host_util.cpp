// host_util.cpp -- see host_util.h.
#include "host_util.h"

#include <charconv>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace bsq {

void set_status(bsq_status* st, int code, const std::string& msg) {
  if (!st) return;
  st->code = code;
  snprintf(st->message, sizeof(st->message), "%s", msg.c_str());
}

std::string fmt_double(double v, unsigned precision, bool scientific) {
  char buf[400];
  const int n = fmt_double_buf(v, precision, scientific, buf, sizeof(buf));
  return std::string(buf, (size_t)(n > 0 ? n : 0));
}

// printf("%.*e") / printf("%.*f") digits through std::to_chars, which is specified to produce the
// same correctly rounded decimal form; NaN prints as NA.
int fmt_double_buf(double v, unsigned precision, bool scientific, char* buf, size_t len) {
  if (std::isnan(v)) {
    if (len < 3) return -1;
    buf[0] = 'N';
    buf[1] = 'A';
    buf[2] = 0;
    return 2;
  }
  const std::to_chars_result r =
      std::to_chars(buf, buf + len - 1, v, scientific ? std::chars_format::scientific : std::chars_format::fixed,
                    (int)precision);
  if (r.ec != std::errc()) return -1;
  int n = (int)(r.ptr - buf);
  // a 3-digit exponent with a leading zero is shortened (only non-glibc runtimes print one)
  if (scientific && n >= 3 && buf[n - 3] == '0') {
    buf[n - 3] = buf[n - 2];
    buf[n - 2] = buf[n - 1];
    n--;
  }
  buf[n] = 0;
  return n;
}

// printf("%.*g") through std::to_chars (general format with a precision is specified as %g)
std::string fmt_g(double v, int precision) {
  char buf[64];
  const std::to_chars_result r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::general, precision);
  if (r.ec != std::errc()) {
    snprintf(buf, sizeof(buf), "%.*g", precision, v);
    return buf;
  }
  return std::string(buf, (size_t)(r.ptr - buf));
}

static bool repeat_at(const char* s, int64_t len, int64_t pos, int64_t size, int64_t num) {
  const int64_t off = size * num;
  for (int64_t i = 0; i < size; i++) {
    if (pos + i >= len || pos + i + off >= len) return false;
    if (s[pos + i] != s[pos + i + off]) return false;
  }
  return true;
}

void host_sequence_trims(const char* s, int64_t len, uint8_t* t) {
  memset(t, 0, 2 * (size_t)len);
  for (int64_t p = 0; p < len; p++) {
    uint32_t mx = 1;
    for (int64_t size = 1; size <= 18; size++) {
      int64_t num = 1;
      while (repeat_at(s, len, p, size, num)) num++;
      if (num > 1) mx = std::max<uint32_t>(mx, (uint32_t)(num * size));
    }
    if (mx > 255) mx = 255;
    for (uint32_t o = 0; o < mx; o++) {
      if (p + o >= len) break;
      uint8_t v = (uint8_t)(o + 1);
      if (v > t[len + p + o]) t[len + p + o] = v;
    }
    for (uint32_t o = 0; o < mx; o++) {
      if (p + o >= len) break;
      uint8_t v = (uint8_t)(mx - o);
      if (v > t[p + o]) t[p + o] = v;
    }
  }
}

static char comp(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
  }
  return c;
}

void host_edge_trims(const std::string& seq, uint32_t& left, uint32_t& right) {
  left = right = 1;
  const int64_t n = (int64_t)seq.size();
  std::string rc(seq.rbegin(), seq.rend());
  for (char& c : rc) c = comp(c);
  for (int64_t size = 1; size <= 18; size++) {
    int64_t num = 1;
    while (repeat_at(seq.data(), n, 0, size, num)) num++;
    if (num > 1) left = std::max<uint32_t>(left, (uint32_t)(num * size));
    num = 1;
    while (repeat_at(rc.data(), n, 0, size, num)) num++;
    if (num > 1) right = std::max<uint32_t>(right, (uint32_t)(num * size));
  }
}

}  // namespace bsq
