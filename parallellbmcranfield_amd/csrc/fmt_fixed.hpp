// fmt_fixed.hpp — the text printf("%f") gives for a double, without printf.
//
// printf("%f", x) prints x rounded to 6 decimals; glibc rounds the exact
// binary value (ties to even) and prints the sign of negative values, -0 and
// values that round to zero included ("-0.000000").  fmt_f6 does the same
// with integer arithmetic: x = m * 2^e exactly (m < 2^53), so
// x * 10^6 = m * 10^6 * 2^e is an exact 128-bit product shifted right by -e,
// rounded half to even on the shifted-out bits.  Values outside that range
// (|x| >= 2^63 / 10^6, inf, nan) go to snprintf.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace lbmfmt {

// Writes the "%f" text of x at p (no terminator); returns its length.  p
// must hold 32 characters; wider values fall back to snprintf with cap.
inline int fmt_f6(double x, char* p, int cap = 64) {
  if (!std::isfinite(x) || std::fabs(x) >= 9.2e12) return std::snprintf(p, (size_t)cap, "%f", x);
  uint64_t bits;
  std::memcpy(&bits, &x, sizeof bits);
  const bool neg = bits >> 63;
  const int be = (int)((bits >> 52) & 0x7ff);
  uint64_t m = bits & ((1ULL << 52) - 1);
  int e;
  if (be == 0) {
    e = -1074;  // subnormal (or zero)
  } else {
    m |= 1ULL << 52;
    e = be - 1075;
  }
  const unsigned __int128 prod = (unsigned __int128)m * 1000000u;
  uint64_t n;  // round(|x| * 10^6), ties to even
  if (e >= 0) {
    n = (uint64_t)(prod << e);  // |x| < 9.2e12: the product stays below 2^63
  } else if (-e >= 128) {
    n = 0;  // prod < 2^73, so |x| * 10^6 < 2^-55
  } else {
    const int s = -e;
    const unsigned __int128 q = prod >> s;
    const unsigned __int128 rem = prod - (q << s);
    const unsigned __int128 half = (unsigned __int128)1 << (s - 1);
    n = (uint64_t)q;
    if (rem > half || (rem == half && (n & 1))) n++;
  }
  const uint64_t ip = n / 1000000u;
  uint32_t fr = (uint32_t)(n % 1000000u);
  char tmp[24];
  int k = 0;
  uint64_t v = ip;
  do {
    tmp[k++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  int len = 0;
  if (neg) p[len++] = '-';
  while (k) p[len++] = tmp[--k];
  p[len++] = '.';
  for (int d = 5; d >= 0; d--) {
    p[len + d] = (char)('0' + fr % 10);
    fr /= 10;
  }
  return len + 6;
}

// "%d"
inline int fmt_int(int x, char* p) {
  char tmp[12];
  int k = 0, len = 0;
  unsigned v = x < 0 ? 0u - (unsigned)x : (unsigned)x;
  do {
    tmp[k++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  if (x < 0) p[len++] = '-';
  while (k) p[len++] = tmp[--k];
  return len;
}

}  // namespace lbmfmt
