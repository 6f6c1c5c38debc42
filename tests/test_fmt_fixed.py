"""The result writers' "%f" formatter (parallellbmcranfield_amd/csrc/fmt_fixed.hpp)
against the C library's printf, byte for byte: ties at the seventh decimal,
values that round to zero with a sign, subnormals, integers, the range limit
and random values at the scales the result files hold."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "parallellbmcranfield_amd", "csrc", "fmt_fixed.hpp")

PROG = r'''
#include <cstdio>
#include <cstring>
#include <cmath>
#include <random>
#include "fmt_fixed.hpp"

static long bad = 0, total = 0;
static void check(double x) {
  char a[512], b[512];
  const int la = std::snprintf(a, sizeof a, "%f", x);
  const int lb = lbmfmt::fmt_f6(x, b, (int)sizeof b);
  total++;
  if (la != lb || std::memcmp(a, b, (size_t)la) != 0) {
    if (bad < 10) std::printf("MISMATCH %.17g: printf '%s' ours '%.*s'\n", x, a, lb, b);
    bad++;
  }
}
static void check_int(int x) {
  char a[32], b[32];
  const int la = std::snprintf(a, sizeof a, "%d", x);
  const int lb = lbmfmt::fmt_int(x, b);
  total++;
  if (la != lb || std::memcmp(a, b, (size_t)la) != 0) {
    if (bad < 10) std::printf("MISMATCH int %d\n", x);
    bad++;
  }
}
int main() {
  const double special[] = {0.0, -0.0, 1.0, -1.0, 0.5, 1e-7, -1e-7, 5e-7, -5e-7, 4.9999999e-7,
                            5.0000001e-7, 0.0000005, 0.0078125, -0.0078125, 0.0000015, 0.0000025,
                            1.0 / 3, 2.0 / 3, 4096.0, 4097.5, 0.1, 0.2, 0.3, 1e-300, -1e-300,
                            4.9e-324, -4.9e-324, 2.2250738585072014e-308, 9.1e12, -9.1e12,
                            9.3e12, 1e300, -1e300, 123456.7890125, 0.9999995, 0.99999949999,
                            1.0 / 0.0, -1.0 / 0.0, std::nan("")};
  for (double x : special) check(x);
  // exact binary ties at the 7th decimal: k / 2^n with n <= 7 and friends
  for (int n = 1; n <= 20; n++)
    for (int k = -3000; k <= 3000; k++) check(std::ldexp((double)k, -n));
  std::mt19937_64 g(12345);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  for (int i = 0; i < 400000; i++) {
    const double r = u(g);
    check(r);                   // velocities, f's
    check(r * 4100.0);          // coordinates
    check(1.0 + r * 1e-3);      // densities
    check(r * 1e-6);            // values that round to (-)0.000000 or 0.000001
    check(std::ldexp(r, (int)(g() % 80) - 60));
    // around a rounding boundary: (j + 0.5) * 1e-6 plus a few ulps
    const double t = (std::floor(r * 1e6) + 0.5) * 1e-6;
    check(t);
    check(std::nextafter(t, 2.0));
    check(std::nextafter(t, -2.0));
  }
  const int ints[] = {0, 1, -1, 7, 10, -10, 123456789, -2147483647 - 1, 2147483647};
  for (int x : ints) check_int(x);
  std::printf("checked %ld, mismatches %ld\n", total, bad);
  return bad != 0;
}
'''


def test_fmt_f6_matches_printf():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cpp")
        exe = os.path.join(d, "t")
        with open(src, "w") as f:
            f.write(PROG)
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.dirname(HDR), src, "-o", exe],
                       check=True, capture_output=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        print(r.stdout)
        assert r.returncode == 0, r.stdout
        assert "mismatches 0" in r.stdout
