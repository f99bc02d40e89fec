// Host check of sem2d::libm (the friction solve's asinh/sinh/cosh decomposition)
// against the C library: with the C library's own log/log1p/exp/expm1 as primitives
// the decomposition must reproduce asinh/sinh/cosh bit for bit on every branch,
// including the high-word branch points.  Prints one line: "<n> <mismatches>".
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "sem2d_libm.hpp"

static uint64_t state = 0x9e3779b97f4a7c15ull;
static uint64_t next() {
  state ^= state << 13;
  state ^= state >> 7;
  state ^= state << 17;
  return state;
}
static double uni() { return (double)(next() >> 11) * 0x1p-53; }
static double from_bits(uint64_t u) {
  double x;
  memcpy(&x, &u, sizeof x);
  return x;
}
static int same(double a, double b) { return memcmp(&a, &b, sizeof a) == 0 || (a != a && b != b); }

int main() {
  long n = 0, bad = 0;
  // values around every branch point (high words of 2^-28, 2, 2^28, ln2/2, 1, 22,
  // log(DBL_MAX), the overflow threshold), a few thousand ulps either side
  const uint64_t pts[] = {0x3e30000000000000ull, 0x4000000000000000ull, 0x41b0000000000000ull,
                          0x41b00000ffffffffull, 0x3fd62e4300000000ull, 0x3ff0000000000000ull,
                          0x4036000000000000ull, 0x40862e4200000000ull, 0x408633ce8fb9f87dull,
                          0x3c80000000000000ull};
  for (uint64_t p : pts)
    for (int64_t d = -3000; d <= 3000; ++d)
      for (int s = 0; s < 2; ++s) {
        const double x = (s ? -1.0 : 1.0) * from_bits(p + (uint64_t)d);
        ++n;
        bad += !same(sem2d::libm::asinh(x), asinh(x));
        bad += !same(sem2d::libm::sinh(x), sinh(x));
        bad += !same(sem2d::libm::cosh(x), cosh(x));
      }
  // log-uniform magnitudes over the whole range, both signs
  for (long i = 0; i < 3000000; ++i) {
    const double x = (next() & 1 ? -1.0 : 1.0) * pow(2.0, uni() * 2000.0 - 1000.0);
    const double y = (next() & 1 ? -1.0 : 1.0) * uni() * 720.0;
    ++n;
    bad += !same(sem2d::libm::asinh(x), asinh(x));
    bad += !same(sem2d::libm::sinh(y), sinh(y));
    bad += !same(sem2d::libm::cosh(y), cosh(y));
  }
  const double special[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 5e-324, 1e308};
  for (double x : special) {
    ++n;
    bad += !same(sem2d::libm::asinh(x), asinh(x));
    bad += !same(sem2d::libm::sinh(x), sinh(x));
    bad += !same(sem2d::libm::cosh(x), cosh(x));
  }
  printf("%ld %ld\n", n, bad);
  return bad != 0;
}
