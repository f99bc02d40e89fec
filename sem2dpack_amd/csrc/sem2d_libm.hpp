// asinh, sinh and cosh of the friction solve, evaluated through the same
// decomposition into log / log1p / exp / expm1 / sqrt as the C library that the
// reference's Fortran runtime calls (fdlibm's algorithms, as shipped in glibc:
// s_asinh.c, e_sinh.c, e_cosh.c; branch points compared on the high word of |x|).
//
// Why: on a creeping rate-and-state fault the slip rate is recomputed from the
// difference of two tractions of ~1e8 Pa (SRC/bc_dynflt.f90:685-687,
// V = dV + CoefA2V*(dA - T/(Z*CoefA2V))), which cancels about ten digits, and the
// traction comes from mu = a*asinh(...) (SRC/bc_dynflt_rsf.f90:165-184).  A one-ulp
// difference in asinh therefore moves the slip rate by ~1e-7 relative.  A
// single-rounding asinh (log(2x) in one step) differs from the C library's
// log(x) + ln2 (two roundings) in a large fraction of calls; with the same
// decomposition the results differ only where the primitives themselves do.
//
// The functions are host+device so that a host build can check them against the C
// library (tests/test_libm_decomposition.py).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

namespace sem2d {
namespace libm {

__host__ __device__ inline uint32_t hi_word(double x) {
  uint64_t u;
  memcpy(&u, &x, sizeof u);
  return (uint32_t)(u >> 32);
}

__host__ __device__ inline uint32_t lo_word(double x) {
  uint64_t u;
  memcpy(&u, &x, sizeof u);
  return (uint32_t)u;
}

#pragma clang fp contract(off)

// ln 2 rounded to double
constexpr double kLn2 = 6.93147180559945286227e-01;

__host__ __device__ inline double asinh(double x) {
  const uint32_t ix = hi_word(x) & 0x7fffffffu;
  if (ix >= 0x7ff00000u) return x + x;          // inf or NaN
  if (ix < 0x3e300000u) return x;               // |x| < 2^-28
  const double xa = fabs(x);
  double w;
  if (ix > 0x41b00000u) {                       // |x| > 2^28
    w = ::log(xa) + kLn2;
  } else if (ix > 0x40000000u) {                // 2 < |x| <= 2^28
    w = ::log(2.0 * xa + 1.0 / (::sqrt(xa * xa + 1.0) + xa));
  } else {                                      // 2^-28 <= |x| <= 2
    const double t = xa * xa;
    w = ::log1p(xa + t / (1.0 + ::sqrt(1.0 + t)));
  }
  return copysign(w, x);
}

__host__ __device__ inline double sinh(double x) {
  const uint32_t jx = hi_word(x), ix = jx & 0x7fffffffu;
  if (ix >= 0x7ff00000u) return x + x;
  const double h = (jx >> 31) ? -0.5 : 0.5;
  if (ix < 0x40360000u) {                       // |x| < 22
    if (ix < 0x3e300000u) return x;             // |x| < 2^-28
    const double t = ::expm1(fabs(x));
    if (ix < 0x3ff00000u) return h * (2.0 * t - t * t / (t + 1.0));
    return h * (t + t / (t + 1.0));
  }
  if (ix < 0x40862e42u) return h * ::exp(fabs(x));  // |x| < log(DBL_MAX)
  const uint32_t lx = lo_word(x);
  if (ix < 0x408633ceu || (ix == 0x408633ceu && lx <= 0x8fb9f87du)) {
    const double w = ::exp(0.5 * fabs(x));
    const double t = h * w;
    return t * w;
  }
  return x * 1.0e307;                           // overflow
}

__host__ __device__ inline double cosh(double x) {
  const uint32_t ix = hi_word(x) & 0x7fffffffu;
  if (ix < 0x40360000u) {                       // |x| < 22
    if (ix < 0x3fd62e43u) {                     // |x| < ln2 / 2
      if (ix < 0x3c800000u) return 1.0;
      const double t = ::expm1(fabs(x));
      const double w = 1.0 + t;
      return 1.0 + (t * t) / (w + w);
    }
    const double t = ::exp(fabs(x));
    return 0.5 * t + 0.5 / t;
  }
  if (ix < 0x40862e42u) return 0.5 * ::exp(fabs(x));
  const uint32_t lx = lo_word(x);
  if (ix < 0x408633ceu || (ix == 0x408633ceu && lx <= 0x8fb9f87du)) {
    const double w = ::exp(0.5 * fabs(x));
    const double t = 0.5 * w;
    return t * w;
  }
  if (ix >= 0x7ff00000u) return x * x;          // inf or NaN
  return 1.0e300 * 1.0e300;                     // overflow
}

#pragma clang fp contract(on)

}  // namespace libm
}  // namespace sem2d
