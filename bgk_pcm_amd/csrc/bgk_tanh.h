// bgk_tanh.h -- double-precision tanh for the phase-field update, written
// from the published fdlibm method (Sun Microsystems fdlibm 5.3, s_tanh.c and
// s_expm1.c: tanh(x) = +-(1 - 2/(expm1(2|x|)+2)) for |x| >= 1 and
// -expm1(-2|x|)/(expm1(-2|x|)+2) below; expm1 by reduction x = k ln2 + r
// and a degree-5 rational correction in r^2/2).  The correction polynomial
// is evaluated in the split form R1 + h2*R2 + h4*R3 that GNU libc 2.35 uses,
// so the result equals glibc's tanh bit for bit (the reference calls glibc
// tanh at STEP_7/phase_field.f90:24-25; tests/test_tanh.py checks the
// equality on the host).  Every operation is a correctly rounded IEEE add,
// multiply or divide, so the value is the same on gfx950 provided nothing is
// contracted into FMA; the functions switch contraction off themselves.
//
// Usable from HIP device code and from plain C++ (the host test).
#ifndef BGK_TANH_H
#define BGK_TANH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define BGK_HD __host__ __device__ inline
#else
#define BGK_HD inline
#endif

namespace bgk {

BGK_HD uint32_t hi_word(double x) { return (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32); }
BGK_HD uint32_t lo_word(double x) { return (uint32_t)__builtin_bit_cast(uint64_t, x); }
BGK_HD double from_words(uint32_t hi, uint32_t lo) {
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint64_t)lo);
}

// expm1(x) for the arguments tanh passes (|x| < 44); the overflow branch is
// kept for completeness.
BGK_HD double expm1_fdlibm(double x) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double invln2 = 1.44269504088896338700e+00;
  const double Q1 = -3.33333333333331316428e-02;
  const double Q2 = 1.58730158725481460165e-03;
  const double Q3 = -7.93650757867487942473e-05;
  const double Q4 = 4.00821782732936239552e-06;
  const double Q5 = -2.01099218183624371326e-07;
  uint32_t hx = hi_word(x);
  const bool neg = (hx & 0x80000000u) != 0;
  hx &= 0x7fffffffu;
  double c = 0.0;
  int k;
  if (hx >= 0x4043687Au) {  // |x| >= 56 ln2
    if (neg) return -1.0;
    return x * 1.0e300;
  }
  if (hx > 0x3fd62e42u) {  // |x| > ln2/2: reduce
    double hi, lo;
    if (hx < 0x3FF0A2B2u) {
      if (!neg) { hi = x - ln2_hi; lo = ln2_lo; k = 1; }
      else { hi = x + ln2_hi; lo = -ln2_lo; k = -1; }
    } else {
      k = (int)(invln2 * x + (neg ? -0.5 : 0.5));
      const double t = (double)k;
      hi = x - t * ln2_hi;
      lo = t * ln2_lo;
    }
    x = hi - lo;
    c = (hi - x) - lo;
  } else if (hx < 0x3c900000u) {  // |x| < 2^-54
    return x;
  } else {
    k = 0;
  }
  const double hfx = 0.5 * x;
  const double hxs = x * hfx;
  const double R1 = 1.0 + hxs * Q1;
  const double h2 = hxs * hxs;
  const double R2 = Q2 + hxs * Q3;
  const double h4 = h2 * h2;
  const double R3 = Q4 + hxs * Q5;
  const double r1 = R1 + h2 * R2 + h4 * R3;
  double t = 3.0 - r1 * hfx;
  double e = hxs * ((r1 - t) / (6.0 - x * t));
  if (k == 0) return x - (x * e - hxs);
  e = (x * (e - c) - c);
  e -= hxs;
  if (k == -1) return 0.5 * (x - e) - 0.5;
  if (k == 1) {
    if (x < -0.25) return -2.0 * (e - (x + 0.5));
    return 1.0 + 2.0 * (x - e);
  }
  double y;
  if (k <= -2 || k > 56) {
    y = 1.0 - (e - x);
    y = from_words(hi_word(y) + ((uint32_t)k << 20), lo_word(y));
    return y - 1.0;
  }
  if (k < 20) {
    t = from_words(0x3ff00000u - (0x200000u >> k), 0u);
    y = t - (e - x);
    y = from_words(hi_word(y) + ((uint32_t)k << 20), lo_word(y));
  } else {
    t = from_words((uint32_t)(0x3ff - k) << 20, 0u);
    y = x - (e + t);
    y += 1.0;
    y = from_words(hi_word(y) + ((uint32_t)k << 20), lo_word(y));
  }
  return y;
}

BGK_HD double tanh_fdlibm(double x) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  const uint32_t jx = hi_word(x);
  const uint32_t ix = jx & 0x7fffffffu;
  double z;
  if (ix >= 0x7ff00000u) {  // inf or nan
    return (jx & 0x80000000u) ? 1.0 / x - 1.0 : 1.0 / x + 1.0;
  }
  if (ix < 0x40360000u) {  // |x| < 22
    if (ix < 0x3c800000u) return x * (1.0 + x);  // |x| < 2^-55
    const double ax = (jx & 0x80000000u) ? -x : x;
    if (ix >= 0x3ff00000u) {  // |x| >= 1
      const double t = expm1_fdlibm(2.0 * ax);
      z = 1.0 - 2.0 / (t + 2.0);
    } else {
      const double t = expm1_fdlibm(-2.0 * ax);
      z = -t / (t + 2.0);
    }
  } else {
    z = 1.0 - 1.0e-300;  // rounds to 1 with the inexact flag
  }
  return (jx & 0x80000000u) ? -z : z;
}

}  // namespace bgk

#endif  // BGK_TANH_H
