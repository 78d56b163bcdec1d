// Deterministic float primitives shared by the gfx950 kernels and the host checker.
//
// The reference (src/kernels.cu) is compiled with -use_fast_math, so its bits are
// hardware-specific. This project instead fixes every rounding step so that the HIP
// build and an IEEE host build produce identical bits:
//   * +, -, *, / and sqrt are IEEE-754 binary32/binary64, round-to-nearest-even,
//     with FMA contraction disabled on both sides (-ffp-contract=off) and
//     subnormals preserved.
//   * max/min follow IEEE 754-2019 maximumNumber/minimumNumber: a NaN operand
//     is ignored, and -0 orders below +0, so the result never depends on operand
//     order. On gfx950 this is exactly v_max_f32 / v_min_f32.
//   * x^2.33f (the Strickler depth exponent of src/kernels.cu:94,121) is the
//     project's own function fip_pow233 below, built only from IEEE basic
//     operations in binary64, so host and device agree bit for bit.
#pragma once
#if defined(__HIPCC_RTC__)
typedef unsigned int uint32_t;
typedef unsigned long long uint64_t;
#else
#include <stdint.h>
#endif

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define FIP_HD __host__ __device__ __forceinline__
#else
#define FIP_HD inline
#endif

FIP_HD uint32_t fip_f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
FIP_HD float fip_u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
FIP_HD uint64_t fip_d2u(double x) { return __builtin_bit_cast(uint64_t, x); }
FIP_HD double fip_u2d(uint64_t u) { return __builtin_bit_cast(double, u); }

// maximumNumber / minimumNumber (IEEE 754-2019, clause 9.6).
FIP_HD float fip_fmaxf(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fmaximum_numf(a, b);
#else
  if (a != a) return b;
  if (b != b) return a;
  if (a > b) return a;
  if (b > a) return b;
  return fip_u2f(fip_f2u(a) & fip_f2u(b));  // equal: +0 wins over -0
#endif
}
FIP_HD float fip_fminf(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fminimum_numf(a, b);
#else
  if (a != a) return b;
  if (b != b) return a;
  if (a < b) return a;
  if (b < a) return b;
  return fip_u2f(fip_f2u(a) | fip_f2u(b));  // equal: -0 wins over +0
#endif
}

// x^p with p = 2.33f = 0x1.2a3d7p+1 exactly, for the friction term
// reib = dt*G/kSt^2 * |q| * (1/max(h_face, DH))^2.33 (src/kernels.cu:93-94,120-121).
//
// Method (Cody & Waite, "Software Manual for the Elementary Functions", 1980):
//   log:  x = 2^e * m with m in [sqrt(1/2), sqrt(2)); s = (m-1)/(m+1);
//         ln m = 2 atanh(s) = 2 s (1 + s^2/3 + s^4/5 + ... + s^16/17)   (|s| <= 0.1716)
//   exp:  y = p ln x; k = nearest(y / ln2); r = (y - k ln2_hi) - k ln2_lo (|r| <= 0.347);
//         e^r by its Taylor polynomial of degree 12; result = e^r * 2^k.
// All steps are binary64 adds, multiplies and one divide, so the value is fixed by
// IEEE-754 alone; the final conversion to binary32 rounds to nearest-even.
// Relative error of the binary64 value is below 1e-14, so the binary32 result is
// the correctly rounded x^2.33f except when x^2.33f lies within ~1e-14 relative of
// a rounding boundary (then it may be 1 ulp off). Special cases: NaN -> NaN,
// x < 0 -> NaN, x == 0 -> +0, x == +inf -> +inf.
FIP_HD float fip_pow233(float x) {
  const double P = 0x1.2a3d7p+1;  // (double)2.33f
  if (!(x > 0.0f)) {
    if (x == 0.0f) return 0.0f;
    return fip_u2f(0x7fc00000u);  // NaN for x < 0 and x = NaN
  }
  if (fip_f2u(x) >= 0x7f800000u) return x;  // +inf
  const double d = (double)x;               // exact; a normal binary64 even for subnormal x
  const uint64_t u = fip_d2u(d);
  int e = (int)((u >> 52) & 0x7ff) - 1023;
  double m = fip_u2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
  if (m > 0x1.6a09e667f3bcdp+0) {
    m = m * 0.5;
    e = e + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double t = 0x1.e1e1e1e1e1e1ep-5;      // 1/17
  t = t * z + 0x1.1111111111111p-4;     // 1/15
  t = t * z + 0x1.3b13b13b13b14p-4;     // 1/13
  t = t * z + 0x1.745d1745d1746p-4;     // 1/11
  t = t * z + 0x1.c71c71c71c71cp-4;     // 1/9
  t = t * z + 0x1.2492492492492p-3;     // 1/7
  t = t * z + 0x1.999999999999ap-3;     // 1/5
  t = t * z + 0x1.5555555555555p-2;     // 1/3
  t = t * z + 1.0;
  const double lnm = (2.0 * s) * t;
  const double lnx = (double)e * 0x1.62e42fefa39efp-1 + lnm;
  const double y = P * lnx;
  if (y > 100.0) return fip_u2f(0x7f800000u);  // beyond FLT_MAX
  if (y < -110.0) return 0.0f;                 // below half the smallest subnormal
  const double kd = __builtin_floor(y * 0x1.71547652b82fep+0 + 0.5);
  const double r = (y - kd * 0x1.62e42fee00000p-1) - kd * 0x1.a39ef35793c76p-33;
  double q = 0x1.1eed8eff8d898p-29;     // 1/12!
  q = q * r + 0x1.ae64567f544e4p-26;    // 1/11!
  q = q * r + 0x1.27e4fb7789f5cp-22;    // 1/10!
  q = q * r + 0x1.71de3a556c734p-19;    // 1/9!
  q = q * r + 0x1.a01a01a01a01ap-16;    // 1/8!
  q = q * r + 0x1.a01a01a01a01ap-13;    // 1/7!
  q = q * r + 0x1.6c16c16c16c17p-10;    // 1/6!
  q = q * r + 0x1.1111111111111p-7;     // 1/5!
  q = q * r + 0x1.5555555555555p-5;     // 1/4!
  q = q * r + 0x1.5555555555555p-3;     // 1/3!
  q = q * r + 0.5;
  q = q * r + 1.0;
  q = q * r + 1.0;
  const int k = (int)kd;
  const double scale = fip_u2d((uint64_t)(k + 1023) << 52);
  return (float)(q * scale);
}
