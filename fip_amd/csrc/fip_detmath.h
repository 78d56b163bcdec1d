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
// Method: table-driven logarithm and exponential in binary64, with the range reductions
// of P. T. P. Tang, "Table-driven implementation of the exponential function in IEEE
// floating-point arithmetic", ACM TOMS 15(2), 1989, and "... of the logarithm function",
// ACM TOMS 16(4), 1990:
//   log:  x = 2^e m with m in [1, 2); i = the top 6 bits of the fraction of m;
//         c_i = 128 / (129 + 2 i) rounded to 28 significant bits, so that m c_i is exact
//         (m has at most 24 significant bits) and r = m c_i - 1 is exact, |r| < 2^-7;
//         ln x = e ln2 + L_i + ln(1 + r) with L_i = -ln c_i and ln(1 + r) by its Taylor
//         polynomial to r^7 (truncation error below 2^-59).
//   exp:  y = p ln x; k = nearest(64 y / ln2) = 64 n + j with 0 <= j < 64;
//         r = (y - k C1) - k C2 with C1 + C2 = ln2 / 64 (C1 with 21 trailing zero bits,
//         so k C1 is exact), |r| <= 0.00542; e^y = 2^n T_j e^r with T_j = 2^(j/64) and
//         e^r by its Taylor polynomial to r^5 (truncation error below 2^-54).
// The tables hold L_i and T_j rounded to binary64 (checked against a 60-digit evaluation
// by the tests). Every step is a binary64 add, multiply or comparison, so the value is
// fixed by IEEE-754 alone; the final conversion to binary32 rounds to nearest-even. For
// 0 < x <= 100 the relative error of the binary64 value is below 2^-48 (mostly the
// rounding of y = p ln x), so the binary32 result is the correctly rounded x^2.33f except
// within 2^-48 relative of a rounding boundary; the tests count the misrounded inputs
// exhaustively. Special cases: NaN -> NaN, x < 0 -> NaN, x == 0 -> +0, x == +inf -> +inf.
//
// The tables are read through pointers: the host passes these arrays, a kernel may pass
// copies it keeps in LDS (same values, so the same bits).
static constexpr double kFipPowC[64] = {
    0x1.fc07f02000000p-1, 0x1.f44659e000000p-1, 0x1.ecc07b4000000p-1, 0x1.e573aca000000p-1,
    0x1.de5d6e4000000p-1, 0x1.d77b654000000p-1, 0x1.d0cb590000000p-1, 0x1.ca4b306000000p-1,
    0x1.c3f8f02000000p-1, 0x1.bdd2b8a000000p-1, 0x1.b7d6c3e000000p-1, 0x1.b203640000000p-1,
    0x1.ac5701a000000p-1, 0x1.a6d01a6000000p-1, 0x1.a16d3fa000000p-1, 0x1.9c2d14e000000p-1,
    0x1.970e4f8000000p-1, 0x1.920fb4a000000p-1, 0x1.8d3018e000000p-1, 0x1.886e5f0000000p-1,
    0x1.83c977a000000p-1, 0x1.7f405fe000000p-1, 0x1.7ad2208000000p-1, 0x1.767dce4000000p-1,
    0x1.7242880000000p-1, 0x1.6e1f76c000000p-1, 0x1.6a13cd2000000p-1, 0x1.661ec6a000000p-1,
    0x1.623fa78000000p-1, 0x1.5e75bb8000000p-1, 0x1.5ac056c000000p-1, 0x1.571ed3c000000p-1,
    0x1.5390948000000p-1, 0x1.5015016000000p-1, 0x1.4cab888000000p-1, 0x1.49539e4000000p-1,
    0x1.460cbc8000000p-1, 0x1.42d6626000000p-1, 0x1.3fb0140000000p-1, 0x1.3c995a4000000p-1,
    0x1.3991c2c000000p-1, 0x1.3698df4000000p-1, 0x1.33ae45c000000p-1, 0x1.30d1902000000p-1,
    0x1.2e025c0000000p-1, 0x1.2b404ae000000p-1, 0x1.288b012000000p-1, 0x1.25e2270000000p-1,
    0x1.2345678000000p-1, 0x1.20b470c000000p-1, 0x1.1e2ef3c000000p-1, 0x1.1bb4a40000000p-1,
    0x1.1945380000000p-1, 0x1.16e068a000000p-1, 0x1.1485f0e000000p-1, 0x1.12358e8000000p-1,
    0x1.0fef010000000p-1, 0x1.0db20a8000000p-1, 0x1.0b7e6ec000000p-1, 0x1.0953f3a000000p-1,
    0x1.073260a000000p-1, 0x1.05197f8000000p-1, 0x1.03091b6000000p-1, 0x1.0101010000000p-1,
};
static constexpr double kFipPowL[64] = {
    0x1.fe02a69106789p-8, 0x1.7b91b1155b11bp-6, 0x1.39e87a97ebd60p-5, 0x1.b42dd605971bfp-5,
    0x1.16536ee637ae1p-4, 0x1.51b074546183fp-4, 0x1.8c345d1319b21p-4, 0x1.c5e5489bbc743p-4,
    0x1.fec912fbbeabbp-4, 0x1.1b72ad33f67a0p-3, 0x1.371fc1f6e8f74p-3, 0x1.526e5e5a1b438p-3,
    0x1.6d60feac9d21dp-3, 0x1.87fa06910c911p-3, 0x1.a23bc1d52b563p-3, 0x1.bc286789d8cd6p-3,
    0x1.d5c216b8fbb91p-3, 0x1.ef0adcaec5936p-3, 0x1.04025929cd041p-2, 0x1.1058bfb6e4ad5p-2,
    0x1.1c898c34199fbp-2, 0x1.2895a113686a3p-2, 0x1.347dd9cf87d55p-2, 0x1.404308716a7e4p-2,
    0x1.4be5f937778a1p-2, 0x1.5767715355a6cp-2, 0x1.62c82f0d1c795p-2, 0x1.6e08eab13a1e4p-2,
    0x1.792a55cfd47a3p-2, 0x1.842d1dc7e8b17p-2, 0x1.8f11e844662c8p-2, 0x1.99d958207e08bp-2,
    0x1.a484093c5bb0ap-2, 0x1.af1292f3f786bp-2, 0x1.b985893f310fbp-2, 0x1.c3dd7a6ddad4dp-2,
    0x1.ce1af0b65f3ebp-2, 0x1.d83e725022f3ep-2, 0x1.e2488197c6c26p-2, 0x1.ec399d3d68cc0p-2,
    0x1.f6123fac028acp-2, 0x1.ffd2e07e7f498p-2, 0x1.04bdf9c9126d2p-1, 0x1.0986f4dfb3521p-1,
    0x1.0e4498651cc8cp-1, 0x1.12f7193dfefbcp-1, 0x1.179eabcc499a1p-1, 0x1.1c3b820513c25p-1,
    0x1.20cdcd2a2ab6ep-1, 0x1.2555bcf50f7cbp-1, 0x1.29d37fd6ab08bp-1, 0x1.2e47437640268p-1,
    0x1.32b133a121d71p-1, 0x1.37117b3eb47b6p-1, 0x1.3b6844a13fc23p-1, 0x1.3fb5b83a16f42p-1,
    0x1.43f9fe4d9ce68p-1, 0x1.48353d2fa88dfp-1, 0x1.4c679b014ee3ap-1, 0x1.50913ca15686bp-1,
    0x1.54b2468259498p-1, 0x1.58cadb57d7989p-1, 0x1.5cdb1dab01765p-1, 0x1.60e32f46788d9p-1,
};
static constexpr double kFipPowT[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};

FIP_HD float fip_pow233_tab(float x, const double* tc, const double* tl, const double* tt) {
  const double P = 0x1.2a3d7p+1;  // (double)2.33f
  if (!(x > 0.0f)) {
    if (x == 0.0f) return 0.0f;
    return fip_u2f(0x7fc00000u);  // NaN for x < 0 and x = NaN
  }
  if (fip_f2u(x) >= 0x7f800000u) return x;  // +inf
  const double d = (double)x;               // exact; a normal binary64 even for subnormal x
  const uint64_t u = fip_d2u(d);
  const int e = (int)((u >> 52) & 0x7ff) - 1023;
  const int i = (int)((u >> 46) & 63);
  const double m = fip_u2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
  const double r = m * tc[i] - 1.0;  // both steps exact
  double t = 0x1.2492492492492p-3;   // 1/7
  t = t * r - 0x1.5555555555555p-3;  // 1/6
  t = t * r + 0x1.999999999999ap-3;  // 1/5
  t = t * r - 0x1.0000000000000p-2;  // 1/4
  t = t * r + 0x1.5555555555555p-2;  // 1/3
  t = t * r - 0x1.0000000000000p-1;  // 1/2
  t = t * r + 1.0;
  const double lnr = t * r;                                        // ln(1 + r)
  const double ed = (double)e;
  const double lnx = (ed * 0x1.62e42fefa3800p-1 + tl[i]) + (ed * 0x1.ef35793c7673p-45 + lnr);
  const double y = P * lnx;
  if (y > 100.0) return fip_u2f(0x7f800000u);  // beyond FLT_MAX
  if (y < -110.0) return 0.0f;                 // below half the smallest subnormal
  const double kd = __builtin_floor(y * 0x1.71547652b82fep+6 + 0.5);  // 64 / ln2
  const double rr = (y - kd * 0x1.62e42fee00000p-7) - kd * 0x1.a39ef35793c76p-39;  // ln2/64 = C1 + C2
  double q = 0x1.1111111111111p-7;  // 1/5!
  q = q * rr + 0x1.5555555555555p-5;  // 1/4!
  q = q * rr + 0x1.5555555555555p-3;  // 1/3!
  q = q * rr + 0.5;
  q = q * rr + 1.0;
  q = q * rr + 1.0;
  const int k = (int)kd;
  const int j = k & 63;
  const int n = (k - j) / 64;
  const double scale = fip_u2d((uint64_t)(n + 1023) << 52);
  return (float)((tt[j] * q) * scale);
}

FIP_HD float fip_pow233(float x) { return fip_pow233_tab(x, kFipPowC, kFipPowL, kFipPowT); }
