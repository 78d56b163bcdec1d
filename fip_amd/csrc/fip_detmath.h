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

// RN(1 / b), the correctly rounded binary32 reciprocal, for b >= 0.01 (the face depths
// of the friction term are max(h, DH) >= DH = 0.01). On the device: the hardware
// reciprocal (v_rcp_f32, within 1 ulp) refined by two Newton steps with fused
// multiply-adds, e = 1 - b y, y' = y + e y; after the first step y is within 2^-45
// relative of 1/b, and the second rounds y + y (1 - b y) once, which gives RN(1/b) when y
// is a faithful approximation (P. Markstein, "IA-64 and elementary functions", 2000,
// theorem 8.4), except possibly for significands of b of all ones. That exception and the
// hardware reciprocal itself are covered by checking every binary32 b in [0.01, 2^64]
// against the IEEE division (fip_check_rcp). Beyond 2^64 (never a face depth of a
// physical state) the IEEE division is taken.
FIP_HD float fip_rcp_rn(float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (!(b <= 0x1p64f)) return 1.0f / b;
  float y = __builtin_amdgcn_rcpf(b);
  float e = __builtin_fmaf(-b, y, 1.0f);
  y = __builtin_fmaf(e, y, y);
  e = __builtin_fmaf(-b, y, 1.0f);
  return __builtin_fmaf(e, y, y);
#else
  return 1.0f / b;
#endif
}

// x^p with p = 2.33f = 0x1.2a3d7p+1 exactly, for the friction term
// reib = dt*G/kSt^2 * |q| * (1/max(h_face, DH))^2.33 (src/kernels.cu:93-94,120-121).
//
// Method: a table-driven power in binary64. The argument reduction is the one of
// P. T. P. Tang's table-driven logarithm ("Table-driven implementation of the logarithm
// function in IEEE floating-point arithmetic", ACM TOMS 16(4), 1990); the power is then
// read from tables instead of going through log and exp:
//   x = 2^e m with m in [1, 2); i = the top 6 bits of the fraction of m;
//   c_i = 128 / (129 + 2 i) rounded to 28 significant bits, so that m c_i is exact
//   (m has at most 24 significant bits) and r = m c_i - 1 is exact, |r| < 2^-7;
//   x^p = 2^(e p) * c_i^-p * (1 + r)^p = E[e] * M[i] * S(r),
//   S(r) = 1 + r (b1 + r (b2 + ... + r b6)), b_k = binomial(p, k), the binomial series
//   of (1 + r)^p cut after r^6 (truncation error below 2^-57.9).
// M_i = c_i^-p, E[e] = 2^(e p) (e = -65..54) and b_k are rounded to binary64 (checked
// against a 60-digit evaluation by the tests). Every step is a binary64 add, multiply or
// comparison, so the value is fixed by IEEE-754 alone; the final conversion to binary32
// rounds to nearest-even. The relative error of the binary64 value is below 2^-50.5
// (two table roundings, two products, the last add of S), so the binary32 result is the
// correctly rounded x^2.33f except within 2^-50.5 relative of a rounding boundary; the
// tests count the misrounded inputs exhaustively. Outside e = -65..54 the result is 0
// (x < 2^-65: x^p < 2^-151.4, below half the smallest subnormal) or +inf (x >= 2^55:
// x^p >= 2^128.1). Special cases: NaN -> NaN, x < 0 -> NaN, x == 0 -> +0, +inf -> +inf.
// The special cases are selected at the end rather than branched on, so a wavefront
// runs one straight path.
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
static constexpr double kFipPowM[64] = {
    0x1.04af29bf4ac18p+0, 0x1.0e32cdac9b0dep+0, 0x1.17e867ec42ac5p+0, 0x1.21d038d68b45bp+0,
    0x1.2bea7fe11a8b8p+0, 0x1.36377bcad2542p+0, 0x1.40b76aa00a381p+0, 0x1.4b6a8a542ab3dp+0,
    0x1.565117e3c22eap+0, 0x1.616b4fadd28fdp+0, 0x1.6cb96dbdc34cdp+0, 0x1.783bad7271197p+0,
    0x1.83f24995b4ea5p+0, 0x1.8fdd7c78c294fp+0, 0x1.9bfd7fc59c1e0p+0, 0x1.a8528d677968fp+0,
    0x1.b4dcdd79d259fp+0, 0x1.c19ca8c58e024p+0, 0x1.ce92270704c01p+0, 0x1.dbbd901d0901fp+0,
    0x1.e91f1a685812bp+0, 0x1.f6b6fc607592ap+0, 0x1.0242b6931b247p+1, 0x1.0945509ee6a77p+1,
    0x1.1063672ebee43p+1, 0x1.179d14d30c5a1p+1, 0x1.1ef273b2d6e86p+1, 0x1.26639df336601p+1,
    0x1.2df0ad0a4b23dp+1, 0x1.3599bb4c0c13bp+1, 0x1.3d5ee1579e60fp+1, 0x1.4540395930ddep+1,
    0x1.4d3ddbef58a39p+1, 0x1.5557e1bfaa361p+1, 0x1.5d8e643bc1329p+1, 0x1.65e17bc0c63a0p+1,
    0x1.6e5140a181edep+1, 0x1.76ddcb1cd8837p+1, 0x1.7f87335ce3b1fp+1, 0x1.884d91804155fp+1,
    0x1.9130fcf07919bp+1, 0x1.9a318d8d7a6ecp+1, 0x1.a34f5abd6e32ep+1, 0x1.ac8a7bf87058dp+1,
    0x1.b5e308a11710bp+1, 0x1.bf5916fe60ce4p+1, 0x1.c8ecbf17781cep+1, 0x1.d29e16bbc51a1p+1,
    0x1.dc6d34eb43a80p+1, 0x1.e65a3009772d9p+1, 0x1.f0651e46381bfp+1, 0x1.fa8e16a768736p+1,
    0x1.026a97449bc38p+2, 0x1.079d3dd362e1cp+2, 0x1.0cdf0a6b19cadp+2, 0x1.1230076a7e75ep+2,
    0x1.17904029fe1b0p+2, 0x1.1cffbec17dcb4p+2, 0x1.227e8e2cc2410p+2, 0x1.280cb8eb38887p+2,
    0x1.2daa4a10dd732p+2, 0x1.33574b71b10d0p+2, 0x1.3913c7bb6fe5dp+2, 0x1.3edfc9ab357e2p+2,
};
static constexpr int kFipPowEMin = -65, kFipPowEMax = 54;
static constexpr double kFipPowE[kFipPowEMax - kFipPowEMin + 1] = {
    0x1.76cea665aa601p-152, 0x1.d7233e739e583p-150, 0x1.281cf828e0223p-147, 0x1.7437daf62d720p-145,
    0x1.d3e21a2bd150ap-143, 0x1.26115591f8452p-140, 0x1.71a5a396328acp-138, 0x1.d0a6b72fba2c4p-136,
    0x1.240950f4f73e3p-133, 0x1.6f17f82d19343p-131, 0x1.cd710b520721fp-129, 0x1.2204e3ec6721cp-126,
    0x1.6c8ed0b0921f5p-124, 0x1.ca410c77658e3p-122, 0x1.2004081e21a39p-119, 0x1.6a0a252485d35p-117,
    0x1.c716b096621ffp-115, 0x1.1e06b73b3c164p-112, 0x1.6789ed9afb888p-110, 0x1.c3f1edb7493dap-108,
    0x1.1c0ceafff38e1p-105, 0x1.650e22340030dp-103, 0x1.c0d2b9f407a29p-101, 0x1.1a169d339926dp-98,
    0x1.6296bb1d8dab8p-96, 0x1.bdb90b780b333p-94, 0x1.1823c7a87e6d0p-91, 0x1.6023b09372250p-89,
    0x1.baa4d8802408fp-87, 0x1.1634643be1e92p-84, 0x1.5db4fadf37a2bp-82, 0x1.b796175a65b48p-80,
    0x1.14486cd5dbcdbp-77, 0x1.5b4a92580bb98p-75, 0x1.b48cbe6608b6ep-73, 0x1.125fdb694ac73p-70,
    0x1.58e46f62a76f6p-68, 0x1.b188c4134c2fep-66, 0x1.107aa9f3c0ee3p-63, 0x1.56828a713747dp-61,
    0x1.ae8a1ee357c2ap-59, 0x1.0e98d27d70db2p-56, 0x1.5424dc03437a6p-54, 0x1.ab90c5681daf5p-52,
    0x1.0cba4f191adb6p-49, 0x1.51cb5ca59853ap-47, 0x1.a89cae443d216p-45, 0x1.0adf19e3fa47ap-42,
    0x1.4f7604f22ebf6p-40, 0x1.a5add02ae4b2dp-38, 0x1.09072d05b2fb6p-35, 0x1.4d24cd9014fcep-33,
    0x1.a2c421dfb5232p-31, 0x1.073282b03eed1p-28, 0x1.4ad7af33577bdp-26, 0x1.9fdf9a36a4427p-24,
    0x1.0561151fdbe6ep-21, 0x1.488ea29ce9e28p-19, 0x1.9d003013e00f9p-17, 0x1.0392de9af95fap-14,
    0x1.4649a09a903cbp-12, 0x1.9a25da6bb209fp-10, 0x1.01c7d97226742p-7, 0x1.4408a206c8528p-5,
    0x1.9750904262b55p-3, 0x1.0000000000000p+0, 0x1.41cb9fc8b327ap+2, 0x1.948048ac1d508p+4,
    0x1.fc7699523dad3p+6, 0x1.3f9292d3fea25p+9, 0x1.91b4faccd3bddp+11, 0x1.f8f373b80c35dp+13,
    0x1.3d5d7428cf59ap+16, 0x1.8eee9dd8229d7p+18, 0x1.f5768422236e1p+20, 0x1.3b2c3cd3aa8b8p+23,
    0x1.8c2d29113598fp+25, 0x1.f1ffbf94c9e4fp+27, 0x1.38fee5ed60392p+30, 0x1.897093caabdfbp+32,
    0x1.ee8f1b27b24ecp+34, 0x1.36d5689af56a6p+37, 0x1.86b8d5667cd45p+39, 0x1.eb248c05d92c5p+41,
    0x1.34afbe0d8e97bp+44, 0x1.8405e555dce9ep+46, 0x1.e7c0076d62af3p+48, 0x1.328ddf825a3a4p+51,
    0x1.8157bb1922b1dp+53, 0x1.e46182af78da5p+55, 0x1.306fc6427b817p+58, 0x1.7eae4e3fac18bp+60,
    0x1.e108f33029de4p+62, 0x1.2e556ba2f52edp+65, 0x1.7c099667c3d2fp+67, 0x1.ddb64e6646b15p+69,
    0x1.2c3ec9049496dp+72, 0x1.79698b3e86f83p+74, 0x1.da6989db41e22p+76, 0x1.2a2bd7d3dcc70p+79,
    0x1.76ce247fcacdap+81, 0x1.d7229b2b0ea5bp+83, 0x1.281c9188f1d0cp+86, 0x1.743759f602be2p+88,
    0x1.d3e17804001f0p+90, 0x1.2610efa784393p+93, 0x1.71a5237a26813p+95, 0x1.d0a61626a8e0dp+97,
    0x1.2408ebbebc8cap+100, 0x1.6f1778f3986efp+102, 0x1.cd706b65baa91p+104, 0x1.22047f6927171p+107,
    0x1.6c8e52580c01dp+109, 0x1.ca406da5e654dp+111, 0x1.2003a44c9fbfep+114, 0x1.6a09a7ab6c856p+116,
    0x1.c71612ddbc0d9p+118, 0x1.1e06541a3e094p+121, 0x1.678970ffc3f18p+123, 0x1.c3f151158bae2p+125,
};

FIP_HD float fip_pow233_tab(float x, const double* tc, const double* tm, const double* te) {
  const double d = (double)x;  // exact; a normal binary64 even for subnormal x
  const uint64_t u = fip_d2u(d);
  const int e = (int)((u >> 52) & 0x7ff) - 1023;  // 1024 for inf and NaN, -1023 for 0
  const int i = (int)((u >> 46) & 63);
  const double m = fip_u2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
  const double r = m * tc[i] - 1.0;  // both steps exact
  double s = -0x1.1614c07b9a345p-8;  // b6
  s = s * r + 0x1.38735c00987a1p-7;  // b5
  s = s * r - 0x1.d3bd8ced2995cp-6;  // b4
  s = s * r + 0x1.5d0f5b19f269ep-3;  // b3
  s = s * r + 0x1.8ca8befd22000p+0;  // b2
  s = s * r + 0x1.2a3d700000000p+1;  // b1 = p
  s = s * r + 1.0;                   // (1 + r)^p
  const int ei = e < kFipPowEMin ? kFipPowEMin : e > kFipPowEMax ? kFipPowEMax : e;
  float v = (float)((s * tm[i]) * te[ei - kFipPowEMin]);
  v = e < kFipPowEMin ? 0.0f : v;
  v = e > kFipPowEMax ? fip_u2f(0x7f800000u) : v;  // also +inf itself
  v = x == 0.0f ? 0.0f : v;
  return (x >= 0.0f) ? v : fip_u2f(0x7fc00000u);  // NaN for x < 0 and x = NaN
}

FIP_HD float fip_pow233(float x) { return fip_pow233_tab(x, kFipPowC, kFipPowM, kFipPowE); }
