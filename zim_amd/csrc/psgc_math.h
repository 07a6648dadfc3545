// psgc_math.h — deterministic scalar routines shared by the HIP kernels and
// the host code, so that host and device produce bit-identical results: the
// FP64 atan2 of the decode and the float exp / acos / cos of the MLS filter.
//
// Every routine uses only IEEE-754 basic operations (+ - * /, fused
// multiply-add where written explicitly, sqrt, rounding to an integer,
// scaling by a power of two, conversions, comparisons), each of which is
// exact or correctly rounded on x86-64 (with FMA3) and on gfx950.
// Translation units that include this header must be compiled without
// floating-point contraction (-ffp-contract=off) so that no mul+add pair is
// fused differently on the two sides.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PSGC_HD __host__ __device__
#else
#define PSGC_HD
#endif

namespace psgc {

// atan(i/64), i = 0..64 (generated by gen_atan_table.py).
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __constant__ static const double kAtanTab[65] = {
#include "atan_table.inc"
};
#else
static const double kAtanTab[65] = {
#include "atan_table.inc"
};
#endif

constexpr double kPi = 3.14159265358979323846;
constexpr double kPiHi = 0x1.921fb54442d18p+1;    // pi rounded to double
constexpr double kPiLo = 0x1.1a62633145c07p-53;   // pi - kPiHi
constexpr double kPio2Hi = 0x1.921fb54442d18p+0;  // pi/2 rounded to double
constexpr double kPio2Lo = 0x1.1a62633145c07p-54; // pi/2 - kPio2Hi

// atan(t) for |t| <= 1/128 by its odd Taylor series through t^7; the first
// omitted term t^9/9 is below 2^-56 * |t|.
PSGC_HD inline double atan_small(double t) {
  const double t2 = t * t;
  const double p = (-1.0 / 3.0) + t2 * ((1.0 / 5.0) + t2 * (-1.0 / 7.0));
  return t + (t * t2) * p;
}

// atan(x) for x in [0, 1]: x = c + d with c = i/64 the nearest table point,
// atan(x) = atan(c) + atan((x - c) / (1 + x c)). The numerator is exact
// (Sterbenz), the argument of atan_small is at most 1/128.
// tab: the 65 table values (kAtanTab, or a copy of it, e.g. in LDS, where a
// per-lane lookup does not wait on the memory hierarchy).
PSGC_HD inline double atan01(double x, const double* tab) {
  const int i = (int)(x * 64.0 + 0.5);
  const double c = (double)i * 0.015625;
  const double t = (x - c) / (1.0 + x * c);
  return tab[i] + atan_small(t);
}

// atan2 with C99 semantics for signed zeros and infinities; within 2 ulp of
// a correctly rounded result (checked against the C library in the tests).
PSGC_HD inline double atan2(double y, double x, const double* tab) {
  if (__builtin_isnan(x) || __builtin_isnan(y)) return x + y;
  const double ax = __builtin_fabs(x);
  const double ay = __builtin_fabs(y);
  const bool xneg = __builtin_signbit(x);
  double a;
  if (ay == 0.0) {
    a = xneg ? kPi : 0.0;
  } else if (__builtin_isinf(ax) && __builtin_isinf(ay)) {
    a = xneg ? 0.75 * kPi : 0.25 * kPi;
  } else {
    const bool swap = ay > ax;
    const double r = swap ? (ax / ay) : (ay / ax);
    a = atan01(r, tab);
    if (swap) a = (kPio2Hi - a) + kPio2Lo;
    if (xneg) a = (kPiHi - a) + kPiLo;
  }
  return __builtin_signbit(y) ? -a : a;
}

PSGC_HD inline double atan2(double y, double x) { return atan2(y, x, kAtanTab); }

// C round(): halves away from zero.
PSGC_HD inline double round_half_away(double v) { return __builtin_round(v); }

// ---- float exp / acos / cos ---------------------------------------------
// Each is evaluated in double and rounded once to float, so the result is
// the correctly rounded float unless the exact value lies within ~1e-13
// (relative) of a rounding boundary; a C library's correctly rounded or
// sub-ulp float routine agrees with it except in such cases.

constexpr double kInvLn2 = 0x1.71547652b82fep+0;  // 1/ln 2
constexpr double kLn2Hi = 0x1.62e42p-1;           // ln 2, 20 significant bits
constexpr double kLn2Lo = 0x1.fdf473de6af28p-22;  // ln 2 - kLn2Hi
constexpr double k2OverPi = 0x1.45f306dc9c883p-1; // 2/pi
constexpr double kPio2_1 = 0x1.921fap+0;          // pi/2 = P1 + P2 + P3, P1 and P2
constexpr double kPio2_2 = 0x1.54442p-20;         // with 20 significant bits so that
constexpr double kPio2_3 = 0x1.a308d313198a3p-41; // n * P1, n * P2 are exact

// Fused multiply-add a * b + c with one rounding. On the device it is the
// three-operand v_fma_f64 (the compiler otherwise copies a loop-invariant
// addend into the destination of the two-operand form before every use);
// the result is the same IEEE fma on both sides.
PSGC_HD inline double fma_rn(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
#else
  return __builtin_fma(a, b, c);
#endif
}

// a * b + c in FP32: with C one fused multiply-add (a single rounding, what a
// compiler that contracts multiply-adds emits), else a rounded product and a
// rounded sum.
template <bool C>
PSGC_HD inline float madd(float a, float b, float c) {
  if constexpr (C) return __builtin_fmaf(a, b, c);
  else return a * b + c;
}

// dx^2 + dy^2 + dz^2 summed left to right, contracted with C.
template <bool C>
PSGC_HD inline float sq3(float dx, float dy, float dz) {
  return madd<C>(dz, dz, madd<C>(dy, dy, dx * dx));
}

// 2^(j/64), j = 0..63, each the double nearest to the exact value.
constexpr double kExp2Tab64[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0,
    0x1.0874518759bc8p+0, 0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0,
    0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0, 0x1.172b83c7d517bp+0,
    0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0,
    0x1.2d285a6e4030bp+0, 0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0,
    0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0, 0x1.3dea64c123422p+0,
    0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0,
    0x1.56f4736b527dap+0, 0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0,
    0x1.6247eb03a5585p+0, 0x1.6623882552225p+0, 0x1.6a09e667f3bcdp+0,
    0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0,
    0x1.868d99b4492edp+0, 0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0,
    0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0, 0x1.9c49182a3f090p+0,
    0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0,
    0x1.bcc1e904bc1d2p+0, 0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0,
    0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0, 0x1.d5818dcfba487p+0,
    0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0,
    0x1.fa7c1819e90d8p+0
};

// e^x: x = (64 e + j) ln2/64 + r with |r| <= ln2/128 (Cody and Waite's split
// of ln 2 / 64; the product of k = 64 e + j with its leading part is exact),
// e^x = 2^e 2^(j/64) e^r, e^r by its Taylor series through r^5 (first omitted
// term below 4e-17 e^r) in Horner form with fused multiply-adds (one rounding
// each, the same on both sides). k = rint(64 x / ln 2) is formed in float (any
// integer within a few units of 64 x / ln 2 keeps |r| far inside the series'
// range). The double result is within a few units of 2^-53 of e^x, and over
// every float its rounding to float equals the correctly rounded e^x.
// `tab` holds kExp2Tab64 (on the device a copy in shared memory, so that the
// per-lane lookup is one LDS read). exp_f32_core takes xc in [-104, 89].
PSGC_HD inline float exp_f32_core(float xc, const double* tab) {
  const float kf = __builtin_rintf(xc * (float)(64.0 * kInvLn2));
  const int ki = (int)kf;
  const double t = tab[ki & 63];
  const double x = xc, k = kf;
  const double r = __builtin_fma(-k, kLn2Lo / 64.0, __builtin_fma(-k, kLn2Hi / 64.0, x));
  double p = fma_rn(1.0 / 120.0, r, 1.0 / 24.0);
  p = fma_rn(p, r, 1.0 / 6.0);
  p = __builtin_fma(p, r, 0.5);
  p = __builtin_fma(p, r, 1.0);
  p = __builtin_fma(p, r, 1.0);
  return (float)__builtin_ldexp(p * t, ki >> 6);
}

// Arguments are clamped to [-104, 89] rather than branched on: e^-104 rounds to
// float 0 and e^89 to float infinity, the values of the unclamped tails.
PSGC_HD inline float exp_f32(float xf, const double* tab) {
  const float res = exp_f32_core(__builtin_fminf(__builtin_fmaxf(xf, -104.0f), 89.0f), tab);
  return __builtin_isnan(xf) ? xf : res;
}

// exp_f32 of an argument known to be <= 0 and not NaN (a Gaussian weight of
// a finite squared distance): the same result with the upper clamp and the
// NaN test left out.
PSGC_HD inline float exp_f32_nonpos(float xf, const double* tab) {
  return exp_f32_core(__builtin_fmaxf(xf, -104.0f), tab);
}

PSGC_HD inline float exp_f32(float xf) { return exp_f32(xf, kExp2Tab64); }

// 2^(j/256), j = 0..255, each the double nearest to the exact value
// (generated by gen_exp2_table.py).
constexpr double kExp2Tab256[256] = {
#include "exp2_table.inc"
};

// p * r + q with the scalar coefficient p in a scalar register on the device
// (one v_fma_f64; the addend q stays where it is).
PSGC_HD inline double fma_rn_s(double p, double r, double q) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "s"(p), "v"(r), "v"(q));
  return d;
#else
  return __builtin_fma(p, r, q);
#endif
}

// exp_f32_nonpos with fewer instructions, for the Gaussian weights of the
// organized MLS kernel: x = (256 e + j) ln2/256 + r with |r| <= ln2/512, where
// k = 256 e + j, the integer nearest to 256 x / ln 2, comes from one fused
// multiply-add with 1.5 * 2^52 (the sum is rounded to an integer, which its
// low word holds); e^r by its Taylor series through r^4 (first omitted term
// below 4e-17 e^r) in Horner form with fused multiply-adds; `tab` holds
// kExp2Tab256 (on the device a copy in shared memory). For every float
// x <= 0 (and NaN, which both clamp to -104) it returns exp_f32_nonpos's
// float, checked exhaustively over the 2^31 non-positive floats.
PSGC_HD inline float exp_f32_w(float xf, const double* tab) {
  const double x = (double)__builtin_fmaxf(xf, -104.0f);
  constexpr double kRound = 0x1.8p52;
  const double s = __builtin_fma(x, 256.0 * kInvLn2, kRound);
  const double k = s - kRound;
  const int ki = (int)(unsigned)__builtin_bit_cast(unsigned long long, s);
  const double r = __builtin_fma(-k, kLn2Lo / 256.0, __builtin_fma(-k, kLn2Hi / 256.0, x));
  double p = fma_rn_s(1.0 / 24.0, r, 1.0 / 6.0);
  p = __builtin_fma(p, r, 0.5);
  p = __builtin_fma(p, r, 1.0);
  p = __builtin_fma(p, r, 1.0);
  return (float)__builtin_ldexp(p * tab[ki & 255], ki >> 8);
}

// acos(x) = atan2(sqrt((1 - x)(1 + x)), x) in double (the product of two
// float-derived factors is exact, the square root correctly rounded).
PSGC_HD inline float acos_f32(float xf) {
  if (__builtin_isnan(xf) || xf > 1.0f || xf < -1.0f) return __builtin_nanf("");
  const double x = xf;
  return (float)atan2(__builtin_sqrt((1.0 - x) * (1.0 + x)), x);
}

// cos(x) for |x| < 2^20: x = n pi/2 + r, |r| <= pi/4, with the three-part
// pi/2 above; then +-cos r or +-sin r by their Taylor series through r^16 /
// r^17 (first omitted terms below 5e-17).
PSGC_HD inline float cos_f32(float xf) {
  if (__builtin_isnan(xf) || __builtin_isinf(xf)) return __builtin_nanf("");
  const double x = __builtin_fabs((double)xf);
  const double n = __builtin_rint(x * k2OverPi);
  const double r = ((x - n * kPio2_1) - n * kPio2_2) - n * kPio2_3;
  const double r2 = r * r;
  const int q = (int)__builtin_fmod(n, 4.0);
  double v;
  if (q == 0 || q == 2) {
    double c = 1.0 / 20922789888000.0;  // 1/16!
    c = __builtin_fma(c, r2, -1.0 / 87178291200.0);
    c = __builtin_fma(c, r2, 1.0 / 479001600.0);
    c = __builtin_fma(c, r2, -1.0 / 3628800.0);
    c = __builtin_fma(c, r2, 1.0 / 40320.0);
    c = __builtin_fma(c, r2, -1.0 / 720.0);
    c = __builtin_fma(c, r2, 1.0 / 24.0);
    c = __builtin_fma(c, r2, -0.5);
    v = __builtin_fma(c, r2, 1.0);
  } else {
    double sn = 1.0 / 355687428096000.0;  // 1/17!
    sn = __builtin_fma(sn, r2, -1.0 / 1307674368000.0);
    sn = __builtin_fma(sn, r2, 1.0 / 6227020800.0);
    sn = __builtin_fma(sn, r2, -1.0 / 39916800.0);
    sn = __builtin_fma(sn, r2, 1.0 / 362880.0);
    sn = __builtin_fma(sn, r2, -1.0 / 5040.0);
    sn = __builtin_fma(sn, r2, 1.0 / 120.0);
    sn = __builtin_fma(sn, r2, -1.0 / 6.0);
    v = __builtin_fma(r * r2, sn, r);
  }
  // cos(n pi/2 + r) = cos r, -sin r, -cos r, sin r for n mod 4 = 0..3
  if (q == 1 || q == 2) v = -v;
  return (float)v;
}

}  // namespace psgc
