// psgc_math.h — deterministic FP64 scalar routines shared by the HIP kernels
// and the host code, so that host and device produce bit-identical results.
//
// Every routine uses only IEEE-754 basic operations (+ - * /, sqrt, round,
// comparisons), each of which is correctly rounded on x86-64 and on gfx950.
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
PSGC_HD inline double atan01(double x) {
  const int i = (int)(x * 64.0 + 0.5);
  const double c = (double)i * 0.015625;
  const double t = (x - c) / (1.0 + x * c);
  return kAtanTab[i] + atan_small(t);
}

// atan2 with C99 semantics for signed zeros and infinities; within 2 ulp of
// a correctly rounded result (checked against the C library in the tests).
PSGC_HD inline double atan2(double y, double x) {
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
    a = atan01(r);
    if (swap) a = (kPio2Hi - a) + kPio2Lo;
    if (xneg) a = (kPiHi - a) + kPiLo;
  }
  return __builtin_signbit(y) ? -a : a;
}

// C round(): halves away from zero.
PSGC_HD inline double round_half_away(double v) { return __builtin_round(v); }

}  // namespace psgc
