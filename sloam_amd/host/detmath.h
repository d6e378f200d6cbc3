// Sine, cosine, arc tangent and arc cosine from IEEE-754 double addition,
// subtraction, multiplication, division and square root alone.  The host's
// libm and the GPU's return different last bits for some arguments; the pose
// composition, the angle-axis conversions and the model gates call these
// instead, so the host build and the device build of the same expression
// give the same bits (both compiled without contraction into fused
// multiply-adds).
//
// Method: Cody-Waite reduction by pi/2 (the reduced argument carried as a
// sum of two doubles) and the Taylor series of sin and cos on [-pi/4, pi/4];
// for the arc tangent, reduction to [0, 1] by atan t = pi/2 - atan(1/t),
// then by the tangent addition formula about 1/2 or 1 above 7/16, and the
// Taylor series; the rounding of y / x in atan2 is recovered and carried to
// first order.  Series truncation lies far below the last place.  Measured
// against a 64-bit-significand libm: sin and cos within 0.8 units in the
// last place, atan2 within 1.3, acos within 1.8 (the rounding of
// (1 - x)(1 + x) and of its square root added to atan2's).
#pragma once

#include <cmath>

#ifndef SLOAM_HD
#if defined(__HIPCC__)
#define SLOAM_HD __host__ __device__
#else
#define SLOAM_HD
#endif
#endif

namespace sloam_dm {

// pi/2 in three parts: the first two carry 33 significant bits, so k * part
// is exact for |k| < 2^20
constexpr double kPio2_1 = 0x1.921fb544p+0;
constexpr double kPio2_2 = 0x1.0b4611a6p-34;
constexpr double kPio2_3 = 0x1.3198a2e037073p-69;
constexpr double kTwoOverPi = 0.6366197723675814;
// largest |x| the reduction handles exactly (|k| < 2^20)
constexpr double kMaxReduce = 1.0e6;
// pi, pi/2 and pi/4 as the nearest double and the rest
constexpr double kPi_hi = 3.141592653589793, kPi_lo = 1.2246467991473532e-16;
constexpr double kPio2_hi = 1.5707963267948966, kPio2_lo = 6.123233995736766e-17;
constexpr double kPio4_hi = 0.7853981633974483, kPio4_lo = 3.061616997868383e-17;
// atan(1/2) as the nearest double and the rest; the interval bounds 7/16
// and 11/16 of the arc tangent's reduction
constexpr double kAtanHalf_hi = 0.4636476090008061, kAtanHalf_lo = 2.2698777452961687e-17;
constexpr double kAtanDirect = 0.4375, kAtanHalfEnd = 0.6875;

// sin(r + rl), |r| <= pi/4 (a little beyond is fine), rl a correction far
// below r's last place: the series through r^19 by Horner's rule in r^2,
// coefficients (-1)^k / (2k+1)! correctly rounded, plus rl cos r to first
// order
SLOAM_HD inline double sin_kernel(double r, double rl) {
  const double z = r * r;
  double p = -8.22063524662433e-18;  // -1/19!
  p = p * z + 2.8114572543455206e-15;
  p = p * z + -7.647163731819816e-13;
  p = p * z + 1.6059043836821613e-10;
  p = p * z + -2.505210838544172e-08;
  p = p * z + 2.7557319223985893e-06;
  p = p * z + -0.0001984126984126984;
  p = p * z + 0.008333333333333333;
  p = p * z + -0.16666666666666666;  // -1/3!
  return r + (r * z * p + rl * (1.0 - 0.5 * z));
}

// cos(r + rl), |r| <= pi/4: 1 - r^2/2 with its rounding error carried, plus
// the series from r^4 through r^18 (coefficients (-1)^k / (2k)!), minus
// rl sin r to first order
SLOAM_HD inline double cos_kernel(double r, double rl) {
  const double z = r * r;
  double p = -1.5619206968586225e-16;  // -1/18!
  p = p * z + 4.779477332387385e-14;
  p = p * z + -1.1470745597729725e-11;
  p = p * z + 2.08767569878681e-09;
  p = p * z + -2.755731922398589e-07;
  p = p * z + 2.48015873015873e-05;
  p = p * z + -0.001388888888888889;
  p = p * z + 0.041666666666666664;  // 1/4!
  const double hz = 0.5 * z;
  const double w = 1.0 - hz;
  return w + (((1.0 - w) - hz) + (z * z * p - r * rl));
}

// sin x and cos x; NaN for |x| > 1e6, infinities and NaN
SLOAM_HD inline void sincos(double x, double* s, double* c) {
  if (!(fabs(x) <= kMaxReduce)) {
    *s = *c = __builtin_nan("");
    return;
  }
  // x = k pi/2 + r + rl, |r| <= pi/4 (up to the rounding of x * 2/pi): x -
  // k kPio2_1 and both products are exact; the rounding of the subtraction
  // of k kPio2_2 is recovered by TwoSum into rl
  const double k = __builtin_rint(x * kTwoOverPi);
  const double a = x - k * kPio2_1, b = -(k * kPio2_2);
  const double r1 = a + b;
  const double bv = r1 - a, av = r1 - bv;
  const double e = (a - av) + (b - bv);
  const double rl1 = e - k * kPio2_3;
  const double r = r1 + rl1;
  const double rl = rl1 - (r - r1);
  const double sr = sin_kernel(r, rl), cr = cos_kernel(r, rl);
  switch ((int)k & 3) {  // k mod 4, also for negative k
    case 0:
      *s = sr;
      *c = cr;
      break;
    case 1:
      *s = cr;
      *c = -sr;
      break;
    case 2:
      *s = -sr;
      *c = -cr;
      break;
    default:
      *s = -cr;
      *c = sr;
      break;
  }
}

SLOAM_HD inline double sin(double x) {
  double s, c;
  sincos(x, &s, &c);
  return s;
}

SLOAM_HD inline double cos(double x) {
  double s, c;
  sincos(x, &s, &c);
  return c;
}

// atan u - u, |u| <= 7/16: the series from u^3 through u^45 by Horner's rule
// in u^2, coefficients (-1)^k / (2k+1) correctly rounded
SLOAM_HD inline double atan_tail(double u) {
  const double z = u * u;
  double p = 0.022222222222222223;  // 1/45
  p = p * z + -0.023255813953488372;
  p = p * z + 0.024390243902439025;
  p = p * z + -0.02564102564102564;
  p = p * z + 0.02702702702702703;
  p = p * z + -0.02857142857142857;
  p = p * z + 0.030303030303030304;
  p = p * z + -0.03225806451612903;
  p = p * z + 0.034482758620689655;
  p = p * z + -0.037037037037037035;
  p = p * z + 0.04;
  p = p * z + -0.043478260869565216;
  p = p * z + 0.047619047619047616;
  p = p * z + -0.05263157894736842;
  p = p * z + 0.058823529411764705;
  p = p * z + -0.06666666666666667;
  p = p * z + 0.07692307692307693;
  p = p * z + -0.09090909090909091;
  p = p * z + 0.1111111111111111;
  p = p * z + -0.14285714285714285;
  p = p * z + 0.2;
  p = p * z + -0.3333333333333333;  // -1/3
  return u * z * p;
}

SLOAM_HD inline double atan_kernel(double u) { return u + atan_tail(u); }

// atan(t + tl) for 0 <= t <= 1, tl a correction far below t's last place:
// the series up to 7/16; above, atan t = atan a + atan((t - a) / (1 + t a))
// with a = 1/2 up to 11/16 (the reduced argument as (2t - 1) / (2 + t), its
// numerator exact) and a = 1 beyond, so that atan a dominates the sum; tl
// enters to first order as tl / (1 + t^2)
SLOAM_HD inline double atan_unit(double t, double tl) {
  const double dt = tl / (1.0 + t * t);
  if (t <= kAtanDirect) return t + (atan_tail(t) + dt);
  if (t <= kAtanHalfEnd) {
    const double u = (2.0 * t - 1.0) / (2.0 + t);
    return kAtanHalf_hi + (atan_kernel(u) + (kAtanHalf_lo + dt));
  }
  const double u = (t - 1.0) / (1.0 + t);
  return kPio4_hi + (atan_kernel(u) + (kPio4_lo + dt));
}

// n / d for 0 <= n <= d as q + ql with ql the quotient of the division's
// remainder: n - q d from Dekker's exact product (Veltkamp splitting, no
// fused multiply-add), exact for operands below 1e290
SLOAM_HD inline double div_rem(double n, double d, double* ql) {
  const double q = n / d;
  *ql = 0.0;
  if (d < 1.0e290) {
    constexpr double kSplit = 134217729.0;  // 2^27 + 1
    const double cq = kSplit * q, qh = cq - (cq - q), qlo = q - qh;
    const double cd = kSplit * d, dh = cd - (cd - d), dlo = d - dh;
    const double p = q * d;
    const double pe = ((qh * dh - p) + qh * dlo + qlo * dh) + qlo * dlo;
    *ql = ((n - p) - pe) / d;
  }
  return q;
}

// atan2(y, x) in [-pi, pi] with the signs of zeros as in C's atan2; NaN when
// either argument is NaN or both are infinite
SLOAM_HD inline double atan2(double y, double x) {
  if (x != x || y != y) return x + y;
  const double ax = fabs(x), ay = fabs(y);
  double r, tl = 0.0;
  if (ay <= ax) {
    r = 0.0;
    if (ax != 0.0) {
      const double t = div_rem(ay, ax, &tl);
      r = atan_unit(t, tl);
    }
  } else {
    const double t = div_rem(ax, ay, &tl);
    r = kPio2_hi - (atan_unit(t, tl) - kPio2_lo);
  }
  if (__builtin_signbit(x)) r = kPi_hi - (r - kPi_lo);
  return __builtin_signbit(y) ? -r : r;
}

// acos x = atan2(sqrt((1 - x)(1 + x)), x); NaN outside [-1, 1]
SLOAM_HD inline double acos(double x) { return sloam_dm::atan2(sqrt((1.0 - x) * (1.0 + x)), x); }

}  // namespace sloam_dm
