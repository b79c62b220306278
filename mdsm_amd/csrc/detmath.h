// detmath.h -- natural log and sin/cos of a fraction of a turn built from
// IEEE +, -, *, / only, for the synthetic data generators.
//
// The library's log/sin/cos (device ocml, host libm) are accurate to about
// one ulp but not bit-identical to each other, so a generator that calls
// them cannot be reproduced bit for bit by a CPU program.  These functions
// use only correctly rounded operations in a fixed order (the library is
// built with -ffp-contract=off), so any CPU evaluation of the same
// expressions gives the same bits.  Accuracy is a few ulp, ample for
// Box-Muller noise and for the RFI sinusoids.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mdsm {

// ln x for a normal x > 0.  x = m * 2^e with m in [sqrt(1/2), sqrt(2)),
// ln x = e ln 2 + 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716; the odd
// series of atanh to s^23 leaves a truncation below 1e-19 relative.
__device__ __forceinline__ double det_log(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  int e = (int)((b >> 52) & 0x7FFu) - 1023;
  double m = __longlong_as_double((long long)((b & 0xFFFFFFFFFFFFFull) | 0x3FF0000000000000ull));
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e = e + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  p = p * s2 + 1.0;
  return (double)e * 0.6931471805599453 + 2.0 * (s * p);
}

// sin and cos of 2 pi u for 0 <= u <= 1.  The quarter turn q = round(4u)
// is split off exactly (r = u - q/4, |r| <= 1/8), x = 2 pi r lies in
// [-pi/4, pi/4] and the Taylor series to x^17 (sin) and x^18 (cos) leave
// truncations below 1e-18; the quadrant then swaps and negates.
__device__ __forceinline__ void det_sincos_turn(double u, double *sn, double *cs) {
  const int q = (int)(u * 4.0 + 0.5);
  const double r = u - (double)q * 0.25;
  const double x = r * 6.283185307179586;
  const double x2 = x * x;
  // sin x = x - x^3/3! + x^5 * ps, ps = 1/5! - x^2/7! + ... + x^12/17!
  double ps = 1.0 / 355687428096000.0;   // 1/17!
  ps = ps * x2 - 1.0 / 1307674368000.0;  // 1/15!
  ps = ps * x2 + 1.0 / 6227020800.0;     // 1/13!
  ps = ps * x2 - 1.0 / 39916800.0;       // 1/11!
  ps = ps * x2 + 1.0 / 362880.0;         // 1/9!
  ps = ps * x2 - 1.0 / 5040.0;           // 1/7!
  ps = ps * x2 + 1.0 / 120.0;            // 1/5!
  const double s = (x - (x * x2) * (1.0 / 6.0)) + ((x * x2) * x2) * ps;
  // cos x = 1 - x^2/2! + x^4/4! + x^6 * pc, pc = -1/6! + x^2/8! - ... - x^12/18!
  double pc = -1.0 / 6402373705728000.0;  // 1/18!
  pc = pc * x2 + 1.0 / 20922789888000.0;   // 1/16!
  pc = pc * x2 - 1.0 / 87178291200.0;      // 1/14!
  pc = pc * x2 + 1.0 / 479001600.0;        // 1/12!
  pc = pc * x2 - 1.0 / 3628800.0;          // 1/10!
  pc = pc * x2 + 1.0 / 40320.0;            // 1/8!
  pc = pc * x2 - 1.0 / 720.0;              // 1/6!
  const double c = ((1.0 - x2 * 0.5) + (x2 * x2) * (1.0 / 24.0)) + ((x2 * x2) * x2) * pc;
  switch (q & 3) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
  }
}

}  // namespace mdsm
