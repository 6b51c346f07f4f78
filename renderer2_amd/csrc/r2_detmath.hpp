// Float sin/cos, exp and log evaluated in double from their series and rounded
// once to float.  Used where the filter must give the same bits on every device
// and on the host: the results are the correctly rounded float values except in
// cases closer than ~1e-16 relative to a rounding boundary.
//   sin/cos: quadrant reduction by pi/2 (two-part constant), Taylor series on |r| <= pi/4
//   exp:     x = n ln2 + r, |r| <= ln2/2, Taylor series to degree 13, scaled by 2^n
//   log:     x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)) by its odd series
// Compile with -ffp-contract=off so that no step is fused.
#pragma once
#include <cmath>
#include "r2_math.hpp"

namespace r2 {

R2_HD void det_sincosf(float x, float* s, float* c) {
  const double xd = (double)x;
  const double q = rint(xd * 0.63661977236758134308);
  const double r = (xd - q * 1.5707963267948966) - q * 6.123233995736766e-17;
  const double r2 = r * r;
  double sp = 1.0 / 355687428096000.0;
  sp = sp * r2 - 1.0 / 1307674368000.0;
  sp = sp * r2 + 1.0 / 6227020800.0;
  sp = sp * r2 - 1.0 / 39916800.0;
  sp = sp * r2 + 1.0 / 362880.0;
  sp = sp * r2 - 1.0 / 5040.0;
  sp = sp * r2 + 1.0 / 120.0;
  sp = sp * r2 - 1.0 / 6.0;
  sp = sp * r2 + 1.0;
  const double sr = sp * r;
  double cp = 1.0 / 20922789888000.0;
  cp = cp * r2 - 1.0 / 87178291200.0;
  cp = cp * r2 + 1.0 / 479001600.0;
  cp = cp * r2 - 1.0 / 3628800.0;
  cp = cp * r2 + 1.0 / 40320.0;
  cp = cp * r2 - 1.0 / 720.0;
  cp = cp * r2 + 1.0 / 24.0;
  cp = cp * r2 - 1.0 / 2.0;
  const double cr = cp * r2 + 1.0;
  const int iq = ((int)q) & 3;
  const double so = iq == 0 ? sr : (iq == 1 ? cr : (iq == 2 ? -sr : -cr));
  const double co = iq == 0 ? cr : (iq == 1 ? -sr : (iq == 2 ? -cr : sr));
  *s = (float)so;
  *c = (float)co;
}

// exp of a double argument in [-745, 709] by the same series (double result)
R2_HD double det_exp_d(double xd) {
  const double n = rint(xd * 1.4426950408889634);
  const double r = xd - n * 0.6931471805599453;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)n);
}

R2_HD float det_expf(float x) {
  if (x != x) return x;
  if (x < -104.f) return 0.f;
  if (x > 89.f) return INFINITY;
  return (float)det_exp_d((double)x);
}

// log of a positive finite double (double result)
R2_HD double det_log_d(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
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
  return 2.0 * s * p + (double)e * 0.6931471805599453;
}

R2_HD float det_logf(float x) {
  if (x != x || x < 0.f) return NAN;
  if (x == 0.f) return -INFINITY;
  if (x == INFINITY) return x;
  return (float)det_log_d((double)x);
}

// x^y for x >= 0 as exp(y log x) in double, rounded once (0^y = 0 for y > 0)
R2_HD float det_powf(float x, float y) {
  if (x != x || y != y || x < 0.f) return NAN;
  if (x == 0.f) return y > 0.f ? 0.f : (y == 0.f ? 1.f : INFINITY);
  if (x == INFINITY) return y > 0.f ? INFINITY : 0.f;
  const double t = (double)y * det_log_d((double)x);
  if (t < -745.0) return 0.f;
  if (t > 709.0) return INFINITY;
  return (float)det_exp_d(t);
}

}  // namespace r2
