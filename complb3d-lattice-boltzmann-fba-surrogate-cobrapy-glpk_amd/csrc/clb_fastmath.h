// clb_fastmath.h -- FP64 10^x and log10(y) for the per-voxel speciation solves.
//
// Written from the textbook reductions, with every multiply-add an explicit fma so the
// host and device builds round identically (the CPU test compares them with long double):
//   10^x = 2^t, t = x log2(10) carried as a double-double; t = n + f, |f| <= 1/2;
//          2^f = e^r, r = f ln 2 (|r| <= 0.347), e^r by its Taylor series to r^13/13!
//          (truncation < 2e-16 relative); the result is scaled by 2^n (ldexp).
//   ln y = e ln 2 + ln m with y = m 2^e, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s),
//          s = (m - 1) / (m + 1), |s| <= 0.1716, by the series s + s^3/3 + ... + s^23/23
//          (truncation < 1e-18); log10 y = ln y / ln 10.
// Both are accurate to a few ulp over the ranges the solver uses (|x| <= 300, y normal),
// and log10 is exact-relative near y = 1 (m - 1 is exact there), which is where the
// convergence judge evaluates it.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define CLB_HD __host__ __device__
#else
#define CLB_HD
#endif

namespace clb {

CLB_HD inline double fast_exp10(double x) {
  const double kL2tHi = 3.321928094887362181708567732130177319049835205078125;  // log2(10), nearest double
  const double kL2tLo = 1.6616175169735920e-16;                                   // log2(10) - kL2tHi
  const double kLn2Hi = 0.693147180559945286226763982995180413126945495605468750;  // ln 2, nearest double
  const double kLn2Lo = 2.319046813846299558e-17;                                  // ln 2 - kLn2Hi
  x = fmin(fmax(x, -320.0), 308.0);
  const double th = x * kL2tHi;
  const double tl = fma(x, kL2tHi, -th) + x * kL2tLo;
  const double n = rint(th);
  const double f = (th - n) + tl;
  const double r = fma(f, kLn2Hi, f * kLn2Lo);
  // e^r, Taylor coefficients 1/k!, Horner from k = 13
  double p = 1.0 / 6227020800.0;
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)n);
}

// y > 0 and finite.
CLB_HD inline double fast_log10(double y) {
  const double kLn2Hi = 0.693147180559945286226763982995180413126945495605468750;
  const double kLn2Lo = 2.319046813846299558e-17;
  const double kInvLn10 = 0.43429448190325182765;  // 1 / ln 10
  const double kSqrtHalf = 0.70710678118654752440;
  int e = 0;
  double m = frexp(y, &e);  // m in [0.5, 1)
  if (m < kSqrtHalf) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  // atanh(s) / s = sum_k s^(2k) / (2k + 1), k = 0..11
  double q = 1.0 / 23.0;
  q = fma(q, s2, 1.0 / 21.0);
  q = fma(q, s2, 1.0 / 19.0);
  q = fma(q, s2, 1.0 / 17.0);
  q = fma(q, s2, 1.0 / 15.0);
  q = fma(q, s2, 1.0 / 13.0);
  q = fma(q, s2, 1.0 / 11.0);
  q = fma(q, s2, 1.0 / 9.0);
  q = fma(q, s2, 1.0 / 7.0);
  q = fma(q, s2, 1.0 / 5.0);
  q = fma(q, s2, 1.0 / 3.0);
  q = fma(q, s2, 1.0);
  const double lnm = 2.0 * s * q;
  const double de = (double)e;
  const double ln = fma(de, kLn2Hi, fma(de, kLn2Lo, lnm));
  return ln * kInvLn10;
}

}  // namespace clb
