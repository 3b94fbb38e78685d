// clb_fastmath.h -- FP64 10^x and log10(y) for the per-voxel speciation solves.
//
// Written from the textbook reductions, with every multiply-add an explicit fma so the
// host and device builds round identically (the CPU test compares them with long double):
//   10^x = 2^t, t = x log2(10) carried as a double-double; t = n + f, |f| <= 1/2;
//          2^f = e^r, r = f ln 2 (|r| <= 0.347), e^r by its Taylor series to r^13/13!
//          (truncation < 2e-16 relative); the result is scaled by 2^n (ldexp).
// The polynomials are evaluated by Estrin's scheme, which keeps the dependency chains
// short (the per-voxel solves are latency-bound).
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

// 1 / v and 1 / sqrt(v) for v > 0: on the device the hardware estimate refined by two
// Newton-Raphson steps (each doubles the correct bits), on the host the IEEE operation.
CLB_HD inline double fast_rcp(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(v);
  r = fma(r, fma(-v, r, 1.0), r);
  r = fma(r, fma(-v, r, 1.0), r);
  return r;
#else
  return 1.0 / v;
#endif
}

CLB_HD inline double fast_rsqrt(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rsq(v);
  y = y * fma(-0.5 * v, y * y, 1.5);
  y = y * fma(-0.5 * v, y * y, 1.5);
  return y;
#else
  return 1.0 / sqrt(v);
#endif
}

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
  // e^r, Taylor coefficients 1/k! to k = 13, evaluated by Estrin's scheme (pairs, then
  // pairs of pairs) so the dependency chain is 5 deep instead of 13
  const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
  const double a0 = fma(r, 1.0, 1.0);
  const double a1 = fma(r, 1.0 / 6.0, 0.5);
  const double a2 = fma(r, 1.0 / 120.0, 1.0 / 24.0);
  const double a3 = fma(r, 1.0 / 5040.0, 1.0 / 720.0);
  const double a4 = fma(r, 1.0 / 362880.0, 1.0 / 40320.0);
  const double a5 = fma(r, 1.0 / 39916800.0, 1.0 / 3628800.0);
  const double a6 = fma(r, 1.0 / 6227020800.0, 1.0 / 479001600.0);
  const double b0 = fma(r2, a1, a0);
  const double b1 = fma(r2, a3, a2);
  const double b2 = fma(r2, a5, a4);
  const double c0 = fma(r4, b1, b0);
  const double c1 = fma(r4, a6, b2);
  const double p = fma(r8, c1, c0);
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
  const double s = (m - 1.0) * fast_rcp(m + 1.0);
  const double s2 = s * s;
  // atanh(s) / s = sum_k s^(2k) / (2k + 1), k = 0..11, by Estrin's scheme in t = s^2
  const double t = s2, t2 = t * t, t4 = t2 * t2, t8 = t4 * t4;
  const double a0 = fma(t, 1.0 / 3.0, 1.0);
  const double a1 = fma(t, 1.0 / 7.0, 1.0 / 5.0);
  const double a2 = fma(t, 1.0 / 11.0, 1.0 / 9.0);
  const double a3 = fma(t, 1.0 / 15.0, 1.0 / 13.0);
  const double a4 = fma(t, 1.0 / 19.0, 1.0 / 17.0);
  const double a5 = fma(t, 1.0 / 23.0, 1.0 / 21.0);
  const double b0 = fma(t2, a1, a0);
  const double b1 = fma(t2, a3, a2);
  const double b2 = fma(t2, a5, a4);
  const double c0 = fma(t4, b1, b0);
  const double q = fma(t8, b2, c0);
  const double lnm = 2.0 * s * q;
  const double de = (double)e;
  const double ln = fma(de, kLn2Hi, fma(de, kLn2Lo, lnm));
  return ln * kInvLn10;
}

}  // namespace clb
