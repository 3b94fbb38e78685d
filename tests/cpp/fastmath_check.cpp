// fastmath_check.cpp -- accuracy of clb_fastmath.h (host build of the same code the
// device runs) against long-double references; prints the worst errors.
#include <cmath>
#include <cstdio>
#include <random>

#include "clb_fastmath.h"

int main() {
  std::mt19937_64 g(1);
  std::uniform_real_distribution<double> ux(-300.0, 250.0), uy(-35.0, 35.0), u1(-1.0, 1.0);
  long double e_exp = 0, e_log = 0, e_near1 = 0;
  for (int k = 0; k < 1000000; ++k) {
    const double x = ux(g);
    const long double r = powl(10.0L, (long double)x);
    e_exp = fmaxl(e_exp, fabsl(((long double)clb::fast_exp10(x) - r) / r));
    const double y = pow(10.0, uy(g));
    const long double l = log10l((long double)y);
    e_log = fmaxl(e_log, fabsl((long double)clb::fast_log10(y) - l) / fmaxl(1.0L, fabsl(l)));
    const double z = 1.0 + u1(g) * 1e-9;
    if (z == 1.0) continue;
    const long double lz = log10l((long double)z);
    e_near1 = fmaxl(e_near1, fabsl(((long double)clb::fast_log10(z) - lz) / lz));
  }
  // special points: exact powers of ten, the clamp ends the solver uses
  long double e_pts = 0;
  const double pts[] = {-30.0, -20.0, -7.0, -1.0, 0.0, 1.0, 2.0, 250.0};
  for (double x : pts) {
    const long double r = powl(10.0L, (long double)x);
    e_pts = fmaxl(e_pts, fabsl(((long double)clb::fast_exp10(x) - r) / r));
  }
  std::printf("%.6Le %.6Le %.6Le %.6Le %.17g\n", e_exp, e_log, e_near1, e_pts, clb::fast_log10(1.0));
  return 0;
}
