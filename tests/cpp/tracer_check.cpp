// tracer_check.cpp -- host-only check of complab_gpu::ValueTracer (no GPU, no library calls).
// Reads "u L epsilon" and then values from stdin; prints, for every value taken, 1 when the
// tracer has converged after it, else 0, one per line.
#include <cstdio>

#include "complab_gpu.hh"

int main() {
  double u = 0, L = 0, eps = 0;
  if (std::scanf("%lf %lf %lf", &u, &L, &eps) != 3) return 2;
  complab_gpu::ValueTracer t(u, L, eps);
  double v = 0;
  while (std::scanf("%lf", &v) == 1) {
    t.takeValue(v);
    std::printf("%d\n", t.hasConverged() ? 1 : 0);
  }
  return 0;
}
