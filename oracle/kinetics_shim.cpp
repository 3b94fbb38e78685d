// kinetics_shim.cpp -- TEST INFRASTRUCTURE: compiles a defineKinetics.hh plugin as
// host C++ and adapts it to the oracle's flattened hook type, the way the reference's
// run_kinetics processor builds std::vector arguments per voxel
// (src/complab3d_processors_part1.hh:79-100). The plugin path comes from
// -DCLB_KINETICS_HEADER.
#include <cstddef>
#include <vector>

namespace plb { typedef long long plint; }

#include CLB_KINETICS_HEADER

extern "C" __attribute__((visibility("default"))) void orc_kinetics_plugin(const double* B, int nb, const double* C,
                                                                            int ns, double* subsR, double* bioR,
                                                                            long long mask) {
  std::vector<double> b(B, B + nb), c(C, C + ns), sr(ns, 0.0), br(nb, 0.0);
  defineRxnKinetics(b, c, sr, br, (plb::plint)mask);
  for (int i = 0; i < ns; ++i) subsR[i] = sr[i];
  for (int i = 0; i < nb; ++i) bioR[i] = br[i];
}

#include CLB_ABIOTIC_HEADER

extern "C" __attribute__((visibility("default"))) void orc_abiotic_plugin(const double* C, int ns, double* subsR,
                                                                           long long mask) {
  std::vector<double> c(C, C + ns), sr(ns, 0.0);
  defineAbioticRxnKinetics(c, sr, (plb::plint)mask);
  for (int i = 0; i < ns; ++i) subsR[i] = sr[i];
}

// defineDissolutionRate (defineAbioticKinetics.hh:292 of the reference) for the oracle.
extern "C" __attribute__((visibility("default"))) void orc_dissolution_plugin(long long phase_id, const double* C,
                                                                               int ns, double mineral, double* subsR,
                                                                               double* mineralR, long long mask) {
  std::vector<double> c(C, C + ns), sr(ns, 0.0);
  double mr = 0.0;
  defineDissolutionRate((plb::plint)phase_id, c, mineral, sr, mr, (plb::plint)mask);
  for (int i = 0; i < ns; ++i) subsR[i] = sr[i];
  *mineralR = mr;
}
