// clb_plugin_check.hip -- compiled for the device only, without code generation, before the
// library (Makefile): every hook of the configured kinetics plugins is called from a kernel,
// so a plugin that cannot run on the GPU fails here, in about a second, and the build says
// why. Nothing of it is linked.
#include "clb_kinetics_device.h"

__global__ void clb_plugin_check_kernel(double* out) {
  clb::KVec B(CLB_MAX_BIOMASS, 1.0), C(CLB_MAX_SOLUTES, 1.0);
  clb::KVec sr(CLB_MAX_SOLUTES, 0.0), br(CLB_MAX_BIOMASS, 0.0), ar(CLB_MAX_SOLUTES, 0.0), dr(CLB_MAX_SOLUTES, 0.0);
  defineRxnKinetics(B, C, sr, br, (plb::plint)2);
  defineAbioticRxnKinetics(C, ar, (plb::plint)2);
  double mineralR = 0.0;
  defineDissolutionRate((plb::plint)1, C, 1.0, dr, mineralR, (plb::plint)2);
  out[threadIdx.x] = sr[0] + br[0] + ar[0] + dr[0] + mineralR;
}
