// clb_ade_spec_f32.hip -- fused transport + speciation step, FP32 population store.
#include "clb_ade_spec.h"

namespace clb {

template <>
bool launch_ade_spec_step<float>(const AdeArgs<float>& a, bool odd, dim3 block, const Tableau& tb,
                                 const FusedSpecIO& io, hipStream_t s) {
  return dispatch_fused<float>(a, odd, block, tb, io, s);
}

long long fused_waves(const Geom& g, dim3 block) {
  const long long blocks = (long long)((g.nx + block.x - 1) / block.x) * ((g.ny + block.y - 1) / block.y) * g.nz;
  return blocks * (block.x * block.y <= 256 ? 4 : 8);
}

}  // namespace clb
