// clb_ade_spec_compact_f32.hip -- fused transport + speciation step on the compact layout,
// FP32 population store.
#include "clb_ade_spec.h"

namespace clb {

template <>
bool launch_ade_spec_step_compact<float>(const AdeArgs<float>& a, bool odd, int bt, const Tableau& tb,
                                         const FusedSpecIO& io, hipStream_t s) {
  return dispatch_fused_compact<float>(a, odd, bt, tb, io, s);
}

long long fused_waves_compact(long long ccap, int bt) { return ((ccap + bt - 1) / bt) * (bt <= 256 ? 4 : 8); }

}  // namespace clb
