// clb_ade_spec_compact_f64.hip -- fused transport + speciation step on the compact layout,
// FP64 population store.
#include "clb_ade_spec.h"

namespace clb {

template <>
bool launch_ade_spec_step_compact<double>(const AdeArgs<double>& a, bool odd, int bt, const Tableau& tb,
                                          const FusedSpecIO& io, hipStream_t s) {
  return dispatch_fused_compact<double>(a, odd, bt, tb, io, s);
}

}  // namespace clb
