// clb_ade_spec_f64.hip -- fused transport + speciation step, FP64 population store.
#include "clb_ade_spec.h"

namespace clb {

template <>
bool launch_ade_spec_step<double>(const AdeArgs<double>& a, bool odd, dim3 block, const Tableau& tb,
                                  const FusedSpecIO& io, hipStream_t s) {
  return dispatch_fused<double>(a, odd, block, tb, io, s);
}

}  // namespace clb
