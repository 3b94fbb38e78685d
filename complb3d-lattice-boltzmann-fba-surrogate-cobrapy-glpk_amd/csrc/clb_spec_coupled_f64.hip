// clb_spec_coupled_f64.hip -- speciation inside the coupled step, double populations:
// run_equilibrium_biotic (src/complab3d_processors_part4_eqsolver.hh:1355-1494).
#include "clb_spec_kernel.h"

namespace clb {

void launch_spec_coupled(const AdeArgs<double>& a, bool odd_done, const Tableau& tb, SpecArgs sa, int grid,
                             hipStream_t s) {
  CoupledIO io;
  io.g = (void*)a.g;
  io.mask = a.mask;
  io.list = (const int*)sa.iters;  // reactive node list travels in the iters slot
  io.cbuf = a.cbuf;
  io.geo = a.geo;
  io.odd_done = odd_done ? 1 : 0;
  for (int s = 0; s < CLB_MAX_SOLUTES; ++s) io.immobile[s] = a.immobile[s];
  io.pend_w = (void*)a.pend_w;
  sa.iters = nullptr;
  dispatch_spec<double, 1>(tb, sa, CLB_WARM_REF, io, grid, s);
}

}  // namespace clb
