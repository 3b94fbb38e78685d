// clb_spec_coupled_f32.hip -- speciation inside the coupled step, float populations:
// run_equilibrium_biotic (src/complab3d_processors_part4_eqsolver.hh:1355-1494).
#include "clb_spec_kernel.h"

namespace clb {

void launch_spec_coupled(const AdeArgs<float>& a, bool odd_done, const Tableau& tb, SpecArgs sa, int grid,
                             hipStream_t s) {
  CoupledIO io{};  // value-initialized: fields a launcher does not set are null / zero
  io.g = (void*)a.g;
  io.mask = a.mask;
  io.list = (const int*)sa.iters;  // node list (reactive voxels, or the first pass's failures)
  io.cbuf = a.cbuf;
  io.geo = a.geo;
  io.odd_done = odd_done ? 1 : 0;
  for (int s = 0; s < CLB_MAX_SOLUTES; ++s) io.immobile[s] = a.immobile[s];
  io.pend_w = (void*)a.pend_w;
  io.blk_stats = sa.blk_stats;
  io.blk_stats_n = sa.blk_stats_n;
  sa.iters = nullptr;
  // the fused step's stragglers (sa.seg_*): a first pass continues their fused solves and
  // leaves what it does not finish in the failure list (io.list, sa.logc, sa.n_dev), for the
  // full iteration; that list is short (empty in most steps), so the full iteration runs on a
  // small grid, whose work queue takes any number
  const bool first = io.pend_w && sa.seg_n && dispatch_straggler<float>(tb, sa, io, grid, s);
  dispatch_spec<float, 1>(tb, sa, CLB_WARM_REF, io, first ? (grid < 8 ? grid : 8) : grid, s);
}

}  // namespace clb
