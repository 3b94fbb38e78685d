// clb_spec_coupled_f32.hip -- speciation inside the coupled step, float populations:
// run_equilibrium_biotic (src/complab3d_processors_part4_eqsolver.hh:1355-1494).
#include "clb_spec_kernel.h"

namespace clb {

void launch_spec_coupled(const AdeArgs<float>& a, bool odd_done, const Tableau& tb, SpecArgs sa, int grid,
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
  io.blk_stats = sa.blk_stats;
  io.blk_stats_n = sa.blk_stats_n;
  io.more = sa.more;  // cleared before the step; the first pass sets it when it leaves a voxel
  sa.iters = nullptr;
  // the fused step's stragglers (sa.logc: the points their solves reached): a first pass
  // continues the fused solve; the full iteration finishes what it leaves, and returns at
  // once when it left none. Without the first pass the full iteration takes every straggler.
  bool first = false;
  if (io.pend_w && sa.logc && sa.n_dev && sa.more) first = dispatch_straggler<float>(tb, sa, io, grid, s);
  if (!first) io.more = nullptr;
  // after a first pass the full iteration sees at most a few voxels (none in most steps):
  // a small grid, whose work queue takes any number; launching the persistent grid's
  // thousand wavefronts of a kernel this size costs about 50 us even when they return at once
  dispatch_spec<float, 1>(tb, sa, CLB_WARM_REF, io, first ? (grid < 8 ? grid : 8) : grid, s);
}

}  // namespace clb
