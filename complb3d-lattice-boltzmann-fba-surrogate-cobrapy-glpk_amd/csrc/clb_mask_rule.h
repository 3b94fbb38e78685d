// clb_mask_rule.h -- the pore <-> biofilm rule of updateLocalMaskNtotalLattices3D
// (src/complab3d_processors_part2.hh:99-210) for one voxel, shared by the step's mask pass
// (clb_ade.hip) and the CA's mask / total pass (clb_ca.hip).
#pragma once

#include "clb_internal.h"

namespace clb {

// The new code of voxel (x, y, z) of code m (neither solid nor bounce-back), whose biomass
// fields above 1e-14 sum to btot and whose label is the sum of those fields' biofilm codes,
// or m when it keeps its type:
//   * of no pore type and btot < thr: pore[0], or with several pore types the type of the
//     first face neighbour of a pore type, in the order +x -x +y -y +z -z, inside the box (no
//     wrap), read from prev (the mask before the pass) (:141-186);
//   * of a pore type and btot >= thr: the label (or bio[0] with one biofilm code), when it is
//     positive (:190-198; the reference leaves the voxel as it is otherwise).
__device__ __forceinline__ int mask_rule_code(const MaskRule& r, const uint8_t* prev, const Geom& g, int x, int y,
                                              int z, int m, double btot, int label, double thr, int solid, int bb) {
  const bool pore = mr_is_pore(r, m);
  if (!pore && btot < thr) {
    int nm = r.pore[0];
    if (r.npore > 1) {
      for (int d = 0; d < 6; ++d) {
        const int xx = x + (d == 0) - (d == 1), yy = y + (d == 2) - (d == 3), zz = z + (d == 4) - (d == 5);
        if (xx < 0 || xx >= g.nx || yy < 0 || yy >= g.ny || zz < 0 || zz >= g.nz) continue;
        const int q = prev[xx + (long long)g.nx * (yy + (long long)g.ny * zz)];
        if (q != solid && q != bb && mr_is_pore(r, q)) {
          nm = q;
          break;
        }
      }
    }
    return nm;
  }
  if (pore && btot >= thr) {
    const int lab = r.sum_labels ? label : r.bio[0];
    if (lab > 0) return lab;
  }
  return m;
}

}  // namespace clb
