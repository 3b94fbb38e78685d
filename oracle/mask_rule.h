// mask_rule.h -- TEST INFRASTRUCTURE: the pore <-> biofilm rule of
// updateLocalMaskNtotalLattices3D (src/complab3d_processors_part2.hh:117-198) for one voxel,
// shared by oracle.cpp (the step) and ca.cpp (the CA's mask / total pass).
#pragma once

#include <cstdint>

struct OrcMaskRule {
  int npore;
  int pore[8];      // pore_dynamics (pore[0] the default type)
  int bio[4];       // bio_dynamics[i][0] of biomass field i
  int sum_labels;   // 1: the label is the sum of the present fields' codes (the reference's)

  // one pore type, one biofilm code (the label is that code)
  static OrcMaskRule single(int pore_code, int biofilm_code) {
    OrcMaskRule r{};
    r.npore = 1;
    r.pore[0] = pore_code;
    for (int i = 0; i < 4; ++i) r.bio[i] = biofilm_code;
    r.sum_labels = 0;
    return r;
  }
  // v = {npore, pore[8], bio[4], sum_labels}, or null for single(pore_code, biofilm_code)
  static OrcMaskRule from(const int* v, int pore_code, int biofilm_code) {
    if (!v || v[0] < 1) return single(pore_code, biofilm_code);
    OrcMaskRule r{};
    r.npore = v[0];
    for (int i = 0; i < 8; ++i) r.pore[i] = v[1 + i];
    for (int i = 0; i < 4; ++i) r.bio[i] = v[9 + i];
    r.sum_labels = v[13];
    return r;
  }
  // the reference's poreflag loop (:138-141)
  bool is_pore(int m) const {
    for (int i = 0; i < npore; ++i)
      if (m == pore[i]) return true;
    return false;
  }
};

// The new code of voxel (x, y, z) of code m (neither solid nor bounce-back) with summed
// biomass bmass and species label newmask (both over the fields above 1e-14, :119-125), or m:
// a voxel of no pore type below the threshold returns to pore[0], or with several pore types
// to the type of the first face neighbour of a pore type, the neighbours listed +x -x +y -y
// +z -z without the ones off the domain (:141-186), their codes read from prev; a pore voxel
// at or above it takes the label when it is positive (:190-198).
inline int orc_mask_rule(const OrcMaskRule& r, const uint8_t* prev, int nx, int ny, int nz, int x, int y, int z,
                         int m, double bmass, int newmask, double thr, int solid, int bb) {
  const bool poreflag = r.is_pore(m);
  if (!poreflag && bmass < thr) {
    int nm = r.pore[0];
    if (r.npore > 1) {
      int del[6][3], nbrs = 0;
      if (x != nx - 1) { del[nbrs][0] = 1; del[nbrs][1] = 0; del[nbrs][2] = 0; ++nbrs; }
      if (x != 0) { del[nbrs][0] = -1; del[nbrs][1] = 0; del[nbrs][2] = 0; ++nbrs; }
      if (y != ny - 1) { del[nbrs][0] = 0; del[nbrs][1] = 1; del[nbrs][2] = 0; ++nbrs; }
      if (y != 0) { del[nbrs][0] = 0; del[nbrs][1] = -1; del[nbrs][2] = 0; ++nbrs; }
      if (z != nz - 1) { del[nbrs][0] = 0; del[nbrs][1] = 0; del[nbrs][2] = 1; ++nbrs; }
      if (z != 0) { del[nbrs][0] = 0; del[nbrs][1] = 0; del[nbrs][2] = -1; ++nbrs; }
      for (int t = 0; t < nbrs; ++t) {
        const long q = (x + del[t][0]) + (long)nx * ((y + del[t][1]) + (long)ny * (z + del[t][2]));
        const int nbrmask = prev[q];
        if (nbrmask != bb && nbrmask != solid && r.is_pore(nbrmask)) {
          nm = nbrmask;
          break;
        }
      }
    }
    return nm;
  }
  if (poreflag && bmass >= thr) {
    const int lab = r.sum_labels ? newmask : r.bio[0];
    if (lab > 0) return lab;
  }
  return m;
}
