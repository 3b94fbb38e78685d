// clb_brick.h -- the brick order of the compact layouts (solutes: clb_compact.hip, flow:
// clb_flow_compact.hip). The lattice is cut into bricks of bx x by x bz voxels, the bricks
// taken in x-fastest order and the voxels x-fastest inside each; the padding of the last
// bricks lies outside the lattice. A workgroup's run of consecutive compact nodes then covers
// a compact block of the lattice, so most of its nodes' y and z neighbours are its own.
#pragma once

#include "clb_internal.h"

namespace clb {

namespace {

// Brick-order keys: key(x, y, z) and its inverse; voxels of the padding are outside the lattice.
__device__ __forceinline__ long long brick_key(int x, int y, int z, const BrickOrder& o) {
  const long long b = (long long)(x / o.bx) + (long long)o.nbx * ((y / o.by) + (long long)o.nby * (z / o.bz));
  return b * o.vol + (x % o.bx) + (long long)o.bx * ((y % o.by) + (long long)o.by * (z % o.bz));
}

__device__ __forceinline__ bool brick_voxel(long long k, const BrickOrder& o, const Geom& g, int& x, int& y, int& z) {
  const long long b = k / o.vol, r = k % o.vol;
  x = (int)((b % o.nbx) * o.bx + r % o.bx);
  y = (int)(((b / o.nbx) % o.nby) * o.by + (r / o.bx) % o.by);
  z = (int)((b / ((long long)o.nbx * o.nby)) * o.bz + r / ((long long)o.bx * o.by));
  return x < g.nx && y < g.ny && z < g.nz;
}

}  // namespace
}  // namespace clb
