// clb_lattice.h -- device-side lattice helpers shared by the D3Q19 and D3Q7 kernels.
//
// In-place AA-pattern streaming (one population array, no second lattice):
//   even step: node x reads f_i from slot (i, x), collides, writes f*_i to slot (opp(i), x);
//   odd step : node x reads f_i from slot (opp(i), x - c_i), collides, writes f*_i to
//              slot (i, x + c_i).
// Each node reads and writes exactly the same slots within a step, so the update is
// race-free with one thread per node. After an odd step the array is in natural order
// (slot (i, x) holds the pre-collision f_i(x)). A node whose collision is the full
// reversal f*_i = f_opp(i) (node-based bounce-back) leaves every slot it owns unchanged
// in both steps, so bounce-back and no-dynamics nodes are simply not visited.
#pragma once

#include "clb_internal.h"

namespace clb {

struct Q19 {
  static constexpr int cx[19] = CLB_D3Q19_CX;
  static constexpr int cy[19] = CLB_D3Q19_CY;
  static constexpr int cz[19] = CLB_D3Q19_CZ;
  static __host__ __device__ constexpr double w(int i) {
    return i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
  }
};

struct Q7 {
  static constexpr int cx[7] = CLB_D3Q7_CX;
  static constexpr int cy[7] = CLB_D3Q7_CY;
  static constexpr int cz[7] = CLB_D3Q7_CZ;
  static __host__ __device__ constexpr double w(int i) { return i == 0 ? 0.25 : 0.125; }
};

// Neighbour addressing for one node with periodic y/z; x neighbours outside [0, nx) are
// reported invalid (the inlet/outlet planes have no neighbour beyond them).
struct Nbr {
  int x, nx;
  int yo[3];   // row offsets for dy = -1, 0, +1
  int zo[3];   // plane offsets for dz = -1, 0, +1
  __device__ Nbr(int x_, int y, int z, const Geom& g) : x(x_), nx(g.nx) {
    const int ym = (y == 0) ? g.ny - 1 : y - 1, yp = (y == g.ny - 1) ? 0 : y + 1;
    const int zm = (z == 0) ? g.nz - 1 : z - 1, zp = (z == g.nz - 1) ? 0 : z + 1;
    yo[0] = ym * g.nx; yo[1] = y * g.nx; yo[2] = yp * g.nx;
    const int pl = g.nx * g.ny;
    zo[0] = zm * pl; zo[1] = z * pl; zo[2] = zp * pl;
  }
  __device__ int at(int dx, int dy, int dz) const { return (x + dx) + yo[dy + 1] + zo[dz + 1]; }
  __device__ bool xok(int dx) const { return (unsigned)(x + dx) < (unsigned)nx; }
};

}  // namespace clb
