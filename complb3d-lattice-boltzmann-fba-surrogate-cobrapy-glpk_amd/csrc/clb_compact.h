// clb_compact.h -- the D3Q7 transport over the compact (fluid-only) layout.
//
// In the dense layout every voxel has seven population slots per solute, and a wavefront's
// loads and stores touch nearly every cache line of the lattice: in a sphere pack of
// porosity 0.35, 99 % of the 32-voxel runs along x hold a fluid voxel. The compact layout
// keeps only the ADE-fluid voxels, in natural order, so the populations of 64 consecutive
// nodes are 64 consecutive words and no byte of a wall is moved.
//
// Node j's slots: its own, and for each direction d (1..6: +x -x +y -y +z -z) the slot at
// +c_d, cnbr[(d-1) ccap + j]: a fluid node's j, or, when the neighbour is a wall, a ghost
// slot (>= ccap) of direction d -- the wall's slot d of the dense layout, which is where the
// AA pattern keeps a population in flight towards a bounce-back wall (full-way bounce-back:
// it is read back reversed one step later); or -1 beyond an x face. Even steps touch only
// the node's own slots; odd steps read A[opp(i)] at -c_i and write A[i] at +c_i, exactly as
// the dense kernels (clb_ade_node.h) do, so the two layouts give the same numbers.
#pragma once

#include "clb_ade_node.h"
#include "clb_brick.h"

namespace clb {

namespace {

// s[0] = j; s[d] = the slot at +c_d (d = 1..6).
struct CSlots {
  int s[7];
};

template <typename R>
__device__ __forceinline__ CSlots cslots(const AdeArgs<R>& a, int j) {
  CSlots L;
  L.s[0] = j;
#pragma unroll
  for (int d = 1; d < 7; ++d) L.s[d] = a.cnbr[(long long)(d - 1) * a.ccap + j];
  return L;
}

// XCHK = false for an interior node (1 <= x <= nx-2): every slot exists.
template <typename R, int S, bool ODD, bool XCHK = true>
__device__ __forceinline__ void cload7(const AdeArgs<R>& a, const CSlots& L, R* p) {
  const long long st = a.cstride;
  const int j = L.s[0];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (a.immobile[s]) {
      p[s * 7] = a.g[(long long)(s * 7) * st + j];
#pragma unroll
      for (int i = 1; i < 7; ++i) p[s * 7 + i] = R(0);
      continue;
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      if (!ODD) {
        p[s * 7 + i] = a.g[(long long)(s * 7 + i) * st + j];
      } else {
        // population i arrives from -c_i = +c_opp(i), where it was stored in array opp(i)
        const int src = L.s[opp_dir(i)];
        if (!XCHK || src >= 0)
          p[s * 7 + i] = a.g[(long long)(s * 7 + opp_dir(i)) * st + src];
        else
          p[s * 7 + i] = R(0);
      }
    }
  }
}

template <typename R, int S, bool ODD, bool XCHK = true>
__device__ __forceinline__ void cstore7(const AdeArgs<R>& a, const CSlots& L, const R* p) {
  const long long st = a.cstride;
  const int j = L.s[0];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (a.immobile[s]) {
      a.g[(long long)(s * 7) * st + j] = p[s * 7];
      continue;
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      if (!ODD) {
        a.g[(long long)(s * 7 + opp_dir(i)) * st + j] = p[s * 7 + i];
      } else {
        const int dst = L.s[i];
        if (!XCHK || dst >= 0) a.g[(long long)(s * 7 + i) * st + dst] = p[s * 7 + i];
      }
    }
  }
}

// The deferred speciation increments of add_pending (clb_ade_node.h) on the compact layout:
// a population from a fluid neighbour takes that neighbour's increment of the previous step
// (pend_r), one reflected by a wall (its slot is a ghost) the node's own of the step before
// (pend_w), one entering through an x face nothing.
template <typename R, int S, bool XCHK = true>
__device__ __forceinline__ void cadd_pending(const AdeArgs<R>& a, const CSlots& L, R* p) {
  const long long cap = a.ccap;
  const int j = L.s[0];
  int src[7];
  bool wall[7], out[7];
#pragma unroll
  for (int i = 1; i < 7; ++i) {
    const int v = L.s[opp_dir(i)];  // the slot at -c_i
    out[i] = XCHK && v < 0;
    wall[i] = v >= cap;
    src[i] = (out[i] || wall[i]) ? j : v;
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const R* P1 = a.pend_r + (long long)s * cap;  // previous step
    const R* P2 = a.pend_w + (long long)s * cap;  // the step before
    const bool imm = a.immobile[s] != 0;
    const R own = P1[j];
    p[s * 7] += imm ? own : R(Q7::w(0)) * own;
#pragma unroll
    for (int i = 1; i < 7; ++i) {
      const R v = out[i] ? R(0) : (wall[i] ? P2[j] : P1[src[i]]);
      p[s * 7 + i] += imm ? R(0) : R(Q7::w(i)) * v;
    }
  }
}

// Dense coordinates of compact node j.
struct CPos {
  int n, x, y, z;
};

template <typename R>
__device__ __forceinline__ CPos cpos(const AdeArgs<R>& a, int j) {
  const Geom& g = a.geo;
  CPos q;
  q.n = a.cnode[j];
  q.x = q.n % g.nx;
  q.y = (q.n / g.nx) % g.ny;
  q.z = q.n / (g.nx * g.ny);
  return q;
}

// ade_step_kernel (clb_ade.hip) over the compact layout: one thread per fluid node. A face
// node (x = 0 or nx-1) whose interior neighbour is fluid is updated by that neighbour's
// thread, after the neighbour has taken its pre-collision density (the Neumann face's
// value); a face node next to a wall updates itself with its own density.
template <typename R, int S, bool ODD, bool KIN>
__global__ void __launch_bounds__(256) ade_step_compact_kernel(AdeArgs<R> a) {
  const int j = xcd_block((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
  if (j >= (int)*a.cnf) return;
  const Geom& g = a.geo;
  const CPos q = cpos(a, j);
  const CSlots L = cslots(a, j);
  R rho_in[S];
  if (q.x == 0 || q.x == g.nx - 1) {
    const int inner = L.s[q.x == 0 ? 1 : 2];
    if (inner >= 0 && inner < a.ccap) return;
    R f[S * 7];
    cload7<R, S, ODD>(a, L, f);
#pragma unroll
    for (int s = 0; s < S; ++s) rho_in[s] = pop_sum(&f[s * 7]) + a.c_ref[s];
    if (q.x == 0) ade_node<R, S, false, false, 1>(a, q.x, q.y, q.z, q.n, a.mask[q.n], f, rho_in, j, a.ccap);
    else ade_node<R, S, false, false, 2>(a, q.x, q.y, q.z, q.n, a.mask[q.n], f, rho_in, j, a.ccap);
    cstore7<R, S, ODD>(a, L, f);
    return;
  }
  {
    R p[S * 7];
    cload7<R, S, ODD, false>(a, L, p);
#pragma unroll
    for (int s = 0; s < S; ++s) rho_in[s] = pop_sum(&p[s * 7]) + a.c_ref[s];
    ade_node<R, S, KIN, false>(a, q.x, q.y, q.z, q.n, a.mask[q.n], p, nullptr, j, a.ccap);
    cstore7<R, S, ODD, false>(a, L, p);
  }
  if (q.x == 1 || q.x == g.nx - 2) {
    const int fj = L.s[q.x == 1 ? 2 : 1];
    if (fj >= 0 && fj < a.ccap) {
      const CSlots F = cslots(a, fj);
      const int xb = (q.x == 1) ? 0 : g.nx - 1;
      const int nb = q.n + xb - q.x;
      R f[S * 7];
      cload7<R, S, ODD>(a, F, f);
      if (xb == 0) ade_node<R, S, false, false, 1>(a, xb, q.y, q.z, nb, a.mask[nb], f, rho_in, fj, a.ccap);
      else ade_node<R, S, false, false, 2>(a, xb, q.y, q.z, nb, a.mask[nb], f, rho_in, fj, a.ccap);
      cstore7<R, S, ODD>(a, F, f);
    }
  }
}

}  // namespace
}  // namespace clb
