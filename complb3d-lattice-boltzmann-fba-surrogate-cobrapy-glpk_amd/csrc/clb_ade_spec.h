// clb_ade_spec.h -- the coupled step with the equilibrium speciation fused into the D3Q7
// transport kernel (run_equilibrium_biotic, src/complab3d_processors_part4_eqsolver.hh:
// 1355-1494, in the step order of src/complab.cpp:1252-1849: collide -> kinetics ->
// abiotic -> speciation -> stream).
//
// The transport kernel hands each reactive voxel's post-kinetics concentrations to the
// near-equilibrium Newton solve (spec_fused) through LDS, inside the same block, and adds
// the limited increments to the populations it has just stored, so speciation adds no HBM
// traffic beyond cache re-reads. A voxel whose solve has not met the judge within the fused
// budget writes its concentrations to cbuf and is appended to a straggler list; the
// persistent work-queue kernel (spec_kernel, MODE 1) then solves those voxels from cbuf and
// adds their increments to the stored populations, in the same step.
#pragma once

#include "clb_ade_node.h"
#include "clb_internal.h"
#include "clb_spec_kernel.h"

namespace clb {

namespace {

// Add dC * t_i of solute s to the slots node n wrote in this step (the AA slot map of
// store7: even step -> (opp(i), n), odd step -> (i, n + c_i); an immobile solute's whole
// increment goes to its rest slot).
template <typename R, bool ODD>
__device__ __forceinline__ void add_stored(const AdeArgs<R>& a, const Nbr& nb, int s, int n, double dC) {
  const long long N = a.geo.N;
  if (a.immobile[s]) {
    R* q = &a.g[(long long)(s * 7) * N + n];
    *q = (R)((double)*q + dC);
    return;
  }
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    R* q = ODD ? &a.g[(long long)(s * 7 + i) * N + nb.at(Q7::cx[i], Q7::cy[i], Q7::cz[i])]
               : &a.g[(long long)(s * 7 + opp_dir(i)) * N + n];
    *q = (R)((double)*q + dC * Q7::w(i));
  }
}

// Block of at most 256 threads, x fastest; every thread reaches the barriers.
//  1. transport: load, collide, kinetics, store (as ade_step_kernel); each reactive voxel
//     leaves its post-kinetics concentrations in an LDS slot, slots compacted over the block;
//  2. solve: the first `count` threads of the block each solve one slot, so every lane of
//     a solving wavefront has work whatever the pore fraction;
//  3. increments: each reactive voxel adds its limited increments to the slots it stored
//     in phase 1 (just written, so they come from cache); a voxel whose solve did not
//     converge writes cbuf and joins the straggler list instead.
template <typename R, int S, bool ODD, int NC>
__global__ void __launch_bounds__(256) ade_spec_step_kernel(AdeArgs<R> a, Tableau tb, FusedSpecIO io) {
  __shared__ double sc[S][256];        // concentrations in, increments out
  __shared__ unsigned char ok[256];    // solve converged
  __shared__ unsigned int woff[5];     // per-wavefront slot offsets, [4] = block total
  __shared__ unsigned long long red[2][4];
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  const Geom& g = a.geo;
  const int lane = threadIdx.x + blockDim.x * threadIdx.y;
  const int wlane = lane & 63, wave = lane >> 6;
  const int nwaves = (blockDim.x * blockDim.y + 63) >> 6;
  const bool valid = x >= 1 && x <= g.nx - 2 && y < g.ny;
  const int n = valid ? x + g.nx * (y + g.ny * z) : 0;
  const uint8_t code = valid ? a.mask[n] : (uint8_t)a.solid_code;
  const bool fluid = valid && a.kind[code] == CLB_KIND_FLUID;
  const bool react = fluid && code != a.solid_code && code != a.bb_code;
  // compacted slot of this voxel
  const unsigned long long rb = __ballot(react);
  if (wlane == 0) woff[wave] = (unsigned int)__popcll(rb);
  __syncthreads();
  if (lane == 0) {
    unsigned int acc = 0;
    for (int w = 0; w < nwaves; ++w) {
      const unsigned int c = woff[w];
      woff[w] = acc;
      acc += c;
    }
    woff[4] = acc;
  }
  __syncthreads();
  const int slot = (int)woff[wave] + __popcll(rb & ((1ull << wlane) - 1ull));
  const int count = (int)woff[4];

  // ---- 1. transport
  R rho_in[S];
  if (fluid) {
    const Nbr nb(x, y, z, g);
    R p[S * 7];
    load7<R, S, ODD>(a, nb, n, p);
#pragma unroll
    for (int s = 0; s < S; ++s) rho_in[s] = pop_sum(&p[s * 7]) + a.c_ref[s];
    ade_node<R, S, true, false>(a, x, y, z, n, code, p, nullptr);
    if (react) {
#pragma unroll
      for (int s = 0; s < S; ++s) sc[s][slot] = fmax((double)pop_sum(&p[s * 7]) + (double)a.c_ref[s], kMinConc);
    }
    store7<R, S, ODD>(a, nb, n, p);
  }
  if (valid && (x == 1 || x == g.nx - 2)) {
    const int xb = (x == 1) ? 0 : g.nx - 1;
    const int nbn = xb + g.nx * (y + g.ny * z);
    const uint8_t cb = a.mask[nbn];
    if (a.kind[cb] == CLB_KIND_FLUID) {
      const Nbr nb(xb, y, z, g);
      R q[S * 7];
      load7<R, S, ODD>(a, nb, nbn, q);
      if (!fluid) {
        // interior neighbour is a wall: the Neumann face keeps its own density
#pragma unroll
        for (int s = 0; s < S; ++s) rho_in[s] = pop_sum(&q[s * 7]) + a.c_ref[s];
      }
      if (xb == 0) ade_node<R, S, false, false, 1>(a, xb, y, z, nbn, cb, q, rho_in);
      else ade_node<R, S, false, false, 2>(a, xb, y, z, nbn, cb, q, rho_in);
      store7<R, S, ODD>(a, nb, nbn, q);
    }
  }
  __syncthreads();

  // ---- 2. solve the block's reactive voxels on its first `count` threads
  bool solved = false;
  int evals = 0;
  if (lane < count) {
    double c[S], xs[NC];
#pragma unroll
    for (int s = 0; s < S; ++s) c[s] = sc[s][lane];
    solved = spec_fused<NC, S>(tb, c, xs, evals);
    if (solved) {
#pragma unroll
      for (int s = 0; s < S; ++s) sc[s][lane] = spec_increment<NC>(tb, s, xs, c[s]);
    }
    ok[lane] = solved ? 1 : 0;
  }
  __syncthreads();

  // ---- 3. increments, or hand the voxel to the persistent kernel
  bool straggler = false;
  if (react) {
    const Nbr nb(x, y, z, g);
    if (ok[slot]) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const double dC = sc[s][slot];
        if (dC > 1e-12 || dC < -1e-12) add_stored<R, ODD>(a, nb, s, n, dC);
      }
    } else {
      straggler = true;
#pragma unroll
      for (int s = 0; s < S; ++s) a.cbuf[(long long)s * g.N + n] = sc[s][slot];
    }
  }
  const unsigned long long sb = __ballot(straggler);
  if (sb) {
    unsigned int base = 0;
    if (wlane == 0) base = atomicAdd(io.scount, (unsigned int)__popcll(sb));
    base = __shfl(base, 0);
    if (straggler) io.slist[base + __popcll(sb & ((1ull << wlane) - 1ull))] = n;
  }
  // ---- statistics: per-block accumulators (solves, evaluations), no atomics
  unsigned long long it = solved ? (unsigned long long)evals : 0ull;
  for (int off = 32; off > 0; off >>= 1) it += __shfl_xor(it, off);
  const unsigned long long nsol = __popcll(__ballot(solved));
  if (wlane == 0) {
    red[0][wave] = nsol;
    red[1][wave] = it;
  }
  __syncthreads();
  if (lane == 0) {
    const long long bid = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    unsigned long long s0 = 0, s1 = 0;
    for (int w = 0; w < nwaves; ++w) { s0 += red[0][w]; s1 += red[1][w]; }
    if (s0) {
      io.blk_stats[2 * bid] += s0;
      io.blk_stats[2 * bid + 1] += s1;
    }
  }
}

template <typename R, int S, int NC>
void launch_fused_shape(const AdeArgs<R>& a, bool odd, dim3 block, const Tableau& tb, const FusedSpecIO& io,
                        hipStream_t st) {
  dim3 grid((a.geo.nx + block.x - 1) / block.x, (a.geo.ny + block.y - 1) / block.y, a.geo.nz);
  if (odd) hipLaunchKernelGGL((ade_spec_step_kernel<R, S, true, NC>), grid, block, 0, st, a, tb, io);
  else hipLaunchKernelGGL((ade_spec_step_kernel<R, S, false, NC>), grid, block, 0, st, a, tb, io);
}

}  // namespace
}  // namespace clb

namespace clb {
namespace {

// Shapes compiled in: S solutes (all of them species of the tableau) and NC components
// (nc padded up to 2, 3 or 4).
template <typename R>
bool dispatch_fused(const AdeArgs<R>& a, bool odd, dim3 block, const Tableau& tb, const FusedSpecIO& io,
                    hipStream_t st) {
  const int S = a.S, nc = tb.nc;
  if (tb.ns != S || nc < 1 || nc > 4) return false;
  const int NC = (nc <= 2) ? 2 : nc;
  if (NC > S) return false;
#define CLB_FUSED(S_, NC_) \
  case S_ * 10 + NC_: launch_fused_shape<R, S_, NC_>(a, odd, block, tb, io, st); return true;
  switch (S * 10 + NC) {
    CLB_FUSED(2, 2)
    CLB_FUSED(3, 2) CLB_FUSED(3, 3)
    CLB_FUSED(4, 2) CLB_FUSED(4, 3) CLB_FUSED(4, 4)
    CLB_FUSED(5, 2) CLB_FUSED(5, 3) CLB_FUSED(5, 4)
    CLB_FUSED(6, 2) CLB_FUSED(6, 3) CLB_FUSED(6, 4)
    CLB_FUSED(7, 2) CLB_FUSED(7, 3) CLB_FUSED(7, 4)
    CLB_FUSED(8, 2) CLB_FUSED(8, 3) CLB_FUSED(8, 4)
    default: return false;
  }
#undef CLB_FUSED
}

}  // namespace
}  // namespace clb
