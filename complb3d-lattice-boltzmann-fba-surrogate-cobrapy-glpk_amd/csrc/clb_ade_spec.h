// clb_ade_spec.h -- the coupled step with the equilibrium speciation fused into the D3Q7
// transport kernel (run_equilibrium_biotic, src/complab3d_processors_part4_eqsolver.hh:
// 1355-1494, in the step order of src/complab.cpp:1252-1849: collide -> kinetics ->
// abiotic -> speciation -> stream).
//
// The transport kernel hands each reactive voxel's post-kinetics concentrations to the
// near-equilibrium Newton solve (spec_fused) through LDS, inside the same block. The limited
// increments are not written into the seven population slots (a read-modify-write of data
// the kernel has just stored): each voxel writes its dC per solute once, and the next step
// adds t_i dC to the populations as it loads them (add_pending). A voxel whose solve has not
// met the judge within the fused budget writes its concentrations to cbuf and is appended to
// a straggler list; the persistent work-queue kernel (spec_kernel, MODE 1) solves those
// voxels from cbuf and writes their increments the same way, in the same step.
#pragma once

#include "clb_ade_node.h"
#include "clb_brick.h"
#include "clb_compact.h"
#include "clb_internal.h"
#include "clb_spec_kernel.h"

namespace clb {

namespace {

// Block of at most BT threads (BT = 256 or 512), x fastest; every thread reaches the barriers.
//  1. transport: load (plus the previous step's deferred increments), collide, kinetics,
//     store, as ade_step_kernel; each reactive voxel leaves its post-kinetics
//     concentrations and its index in an LDS slot, slots compacted over the block;
//  2. solve: the first `count` threads of the block each solve one slot, so every lane of a
//     solving wavefront has work whatever the pore fraction, and write the limited
//     increments to pend_w (added when the next step loads the populations). A voxel whose
//     solve did not converge within io.evals evaluations writes zeros there, its
//     concentrations to cbuf and its next trial point to io.sx, and joins the straggler
//     list for the persistent kernel.
//
// With FP32 populations, up to six solutes and three components, the kernel is held to four
// wavefronts per SIMD (128 VGPRs; the compiler spills about one register): the transport
// phase is latency-bound on HBM and the solve on FP64 dependences, so occupancy matters
// more than that spill. Other shapes would spill heavily at 128 and keep the compiler's
// choice.
template <typename R, int S, int NC>
constexpr int fused_waves_per_eu() { return (sizeof(R) == 4 && S <= 6 && NC <= 3) ? 4 : 1; }

// Phase 2 of the fused step, shared by the dense and compact kernels: the first `count`
// threads of the block each solve one slot (the post-kinetics population sums sps, so the
// concentration is max(sps + c_ref, kMinConc) exactly as the transport formed it, in
// R to halve the LDS the block holds at FP32; node index snode), write the limited
// increments to pend_w (stride: N dense, ccap compact), or, for a
// voxel that has not met the judge within io.evals evaluations, zeros there, its
// concentrations to cbuf and its next trial point to io.sx, and append it to block bid's
// segment of the straggler list; then the per-wavefront statistics of block bid.
template <typename R, int S, int NC, int BT>
__device__ __forceinline__ void fused_solve_phase(const AdeArgs<R>& a, const Tableau& tb, const FusedSpecIO& io,
                                                  int lane, int count, R (&sps)[S][BT], double (&scr)[S][BT],
                                                  const int (&snode)[BT], long long stride, long long bid) {
  constexpr int NW = BT / 64;
  const int wlane = lane & 63, wave = lane >> 6;
  bool solved = false, straggler = false;
  int evals = 0, node = 0;
  double xs[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) xs[j] = 0.0;
  if (lane < count) {
    double c[S];
#pragma unroll
    for (int s = 0; s < S; ++s) c[s] = fmax((double)sps[s][lane] + (double)a.c_ref[s], kMinConc);
    node = snode[lane];
    solved = spec_fused<NC, S, BT>(tb, c, xs, evals, io.evals, &scr[0][lane]);
    straggler = !solved;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      double d = 0.0;
      const double cs = fmax((double)sps[s][lane] + (double)a.c_ref[s], kMinConc);
      if (solved) {
        // the judged evaluation that converged was at xs: its species concentrations are
        // still in scr (clamping them equals calc_species' clamp of log c to [-30, 1])
        d = spec_increment_eq(tb.is_eq[s], scr[s][lane], cs);
        if (!(d > 1e-12 || d < -1e-12)) d = 0.0;
      } else {
        a.cbuf[(long long)s * stride + node] = cs;
      }
      a.pend_w[(long long)s * stride + node] = (R)d;
    }
  }
  // appended to the block's segment of the list (one counter per segment)
  const unsigned long long sb = __ballot(straggler);
  if (sb) {
    const int seg = (int)(bid % kStragglerSegments);
    unsigned int base = 0;
    if (wlane == 0) base = atomicAdd(&io.scount[seg], (unsigned int)__popcll(sb));
    base = __shfl(base, 0);
    if (straggler) {
      const long long k = seg * io.segcap + base + __popcll(sb & ((1ull << wlane) - 1ull));
      io.slist[k] = node;
#pragma unroll
      for (int j = 0; j < 4; ++j) io.sx[k * 4 + j] = (j < NC) ? xs[j] : 0.0;
    }
  }
  // ---- statistics: per-wavefront accumulators (solves, evaluations), no atomics
  if (lane < ((count + 63) & ~63)) {
    unsigned long long it = solved ? (unsigned long long)evals : 0ull;
    for (int off = 32; off > 0; off >>= 1) it += __shfl_xor(it, off);
    const unsigned long long nsol = __popcll(__ballot(solved));
    if (wlane == 0 && nsol) {
      unsigned long long* st = io.blk_stats + 2 * (bid * NW + wave);
      st[0] += nsol;
      st[1] += it;
    }
  }
}

template <typename R, int S, bool ODD, int NC, int BT>
__global__ void __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(fused_waves_per_eu<R, S, NC>(), 8)))
ade_spec_step_kernel(AdeArgs<R> a, Tableau tb, FusedSpecIO io) {
  constexpr int NW = BT / 64;
  __shared__ R sps[S][BT];             // post-kinetics population sums of the block's reactive voxels
  __shared__ double scr[S][BT];        // species concentrations of the solve's latest evaluation
  __shared__ int snode[BT];            // node indices of the slots
  __shared__ unsigned int woff[NW + 1];  // per-wavefront slot offsets, [NW] = block total
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
    woff[NW] = acc;
  }
  __syncthreads();
  const int slot = (int)woff[wave] + __popcll(rb & ((1ull << wlane) - 1ull));
  const int count = (int)woff[NW];

  // ---- 1. transport
  R rho_in[S];
  if (fluid) {
    const Nbr nb(x, y, z, g);
    R p[S * 7];
    load7<R, S, ODD, false>(a, nb, n, p);
    add_pending<R, S, false>(a, nb, n, p);
#pragma unroll
    for (int s = 0; s < S; ++s) rho_in[s] = pop_sum(&p[s * 7]) + a.c_ref[s];
    ade_node<R, S, true, false>(a, x, y, z, n, code, p, nullptr, n, g.N);
    if (react) {
#pragma unroll
      for (int s = 0; s < S; ++s) sps[s][slot] = pop_sum(&p[s * 7]);
      snode[slot] = n;
    }
    store7<R, S, ODD, false>(a, nb, n, p);
  }
  if (valid && (x == 1 || x == g.nx - 2)) {
    const int xb = (x == 1) ? 0 : g.nx - 1;
    const int nbn = xb + g.nx * (y + g.ny * z);
    const uint8_t cb = a.mask[nbn];
    if (a.kind[cb] == CLB_KIND_FLUID) {
      const Nbr nb(xb, y, z, g);
      R q[S * 7];
      load7<R, S, ODD>(a, nb, nbn, q);
      add_pending<R, S>(a, nb, nbn, q);
      if (!fluid) {
        // interior neighbour is a wall: the Neumann face keeps its own density
#pragma unroll
        for (int s = 0; s < S; ++s) rho_in[s] = pop_sum(&q[s * 7]) + a.c_ref[s];
      }
      if (xb == 0) ade_node<R, S, false, false, 1>(a, xb, y, z, nbn, cb, q, rho_in, nbn, g.N);
      else ade_node<R, S, false, false, 2>(a, xb, y, z, nbn, cb, q, rho_in, nbn, g.N);
      store7<R, S, ODD>(a, nb, nbn, q);
    }
  }
  __syncthreads();
  fused_solve_phase<R, S, NC, BT>(a, tb, io, lane, count, sps, scr, snode, g.N,
                                  blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z));
}

// ade_spec_step_kernel over the compact layout (clb_compact.h): one thread per fluid node
// of a 1-D block; the same two phases. A face node whose interior neighbour is fluid is
// updated by that neighbour's thread; one next to a wall updates itself with its own density.
template <typename R, int S, bool ODD, int NC, int BT>
__global__ void __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(fused_waves_per_eu<R, S, NC>(), 8)))
ade_spec_step_compact_kernel(AdeArgs<R> a, Tableau tb, FusedSpecIO io) {
  constexpr int NW = BT / 64;
  __shared__ R sps[S][BT];
  __shared__ double scr[S][BT];
  __shared__ int snode[BT];
  __shared__ unsigned int woff[NW + 1];
  const Geom& g = a.geo;
  const int lane = threadIdx.x;
  const int wlane = lane & 63, wave = lane >> 6;
  const int nwaves = (blockDim.x + 63) >> 6;
  const int lb = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const int j = lb * blockDim.x + threadIdx.x;
  const bool valid = j < (int)*a.cnf;
  CPos q{0, 0, 0, 0};
  uint8_t code = (uint8_t)a.solid_code;
  if (valid) {
    q = cpos(a, j);
    code = a.mask[q.n];
  }
  const bool face = valid && (q.x == 0 || q.x == g.nx - 1);
  const bool react = valid && !face && code != a.solid_code && code != a.bb_code;
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
    woff[NW] = acc;
  }
  __syncthreads();
  const int slot = (int)woff[wave] + __popcll(rb & ((1ull << wlane) - 1ull));
  const int count = (int)woff[NW];

  // ---- 1. transport
  if (valid) {
    const CSlots L = cslots(a, j);
    R rho_in[S];
    if (face) {
      const int inner = L.s[q.x == 0 ? 1 : 2];
      if (!(inner >= 0 && inner < a.ccap)) {
        R f[S * 7];
        cload7<R, S, ODD>(a, L, f);
        cadd_pending<R, S>(a, L, f);
#pragma unroll
        for (int s = 0; s < S; ++s) rho_in[s] = pop_sum(&f[s * 7]) + a.c_ref[s];
        if (q.x == 0) ade_node<R, S, false, false, 1>(a, q.x, q.y, q.z, q.n, code, f, rho_in, j, a.ccap);
        else ade_node<R, S, false, false, 2>(a, q.x, q.y, q.z, q.n, code, f, rho_in, j, a.ccap);
        cstore7<R, S, ODD>(a, L, f);
      }
    } else {
      {
        R p[S * 7];
        cload7<R, S, ODD, false>(a, L, p);
        cadd_pending<R, S, false>(a, L, p);
#pragma unroll
        for (int s = 0; s < S; ++s) rho_in[s] = pop_sum(&p[s * 7]) + a.c_ref[s];
        ade_node<R, S, true, false>(a, q.x, q.y, q.z, q.n, code, p, nullptr, j, a.ccap);
        if (react) {
#pragma unroll
          for (int s = 0; s < S; ++s) sps[s][slot] = pop_sum(&p[s * 7]);
          snode[slot] = j;
        }
        cstore7<R, S, ODD, false>(a, L, p);
      }
      if (q.x == 1 || q.x == g.nx - 2) {
        const int fj = L.s[q.x == 1 ? 2 : 1];
        if (fj >= 0 && fj < a.ccap) {
          const CSlots F = cslots(a, fj);
          const int xb = (q.x == 1) ? 0 : g.nx - 1;
          const int nb = q.n + xb - q.x;
          const uint8_t cb = a.mask[nb];
          R f[S * 7];
          cload7<R, S, ODD>(a, F, f);
          cadd_pending<R, S>(a, F, f);
          if (xb == 0) ade_node<R, S, false, false, 1>(a, xb, q.y, q.z, nb, cb, f, rho_in, fj, a.ccap);
          else ade_node<R, S, false, false, 2>(a, xb, q.y, q.z, nb, cb, f, rho_in, fj, a.ccap);
          cstore7<R, S, ODD>(a, F, f);
        }
      }
    }
  }
  __syncthreads();
  fused_solve_phase<R, S, NC, BT>(a, tb, io, lane, count, sps, scr, snode, a.ccap, lb);
}

template <typename R, int S, int NC>
void launch_fused_shape_compact(const AdeArgs<R>& a, bool odd, int bt, const Tableau& tb, const FusedSpecIO& io,
                                hipStream_t st) {
  const unsigned blocks = (unsigned)((a.ccap + bt - 1) / bt);
  if (bt <= 256) {
    if (odd) hipLaunchKernelGGL((ade_spec_step_compact_kernel<R, S, true, NC, 256>), dim3(blocks), dim3(bt), 0, st, a, tb, io);
    else hipLaunchKernelGGL((ade_spec_step_compact_kernel<R, S, false, NC, 256>), dim3(blocks), dim3(bt), 0, st, a, tb, io);
  } else {
    if (odd) hipLaunchKernelGGL((ade_spec_step_compact_kernel<R, S, true, NC, 512>), dim3(blocks), dim3(bt), 0, st, a, tb, io);
    else hipLaunchKernelGGL((ade_spec_step_compact_kernel<R, S, false, NC, 512>), dim3(blocks), dim3(bt), 0, st, a, tb, io);
  }
}

template <typename R, int S, int NC>
void launch_fused_shape(const AdeArgs<R>& a, bool odd, dim3 block, const Tableau& tb, const FusedSpecIO& io,
                        hipStream_t st) {
  dim3 grid((a.geo.nx + block.x - 1) / block.x, (a.geo.ny + block.y - 1) / block.y, a.geo.nz);
  if (block.x * block.y <= 256) {
    if (odd) hipLaunchKernelGGL((ade_spec_step_kernel<R, S, true, NC, 256>), grid, block, 0, st, a, tb, io);
    else hipLaunchKernelGGL((ade_spec_step_kernel<R, S, false, NC, 256>), grid, block, 0, st, a, tb, io);
  } else {
    if (odd) hipLaunchKernelGGL((ade_spec_step_kernel<R, S, true, NC, 512>), grid, block, 0, st, a, tb, io);
    else hipLaunchKernelGGL((ade_spec_step_kernel<R, S, false, NC, 512>), grid, block, 0, st, a, tb, io);
  }
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
  if (block.x * block.y > 512 || block.z != 1 || io.evals < 2) return false;
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

// The compiled shapes of dispatch_fused, on the compact layout (bt threads per 1-D block).
template <typename R>
bool dispatch_fused_compact(const AdeArgs<R>& a, bool odd, int bt, const Tableau& tb, const FusedSpecIO& io,
                            hipStream_t st) {
  const int S = a.S, nc = tb.nc;
  if (tb.ns != S || nc < 1 || nc > 4) return false;
  if (bt > 512 || bt < 64 || (bt & 63) || io.evals < 2) return false;
  const int NC = (nc <= 2) ? 2 : nc;
  if (NC > S) return false;
#define CLB_FUSED(S_, NC_) \
  case S_ * 10 + NC_: launch_fused_shape_compact<R, S_, NC_>(a, odd, bt, tb, io, st); return true;
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
