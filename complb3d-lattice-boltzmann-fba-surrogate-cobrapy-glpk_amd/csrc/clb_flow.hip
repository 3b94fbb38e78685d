// clb_flow.hip -- D3Q19 incompressible BGK flow through the pore mask.
//
// Replaces Palabos IncBGKdynamics + BounceBack + NoDynamics and the regularized local
// pressure boundaries that NSdomainSetup installs (src/complab_functions.hh:245-294),
// stepped by the collideAndStream loops of src/complab.cpp:597-667.
//
// HBM traffic per fluid node and step: 19 populations read + 19 written
// (2*19*sizeof(R)) + 1 mask byte; bounce-back / no-dynamics nodes cost the mask byte.
//
// Storage: each slot holds f_i - t_i (1 + rref(x)) for the x of the slot's own voxel, where
// rref is the linear density profile of the initial state (clb_flow_init). The pressure drop
// along x makes the density, not the velocity, the largest part of f_i - t_i; taking the
// profile out leaves the local deviation from it and the momentum, so an FP32 store keeps
// about as many bits of the velocity as the profile is larger than that deviation. In a
// node's registers the populations are relative to the node's own rref(x); a population
// that moves one voxel along x changes reference by t_i c_ix rdx. The collision is unchanged
// in form: with q_i = f_i - t_i (1 + rref(x)), sum q_i = rho - 1 - rref(x) and
// sum c_i q_i = j (sum c_i t_i = 0), and q_eq has the f_eq - t_i form with that density.
#include "clb_flow_node.h"

namespace clb {

namespace {

// XCHK = false for a node with 1 <= x <= nx-2, whose x neighbours always exist: no checks,
// so the loads and stores need no branches. Odd steps move populations between voxels: one
// arriving from x - c_ix changes reference by -t_i c_ix rdx (and one leaving for x + c_ix,
// in store19, by the same).
template <typename R, bool ODD, bool XCHK = true>
__device__ __forceinline__ void load19(const R* __restrict__ f, long long N, const Nbr& nb, int n, R* p, R rdx) {
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    if (!ODD) {
      p[i] = f[(long long)i * N + n];
    } else {
      const int o = opp_dir(i);
      if (!XCHK || nb.xok(-Q19::cx[i])) {
        p[i] = f[(long long)o * N + nb.at(-Q19::cx[i], -Q19::cy[i], -Q19::cz[i])];
        if (Q19::cx[i] != 0) p[i] = fma(R(-Q19::w(i) * Q19::cx[i]), rdx, p[i]);
      } else {
        p[i] = R(0);
      }
    }
  }
}

template <typename R, bool ODD, bool XCHK = true>
__device__ __forceinline__ void store19(R* __restrict__ f, long long N, const Nbr& nb, int n, const R* p, R rdx) {
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    // non-temporal stores: the 19 populations of the whole lattice (1 GB at 384 x 192 x 192)
    // exceed the last-level cache, so a stored line is evicted before the next step reads it
    // and would only displace lines that this step's neighbour loads still use
    if (!ODD) {
      __builtin_nontemporal_store(p[i], &f[(long long)opp_dir(i) * N + n]);
    } else {
      if (!XCHK || nb.xok(Q19::cx[i])) {
        const R v = (Q19::cx[i] != 0) ? fma(R(-Q19::w(i) * Q19::cx[i]), rdx, p[i]) : p[i];
        __builtin_nontemporal_store(v, &f[(long long)i * N + nb.at(Q19::cx[i], Q19::cy[i], Q19::cz[i])]);
      }
    }
  }
}

// One fluid node: load, boundary regularization on an x face (FACE -1: x = 0, +1: x = nx-1,
// 0: interior), BGK collision, store. Returns |u|^2 / 2 for the energy tracer.
template <typename R, bool ODD, int FACE>
__device__ __forceinline__ double flow_node(const FlowArgs<R>& a, int x, int y, int z, int n, uint8_t code) {
  const Geom& g = a.g;
  const Nbr nb(x, y, z, g);
  R p[19];
  load19<R, ODD, FACE != 0>(a.f, g.N, nb, n, p, a.rdx);
  if (FACE == -1) regularize<R, -1>(p, a.dface_in);
  if (FACE == 1) regularize<R, 1>(p, a.dface_out);
  const R jsq = collide19<R>(p, a.omega[code]);
  store19<R, ODD, FACE != 0>(a.f, g.N, nb, n, p, a.rdx);
  return 0.5 * (double)jsq;
}

template <typename R, bool ODD, bool ENERGY>
__global__ void __launch_bounds__(256) flow_step_kernel(FlowArgs<R> a) {
  if (a.stop && *a.stop) return;  // the device tracer has converged
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  const Geom& g = a.g;
  double e = 0.0, cnt = 0.0;
  if (x < g.nx && y < g.ny) {
    const int n = x + g.nx * (y + g.ny * z);
    const uint8_t code = a.mask[n];
    const uint8_t k = a.kind[code];
    if (k == CLB_KIND_FLUID) {
      double ek;
      if (x == 0) ek = flow_node<R, ODD, -1>(a, x, y, z, n, code);
      else if (x == g.nx - 1) ek = flow_node<R, ODD, 1>(a, x, y, z, n, code);
      else ek = flow_node<R, ODD, 0>(a, x, y, z, n, code);
      if (ENERGY) { e = ek; cnt = 1.0; }
    } else if (k == CLB_KIND_NONE && a.stream_none && x > 0 && x < g.nx - 1) {
      // NoDynamics: no collision; the AA pattern's load and store alone are the streaming
      const Nbr nb(x, y, z, g);
      R p[19];
      load19<R, ODD, false>(a.f, g.N, nb, n, p, a.rdx);
      store19<R, ODD, false>(a.f, g.N, nb, n, p, a.rdx);
    }
  }
  if (ENERGY) {
    // within each wavefront (xor butterfly), then the wavefronts' partials in a fixed
    // order: right for any block of whole wavefronts (64 x 3 as well as 64 x 4)
    for (int off = 32; off > 0; off >>= 1) {
      e += __shfl_xor(e, off);
      cnt += __shfl_xor(cnt, off);
    }
    __shared__ double se[4], sc[4];
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    const int nw = (blockDim.x * blockDim.y) >> 6;
    if ((t & 63) == 0) { se[t >> 6] = e; sc[t >> 6] = cnt; }
    __syncthreads();
    if (t == 0) {
      double es = 0.0, cs = 0.0;
      for (int w = 0; w < nw; ++w) { es += se[w]; cs += sc[w]; }
      // one partial per block (a deterministic sum, and no atomics on one address)
      const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
      a.energy[2 + 2 * b] = es;
      a.energy[3 + 2 * b] = cs;
    }
  }
}

// energy[0..1] = the sums of the per-block partials energy[2 + 2b], energy[3 + 2b], in a
// fixed order (one block of 1024 threads: thread t sums the partials t, t + 1024, ... into
// four accumulators in turn, so four loads are in flight; then a tree over the threads).
constexpr int kEnergyThreads = 1024;
__device__ __forceinline__ void reduce_partials(const double* energy, long long nblocks, double* se, double* sc) {
  double e[4] = {0.0, 0.0, 0.0, 0.0}, c[4] = {0.0, 0.0, 0.0, 0.0};
  const long long t = threadIdx.x;
  long long b = t;
  for (; b + 3 * kEnergyThreads < nblocks; b += 4 * kEnergyThreads) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      e[k] += energy[2 + 2 * (b + k * kEnergyThreads)];
      c[k] += energy[3 + 2 * (b + k * kEnergyThreads)];
    }
  }
  for (int k = 0; b < nblocks; b += kEnergyThreads, ++k) {
    e[k & 3] += energy[2 + 2 * b];
    c[k & 3] += energy[3 + 2 * b];
  }
  se[t] = (e[0] + e[1]) + (e[2] + e[3]);
  sc[t] = (c[0] + c[1]) + (c[2] + c[3]);
  __syncthreads();
  for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
    if (t < s) {
      se[t] += se[t + s];
      sc[t] += sc[t + s];
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kEnergyThreads) energy_reduce_kernel(double* energy, long long nblocks) {
  __shared__ double se[kEnergyThreads], sc[kEnergyThreads];
  reduce_partials(energy, nblocks, se, sc);
  if (threadIdx.x == 0) {
    energy[0] = se[0];
    energy[1] = sc[0];
  }
}

// The device ValueTracer (Palabos' util::ValueTracer as the driver's flow loops use it,
// src/complab.cpp:607-620, 1794-1800, 1836-1840; the host restatements clb.ValueTracer and
// complab_gpu::ValueTracer do the same): after a flow step, the step's stored average energy
// (the per-block partials summed as energy_reduce_kernel sums them) joins a window of the
// last tr->w values; once the window is full the tracer has converged when |stddev / mean| <
// eps, stddev with w - 1, or when that ratio is NaN. The sums run in the window's order, one
// thread, with no fused multiply-adds, as the host restatements compute them. Once it has
// converged (tr->stop) the flow steps and this kernel do nothing until the next resolve.
constexpr int kTracerLds = 4096;  // windows up to this many values are staged in LDS
__global__ void __launch_bounds__(kEnergyThreads) tracer_step_kernel(const double* energy, long long nblocks,
                                                                      FlowTracer* tr, double* win, double* elog) {
#pragma clang fp contract(off)
  if (tr->stop) return;
  __shared__ double se[kEnergyThreads], sc[kEnergyThreads];
  __shared__ double sw[kTracerLds];
  reduce_partials(energy, nblocks, se, sc);
  const int t = threadIdx.x;
  const int w = tr->w;
  if (t == 0) {
    const double v = sc[0] > 0 ? se[0] / sc[0] : 0.0;
    if (tr->n < w) {
      win[(tr->head + tr->n) % w] = v;
      tr->n += 1;
    } else {
      win[tr->head] = v;
      tr->head = (tr->head + 1) % w;
    }
    if (elog) elog[tr->steps - tr->call0] = v;
    tr->steps += 1;
  }
  __syncthreads();
  if (tr->n < w) return;
  const int head = tr->head;
  const bool lds = w <= kTracerLds;
  if (lds) {
    for (int k = t; k < w; k += kEnergyThreads) sw[k] = win[(head + k) % w];
    __syncthreads();
  }
  if (t != 0) return;
  // one thread, in the window's order; from LDS the values are loaded 32 at a time so that
  // the sequential adds do not wait on each load
  double s = 0.0, s2 = 0.0, avg = 0.0;
  if (lds) {
    int k = 0;
    for (; k + 32 <= w; k += 32) {
      double v[32];
#pragma unroll
      for (int q = 0; q < 32; ++q) v[q] = sw[k + q];
#pragma unroll
      for (int q = 0; q < 32; ++q) s += v[q];
    }
    for (; k < w; ++k) s += sw[k];
    avg = s / (double)w;
    for (k = 0; k + 32 <= w; k += 32) {
      double v[32];
#pragma unroll
      for (int q = 0; q < 32; ++q) v[q] = sw[k + q];
#pragma unroll
      for (int q = 0; q < 32; ++q) s2 += (v[q] - avg) * (v[q] - avg);
    }
    for (; k < w; ++k) s2 += (sw[k] - avg) * (sw[k] - avg);
  } else {
    for (int k = 0; k < w; ++k) s += win[(head + k) % w];
    avg = s / (double)w;
    for (int k = 0; k < w; ++k) {
      const double v = win[(head + k) % w];
      s2 += (v - avg) * (v - avg);
    }
  }
  const double ratio = sqrt(s2 / (double)(w - 1)) / avg;
  if (isnan(ratio) || fabs(ratio) < tr->eps) tr->stop = 1;
}

// Start of a resolve: the steps run again (at least one, as the driver's loop takes one
// before it asks the tracer), counted from here.
__global__ void tracer_begin_kernel(FlowTracer* tr) {
  tr->stop = 0;
  tr->call0 = tr->steps;
}

// flag[0] = 1 when a no-dynamics voxel has a fluid neighbour among the Q - 1 of the stencil.
// The kind table sits in LDS, and every neighbour's mask byte is loaded before any is tested,
// so a voxel's loads overlap instead of following one another (most voxels test them all).
template <int Q>
__global__ void __launch_bounds__(256) none_reach_kernel(const uint8_t* mask, const uint8_t* kind, Geom g,
                                                         unsigned int* flag) {
  __shared__ uint8_t sk[256];
  sk[threadIdx.x] = kind[threadIdx.x];
  __syncthreads();
  const long long n = blockIdx.x * 256LL + threadIdx.x;
  if (n >= g.N || sk[mask[n]] != CLB_KIND_NONE) return;
  const int x = (int)(n % g.nx), y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
  const int ym = (y == 0) ? g.ny - 1 : y - 1, yp = (y == g.ny - 1) ? 0 : y + 1;
  const int zm = (z == 0) ? g.nz - 1 : z - 1, zp = (z == g.nz - 1) ? 0 : z + 1;
  uint8_t m[19];
  bool ok[19];
#pragma unroll
  for (int i = 1; i < 19; ++i) {
    ok[i] = false;
    m[i] = 0;
    if (Q == 7 && Q19::cx[i] * Q19::cx[i] + Q19::cy[i] * Q19::cy[i] + Q19::cz[i] * Q19::cz[i] != 1) continue;
    const int xx = x + Q19::cx[i];
    const int yy = (Q19::cy[i] == 0) ? y : ((Q19::cy[i] > 0) ? yp : ym);
    const int zz = (Q19::cz[i] == 0) ? z : ((Q19::cz[i] > 0) ? zp : zm);
    ok[i] = xx >= 0 && xx < g.nx;
    m[i] = mask[ok[i] ? xx + (long long)g.nx * (yy + (long long)g.ny * zz) : n];
  }
  bool hit = false;
#pragma unroll
  for (int i = 1; i < 19; ++i) hit = hit || (ok[i] && sk[m[i]] == CLB_KIND_FLUID);
  if (hit) flag[0] = 1u;
}

// The initial state, at rest with rho linear in x from rho_in to rho_out, is the reference
// profile itself: every slot holds zero.
template <typename R>
__global__ void flow_init_kernel(R* f, Geom g) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= g.N) return;
#pragma unroll
  for (int i = 0; i < 19; ++i) f[(long long)i * g.N + n] = R(0);
}

// Velocity (u = j, rho0 = 1) and density from the pre-collision populations of the next
// step; boundary fluid nodes use the pressure-boundary mass relation. Optionally writes
// the advecting-velocity planes of the solute lattices (latticeToPassiveAdvDiff).
template <typename R, bool ODD>
__global__ void flow_velocity_kernel(FlowArgs<R> a, double* u, double* rho, R* ade_u) {
  const long long n64 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const Geom& g = a.g;
  if (n64 >= g.N) return;
  const int n = (int)n64;
  const int x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
  double jx = 0, jy = 0, jz = 0, rb = 0;
  if (a.kind[a.mask[n]] == CLB_KIND_FLUID) {
    const Nbr nb(x, y, z, g);
    R p[19];
    load19<R, ODD>(a.f, g.N, nb, n, p, a.rdx);
    if (x == 0 || x == g.nx - 1) {
      const int side = (x == 0) ? -1 : 1;
      rb = (double)((x == 0) ? a.rhobar_in : a.rhobar_out);
      const double d = (double)((x == 0) ? a.dface_in : a.dface_out);
      double onwall = 0, normal = 0;
#pragma unroll
      for (int i = 0; i < 19; ++i) {
        if (Q19::cx[i] == 0) onwall += (double)p[i];
        else if (Q19::cx[i] == side) normal += (double)p[i];
      }
      jx = (side < 0) ? (d - onwall - 2.0 * normal) : -(d - onwall - 2.0 * normal);
    } else {
      rb = a.rref0 + a.rrdx * (double)x;
#pragma unroll
      for (int i = 0; i < 19; ++i) {
        rb += (double)p[i];
        jx += Q19::cx[i] * (double)p[i];
        jy += Q19::cy[i] * (double)p[i];
        jz += Q19::cz[i] * (double)p[i];
      }
    }
  }
  if (u) { u[3 * n64] = jx; u[3 * n64 + 1] = jy; u[3 * n64 + 2] = jz; }
  if (rho) rho[n64] = rb + 1.0;
  if (ade_u) { ade_u[n64] = (R)jx; ade_u[g.N + n64] = (R)jy; ade_u[2 * g.N + n64] = (R)jz; }
}

// Velocity statistics used by the driver after a flow solve (src/complab.cpp:625-671):
// acc[0] += sum |u| over fluid nodes, acc[1] = max |u| (as ordered bits), acc[2] += sum of
// u_x over every node of the plane x = nx-2 (the outlet reading), acc[3] += fluid count.
template <typename R, bool ODD>
__global__ void flow_stats_kernel(FlowArgs<R> a, double* acc) {
  const long long n64 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const Geom& g = a.g;
  double su = 0, ux_out = 0, cnt = 0, umax = 0;
  if (n64 < g.N) {
    const int n = (int)n64;
    const int x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
    if (a.kind[a.mask[n]] == CLB_KIND_FLUID && x > 0 && x < g.nx - 1) {
      const Nbr nb(x, y, z, g);
      R p[19];
      load19<R, ODD>(a.f, g.N, nb, n, p, a.rdx);
      double jx = 0, jy = 0, jz = 0;
#pragma unroll
      for (int i = 0; i < 19; ++i) {
        jx += Q19::cx[i] * (double)p[i];
        jy += Q19::cy[i] * (double)p[i];
        jz += Q19::cz[i] * (double)p[i];
      }
      const double um = sqrt(jx * jx + jy * jy + jz * jz);
      su = um;
      umax = um;
      cnt = 1;
      if (x == g.nx - 2) ux_out = jx;
    }
  }
  __shared__ double s0[256], s1[256], s2[256], s3[256];
  const int t = threadIdx.x;
  s0[t] = su; s1[t] = umax; s2[t] = ux_out; s3[t] = cnt;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if (t < k) { s0[t] += s0[t + k]; s1[t] = fmax(s1[t], s1[t + k]); s2[t] += s2[t + k]; s3[t] += s3[t + k]; }
    __syncthreads();
  }
  if (t == 0) {
    atomicAdd(&acc[0], s0[0]);
    atomicMax((unsigned long long*)&acc[1], (unsigned long long)__double_as_longlong(s1[0]));
    atomicAdd(&acc[2], s2[0]);
    atomicAdd(&acc[3], s3[0]);
  }
}

}  // namespace

void launch_energy_reduce(double* energy, long long nblocks, hipStream_t s) {
  hipLaunchKernelGGL(energy_reduce_kernel, dim3(1), dim3(kEnergyThreads), 0, s, energy, nblocks);
}

void launch_tracer_step(const double* energy, long long nblocks, FlowTracer* tr, double* win, double* elog,
                        hipStream_t s) {
  hipLaunchKernelGGL(tracer_step_kernel, dim3(1), dim3(kEnergyThreads), 0, s, energy, nblocks, tr, win, elog);
}

void launch_tracer_begin(FlowTracer* tr, hipStream_t s) {
  hipLaunchKernelGGL(tracer_begin_kernel, dim3(1), dim3(1), 0, s, tr);
}

template <typename R>
void launch_flow_stats(const FlowArgs<R>& a, bool odd_next, double* acc, hipStream_t s) {
  const int bs = 256;
  dim3 grid((unsigned)((a.g.N + bs - 1) / bs));
  if (odd_next) hipLaunchKernelGGL((flow_stats_kernel<R, true>), grid, dim3(bs), 0, s, a, acc);
  else hipLaunchKernelGGL((flow_stats_kernel<R, false>), grid, dim3(bs), 0, s, a, acc);
}
template void launch_flow_stats<float>(const FlowArgs<float>&, bool, double*, hipStream_t);
template void launch_flow_stats<double>(const FlowArgs<double>&, bool, double*, hipStream_t);

template <typename R>
void launch_flow_init(const Geom& g, R* f, hipStream_t s) {
  const int bs = 256;
  hipLaunchKernelGGL(flow_init_kernel<R>, dim3((unsigned)((g.N + bs - 1) / bs)), dim3(bs), 0, s, f, g);
}

void launch_none_reach(const uint8_t* mask, const uint8_t* kind, const Geom& g, int q, unsigned int* flag,
                       hipStream_t s) {
  const dim3 grid((unsigned)((g.N + 255) / 256));
  if (q == 7) hipLaunchKernelGGL((none_reach_kernel<7>), grid, dim3(256), 0, s, mask, kind, g, flag);
  else hipLaunchKernelGGL((none_reach_kernel<19>), grid, dim3(256), 0, s, mask, kind, g, flag);
}

template <typename R>
void launch_flow_step(const FlowArgs<R>& a, bool odd, dim3 block, hipStream_t s, bool reduce) {
  dim3 grid((a.g.nx + block.x - 1) / block.x, (a.g.ny + block.y - 1) / block.y, a.g.nz);
  if (a.energy) {
    if (odd) hipLaunchKernelGGL((flow_step_kernel<R, true, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((flow_step_kernel<R, false, true>), grid, block, 0, s, a);
    if (reduce) launch_energy_reduce(a.energy, (long long)grid.x * grid.y * grid.z, s);
  } else {
    if (odd) hipLaunchKernelGGL((flow_step_kernel<R, true, false>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((flow_step_kernel<R, false, false>), grid, block, 0, s, a);
  }
}

template <typename R>
void launch_flow_velocity(const FlowArgs<R>& a, bool odd_next, double* u, double* rho, R* ade_u, hipStream_t s) {
  const int bs = 256;
  dim3 grid((unsigned)((a.g.N + bs - 1) / bs));
  if (odd_next) hipLaunchKernelGGL((flow_velocity_kernel<R, true>), grid, dim3(bs), 0, s, a, u, rho, ade_u);
  else hipLaunchKernelGGL((flow_velocity_kernel<R, false>), grid, dim3(bs), 0, s, a, u, rho, ade_u);
}

template void launch_flow_init<float>(const Geom&, float*, hipStream_t);
template void launch_flow_init<double>(const Geom&, double*, hipStream_t);
long long flow_blocks(const Geom& g, dim3 block) {
  return (long long)((g.nx + block.x - 1) / block.x) * ((g.ny + block.y - 1) / block.y) * g.nz;
}

template void launch_flow_step<float>(const FlowArgs<float>&, bool, dim3, hipStream_t, bool);
template void launch_flow_step<double>(const FlowArgs<double>&, bool, dim3, hipStream_t, bool);
template void launch_flow_velocity<float>(const FlowArgs<float>&, bool, double*, double*, float*, hipStream_t);
template void launch_flow_velocity<double>(const FlowArgs<double>&, bool, double*, double*, double*, hipStream_t);

}  // namespace clb
