// clb_flow_compact.hip -- the D3Q19 flow over the compact layout: only the voxels of flow kind
// FLUID, numbered in brick order (clb_brick.h).
//
// In a sphere pack of porosity 0.35 walls and pores alternate at the scale of one sphere, so
// nearly every 128-byte line of the dense lattice holds a fluid voxel and the dense kernel
// (clb_flow.hip) moves about the whole lattice's bytes. Here the populations of consecutive
// fluid nodes are consecutive words and no wall slot is moved unless a population is in flight
// in it.
//
// Node j's slots: its own, and for each direction d (1..18) the slot at +c_d,
// fnbr[(d-1) fcap + j]: a flow node's j; a ghost slot (>= fcap, plane d) when the neighbour is
// a wall -- the wall's slot d of the dense layout, where the AA pattern keeps the population
// the node sent into the wall until it reads it back reversed (full-way bounce-back); -1
// beyond an x face. Even steps touch only the node's own slots and read no table; odd steps
// read slot (opp(i), -c_i) and write slot (i, +c_i) exactly as the dense kernel does, with the
// same arithmetic, so the two layouts hold the same numbers (the energy tally sums the same
// per-node terms in another order).
//
// A node whose voxel stops being fluid after the table was built (a precipitated or biofilm
// voxel whose code is a wall for the flow) reads its code from the mask every step and does
// nothing: its own slots then serve as the wall slots its neighbours send into and read back
// from, as the dense layout's would. A voxel that becomes fluid needs a new table.
//
// HBM traffic per flow node and step: 2 * 19 * sizeof(R) of populations, the node's voxel
// index (4 B) and mask byte, and on odd steps the 18 table entries (72 B).
#include <hipcub/hipcub.hpp>

#include "clb_brick.h"
#include "clb_flow_node.h"

namespace clb {

namespace {

__device__ __forceinline__ bool nb19_in_x(int x, int d, int nx) {
  const int xx = x + Q19::cx[d];
  return xx >= 0 && xx < nx;
}

__device__ __forceinline__ long long nb19_index(int x, int y, int z, int d, const Geom& g) {
  int yy = y + Q19::cy[d], zz = z + Q19::cz[d];
  if (yy < 0) yy += g.ny;
  if (yy >= g.ny) yy -= g.ny;
  if (zz < 0) zz += g.nz;
  if (zz >= g.nz) zz -= g.nz;
  return (long long)(x + Q19::cx[d]) + (long long)g.nx * (yy + (long long)g.ny * zz);
}

// 1: flow node (brick order); per direction d: 1 if a flow node with a wall at +c_d (a link
// whose population waits in a ghost slot, direction-major).
__global__ void fc_flags_kernel(const uint8_t* mask, const uint8_t* kind, Geom g, BrickOrder o, unsigned int* flag,
                                unsigned int* lk) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= o.nkey) return;
  int x, y, z;
  const bool in = brick_voxel(k, o, g, x, y, z);
  const long long n = in ? (long long)x + (long long)g.nx * (y + (long long)g.ny * z) : 0;
  const bool f = in && kind[mask[n]] == CLB_KIND_FLUID;
  flag[k] = f ? 1u : 0u;
#pragma unroll
  for (int d = 1; d < 19; ++d) {
    bool link = false;
    if (f && nb19_in_x(x, d, g.nx)) link = kind[mask[nb19_index(x, y, z, d, g)]] != CLB_KIND_FLUID;
    lk[(long long)(d - 1) * o.nkey + k] = link ? 1u : 0u;
  }
}

__global__ void fc_table_kernel(const unsigned int* flag, const unsigned int* cj, const unsigned int* cg, Geom g,
                                BrickOrder o, int* fnode, int* fnbr, long long fcap, unsigned int* fnf) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= o.nkey) return;
  if (k == o.nkey - 1) *fnf = cj[k] + flag[k];
  if (!flag[k]) return;
  int x, y, z;
  (void)brick_voxel(k, o, g, x, y, z);
  const long long j = cj[k];
  fnode[j] = (int)((long long)x + (long long)g.nx * (y + (long long)g.ny * z));
#pragma unroll
  for (int d = 1; d < 19; ++d) {
    int v = -1;
    if (nb19_in_x(x, d, g.nx)) {
      const long long m = nb19_index(x, y, z, d, g);
      const int mx = (int)(m % g.nx), my = (int)((m / g.nx) % g.ny), mz = (int)(m / ((long long)g.nx * g.ny));
      const long long km = brick_key(mx, my, mz, o);
      const long long base = (long long)(d - 1) * o.nkey;
      v = flag[km] ? (int)cj[km] : (int)(fcap + (cg[base + k] - cg[base]));
    }
    fnbr[(long long)(d - 1) * fcap + j] = v;
  }
}

// Dense <-> compact: every node's 19 own slots, and behind each of its wall links the wall's
// slot d (the only dense slot of that wall the node's populations pass through).
template <typename R, bool TO_COMPACT>
__global__ void fc_move_kernel(FlowArgs<R> a, R* fd) {
  const long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (j >= (long long)*a.fnf) return;
  const Geom& g = a.g;
  const long long N = g.N, st = a.fstride, cap = a.fcap;
  const int n = a.fnode[j];
  const int x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    R* dv = &fd[(long long)i * N + n];
    R* cv = &a.f[(long long)i * st + j];
    if (TO_COMPACT) *cv = *dv;
    else *dv = *cv;
  }
#pragma unroll
  for (int d = 1; d < 19; ++d) {
    const int v = a.fnbr[(long long)(d - 1) * cap + j];
    if (v < cap) continue;  // a flow node's slot, or none
    R* dv = &fd[(long long)d * N + nb19_index(x, y, z, d, g)];
    R* cv = &a.f[(long long)d * st + v];
    if (TO_COMPACT) *cv = *dv;
    else *dv = *cv;
  }
}

// The table entries of node j (slot at +c_d, d = 1..18), read on odd steps only.
template <typename R>
__device__ __forceinline__ void fc_slots(const FlowArgs<R>& a, int j, int* L) {
#pragma unroll
  for (int d = 1; d < 19; ++d) L[d] = a.fnbr[(long long)(d - 1) * a.fcap + j];
}

// load19 / store19 of clb_flow.hip over the compact slots. XCHK = false for a node with
// 1 <= x <= nx-2: every slot exists (a flow node's or a ghost).
template <typename R, bool ODD, bool XCHK>
__device__ __forceinline__ void fc_load19(const FlowArgs<R>& a, int j, const int* L, R* p) {
  const long long st = a.fstride;
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    if (!ODD || i == 0) {
      p[i] = a.f[(long long)i * st + j];
    } else {
      const int src = L[opp_dir(i)];  // the slot at -c_i
      if (!XCHK || src >= 0) {
        p[i] = a.f[(long long)opp_dir(i) * st + src];
        if (Q19::cx[i] != 0) p[i] = fma(R(-Q19::w(i) * Q19::cx[i]), a.rdx, p[i]);
      } else {
        p[i] = R(0);
      }
    }
  }
}

template <typename R, bool ODD, bool XCHK>
__device__ __forceinline__ void fc_store19(const FlowArgs<R>& a, int j, const int* L, const R* p) {
  const long long st = a.fstride;
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    // plain stores: with the blocks in runs per XCD, neighbouring wavefronts' stores into the
    // same lines meet in one L2 (non-temporal stores, which the dense kernel keeps, measured
    // 0.236 against 0.233 ms here)
    if (!ODD) {
      a.f[(long long)opp_dir(i) * st + j] = p[i];
    } else if (i == 0) {
      a.f[j] = p[0];
    } else {
      const int dst = L[i];
      if (!XCHK || dst >= 0) {
        const R v = (Q19::cx[i] != 0) ? fma(R(-Q19::w(i) * Q19::cx[i]), a.rdx, p[i]) : p[i];
        a.f[(long long)i * st + dst] = v;
      }
    }
  }
}

// flow_node of clb_flow.hip on compact node j (FACE -1: x = 0, +1: x = nx-1, 0: interior).
template <typename R, bool ODD, int FACE>
__device__ __forceinline__ double fc_node(const FlowArgs<R>& a, int j, uint8_t code) {
  int L[19];
  if (ODD) fc_slots(a, j, L);
  R p[19];
  fc_load19<R, ODD, FACE != 0>(a, j, L, p);
  if (FACE == -1) regularize<R, -1>(p, a.dface_in);
  if (FACE == 1) regularize<R, 1>(p, a.dface_out);
  const R jsq = collide19<R>(p, a.omega[code]);
  fc_store19<R, ODD, FACE != 0>(a, j, L, p);
  return 0.5 * (double)jsq;
}

// One thread per compact node, 1-D blocks of at most 256 threads; the energy tally as the
// dense kernel's (per-block partials, summed by energy_reduce_kernel in a fixed order).
template <typename R, bool ODD, bool ENERGY>
__global__ void __launch_bounds__(256) flow_step_compact_kernel(FlowArgs<R> a) {
  if (a.stop && *a.stop) return;  // the device tracer has converged
  const int lb = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const int j = lb * blockDim.x + threadIdx.x;
  double e = 0.0, cnt = 0.0;
  if (j < (int)*a.fnf) {
    const int n = a.fnode[j];
    const uint8_t code = a.mask[n];
    if (a.kind[code] == CLB_KIND_FLUID) {
      const int x = n % a.g.nx;
      double ek;
      if (x == 0) ek = fc_node<R, ODD, -1>(a, j, code);
      else if (x == a.g.nx - 1) ek = fc_node<R, ODD, 1>(a, j, code);
      else ek = fc_node<R, ODD, 0>(a, j, code);
      if (ENERGY) { e = ek; cnt = 1.0; }
    }
  }
  if (ENERGY) {
    for (int off = 32; off > 0; off >>= 1) {
      e += __shfl_xor(e, off);
      cnt += __shfl_xor(cnt, off);
    }
    __shared__ double se[4], sc[4];
    const int t = threadIdx.x;
    const int nw = blockDim.x >> 6;
    if ((t & 63) == 0) { se[t >> 6] = e; sc[t >> 6] = cnt; }
    __syncthreads();
    if (t == 0) {
      double es = 0.0, cs = 0.0;
      for (int w = 0; w < nw; ++w) { es += se[w]; cs += sc[w]; }
      a.energy[2 + 2 * (long long)lb] = es;
      a.energy[3 + 2 * (long long)lb] = cs;
    }
  }
}

// flow_velocity_kernel's advecting velocity (u = j, rho0 = 1; on an x face j_x from the
// pressure-boundary mass relation and no tangential part) from the pre-collision populations of
// the next step, at every fluid node of the table.
template <typename R, bool ODD>
__global__ void __launch_bounds__(256) flow_velocity_compact_kernel(FlowArgs<R> a, R* ade_u) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= (int)*a.fnf) return;
  const Geom& g = a.g;
  const int n = a.fnode[j];
  if (a.kind[a.mask[n]] != CLB_KIND_FLUID) return;
  const int x = n % g.nx;
  int L[19];
  if (ODD) fc_slots(a, j, L);
  R p[19];
  fc_load19<R, ODD, true>(a, j, L, p);
  double jx = 0, jy = 0, jz = 0;
  if (x == 0 || x == g.nx - 1) {
    const int side = (x == 0) ? -1 : 1;
    const double d = (double)((x == 0) ? a.dface_in : a.dface_out);
    double onwall = 0, normal = 0;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
      if (Q19::cx[i] == 0) onwall += (double)p[i];
      else if (Q19::cx[i] == side) normal += (double)p[i];
    }
    jx = (side < 0) ? (d - onwall - 2.0 * normal) : -(d - onwall - 2.0 * normal);
  } else {
#pragma unroll
    for (int i = 0; i < 19; ++i) {
      jx += Q19::cx[i] * (double)p[i];
      jy += Q19::cy[i] * (double)p[i];
      jz += Q19::cz[i] * (double)p[i];
    }
  }
  ade_u[n] = (R)jx;
  ade_u[g.N + n] = (R)jy;
  ade_u[2 * g.N + n] = (R)jz;
}

}  // namespace

size_t flow_compact_scan_bytes(long long nkey) {
  size_t a = 0, b = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, a, (const unsigned int*)nullptr, (unsigned int*)nullptr, (int)nkey,
                                         (hipStream_t)0);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const unsigned int*)nullptr, (unsigned int*)nullptr,
                                         (int)(18 * nkey + 1), (hipStream_t)0);
  return a > b ? a : b;
}

void launch_flow_compact_build(const uint8_t* mask, const uint8_t* kind, const Geom& g, const BrickOrder& o,
                               const CompactBuild& b, hipStream_t s) {
  hipLaunchKernelGGL(fc_flags_kernel, dim3((unsigned)((o.nkey + 255) / 256)), dim3(256), 0, s, mask, kind, g, o, b.flag,
                     b.lk);
  size_t bytes = b.scan_tmp_bytes;
  (void)hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, bytes, b.flag, b.cj, (int)o.nkey, s);
  bytes = b.scan_tmp_bytes;
  // lk holds 18 nkey flags and one zero, so cg[18 nkey] is the total number of links
  (void)hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, bytes, b.lk, b.cg, (int)(18 * o.nkey + 1), s);
}

void launch_flow_compact_table(const CompactBuild& b, const Geom& g, const BrickOrder& o, int* fnode, int* fnbr,
                               long long fcap, unsigned int* fnf, hipStream_t s) {
  hipLaunchKernelGGL(fc_table_kernel, dim3((unsigned)((o.nkey + 255) / 256)), dim3(256), 0, s, b.flag, b.cj, b.cg, g,
                     o, fnode, fnbr, fcap, fnf);
}

template <typename R>
void launch_flow_compact_move(bool to_compact, const FlowArgs<R>& a, R* f_dense, hipStream_t s) {
  const unsigned blocks = (unsigned)((a.fcap + 255) / 256);
  if (to_compact) hipLaunchKernelGGL((fc_move_kernel<R, true>), dim3(blocks), dim3(256), 0, s, a, f_dense);
  else hipLaunchKernelGGL((fc_move_kernel<R, false>), dim3(blocks), dim3(256), 0, s, a, f_dense);
}

long long flow_compact_blocks(long long fcap, int bt) { return (fcap + bt - 1) / bt; }

template <typename R>
void launch_flow_step_compact(const FlowArgs<R>& a, bool odd, int bt, hipStream_t s, bool reduce) {
  const long long blocks = flow_compact_blocks(a.fcap, bt);
  if (a.energy) {
    if (odd) hipLaunchKernelGGL((flow_step_compact_kernel<R, true, true>), dim3((unsigned)blocks), dim3(bt), 0, s, a);
    else hipLaunchKernelGGL((flow_step_compact_kernel<R, false, true>), dim3((unsigned)blocks), dim3(bt), 0, s, a);
    if (reduce) launch_energy_reduce(a.energy, blocks, s);
  } else {
    if (odd) hipLaunchKernelGGL((flow_step_compact_kernel<R, true, false>), dim3((unsigned)blocks), dim3(bt), 0, s, a);
    else hipLaunchKernelGGL((flow_step_compact_kernel<R, false, false>), dim3((unsigned)blocks), dim3(bt), 0, s, a);
  }
}

template <typename R>
void launch_flow_velocity_compact(const FlowArgs<R>& a, bool odd_next, R* ade_u, hipStream_t s) {
  const unsigned blocks = (unsigned)((a.fcap + 255) / 256);
  if (odd_next) hipLaunchKernelGGL((flow_velocity_compact_kernel<R, true>), dim3(blocks), dim3(256), 0, s, a, ade_u);
  else hipLaunchKernelGGL((flow_velocity_compact_kernel<R, false>), dim3(blocks), dim3(256), 0, s, a, ade_u);
}

template void launch_flow_compact_move<float>(bool, const FlowArgs<float>&, float*, hipStream_t);
template void launch_flow_compact_move<double>(bool, const FlowArgs<double>&, double*, hipStream_t);
template void launch_flow_step_compact<float>(const FlowArgs<float>&, bool, int, hipStream_t, bool);
template void launch_flow_step_compact<double>(const FlowArgs<double>&, bool, int, hipStream_t, bool);
template void launch_flow_velocity_compact<float>(const FlowArgs<float>&, bool, float*, hipStream_t);
template void launch_flow_velocity_compact<double>(const FlowArgs<double>&, bool, double*, hipStream_t);

}  // namespace clb
