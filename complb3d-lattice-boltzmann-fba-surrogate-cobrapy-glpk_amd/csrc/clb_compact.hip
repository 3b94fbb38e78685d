// clb_compact.hip -- the compact (fluid-only) solute layout: building the node table on the
// device, moving the solute state between the dense and compact layouts, and the transport
// step without fused speciation on it (clb_compact.h).
//
// The fluid nodes are numbered in brick order: the lattice cut into bricks of bx x by x bz
// voxels, the bricks in x-fastest order and the voxels natural inside each. A workgroup's run
// of nodes then covers a compact block of the lattice, so most y and z neighbours of its nodes
// are its own nodes (bx = nx, by = bz = 1 is the natural order).
//
// Build, all on the device (it runs after every geometry change, inside a timed loop), over
// the brick-order key of every voxel (padded to whole bricks):
//  1. 1 if ADE-fluid; per direction d: 1 if fluid with a wall at +c_d (a bounce-back link,
//     direction-major);
//  2. exclusive scans of both (the fluid voxel's compact index, each link's ghost slot);
//  3. the node table: dense index per node, and per direction the slot at +c_d.
#include <hipcub/hipcub.hpp>

#include "clb_brick.h"
#include "clb_compact.h"

namespace clb {

namespace {

__device__ __forceinline__ bool nb_in_x(int x, int d, int nx) {
  const int xx = x + Q7::cx[d];
  return xx >= 0 && xx < nx;
}

__device__ __forceinline__ long long nb_index(int x, int y, int z, int d, const Geom& g) {
  int yy = y + Q7::cy[d], zz = z + Q7::cz[d];
  if (yy < 0) yy += g.ny;
  if (yy >= g.ny) yy -= g.ny;
  if (zz < 0) zz += g.nz;
  if (zz >= g.nz) zz -= g.nz;
  return (long long)(x + Q7::cx[d]) + (long long)g.nx * (yy + (long long)g.ny * zz);
}

__global__ void compact_flags_kernel(const uint8_t* mask, const uint8_t* kind, Geom g, BrickOrder o, unsigned int* flag,
                                     unsigned int* lk) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= o.nkey) return;
  int x, y, z;
  const bool in = brick_voxel(k, o, g, x, y, z);
  const long long n = in ? (long long)x + (long long)g.nx * (y + (long long)g.ny * z) : 0;
  const bool f = in && kind[mask[n]] == CLB_KIND_FLUID;
  flag[k] = f ? 1u : 0u;
#pragma unroll
  for (int d = 1; d < 7; ++d) {
    bool link = false;
    if (f && nb_in_x(x, d, g.nx)) link = kind[mask[nb_index(x, y, z, d, g)]] != CLB_KIND_FLUID;
    lk[(long long)(d - 1) * o.nkey + k] = link ? 1u : 0u;
  }
}

__global__ void compact_table_kernel(const unsigned int* flag, const unsigned int* cj, const unsigned int* cg, Geom g,
                                     BrickOrder o, int* cnode, int* cnbr, long long ccap, unsigned int* cnf) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= o.nkey) return;
  if (k == o.nkey - 1) *cnf = cj[k] + flag[k];
  if (!flag[k]) return;
  int x, y, z;
  (void)brick_voxel(k, o, g, x, y, z);
  const long long j = cj[k];
  cnode[j] = (int)((long long)x + (long long)g.nx * (y + (long long)g.ny * z));
#pragma unroll
  for (int d = 1; d < 7; ++d) {
    int v = -1;
    if (nb_in_x(x, d, g.nx)) {
      const long long m = nb_index(x, y, z, d, g);
      const int mx = (int)(m % g.nx), my = (int)((m / g.nx) % g.ny), mz = (int)(m / ((long long)g.nx * g.ny));
      const long long km = brick_key(mx, my, mz, o);
      const long long base = (long long)(d - 1) * o.nkey;
      v = flag[km] ? (int)cj[km] : (int)(ccap + (cg[base + k] - cg[base]));
    }
    cnbr[(long long)(d - 1) * ccap + j] = v;
  }
}

// to_compact: dense -> compact; else compact -> dense. Populations of every fluid node and
// the dense wall slot behind each of its bounce-back links; both deferred-increment buffers;
// towards the compact layout, the velocity.
// FLUSH (compact -> dense, the next step even): the deferred increments (ca.pend_r, ca.pend_w)
// are folded in on the way, as compact_flush_kernel<R, false> would fold them in first (each
// node's own slots and its ghost slots, the same sums), and the dense increments are zero.
template <typename R, bool TO_COMPACT, bool FLUSH = false>
__global__ void compact_move_kernel(AdeArgs<R> ca, R* gd, const R* ud, R* pd, R* pc) {
  const long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (j >= (long long)*ca.cnf) return;
  const Geom& g = ca.geo;
  const long long N = g.N, st = ca.cstride, cap = ca.ccap;
  const int n = ca.cnode[j];
  const int x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
  if (gd) {
    CSlots L;
    if (FLUSH) L = cslots(ca, (int)j);
    for (int s = 0; s < ca.S; ++s) {
      R own = R(0), prev = R(0);  // the node's increments of the previous step and the one before
      if (FLUSH) {
        own = ca.pend_r[(long long)s * cap + j];
        prev = ca.pend_w[(long long)s * cap + j];
      }
      const bool imm = ca.immobile[s] != 0;
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        R* dv = &gd[(long long)(s * 7 + i) * N + n];
        R* cv = &ca.g[(long long)(s * 7 + i) * st + j];
        if (TO_COMPACT) {
          *cv = *dv;
        } else if (FLUSH) {
          // add_pending's increment of the population read from slot (i, j) next step (slots it
          // leaves alone are copied as they are)
          const int src = L.s[opp_dir(i)];
          const bool has = imm ? (i == 0) : (src >= 0);
          const R v = (i == 0) ? own : (src >= 0 && src < cap ? ca.pend_r[(long long)s * cap + src] : prev);
          const R add = imm ? v : R(Q7::w(i)) * v;
          *dv = has ? *cv + add : *cv;
        } else {
          *dv = *cv;
        }
      }
#pragma unroll
      for (int d = 1; d < 7; ++d) {
        const int v = ca.cnbr[(long long)(d - 1) * cap + j];
        if (v < cap) continue;  // a fluid neighbour's slot, or none
        R* dv = &gd[(long long)(s * 7 + d) * N + nb_index(x, y, z, d, g)];
        R* cv = &ca.g[(long long)(s * 7 + d) * st + v];
        if (TO_COMPACT) *cv = *dv;
        else if (FLUSH && !imm) *dv = *cv + R(Q7::w(d)) * own;  // sent towards the wall last step
        else *dv = *cv;
      }
    }
  }
  if (TO_COMPACT && ud) {
    R* cu = const_cast<R*>(ca.u);
#pragma unroll
    for (int k = 0; k < 3; ++k) cu[(long long)k * cap + j] = ud[(long long)k * N + n];
  }
  if (pd && pc) {
    for (int q = 0; q < 2 * ca.S; ++q) {
      R* dv = &pd[(long long)q * N + n];
      R* cv = &pc[(long long)q * cap + j];
      if (TO_COMPACT) *cv = *dv;
      else *dv = FLUSH ? R(0) : *cv;
    }
  }
}

// pending_flush_kernel (clb_ade.hip) on the compact layout: for a next step of parity ODD,
// each fluid node adds to the slots it reads then the increments add_pending would add, and
// to each population it sent towards a wall (its ghost slot) t_i of its own previous-step
// increment. Each slot belongs to one node: no races.
template <typename R, bool ODD>
__global__ void compact_flush_kernel(AdeArgs<R> a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= (int)*a.cnf) return;
  const long long st = a.cstride, cap = a.ccap;
  const CSlots L = cslots(a, j);
  for (int s = 0; s < a.S; ++s) {
    const R* P1 = a.pend_r + (long long)s * cap;
    const R* P2 = a.pend_w + (long long)s * cap;
    if (a.immobile[s]) {
      a.g[(long long)(s * 7) * st + j] += P1[j];
      continue;
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const int src = L.s[opp_dir(i)];  // the slot at -c_i (i = 0: the node's own)
      if (src >= 0) {
        const R v = (i == 0) ? P1[j] : (src < cap ? P1[src] : P2[j]);
        R* q = ODD ? &a.g[(long long)(s * 7 + opp_dir(i)) * st + src] : &a.g[(long long)(s * 7 + i) * st + j];
        *q += R(Q7::w(i)) * v;
      }
      if (i > 0 && L.s[i] >= cap) {
        R* q = ODD ? &a.g[(long long)(s * 7 + opp_dir(i)) * st + j] : &a.g[(long long)(s * 7 + i) * st + L.s[i]];
        *q += R(Q7::w(i)) * P1[j];
      }
    }
  }
}

template <typename R, int S>
void launch_ade_step_compact_s(const AdeArgs<R>& a, bool odd, int bt, hipStream_t st) {
  const unsigned blocks = (unsigned)((a.ccap + bt - 1) / bt);
  const bool kin = (a.kinetics != 0) || (a.abiotic != 0);
  if (odd) {
    if (kin) hipLaunchKernelGGL((ade_step_compact_kernel<R, S, true, true>), dim3(blocks), dim3(bt), 0, st, a);
    else hipLaunchKernelGGL((ade_step_compact_kernel<R, S, true, false>), dim3(blocks), dim3(bt), 0, st, a);
  } else {
    if (kin) hipLaunchKernelGGL((ade_step_compact_kernel<R, S, false, true>), dim3(blocks), dim3(bt), 0, st, a);
    else hipLaunchKernelGGL((ade_step_compact_kernel<R, S, false, false>), dim3(blocks), dim3(bt), 0, st, a);
  }
}

}  // namespace

size_t compact_scan_bytes(long long n) {
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const unsigned int*)nullptr, (unsigned int*)nullptr,
                                         (int)n, (hipStream_t)0);
  return bytes;
}

void launch_compact_flags(const uint8_t* mask, const uint8_t* kind, const Geom& g, const BrickOrder& o,
                          const CompactBuild& b, hipStream_t s) {
  hipLaunchKernelGGL(compact_flags_kernel, dim3((unsigned)((o.nkey + 255) / 256)), dim3(256), 0, s, mask, kind, g, o,
                     b.flag, b.lk);
}

void launch_compact_scan(const CompactBuild& b, long long nkey, hipStream_t s) {
  size_t bytes = b.scan_tmp_bytes;
  (void)hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, bytes, b.flag, b.cj, (int)nkey, s);
  bytes = b.scan_tmp_bytes;
  // lk holds 6 nkey flags and one zero, so cg[6 nkey] is the total number of links
  (void)hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, bytes, b.lk, b.cg, (int)(6 * nkey + 1), s);
}

void launch_compact_table(const CompactBuild& b, const Geom& g, const BrickOrder& o, int* cnode, int* cnbr,
                          long long ccap, unsigned int* cnf, hipStream_t s) {
  hipLaunchKernelGGL(compact_table_kernel, dim3((unsigned)((o.nkey + 255) / 256)), dim3(256), 0, s, b.flag, b.cj,
                     b.cg, g, o, cnode, cnbr, ccap, cnf);
}

template <typename R>
void launch_compact_move(bool to_compact, const AdeArgs<R>& ca, R* g_dense, const R* u_dense, R* pend_dense,
                         R* pend_compact, hipStream_t s) {
  const unsigned blocks = (unsigned)((ca.ccap + 255) / 256);
  if (to_compact)
    hipLaunchKernelGGL((compact_move_kernel<R, true>), dim3(blocks), dim3(256), 0, s, ca, g_dense, u_dense, pend_dense,
                       pend_compact);
  else
    hipLaunchKernelGGL((compact_move_kernel<R, false>), dim3(blocks), dim3(256), 0, s, ca, g_dense, nullptr,
                       pend_dense, pend_compact);
}

template <typename R>
void launch_compact_move_flush(const AdeArgs<R>& ca, R* g_dense, R* pend_dense, R* pend_compact, hipStream_t s) {
  const unsigned blocks = (unsigned)((ca.ccap + 255) / 256);
  hipLaunchKernelGGL((compact_move_kernel<R, false, true>), dim3(blocks), dim3(256), 0, s, ca, g_dense, nullptr,
                     pend_dense, pend_compact);
}

template <typename R>
void launch_compact_flush(const AdeArgs<R>& a, bool odd_next, hipStream_t s) {
  const unsigned blocks = (unsigned)((a.ccap + 255) / 256);
  if (odd_next) hipLaunchKernelGGL((compact_flush_kernel<R, true>), dim3(blocks), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((compact_flush_kernel<R, false>), dim3(blocks), dim3(256), 0, s, a);
}

template <typename R>
void launch_ade_step_compact(const AdeArgs<R>& a, bool odd, int bt, hipStream_t s) {
  switch (a.S) {
    case 1: launch_ade_step_compact_s<R, 1>(a, odd, bt, s); break;
    case 2: launch_ade_step_compact_s<R, 2>(a, odd, bt, s); break;
    case 3: launch_ade_step_compact_s<R, 3>(a, odd, bt, s); break;
    case 4: launch_ade_step_compact_s<R, 4>(a, odd, bt, s); break;
    case 5: launch_ade_step_compact_s<R, 5>(a, odd, bt, s); break;
    case 6: launch_ade_step_compact_s<R, 6>(a, odd, bt, s); break;
    case 7: launch_ade_step_compact_s<R, 7>(a, odd, bt, s); break;
    case 8: launch_ade_step_compact_s<R, 8>(a, odd, bt, s); break;
    default: break;
  }
}

template void launch_compact_move<float>(bool, const AdeArgs<float>&, float*, const float*, float*, float*,
                                         hipStream_t);
template void launch_compact_move<double>(bool, const AdeArgs<double>&, double*, const double*, double*, double*,
                                          hipStream_t);
template void launch_compact_move_flush<float>(const AdeArgs<float>&, float*, float*, float*, hipStream_t);
template void launch_compact_move_flush<double>(const AdeArgs<double>&, double*, double*, double*, hipStream_t);
template void launch_compact_flush<float>(const AdeArgs<float>&, bool, hipStream_t);
template void launch_compact_flush<double>(const AdeArgs<double>&, bool, hipStream_t);
template void launch_ade_step_compact<float>(const AdeArgs<float>&, bool, int, hipStream_t);
template void launch_ade_step_compact<double>(const AdeArgs<double>&, bool, int, hipStream_t);

}  // namespace clb
