// clb_compact.hip -- the compact (fluid-only) solute layout: building the node table on the
// device, moving the solute state between the dense and compact layouts, and the transport
// step without fused speciation on it (clb_compact.h).
//
// Build, all on the device (it runs after every geometry change, inside a timed loop):
//  1. per voxel: 1 if ADE-fluid; per voxel and direction d: 1 if fluid with a wall at +c_d
//     (a bounce-back link, direction-major);
//  2. exclusive scans of both (the fluid voxel's compact index, each link's ghost slot);
//  3. the node table: dense index per node, and per direction the slot at +c_d.
#include <hipcub/hipcub.hpp>

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

__global__ void compact_flags_kernel(const uint8_t* mask, const uint8_t* kind, Geom g, unsigned int* flag,
                                     unsigned int* lk) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  const int x = (int)(n % g.nx), y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
  const bool f = kind[mask[n]] == CLB_KIND_FLUID;
  flag[n] = f ? 1u : 0u;
#pragma unroll
  for (int d = 1; d < 7; ++d) {
    bool link = false;
    if (f && nb_in_x(x, d, g.nx)) link = kind[mask[nb_index(x, y, z, d, g)]] != CLB_KIND_FLUID;
    lk[(long long)(d - 1) * g.N + n] = link ? 1u : 0u;
  }
}

__global__ void compact_table_kernel(const unsigned int* flag, const unsigned int* cj, const unsigned int* cg, Geom g,
                                     int* cnode, int* cnbr, long long ccap, unsigned int* cnf) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  if (n == g.N - 1) *cnf = cj[n] + flag[n];
  if (!flag[n]) return;
  const int x = (int)(n % g.nx), y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
  const long long j = cj[n];
  cnode[j] = (int)n;
#pragma unroll
  for (int d = 1; d < 7; ++d) {
    int v = -1;
    if (nb_in_x(x, d, g.nx)) {
      const long long m = nb_index(x, y, z, d, g);
      const long long base = (long long)(d - 1) * g.N;
      v = flag[m] ? (int)cj[m] : (int)(ccap + (cg[base + n] - cg[base]));
    }
    cnbr[(long long)(d - 1) * ccap + j] = v;
  }
}

// to_compact: dense -> compact; else compact -> dense. Populations of every fluid node and
// the dense wall slot behind each of its bounce-back links; both deferred-increment buffers;
// towards the compact layout, the velocity.
template <typename R, bool TO_COMPACT>
__global__ void compact_move_kernel(AdeArgs<R> ca, R* gd, const R* ud, R* pd, R* pc) {
  const long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (j >= (long long)*ca.cnf) return;
  const Geom& g = ca.geo;
  const long long N = g.N, st = ca.cstride, cap = ca.ccap;
  const int n = ca.cnode[j];
  const int x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
  if (gd) {
    for (int s = 0; s < ca.S; ++s) {
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        R* dv = &gd[(long long)(s * 7 + i) * N + n];
        R* cv = &ca.g[(long long)(s * 7 + i) * st + j];
        if (TO_COMPACT) *cv = *dv;
        else *dv = *cv;
      }
#pragma unroll
      for (int d = 1; d < 7; ++d) {
        const int v = ca.cnbr[(long long)(d - 1) * cap + j];
        if (v < cap) continue;  // a fluid neighbour's slot, or none
        R* dv = &gd[(long long)(s * 7 + d) * N + nb_index(x, y, z, d, g)];
        R* cv = &ca.g[(long long)(s * 7 + d) * st + v];
        if (TO_COMPACT) *cv = *dv;
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
      else *dv = *cv;
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

void launch_compact_flags(const uint8_t* mask, const uint8_t* kind, const Geom& g, const CompactBuild& b,
                          hipStream_t s) {
  hipLaunchKernelGGL(compact_flags_kernel, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, mask, kind, g,
                     b.flag, b.lk);
}

void launch_compact_scan(const CompactBuild& b, long long N, hipStream_t s) {
  size_t bytes = b.scan_tmp_bytes;
  (void)hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, bytes, b.flag, b.cj, (int)N, s);
  bytes = b.scan_tmp_bytes;
  // lk holds 6N flags and one zero, so cg[6N] is the total number of links
  (void)hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, bytes, b.lk, b.cg, (int)(6 * N + 1), s);
}

void launch_compact_table(const CompactBuild& b, const Geom& g, int* cnode, int* cnbr, long long ccap,
                          unsigned int* cnf, hipStream_t s) {
  hipLaunchKernelGGL(compact_table_kernel, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, b.flag, b.cj, b.cg,
                     g, cnode, cnbr, ccap, cnf);
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
template void launch_ade_step_compact<float>(const AdeArgs<float>&, bool, int, hipStream_t);
template void launch_ade_step_compact<double>(const AdeArgs<double>&, bool, int, hipStream_t);

}  // namespace clb
