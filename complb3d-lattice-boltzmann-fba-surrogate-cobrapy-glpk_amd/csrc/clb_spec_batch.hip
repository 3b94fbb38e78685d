// clb_spec_batch.hip -- batched speciation launcher (totals in, log activities out),
// occupancy sizing of the persistent grid, and small fill kernels.
#include "clb_spec_group.h"
#include "clb_spec_kernel.h"

namespace clb {

void launch_spec_batch(const Tableau& tb, SpecArgs sa, int warm, int method, int grid, int grid8, hipStream_t s) {
  if (tb.nc > 4) {
    // one voxel per group of eight lanes (clb_spec_group.h)
    const dim3 blk(256);
    const bool pol = method == CLB_BATCH_POLISH || (method == CLB_BATCH_AUTO && warm != CLB_WARM_REF);
    if (pol) {
      if (tb.ns <= 16) hipLaunchKernelGGL((spec_group_kernel<16, true>), dim3(grid8), blk, 0, s, tb, sa, warm, method);
      else if (tb.ns <= 24) hipLaunchKernelGGL((spec_group_kernel<24, true>), dim3(grid8), blk, 0, s, tb, sa, warm, method);
      else hipLaunchKernelGGL((spec_group_kernel<32, true>), dim3(grid8), blk, 0, s, tb, sa, warm, method);
    } else {
      if (tb.ns <= 16) hipLaunchKernelGGL((spec_group_kernel<16, false>), dim3(grid8), blk, 0, s, tb, sa, warm, method);
      else if (tb.ns <= 24) hipLaunchKernelGGL((spec_group_kernel<24, false>), dim3(grid8), blk, 0, s, tb, sa, warm, method);
      else hipLaunchKernelGGL((spec_group_kernel<32, false>), dim3(grid8), blk, 0, s, tb, sa, warm, method);
    }
    return;
  }
  CoupledIO io{};
  dispatch_spec<double, 0>(tb, sa, warm, io, grid, s);
}

int spec_group_grid(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 1024;
  int blocks = 0;
  // the largest instantiation bounds the others
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, spec_group_kernel<32, true>, 256, 0) != hipSuccess ||
      blocks < 1)
    blocks = 2;
  return prop.multiProcessorCount * blocks;
}

int spec_occupancy_grid(int nc, int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 1024;
  int blocks = 0;
  // occupancy of the largest instantiation bounds all of them
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, spec_kernel<double, 8, 24, 0>, 256, 0) != hipSuccess ||
      blocks < 1)
    blocks = 2;
  (void)nc;
  return prop.multiProcessorCount * blocks;
}

// Small fill kernels (no hipMemset, so the same code is safe inside captured graphs).
namespace {
__global__ void fill_u32_kernel(unsigned int* p, long long n, unsigned int v) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_u64_kernel(unsigned long long* p, long long n, unsigned long long v) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_f64_kernel(double* p, long long n, double v) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
}  // namespace

void launch_fill_u32(unsigned int* p, long long n, unsigned int v, hipStream_t s) {
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}
void launch_fill_u64(unsigned long long* p, long long n, unsigned long long v, hipStream_t s) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}
void launch_fill_f64(double* p, long long n, double v, hipStream_t s) {
  hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

}  // namespace clb
