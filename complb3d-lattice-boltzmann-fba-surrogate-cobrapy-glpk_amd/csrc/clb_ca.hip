// clb_ca.hip -- cellular-automaton biomass redistribution (the CA biofilm solver):
//   pushExcessBiomass3D / halfPushExcessBiomass3D  src/complab3d_processors_part1.hh:432-894
//   updateLocalMaskNtotalLattices3D                src/complab3d_processors_part2.hh:99-210
//   updateAgeDistance3D                            src/complab3d_processors_part2.hh:718-842
//   createAgeDomain3D                              src/complab3d_processors_part3.hh:347-367
//   the sweep loop                                 src/complab.cpp:1453-1731
//   fdDiffusion3D (the FD biofilm solver)          src/complab3d_processors_part2.hh:257-430
//
// The reference sweeps voxels in place, in loop order, with std::random_shuffle and
// std::rand: its result depends on the visiting order and on the clock-seeded generator.
// Here every sweep has two phases over a frozen state, so it does not depend on thread
// order. The rule is the reference's, restated with the frozen state:
//  1. push decision (ca_push_kernel): each overfull voxel splits its excess over its
//     non-wall face neighbours. A random order of the neighbours comes from a counter-based
//     generator keyed by (seed, sweep, voxel, draw). Neighbours are filled up to Bmax in
//     that order, judged on the sweep's starting totals, as the reference's stale total
//     lattice does. The rest goes to one neighbour chosen by the age rule. Only the six
//     amounts are written;
//  2. gather (ca_gather_kernel): each voxel subtracts what it sends and adds what its
//     neighbours send it, species by species in proportion to the sender's starting
//     biomass. Sender and receiver evaluate the same products, so mass is conserved to
//     the last bit.
// FP contraction is off in this file: the CPU restatement reproduces every operation, so
// with FP64 storage the two agree bit for bit.
#include "clb_internal.h"
#include "clb_mask_rule.h"

namespace clb {
namespace {

#pragma clang fp contract(off)

constexpr double kCaThrd = 1e-12;     // thrd (src/complab3d_processors.hh:60)
constexpr double kCaThrdMask = 1e-14; // COMPLAB_THRD (src/complab_functions.hh:64)

__device__ __forceinline__ unsigned long long ca_mix(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Draw k (0..7) of voxel n in sweep `sweep`.
__device__ __forceinline__ unsigned long long ca_rand(unsigned long long seed, unsigned long long sweep, long long n,
                                                      int k) {
  return ca_mix(seed ^ ca_mix(sweep ^ ca_mix((unsigned long long)n * 8ull + (unsigned long long)k)));
}

// Face neighbour d (0 +x, 1 -x, 2 +y, 3 -y, 4 +z, 5 -z: the reference's order) of (x,y,z),
// inside [x0, x1] along x and the whole box along y and z (no wrap); -1 outside.
__device__ __forceinline__ long long ca_nbr(const Geom& g, int x, int y, int z, int d, int x0, int x1) {
  const int dx = (d == 0) - (d == 1), dy = (d == 2) - (d == 3), dz = (d == 4) - (d == 5);
  const int xx = x + dx, yy = y + dy, zz = z + dz;
  if (xx < x0 || xx > x1 || yy < 0 || yy >= g.ny || zz < 0 || zz >= g.nz) return -1;
  return xx + (long long)g.nx * (yy + (long long)g.ny * zz);
}

// Total biomass and mask (updateLocalMaskNtotalLattices3D): bmass = sum of species above
// 1e-14; the stored total follows it when they differ by more than 1e-14 (always at
// `fresh`). The pore <-> biofilm rule (mask_rule_code) on the interior planes
// 1 <= x <= nx-2 (as the transport kernel's mask rule), each voxel deciding from prev (the
// mask itself with one pore type, a copy of it taken before the pass with several).
// red[0]: max total over those planes (bits; the inlet / outlet planes take no part in the
// CA), red[1]: mask changes.
// Grid-stride over a fixed grid with one atomic per block: per-wavefront atomics on one
// address serialise (65 k of them took 0.8 ms at 4 M voxels).
template <typename R>
__global__ void __launch_bounds__(256) ca_mask_total_kernel(const R* b, int nb, const uint8_t* prev, uint8_t* mask,
                                                            double* btot, Geom g, int solid, int bb, MaskRule r,
                                                            double bthr, bool fresh, unsigned long long* red) {
  __shared__ unsigned long long smax[4];
  __shared__ unsigned int schg[4];
  unsigned long long bits = 0;  // max over this thread's voxels (non-negative doubles order as bits)
  unsigned int changed = 0;
  for (long long n = blockIdx.x * 256LL + threadIdx.x; n < g.N; n += (long long)gridDim.x * 256) {
    const int m = prev[n];
    const int xn = (int)(n % g.nx);
    const bool inner = xn >= 1 && xn <= g.nx - 2;
    double t = btot[n];
    if (m != solid && m != bb) {
      double bm = 0.0;
      int label = 0;
      for (int i = 0; i < nb; ++i) {
        const double v = (double)b[(long long)i * g.N + n];
        if (v > kCaThrdMask) {
          bm += v;
          label += r.bio[i];
        }
      }
      if (fresh || fabs(bm - t) > kCaThrdMask) {
        t = bm;
        btot[n] = bm;
      }
      if (inner) {
        const int y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
        const int nm = mask_rule_code(r, prev, g, xn, y, z, m, bm, label, bthr, solid, bb);
        if (nm != m) {
          mask[n] = (uint8_t)nm;
          ++changed;
        }
      }
    }
    if (inner && (t > 0.0 || t != t)) {
      const unsigned long long tb = (unsigned long long)__double_as_longlong(t);
      bits = tb > bits ? tb : bits;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(bits, off);
    bits = o > bits ? o : bits;
    changed += __shfl_xor(changed, off);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smax[w] = bits;
    schg[w] = changed;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long mb = smax[0];
    unsigned int ch = schg[0];
    for (int k = 1; k < 4; ++k) {
      mb = smax[k] > mb ? smax[k] : mb;
      ch += schg[k];
    }
    if (mb) atomicMax(red, mb);
    if (ch) atomicAdd(&red[1], (unsigned long long)ch);
  }
}

// Phase 1: out[d * N + n] = total biomass voxel n sends to its neighbour d this sweep,
// written only where it is not zero; bit d of flags[n] says which (flags are written for
// every voxel, so the amounts need no clearing).
// half: halfPushExcessBiomass3D (biofilm voxels send half their biomass, the age rule
// accepts neighbours of equal age) instead of pushExcessBiomass3D (the excess over Bmax).
__global__ void ca_push_kernel(const uint8_t* mask, const double* btot, const int* age, double* out, uint8_t* flags,
                               Geom g, int solid, int bb, MaskRule r, double bmax, int half, unsigned long long seed,
                               unsigned long long sweep) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  double o[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int x = (int)(n % g.nx), y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
  const int m = mask[n];
  const double T = btot[n];
  bool push = x >= 1 && x <= g.nx - 2 && m != solid && m != bb && T > bmax;
  if (half) push = push && !mr_is_pore(r, m);
  double bMd = half ? T * 0.5 : T - bmax;
  if (!half) push = push && bMd > kCaThrd;
  if (push) {
    int kept[6];
    long long nn[6];
    int nk = 0;
    for (int d = 0; d < 6; ++d) {
      const long long q = ca_nbr(g, x, y, z, d, 1, g.nx - 2);
      if (q >= 0 && mask[q] != solid && mask[q] != bb) {
        kept[nk] = d;
        nn[nk] = q;
        ++nk;
      }
    }
    if (nk > 0) {
      int idx[6] = {0, 1, 2, 3, 4, 5};
      for (int i = nk - 1; i >= 1; --i) {  // Fisher-Yates over the kept neighbours
        const int j = (int)(ca_rand(seed, sweep, n, i) % (unsigned long long)(i + 1));
        const int t = idx[i];
        idx[i] = idx[j];
        idx[j] = t;
      }
      bool done = false;
      for (int t = 0; t < nk && !done; ++t) {
        const int k = idx[t];
        const long long q = nn[k];
        const double nt = mr_is_pore(r, mask[q]) ? 0.0 : btot[q];
        if (nt < bmax) {
          const double hold = bmax - nt;
          double part;
          if (bMd > hold) {
            part = hold;
            bMd -= hold;
          } else {
            part = bMd;
            bMd = 0.0;
            done = true;
          }
          o[kept[k]] += part;
        }
      }
      if (!done) {
        // the rest to a neighbour nearer the biofilm surface (lower age), else any
        const int a0 = age[n];
        int cand[6], nc = 0;
        for (int k = 0; k < nk; ++k) {
          const int a1 = age[nn[k]];
          if (half ? (a0 >= a1) : (a0 > a1)) cand[nc++] = kept[k];
        }
        int d;
        if (nc > 1) d = cand[(int)(ca_rand(seed, sweep, n, 6) % (unsigned long long)nc)];
        else if (nc == 1) d = cand[0];
        else d = kept[nk > 1 ? (int)(ca_rand(seed, sweep, n, 7) % (unsigned long long)nk) : 0];
        o[d] += bMd;
      }
    }
  }
  unsigned int fl = 0;
  for (int d = 0; d < 6; ++d) {
    if (o[d] != 0.0) {
      fl |= 1u << d;
      out[(long long)d * g.N + n] = o[d];
    }
  }
  flags[n] = (uint8_t)fl;
}

// Phase 2: species i of voxel n loses b_i / T * out for each of its sends and gains the
// same expression, evaluated with the sender's values, for each send it receives. An
// amount at or below 1e-12 is not moved (the reference's shove threshold).
template <typename R>
__global__ void ca_gather_kernel(const R* b_old, R* b, int nb, const uint8_t* mask, const double* btot,
                                 const double* out, const uint8_t* flags, Geom g, int solid, int bb) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  const int m = mask[n];
  if (m == solid || m == bb) return;
  const int x = (int)(n % g.nx), y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
  double os[6];
  long long src[6];
  const unsigned int fl = flags[n];
  bool any = fl != 0;
  for (int d = 0; d < 6; ++d) {
    os[d] = ((fl >> d) & 1u) ? out[(long long)d * g.N + n] : 0.0;
    // the neighbour across face d sends towards n through its face opp(d)
    const long long q = ca_nbr(g, x, y, z, d, 0, g.nx - 1);
    src[d] = (q >= 0 && ((flags[q] >> (d ^ 1)) & 1u)) ? q : -1;
    any = any || src[d] >= 0;
  }
  if (!any) return;
  const double T = btot[n];
  for (int i = 0; i < nb; ++i) {
    const double bi = (double)b_old[(long long)i * g.N + n];
    double v = bi;
    for (int d = 0; d < 6; ++d) {
      if (os[d] != 0.0) {
        const double s = bi / T * os[d];
        if (s > kCaThrd) v -= s;
      }
    }
    for (int d = 0; d < 6; ++d) {
      const long long q = src[d];
      if (q >= 0) {
        const double s = (double)b_old[(long long)i * g.N + q] / btot[q] * out[(long long)(d ^ 1) * g.N + q];
        if (s > kCaThrd) v += s;
      }
    }
    b[(long long)i * g.N + n] = (R)v;
  }
}

// One Jacobi pass of updateAgeDistance3D over voxels holding biomass (total > 1e-14);
// neighbours count when their distance flag is set (not wall in the distance field).
// Grid-stride, one atomic per block for the change count.
__global__ void __launch_bounds__(256) ca_age_kernel(const int* age_in, int* age_out, const uint8_t* dpos,
                                                     const double* btot, Geom g, double bmax,
                                                     unsigned long long* changed) {
  __shared__ unsigned int schg[4];
  unsigned int ch = 0;
  for (long long n = blockIdx.x * 256LL + threadIdx.x; n < g.N; n += (long long)gridDim.x * 256) {
    const int age = age_in[n];
    int na = age;
    const double B = btot[n];
    if (B > kCaThrdMask) {
      const int x = (int)(n % g.nx), y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
      long long nn[6];
      int nk = 0;
      for (int d = 0; d < 6; ++d) {
        const long long q = ca_nbr(g, x, y, z, d, 0, g.nx - 1);
        if (q >= 0 && dpos[q]) nn[nk++] = q;
      }
      if (age == 0) {
        na = 1;
      } else if (age == 1 && (B - bmax) > -kCaThrdMask) {
        bool ok = true;
        for (int k = 0; k < nk; ++k)
          if (age_in[nn[k]] == 0) ok = false;
        if (ok) na = 2;
      } else {
        int cnt = 0;
        for (int k = 0; k < nk; ++k)
          if (age_in[nn[k]] >= age && (btot[nn[k]] - bmax) > -kCaThrdMask) ++cnt;
        if (cnt == nk) na = age + 1;
      }
    }
    age_out[n] = na;
    ch += na != age ? 1u : 0u;
  }
  for (int off = 32; off > 0; off >>= 1) ch += __shfl_xor(ch, off);
  if ((threadIdx.x & 63) == 0) schg[threadIdx.x >> 6] = ch;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int t = schg[0] + schg[1] + schg[2] + schg[3];
    if (t) atomicAdd(changed, (unsigned long long)t);
  }
}

// createAgeDomain3D (-1 wall, 0 pore, 1 other) and the distance field's sign (the
// distance is -1 on solid, 0 on bounce-back, positive elsewhere: calculateDistanceFromSolid,
// src/complab_functions.hh:169-230).
__global__ void ca_init_kernel(const uint8_t* mask, int* age, uint8_t* dpos, Geom g, int solid, int bb, MaskRule r) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  const int m = mask[n];
  const bool wall = m == solid || m == bb;
  age[n] = wall ? -1 : (mr_is_pore(r, m) ? 0 : 1);
  dpos[n] = wall ? 0 : 1;
}

// fdDiffusion3D: one explicit step B += nu * (the 7-point Laplacian of B) for every biomass
// field, read from the step's starting copy b_old, written to b, at the voxels of
// 1 <= x <= nx-2 whose mask code is above 1. Per axis, as the reference decides it:
//  * at the low end of the axis (x = 1, y = 0, z = 0), or when the low neighbour's code is
//    below 2: the low neighbour is replaced by the voxel itself and the high one is read;
//  * else at the high end (x = nx-2, y = ny-1, z = nz-1), or when the high neighbour's code
//    is below 2: the high neighbour is the voxel itself and the low one is read;
//  * else both are read.
// (A voxel with walls on both sides of an axis therefore reads its high neighbour, a wall;
// at y = ny-1 or z = nz-1 that read wraps to the other side, as the lattice does in y and
// z.) Values below 1e-14 count as 0; a result at or below 1e-14 is not written, so the
// voxel keeps its value.
constexpr double kFdThrd = 1e-14;  // COMPLAB_THRD, src/complab_functions.hh:64

__device__ __forceinline__ int fd_axis_mode(int c, int lo, int hi, int m_lo, int m_hi) {
  // m_lo / m_hi: the neighbours' codes, 0 where the reference does not read them
  if (c == lo || m_lo < 2) return 1;
  if (c == hi || m_hi < 2) return 2;
  return 0;
}

template <typename R>
__global__ void __launch_bounds__(256) fd_biomass_kernel(const R* b_old, R* b, int nb, const uint8_t* mask, Geom g,
                                                         double nu) {
  const long long n = blockIdx.x * 256LL + threadIdx.x;
  if (n >= g.N) return;
  const int x = (int)(n % g.nx), y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
  if (x < 1 || x > g.nx - 2 || mask[n] < 2) return;
  const long long sy = g.nx, sz = (long long)g.nx * g.ny;
  const long long xp = n + 1, xm = n - 1;
  const long long yp = (y == g.ny - 1) ? n - (g.ny - 1) * sy : n + sy;
  const long long ym = (y == 0) ? n + (g.ny - 1) * sy : n - sy;
  const long long zp = (z == g.nz - 1) ? n - (g.nz - 1) * sz : n + sz;
  const long long zm = (z == 0) ? n + (g.nz - 1) * sz : n - sz;
  const int x_hi = g.nx - 2;
  const int mx = fd_axis_mode(x, 1, x_hi, (x == 1) ? 0 : mask[xm], (x == x_hi && x != 1) ? 0 : mask[xp]);
  const int my = fd_axis_mode(y, 0, g.ny - 1, (y == 0) ? 0 : mask[ym], (y == g.ny - 1 && y != 0) ? 0 : mask[yp]);
  const int mz = fd_axis_mode(z, 0, g.nz - 1, (z == 0) ? 0 : mask[zm], (z == g.nz - 1 && z != 0) ? 0 : mask[zp]);
  for (int k = 0; k < nb; ++k) {
    const R* o = b_old + (long long)k * g.N;
    const double b0r = (double)o[n];
    double bxp = (mx == 2) ? b0r : (double)o[xp], bxn = (mx == 1) ? b0r : (double)o[xm];
    double byp = (my == 2) ? b0r : (double)o[yp], byn = (my == 1) ? b0r : (double)o[ym];
    double bzp = (mz == 2) ? b0r : (double)o[zp], bzn = (mz == 1) ? b0r : (double)o[zm];
    double b0 = b0r;
    if (b0 < kFdThrd) b0 = 0.0;
    if (bxp < kFdThrd) bxp = 0.0;
    if (bxn < kFdThrd) bxn = 0.0;
    if (byp < kFdThrd) byp = 0.0;
    if (byn < kFdThrd) byn = 0.0;
    if (bzp < kFdThrd) bzp = 0.0;
    if (bzn < kFdThrd) bzn = 0.0;
    const double nb_new = b0 + nu * ((bxp - 2 * b0 + bxn) + (byp - 2 * b0 + byn) + (bzp - 2 * b0 + bzn));
    if (nb_new > kFdThrd) b[(long long)k * g.N + n] = (R)nb_new;
  }
}

inline unsigned blocks_of(long long N) { return (unsigned)((N + 255) / 256); }
// fixed grid of the reducing kernels: a few blocks per CU, each thread over several voxels
inline unsigned reduce_blocks(long long N) {
  const long long b = (N + 255) / 256;
  return (unsigned)(b < 2048 ? b : 2048);
}

}  // namespace

void launch_ca_init(const uint8_t* mask, int* age, uint8_t* dpos, const Geom& g, int solid, int bb, const MaskRule& r,
                    hipStream_t s) {
  hipLaunchKernelGGL(ca_init_kernel, dim3(blocks_of(g.N)), dim3(256), 0, s, mask, age, dpos, g, solid, bb, r);
}

template <typename R>
void launch_ca_mask_total(const R* b, int nb, const uint8_t* prev, uint8_t* mask, double* btot, const Geom& g,
                          int solid, int bb, const MaskRule& r, double bthr, bool fresh, unsigned long long* red,
                          hipStream_t s) {
  hipLaunchKernelGGL((ca_mask_total_kernel<R>), dim3(reduce_blocks(g.N)), dim3(256), 0, s, b, nb, prev, mask, btot,
                     g, solid, bb, r, bthr, fresh, red);
}

void launch_ca_push(const uint8_t* mask, const double* btot, const int* age, double* out, uint8_t* flags,
                    const Geom& g, int solid, int bb, const MaskRule& r, double bmax, int half,
                    unsigned long long seed, unsigned long long sweep, hipStream_t s) {
  hipLaunchKernelGGL(ca_push_kernel, dim3(blocks_of(g.N)), dim3(256), 0, s, mask, btot, age, out, flags, g, solid, bb,
                     r, bmax, half, seed, sweep);
}

template <typename R>
void launch_ca_gather(const R* b_old, R* b, int nb, const uint8_t* mask, const double* btot, const double* out,
                      const uint8_t* flags, const Geom& g, int solid, int bb, hipStream_t s) {
  hipLaunchKernelGGL((ca_gather_kernel<R>), dim3(blocks_of(g.N)), dim3(256), 0, s, b_old, b, nb, mask, btot, out,
                     flags, g, solid, bb);
}

void launch_ca_age(const int* age_in, int* age_out, const uint8_t* dpos, const double* btot, const Geom& g,
                   double bmax, unsigned long long* changed, hipStream_t s) {
  hipLaunchKernelGGL(ca_age_kernel, dim3(reduce_blocks(g.N)), dim3(256), 0, s, age_in, age_out, dpos, btot, g, bmax,
                     changed);
}

template <typename R>
void launch_fd_biomass(const R* b_old, R* b, int nb, const uint8_t* mask, const Geom& g, double nu, hipStream_t s) {
  hipLaunchKernelGGL((fd_biomass_kernel<R>), dim3(blocks_of(g.N)), dim3(256), 0, s, b_old, b, nb, mask, g, nu);
}

template void launch_fd_biomass<float>(const float*, float*, int, const uint8_t*, const Geom&, double, hipStream_t);
template void launch_fd_biomass<double>(const double*, double*, int, const uint8_t*, const Geom&, double,
                                        hipStream_t);
template void launch_ca_mask_total<float>(const float*, int, const uint8_t*, uint8_t*, double*, const Geom&, int, int,
                                          const MaskRule&, double, bool, unsigned long long*, hipStream_t);
template void launch_ca_mask_total<double>(const double*, int, const uint8_t*, uint8_t*, double*, const Geom&, int,
                                           int, const MaskRule&, double, bool, unsigned long long*, hipStream_t);
template void launch_ca_gather<float>(const float*, float*, int, const uint8_t*, const double*, const double*,
                                      const uint8_t*, const Geom&, int, int, hipStream_t);
template void launch_ca_gather<double>(const double*, double*, int, const uint8_t*, const double*, const double*,
                                       const uint8_t*, const Geom&, int, int, hipStream_t);

}  // namespace clb
