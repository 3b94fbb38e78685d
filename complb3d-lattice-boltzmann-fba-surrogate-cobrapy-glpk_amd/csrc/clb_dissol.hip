// clb_dissol.hip -- mineral dissolution at the wetted surface of declared solid phases:
// surfaceDissolutionKinetics3D (src/dissolutionVOP.hh:229-459) with the compiled
// defineDissolutionRate hook of the abiotic kinetics plugin (defineAbioticKinetics.hh:292).
//
// A phase voxel (phase id > 0 with an inventory above 1e-12 in its immobile substrate)
// sees the water of its open face neighbours (pore or biofilm voxels, 1 <= x <= nx-2, no
// wrap in y or z), averaged; the hook returns the mineral's rate and the released
// substrates; one step removes at most half the inventory, and the products are scaled by
// the same factor and split evenly among the open neighbours.
//
// Two kernels, so that nothing reads what another thread writes:
//  1. dissol_products_kernel: every open voxel recomputes the reaction of each phase
//     neighbour (the same inputs give the same numbers in every thread, so no atomics)
//     and adds its share of the products to its increments: to the deferred increments
//     of the fused step, or straight into its post-collision populations;
//  2. dissol_mineral_kernel: every phase voxel removes its dissolved inventory from the
//     rest population of its substrate (a wall's slots of an immobile solute are touched
//     by no other kernel).
// The water is written by the transport kernel after the biotic kinetics, so the hook sees
// what the reference's abiotic block starts from; the products join the step's speciation
// increments, applied before the populations stream.
#include "clb_kinetics_device.h"
#include "clb_internal.h"
#include "clb_lattice.h"
#include "clb_spec_kernel.h"

namespace clb {
namespace {

constexpr double kDissThrd = 1e-12;         // thrd (src/complab3d_processors.hh:60)
constexpr double kMaxDissolFraction = 0.5;  // MAX_DISSOL_FRACTION (src/dissolutionVOP.hh:373)

struct Reaction {
  bool on = false;
  int n_open = 0;
  int open[6];        // open face directions, +x, -x, +y, -y, +z, -z order
  double d_mineral = 0.0;
  double released[CLB_MAX_SOLUTES];  // per substrate (0 for the mineral's own), before the split
};

template <typename R>
__device__ __forceinline__ long long face_nbr(const Geom& g, int x, int y, int z, int d) {
  const int dx = (d == 0) - (d == 1), dy = (d == 2) - (d == 3), dz = (d == 4) - (d == 5);
  const int xx = x + dx, yy = y + dy, zz = z + dz;
  if (xx < 1 || xx > g.nx - 2 || yy < 0 || yy >= g.ny || zz < 0 || zz >= g.nz) return -1;
  return xx + (long long)g.nx * (yy + (long long)g.ny * zz);
}

template <typename R>
__device__ __forceinline__ bool open_voxel(const AdeArgs<R>& a, long long q) {
  const uint8_t m = a.mask[q];
  return m != a.solid_code && m != a.bb_code && a.kind[m] == CLB_KIND_FLUID;
}

// The reaction of phase voxel n (surfaceDissolutionKinetics3D's loop body).
template <typename R, int S>
__device__ Reaction phase_reaction(const AdeArgs<R>& a, const DissolArgs& d, long long n) {
  Reaction r;
  const Geom& g = a.geo;
  const int x = (int)(n % g.nx), y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
  if (x < 1 || x > g.nx - 2) return r;
  const int pid = d.phase[n];
  if (pid <= 0 || pid >= CLB_MAX_PHASES) return r;
  const int sm = d.phase_sub[pid];
  if (sm < 0 || sm >= S) return r;
  const double mineral = (double)a.g[(long long)(sm * 7) * g.N + n] + (double)a.c_ref[sm];
  if (!(mineral > kDissThrd)) return r;
  for (int dd = 0; dd < 6; ++dd) {
    const long long q = face_nbr<R>(g, x, y, z, dd);
    if (q >= 0 && open_voxel(a, q)) r.open[r.n_open++] = dd;
  }
  if (r.n_open == 0 && d.surface_only) return r;
  KVec water(S, 0.0), sr(S, 0.0);
  for (int k = 0; k < r.n_open; ++k) {
    const long long q = face_nbr<R>(g, x, y, z, r.open[k]);
#pragma unroll
    for (int s = 0; s < S; ++s) water[s] += (double)a.water[(long long)s * g.N + q];
  }
  if (r.n_open > 0) {
#pragma unroll
    for (int s = 0; s < S; ++s) water[s] /= (double)r.n_open;
  }
#pragma unroll
  for (int s = 0; s < S; ++s)
    if (water[s] < 0.0) water[s] = 0.0;
  double mineral_rate = 0.0;
  defineDissolutionRate((plb::plint)pid, water, mineral, sr, mineral_rate, (plb::plint)a.mask[n]);
  if (!(mineral_rate < 0.0)) return r;
  double dm = mineral_rate * a.dt;
  double scale = 1.0;
  const double max_remove = kMaxDissolFraction * mineral;
  if (-dm > max_remove) scale = max_remove / (-dm);
  dm *= scale;
  if (!(-dm > kDissThrd)) return r;
  r.on = true;
  r.d_mineral = dm;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double rel = (s == sm) ? 0.0 : sr[s] * a.dt * scale;
    r.released[s] = (fabs(rel) <= kDissThrd) ? 0.0 : rel;
  }
  return r;
}

template <typename R, int S>
__global__ void dissol_products_kernel(AdeArgs<R> a, DissolArgs d, CoupledIO io) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const Geom& g = a.geo;
  if (n >= g.N) return;
  const int x = (int)(n % g.nx), y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
  if (x < 1 || x > g.nx - 2 || !open_voxel(a, n)) return;
  // any phase neighbour at all? (most voxels have none)
  long long nbr[6];
  bool any = false;
  for (int dd = 0; dd < 6; ++dd) {
    const long long q = face_nbr<R>(g, x, y, z, dd);
    nbr[dd] = (q >= 0 && d.phase[q] > 0) ? q : -1;
    any = any || nbr[dd] >= 0;
  }
  if (!any) return;
  double prod[S];
#pragma unroll
  for (int s = 0; s < S; ++s) prod[s] = 0.0;
  for (int dd = 0; dd < 6; ++dd) {
    if (nbr[dd] < 0) continue;
    const Reaction r = phase_reaction<R, S>(a, d, nbr[dd]);
    if (!r.on || r.n_open == 0) continue;
#pragma unroll
    for (int s = 0; s < S; ++s) prod[s] += r.released[s] / (double)r.n_open;
  }
#pragma unroll
  for (int s = 0; s < S; ++s)
    if (prod[s] != 0.0) {
      if (io.pend_w) {
        R* q = (R*)io.pend_w + (long long)s * g.N + n;  // the fused step's deferred increment
        *q = (R)((double)*q + prod[s]);
      } else {
        add_increment<R>(io, s, (int)n, prod[s]);
      }
    }
}

template <typename R, int S>
__global__ void dissol_mineral_kernel(AdeArgs<R> a, DissolArgs d) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= a.geo.N || d.phase[n] == 0) return;
  const Reaction r = phase_reaction<R, S>(a, d, n);
  if (!r.on) return;
  R* q = &a.g[(long long)(d.phase_sub[d.phase[n]] * 7) * a.geo.N + n];
  *q = (R)((double)*q + r.d_mineral);
}

template <typename R, int S>
void launch_dissol_shape(const AdeArgs<R>& a, const DissolArgs& d, const CoupledIO& io, hipStream_t s) {
  const dim3 grid((unsigned)((a.geo.N + 255) / 256));
  hipLaunchKernelGGL((dissol_products_kernel<R, S>), grid, dim3(256), 0, s, a, d, io);
  hipLaunchKernelGGL((dissol_mineral_kernel<R, S>), grid, dim3(256), 0, s, a, d);
}

}  // namespace

template <typename R>
void launch_dissolution(const AdeArgs<R>& a, const DissolArgs& d, bool odd_done, hipStream_t s) {
  CoupledIO io{};  // value-initialized: fields a launcher does not set are null / zero
  io.g = (void*)a.g;
  io.mask = a.mask;
  io.list = nullptr;
  io.cbuf = nullptr;
  io.geo = a.geo;
  io.odd_done = odd_done ? 1 : 0;
  for (int k = 0; k < CLB_MAX_SOLUTES; ++k) io.immobile[k] = a.immobile[k];
  io.pend_w = (void*)a.pend_w;
  switch (a.S) {
    case 1: launch_dissol_shape<R, 1>(a, d, io, s); break;
    case 2: launch_dissol_shape<R, 2>(a, d, io, s); break;
    case 3: launch_dissol_shape<R, 3>(a, d, io, s); break;
    case 4: launch_dissol_shape<R, 4>(a, d, io, s); break;
    case 5: launch_dissol_shape<R, 5>(a, d, io, s); break;
    case 6: launch_dissol_shape<R, 6>(a, d, io, s); break;
    case 7: launch_dissol_shape<R, 7>(a, d, io, s); break;
    default: launch_dissol_shape<R, 8>(a, d, io, s); break;
  }
}

template void launch_dissolution<float>(const AdeArgs<float>&, const DissolArgs&, bool, hipStream_t);
template void launch_dissolution<double>(const AdeArgs<double>&, const DissolArgs&, bool, hipStream_t);

}  // namespace clb
