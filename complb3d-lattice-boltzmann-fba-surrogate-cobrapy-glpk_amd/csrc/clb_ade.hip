// clb_ade.hip -- D3Q7 advection-diffusion for every solute, fused with the per-voxel
// biotic and abiotic kinetics, in one pass over the populations (the per-node update is in
// clb_ade_node.h), plus the initialisation, concentration, node-conversion and reset
// kernels of the transport state.
//
// HBM traffic per fluid node and step: 7 populations read + written per mobile solute
// (1 per immobile one), 3 velocity components, 1 mask byte (+ biomass read/write and the
// post-kinetics concentrations when the separate speciation kernel follows).
#include "clb_ade_node.h"
#include "clb_mask_rule.h"

namespace clb {

namespace {

// One thread per (x, y, z) with 1 <= x <= nx-2. The thread of x = 1 also updates the
// inlet node x = 0 and the thread of x = nx-2 the outlet node x = nx-1, after it has taken
// the interior node's pre-collision density, which the Neumann face needs (the density
// FlatAdiabaticBoundaryFunctional3D copies after the previous stream).
template <typename R, int S, bool ODD, bool KIN, bool CBUF>
__global__ void __launch_bounds__(256) ade_step_kernel(AdeArgs<R> a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  const Geom& g = a.geo;
  if (x < 1 || x > g.nx - 2 || y >= g.ny) return;
  const int n = x + g.nx * (y + g.ny * z);
  const uint8_t code = a.mask[n];
  const bool fluid = a.kind[code] == CLB_KIND_FLUID;
  R rho_in[S];
  if (fluid) {
    const Nbr nb(x, y, z, g);
    R p[S * 7];
    load7<R, S, ODD, false>(a, nb, n, p);
#pragma unroll
    for (int s = 0; s < S; ++s) rho_in[s] = pop_sum(&p[s * 7]) + a.c_ref[s];
    ade_node<R, S, KIN, CBUF>(a, x, y, z, n, code, p, nullptr, n, g.N);
    store7<R, S, ODD, false>(a, nb, n, p);
  } else if (a.stream_none && a.kind[code] == CLB_KIND_NONE) {
    // NoDynamics: no collision; the AA pattern's load and store alone are the streaming
    const Nbr nb(x, y, z, g);
    R p[S * 7];
    load7<R, S, ODD, false>(a, nb, n, p);
    store7<R, S, ODD, false>(a, nb, n, p);
  }
  if (x == 1 || x == g.nx - 2) {
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
      if (xb == 0) ade_node<R, S, false, false, 1>(a, xb, y, z, nbn, cb, q, rho_in, nbn, g.N);
      else ade_node<R, S, false, false, 2>(a, xb, y, z, nbn, cb, q, rho_in, nbn, g.N);
      store7<R, S, ODD>(a, nb, nbn, q);
    }
  }
}

template <typename R>
__global__ void ade_init_kernel(R* g, Geom geo, int S, const double* c0, const R* c_ref, const int* immobile) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= geo.N) return;
  for (int s = 0; s < S; ++s)
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const double v = immobile[s] ? ((i == 0) ? c0[s] - (double)c_ref[s] : 0.0)
                                   : Q7::w(i) * (c0[s] - (double)c_ref[s]);
      g[(long long)(s * 7 + i) * geo.N + n] = (R)v;
    }
}

template <typename R, bool ODD>
__device__ double node_density(const AdeArgs<R>& a, int s, int x, int y, int z) {
  const Geom& g = a.geo;
  const int n = x + g.nx * (y + g.ny * z);
  if (a.immobile[s]) return (double)a.g[(long long)(s * 7) * g.N + n] + (double)a.c_ref[s];
  const Nbr nb(x, y, z, g);
  double r = 0;
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    R v;
    if (!ODD) v = a.g[(long long)(s * 7 + i) * g.N + n];
    else v = nb.xok(-Q7::cx[i]) ? a.g[(long long)(s * 7 + opp_dir(i)) * g.N + nb.at(-Q7::cx[i], -Q7::cy[i], -Q7::cz[i])]
                                : R(0);
    r += (double)v;
  }
  return r + (double)a.c_ref[s];
}

// Concentration of every node (sum of the pre-collision populations of the next step).
// Fluid nodes of the x faces report what their boundary prescribes: the Dirichlet value,
// or for a Neumann face the density of the interior neighbour (its own when that is a
// wall), since their entering population is only rebuilt at the next collision.
// Wall nodes (bounce-back / no-dynamics) report 0 for mobile solutes: their slots hold the
// neighbours' reflected populations (Palabos' BounceBack::computeDensity returns a fixed
// fictitious density, 0 by default, not a population sum). Immobile solutes report their
// amount everywhere, so a precipitated voxel still shows its mineral.
template <typename R, bool ODD>
__global__ void ade_conc_kernel(AdeArgs<R> a, double* c) {
  const long long n64 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const Geom& g = a.geo;
  if (n64 >= g.N) return;
  const int n = (int)n64;
  const int x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
  const bool fluid = a.kind[a.mask[n]] == CLB_KIND_FLUID;
  const bool face = (x == 0 || x == g.nx - 1) && fluid;
  for (int s = 0; s < a.S; ++s) {
    double r;
    if (!fluid && !a.immobile[s]) {
      r = 0.0;
    } else if (!face || a.immobile[s]) {
      r = node_density<R, ODD>(a, s, x, y, z);
    } else {
      const int side = (x == 0) ? 0 : 1;
      if (a.bc_type[2 * s + side] == CLB_BC_DIRICHLET) {
        r = (double)a.bc_value[2 * s + side];
      } else {
        const int xi = (x == 0) ? 1 : g.nx - 2;
        const bool inner_fluid = a.kind[a.mask[xi + g.nx * (y + g.ny * z)]] == CLB_KIND_FLUID;
        r = node_density<R, ODD>(a, s, inner_fluid ? xi : x, y, z);
      }
    }
    c[(long long)s * g.N + n] = r;
  }
}

// Write equilibrium-at-rest pre-collision populations (concentration c0[s]) for the next
// step into the slots node n owns for that step's parity; next_odd selects the layout.
template <typename R>
__device__ __forceinline__ void write_rest_state(const AdeArgs<R>& a, bool next_odd, int x, int y, int z, int n,
                                                 const double* c0) {
  const Geom& g = a.geo;
  const Nbr nb(x, y, z, g);
  for (int s = 0; s < a.S; ++s) {
    const double v = (c0 ? c0[s] : 0.0) - (double)a.c_ref[s];
    if (a.immobile[s]) continue;  // an immobile solute keeps its amount
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const R val = (R)(Q7::w(i) * v);
      if (!next_odd) a.g[(long long)(s * 7 + i) * g.N + n] = val;
      else if (nb.xok(-Q7::cx[i]))
        a.g[(long long)(s * 7 + opp_dir(i)) * g.N + nb.at(-Q7::cx[i], -Q7::cy[i], -Q7::cz[i])] = val;
    }
  }
}

// Precipitation / dissolution node conversion (precipNodeConversion3D,
// src/precipitationVOP.hh:176-204; dissolNodeConversion3D, src/dissolutionVOP.hh:469-511).
// The mineral is the concentration of solute `ms` (normally immobile). A voxel that is
// neither solid nor bounce-back and whose mineral reaches precip_max becomes precip_code,
// stamped with phase_id; any other voxel stamped with phase_id whose mineral falls below
// reopen becomes pore_code with phase 0, whatever its mask was. changed[0] counts the
// voxels whose mask code changed (a flow re-solve trigger), changed[1] those of them that
// were walls for the flow and are fluid now; a wall voxel that reopens restarts at rest. The flow's slots are stored relative to the reference profile
// 1 + fref0 + x * frdx of their own voxel (clb_flow.hip), so a population at rest at
// density 1 in a slot of voxel x is -t_i (fref0 + x * frdx).
template <typename R, bool ODD>
__global__ void node_conversion_kernel(AdeArgs<R> a, uint8_t* phase, int ms, int pore_code, int precip_code,
                                       int phase_id, double precip_max, double reopen, unsigned int* changed,
                                       R* flowf, bool flow_next_odd, double fref0, double frdx) {
  const long long n64 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const Geom& g = a.geo;
  bool flip = false, flow_open = false;
  if (n64 < g.N) {
    const int n = (int)n64;
    const int x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
    const uint8_t m = a.mask[n];
    const double mineral = node_density<R, ODD>(a, ms, x, y, z);
    bool precip = false;
    if (m != a.solid_code && m != a.bb_code && precip_max > 0.0 && mineral >= precip_max) {
      a.mask[n] = (uint8_t)precip_code;
      if (phase_id > 0) phase[n] = (uint8_t)phase_id;
      precip = true;
      flip = (int)m != precip_code;
    }
    if (!precip && reopen > 0.0 && phase_id > 0 && phase[n] == phase_id && mineral < reopen) {
      a.mask[n] = (uint8_t)pore_code;
      phase[n] = 0;
      flip = (int)m != pore_code;
    }
    if (flip && a.mask[n] == (uint8_t)pore_code) {
      // a reopened wall voxel starts at rest: flow at the reference density, mobile
      // solutes at zero (its wall slots held only bounced populations); each lattice
      // judges "wall" by its own dynamics table
      if (a.kind[m] != CLB_KIND_FLUID) write_rest_state(a, ODD, x, y, z, n, nullptr);
      if (flowf && a.flow_kind[m] != CLB_KIND_FLUID) {
        flow_open = true;
        const Nbr nb(x, y, z, g);
#pragma unroll
        for (int i = 0; i < 19; ++i) {
          if (!flow_next_odd) {
            flowf[(long long)i * g.N + n] = (R)(-Q19::w(i) * (fref0 + frdx * (double)x));
          } else if (nb.xok(-Q19::cx[i])) {
            const int xs = x - Q19::cx[i];
            flowf[(long long)opp_dir(i) * g.N + nb.at(-Q19::cx[i], -Q19::cy[i], -Q19::cz[i])] =
                (R)(-Q19::w(i) * (fref0 + frdx * (double)xs));
          }
        }
      }
    }
  }
  const unsigned long long bal = __ballot(flip);
  if ((threadIdx.x & 63) == 0 && bal) atomicAdd(changed, (unsigned int)__popcll(bal));
  const unsigned long long fo = __ballot(flow_open);
  if ((threadIdx.x & 63) == 0 && fo) atomicAdd(changed + 1, (unsigned int)__popcll(fo));
}

// updateLocalMaskNtotalLattices3D with several pore types, after the step: every solute-fluid
// voxel of the planes 1 <= x <= nx-2 that is neither solid nor bounce-back decides from prev,
// a copy of the mask before the pass, as the transport kernel's own rule (one pore type)
// decides for its voxel.
template <typename R>
__global__ void __launch_bounds__(256) mask_rule_kernel(const R* b, int nb, const uint8_t* prev, uint8_t* mask,
                                                        const uint8_t* kind, Geom g, int solid, int bb, MaskRule r,
                                                        double thr) {
  const long long n = blockIdx.x * 256LL + threadIdx.x;
  if (n >= g.N) return;
  const int x = (int)(n % g.nx);
  if (x < 1 || x > g.nx - 2) return;
  const int m = prev[n];
  if (m == solid || m == bb || kind[m] != CLB_KIND_FLUID) return;
  double btot = 0.0;
  int label = 0;
  for (int i = 0; i < nb; ++i) {
    const double v = (double)b[(long long)i * g.N + n];
    if (v > 1e-14) {
      btot += v;
      label += r.bio[i];
    }
  }
  const int y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long long)g.nx * g.ny));
  const int nm = mask_rule_code(r, prev, g, x, y, z, m, btot, label, thr, solid, bb);
  if (nm != m) mask[n] = (uint8_t)nm;
}

// Fold the deferred speciation increments into the stored populations (before anything
// other than the next step reads them). For a next step of parity ODD, every fluid node
// adds to the slots load7 will read the contributions add_pending would add (pend_r: the
// previous step's increments, pend_w: the step before), and to each population it has sent
// towards a wall (read back one step later, so add_pending never sees it) t_j dC of its own
// previous-step increment. Each slot belongs to exactly one node, so there are no races; the
// caller clears both buffers afterwards.
template <typename R, bool ODD>
__global__ void pending_flush_kernel(AdeArgs<R> a) {
  const long long n64 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const Geom& g = a.geo;
  if (n64 >= g.N) return;
  const int n = (int)n64;
  if (a.kind[a.mask[n]] != CLB_KIND_FLUID) return;
  const int x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
  const Nbr nb(x, y, z, g);
  for (int s = 0; s < a.S; ++s) {
    const R* P1 = a.pend_r + (long long)s * g.N;
    const R* P2 = a.pend_w + (long long)s * g.N;
    if (a.immobile[s]) {
      a.g[(long long)(s * 7) * g.N + n] += P1[n];
      continue;
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      // the population i that node n reads in the next step
      if (nb.xok(-Q7::cx[i])) {
        R v = P1[n];
        if (i > 0) {
          const int m = nb.at(-Q7::cx[i], -Q7::cy[i], -Q7::cz[i]);
          v = (a.kind[a.mask[m]] == CLB_KIND_FLUID) ? P1[m] : P2[n];
        }
        R* q = ODD ? &a.g[(long long)(s * 7 + opp_dir(i)) * g.N + nb.at(-Q7::cx[i], -Q7::cy[i], -Q7::cz[i])]
                   : &a.g[(long long)(s * 7 + i) * g.N + n];
        *q += R(Q7::w(i)) * v;
      }
      // the population i that node n sent towards a wall in the previous step
      if (i > 0 && nb.xok(Q7::cx[i])) {
        const int w = nb.at(Q7::cx[i], Q7::cy[i], Q7::cz[i]);
        if (a.kind[a.mask[w]] != CLB_KIND_FLUID) {
          R* q = ODD ? &a.g[(long long)(s * 7 + opp_dir(i)) * g.N + n] : &a.g[(long long)(s * 7 + i) * g.N + w];
          *q += R(Q7::w(i)) * P1[n];
        }
      }
    }
  }
}

// stabilizeADElattice3D (src/complab3d_processors_part3.hh:261-310): every fluid (pore /
// biofilm) node back to the equilibrium of c0 at rest; wall nodes keep their populations.
template <typename R>
__global__ void ade_reset_kernel(AdeArgs<R> a, bool next_odd, const double* c0) {
  const long long n64 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const Geom& g = a.geo;
  if (n64 >= g.N) return;
  const int n = (int)n64;
  if (a.kind[a.mask[n]] != CLB_KIND_FLUID) return;
  const int x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
  write_rest_state(a, next_odd, x, y, z, n, c0);
}

// One solute's field at rest: an immobile solute's amount in every voxel's rest slot; a
// mobile solute's at-rest equilibrium of c[n] in every fluid voxel's populations, in the
// layout the next step (parity next_odd) reads.
template <typename R>
__global__ void solute_field_kernel(AdeArgs<R> a, bool next_odd, int s, const double* c) {
  const long long n64 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const Geom& g = a.geo;
  if (n64 >= g.N) return;
  const int n = (int)n64;
  const double v = c[n] - (double)a.c_ref[s];
  if (a.immobile[s]) {
    a.g[(long long)(s * 7) * g.N + n] = (R)v;
    return;
  }
  if (a.kind[a.mask[n]] != CLB_KIND_FLUID) return;
  const int x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
  const Nbr nb(x, y, z, g);
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    const R val = (R)(Q7::w(i) * v);
    if (!next_odd) a.g[(long long)(s * 7 + i) * g.N + n] = val;
    else if (nb.xok(-Q7::cx[i]))
      a.g[(long long)(s * 7 + opp_dir(i)) * g.N + nb.at(-Q7::cx[i], -Q7::cy[i], -Q7::cz[i])] = val;
  }
}

template <typename R, int S>
void launch_ade_step_s(const AdeArgs<R>& a, bool odd, dim3 block, hipStream_t st) {
  dim3 grid((a.geo.nx + block.x - 1) / block.x, (a.geo.ny + block.y - 1) / block.y, a.geo.nz);
  const bool kin = (a.kinetics != 0) || (a.abiotic != 0) || (a.srg_inc != nullptr);
  const bool cb = a.cbuf != nullptr;
#define CLB_ADE_LAUNCH(O, K, C) hipLaunchKernelGGL((ade_step_kernel<R, S, O, K, C>), grid, block, 0, st, a)
  if (odd) {
    if (kin) { if (cb) CLB_ADE_LAUNCH(true, true, true); else CLB_ADE_LAUNCH(true, true, false); }
    else { if (cb) CLB_ADE_LAUNCH(true, false, true); else CLB_ADE_LAUNCH(true, false, false); }
  } else {
    if (kin) { if (cb) CLB_ADE_LAUNCH(false, true, true); else CLB_ADE_LAUNCH(false, true, false); }
    else { if (cb) CLB_ADE_LAUNCH(false, false, true); else CLB_ADE_LAUNCH(false, false, false); }
  }
#undef CLB_ADE_LAUNCH
}

}  // namespace

template <typename R>
void launch_ade_init(const Geom& g, int S, R* gpop, const double* c0, const R* c_ref, const int* immobile,
                     hipStream_t s) {
  const int bs = 256;
  hipLaunchKernelGGL(ade_init_kernel<R>, dim3((unsigned)((g.N + bs - 1) / bs)), dim3(bs), 0, s, gpop, g, S, c0,
                     c_ref, immobile);
}

template <typename R>
void launch_ade_step(const AdeArgs<R>& a, bool odd, dim3 block, hipStream_t s) {
  switch (a.S) {
    case 1: launch_ade_step_s<R, 1>(a, odd, block, s); break;
    case 2: launch_ade_step_s<R, 2>(a, odd, block, s); break;
    case 3: launch_ade_step_s<R, 3>(a, odd, block, s); break;
    case 4: launch_ade_step_s<R, 4>(a, odd, block, s); break;
    case 5: launch_ade_step_s<R, 5>(a, odd, block, s); break;
    case 6: launch_ade_step_s<R, 6>(a, odd, block, s); break;
    case 7: launch_ade_step_s<R, 7>(a, odd, block, s); break;
    case 8: launch_ade_step_s<R, 8>(a, odd, block, s); break;
    default: break;
  }
}

template <typename R>
void launch_ade_conc(const AdeArgs<R>& a, bool odd_next, double* c, hipStream_t s) {
  const int bs = 256;
  dim3 grid((unsigned)((a.geo.N + bs - 1) / bs));
  if (odd_next) hipLaunchKernelGGL((ade_conc_kernel<R, true>), grid, dim3(bs), 0, s, a, c);
  else hipLaunchKernelGGL((ade_conc_kernel<R, false>), grid, dim3(bs), 0, s, a, c);
}

template <typename R>
void launch_node_conversion(const AdeArgs<R>& a, bool odd_next, uint8_t* phase, int ms, int pore_code,
                            int precip_code, int phase_id, double precip_max, double reopen, unsigned int* changed,
                            R* flowf, bool flow_next_odd, double fref0, double frdx, hipStream_t s) {
  const int bs = 256;
  dim3 grid((unsigned)((a.geo.N + bs - 1) / bs));
  if (odd_next)
    hipLaunchKernelGGL((node_conversion_kernel<R, true>), grid, dim3(bs), 0, s, a, phase, ms, pore_code, precip_code,
                       phase_id, precip_max, reopen, changed, flowf, flow_next_odd, fref0, frdx);
  else
    hipLaunchKernelGGL((node_conversion_kernel<R, false>), grid, dim3(bs), 0, s, a, phase, ms, pore_code,
                       precip_code, phase_id, precip_max, reopen, changed, flowf, flow_next_odd, fref0, frdx);
}

template <typename R>
void launch_mask_rule(const R* b, int nb, const uint8_t* prev, uint8_t* mask, const uint8_t* kind, const Geom& g,
                      int solid, int bb, const MaskRule& r, double thr, hipStream_t s) {
  hipLaunchKernelGGL((mask_rule_kernel<R>), dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, b, nb, prev, mask,
                     kind, g, solid, bb, r, thr);
}

template <typename R>
void launch_ade_reset(const AdeArgs<R>& a, bool next_odd, const double* c0, hipStream_t s) {
  const int bs = 256;
  hipLaunchKernelGGL(ade_reset_kernel<R>, dim3((unsigned)((a.geo.N + bs - 1) / bs)), dim3(bs), 0, s, a, next_odd, c0);
}

template <typename R>
void launch_solute_field(const AdeArgs<R>& a, bool next_odd, int solute, const double* c, hipStream_t s) {
  hipLaunchKernelGGL(solute_field_kernel<R>, dim3((unsigned)((a.geo.N + 255) / 256)), dim3(256), 0, s, a, next_odd,
                     solute, c);
}

template <typename R>
void launch_pending_flush(const AdeArgs<R>& a, bool odd_next, hipStream_t s) {
  const int bs = 256;
  const dim3 grid((unsigned)((a.geo.N + bs - 1) / bs));
  if (odd_next) hipLaunchKernelGGL((pending_flush_kernel<R, true>), grid, dim3(bs), 0, s, a);
  else hipLaunchKernelGGL((pending_flush_kernel<R, false>), grid, dim3(bs), 0, s, a);
}

template void launch_pending_flush<float>(const AdeArgs<float>&, bool, hipStream_t);
template void launch_pending_flush<double>(const AdeArgs<double>&, bool, hipStream_t);
template void launch_ade_reset<float>(const AdeArgs<float>&, bool, const double*, hipStream_t);
template void launch_ade_reset<double>(const AdeArgs<double>&, bool, const double*, hipStream_t);
template void launch_solute_field<float>(const AdeArgs<float>&, bool, int, const double*, hipStream_t);
template void launch_solute_field<double>(const AdeArgs<double>&, bool, int, const double*, hipStream_t);
template void launch_ade_init<float>(const Geom&, int, float*, const double*, const float*, const int*, hipStream_t);
template void launch_ade_init<double>(const Geom&, int, double*, const double*, const double*, const int*,
                                      hipStream_t);
template void launch_ade_step<float>(const AdeArgs<float>&, bool, dim3, hipStream_t);
template void launch_ade_step<double>(const AdeArgs<double>&, bool, dim3, hipStream_t);
template void launch_ade_conc<float>(const AdeArgs<float>&, bool, double*, hipStream_t);
template void launch_ade_conc<double>(const AdeArgs<double>&, bool, double*, hipStream_t);
template void launch_mask_rule<float>(const float*, int, const uint8_t*, uint8_t*, const uint8_t*, const Geom&, int,
                                      int, const MaskRule&, double, hipStream_t);
template void launch_mask_rule<double>(const double*, int, const uint8_t*, uint8_t*, const uint8_t*, const Geom&, int,
                                       int, const MaskRule&, double, hipStream_t);
template void launch_node_conversion<float>(const AdeArgs<float>&, bool, uint8_t*, int, int, int, int, double,
                                            double, unsigned int*, float*, bool, double, double, hipStream_t);
template void launch_node_conversion<double>(const AdeArgs<double>&, bool, uint8_t*, int, int, int, int, double,
                                             double, unsigned int*, double*, bool, double, double, hipStream_t);

}  // namespace clb
