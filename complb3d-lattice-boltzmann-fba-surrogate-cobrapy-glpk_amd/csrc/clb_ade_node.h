// clb_ade_node.h -- per-node D3Q7 update shared by the transport kernels (clb_ade.hip) and
// the transport + fused speciation kernels (clb_ade_spec_*.hip): AA-pattern load / store of
// every solute's populations, boundary completion, BGK collision, the compiled biotic and
// abiotic kinetics hooks, and the biomass threshold rule.
//
// Replaces, per step of the loop at src/complab.cpp:1152-1860:
//   * Palabos AdvectionDiffusionBGKdynamics collide (:1252) and stream (:1849) of every
//     mobile solute, with Dirichlet / Neumann x faces (soluteDomainSetup,
//     src/complab_functions.hh:296-340) and bounce-back / no-dynamics nodes; immobile
//     solutes are neither collided nor streamed (src/complab.cpp:248-266);
//   * dC = dC0 resets and run_kinetics + update_rxnLattices
//     (src/complab3d_processors_part1.hh:72-127, 345-415), which call the compiled
//     defineRxnKinetics hook per pore voxel;
//   * the abiotic block: run_abiotic_kinetics / surfaceAbioticKinetics3D +
//     update_abiotic_rxnLattices (src/complab3d_processors_part1.hh:158-336,
//     src/precipitationVOP.hh:53-145) with the compiled defineAbioticRxnKinetics hook;
//   * the biomass threshold reclassification of updateLocalMaskNtotalLattices3D
//     (src/complab3d_processors_part2.hh:99-210, one pore type).
// Populations are stored as g_i = f_i - t_i * c_ref[s]. An immobile solute keeps its whole
// concentration in the rest population, whose AA slot is the node's own in both step
// parities, so it costs one read and one write per node and never moves.
#pragma once

#include "clb_kinetics_device.h"
#include "clb_lattice.h"

namespace clb {

namespace {

// XCHK = false for a node with 1 <= x <= nx-2, whose x neighbours always exist: no checks,
// so the loads and stores need no branches.
template <typename R, int S, bool ODD, bool XCHK = true>
__device__ __forceinline__ void load7(const AdeArgs<R>& a, const Nbr& nb, int n, R* p) {
  const long long N = a.geo.N;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (a.immobile[s]) {
      p[s * 7] = a.g[(long long)(s * 7) * N + n];
#pragma unroll
      for (int i = 1; i < 7; ++i) p[s * 7 + i] = R(0);
      continue;
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      if (!ODD) {
        p[s * 7 + i] = a.g[(long long)(s * 7 + i) * N + n];
      } else {
        if (!XCHK || nb.xok(-Q7::cx[i]))
          p[s * 7 + i] = a.g[(long long)(s * 7 + opp_dir(i)) * N + nb.at(-Q7::cx[i], -Q7::cy[i], -Q7::cz[i])];
        else
          p[s * 7 + i] = R(0);
      }
    }
  }
}

template <typename R, int S, bool ODD, bool XCHK = true>
__device__ __forceinline__ void store7(const AdeArgs<R>& a, const Nbr& nb, int n, const R* p) {
  const long long N = a.geo.N;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (a.immobile[s]) {
      a.g[(long long)(s * 7) * N + n] = p[s * 7];
      continue;
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      if (!ODD) {
        a.g[(long long)(s * 7 + opp_dir(i)) * N + n] = p[s * 7 + i];
      } else {
        if (!XCHK || nb.xok(Q7::cx[i]))
          a.g[(long long)(s * 7 + i) * N + nb.at(Q7::cx[i], Q7::cy[i], Q7::cz[i])] = p[s * 7 + i];
      }
    }
  }
}

// Deferred speciation increments (fused coupled step): the voxel that solved writes its
// limited dC per solute once instead of adding dC t_i to the seven slots it had stored.
// Streaming is a permutation and the increment is additive, so the same state results when
// the loader adds to population i of node n the t_i dC its source received:
//  * the rest population and a population from a fluid neighbour n - c_i left its source in
//    the previous step: dC of the source from pend_r;
//  * a population from a wall is n's own, reflected: with bounce-back it left n two steps
//    ago, so it takes n's dC of that step, which is still in pend_w (the buffer this step
//    writes; n's entry is read here, before the solve phase rewrites it);
//  * a population entering through an x face is rebuilt by the boundary completion.
// Face nodes and walls never hold an increment. An immobile solute takes its whole dC in
// the rest population. Branch-free, like inject.
template <typename R, int S, bool XCHK = true>
__device__ __forceinline__ void add_pending(const AdeArgs<R>& a, const Nbr& nb, int n, R* p) {
  const long long N = a.geo.N;
  int src[7];
  bool wall[7];
  src[0] = n;
  wall[0] = false;
#pragma unroll
  for (int i = 1; i < 7; ++i) {
    wall[i] = false;
    if (!XCHK || nb.xok(-Q7::cx[i])) {
      const int m = nb.at(-Q7::cx[i], -Q7::cy[i], -Q7::cz[i]);
      wall[i] = a.kind[a.mask[m]] != CLB_KIND_FLUID;
      src[i] = wall[i] ? n : m;
    } else {
      src[i] = -1;
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const R* P1 = a.pend_r + (long long)s * N;  // previous step
    const R* P2 = a.pend_w + (long long)s * N;  // the step before
    const bool imm = a.immobile[s] != 0;
    const R own = P1[n];
    p[s * 7] += imm ? own : R(Q7::w(0)) * own;
#pragma unroll
    for (int i = 1; i < 7; ++i) {
      const R v = (src[i] < 0) ? R(0) : (wall[i] ? P2[n] : P1[src[i]]);
      p[s * 7 + i] += imm ? R(0) : R(Q7::w(i)) * v;
    }
  }
}

template <typename R>
__device__ __forceinline__ R pop_sum(const R* q) {
  R r = 0;
#pragma unroll
  for (int i = 0; i < 7; ++i) r += q[i];
  return r;
}

// add dC to solute s of the node (dC * t_i per population; all of it in the rest slot
// for an immobile solute). Increments with |dC| <= 1e-12 are not applied, as in
// update_rxnLattices (src/complab3d_processors_part1.hh:345-415). Written without branches:
// adding an exact zero leaves a population unchanged, and branch-free updates keep the
// population array in registers (conditional stores to different elements get merged into
// one store through a selected address, which forces the array into scratch memory).
template <typename R>
__device__ __forceinline__ void inject(const AdeArgs<R>& a, int s, R* p, double dC) {
  const double d = (dC > 1e-12 || dC < -1e-12) ? dC : 0.0;
  const bool imm = a.immobile[s] != 0;
  p[s * 7] += (R)(imm ? d : d * Q7::w(0));
#pragma unroll
  for (int i = 1; i < 7; ++i) p[s * 7 + i] += (R)(imm ? 0.0 : d * Q7::w(i));
}

// The compiled biotic hook for NB biomass pools (when set), plus the surrogate processor's
// increments (when on): the vectors the plugin receives have
// compile-time sizes, so its loops over .size() unroll and every element stays in a
// register (a run-time size would force the vectors, and with them the populations, into
// scratch memory).
template <typename R, int S, int NB>
__device__ __forceinline__ void biotic_kinetics(const AdeArgs<R>& a, int n, uint8_t code, R* p) {
  const Geom& g = a.geo;
  KVec C(S, 0.0), Bv(NB, 0.0), sr(S, 0.0), br(NB, 0.0);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double c0 = (double)pop_sum(&p[s * 7]) + (double)a.c_ref[s];
    if (c0 < 1e-12) c0 = 0;
    C[s] = c0;
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    double b0 = (double)a.bmass[(long long)b * g.N + n];
    if (b0 < 1e-12) b0 = 0;
    Bv[b] = b0;
  }
  if (a.kinetics) defineRxnKinetics(Bv, C, sr, br, (plb::plint)code);
  // the surrogate processor's increments join the kinetics' in the same dC / dB lattices,
  // which update_rxnLattices applies when their sum exceeds 1e-12
  // (src/complab3d_processors_surrogate.hh:262-283, src/complab.cpp:1294-1313); it left
  // them where a surrogate organism is present, which Bv tells (same field, same threshold)
  bool srg = false;
  if (a.srg_inc) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if ((a.srg_pools >> b) & 1) srg = srg || Bv[b] > 1e-12;
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double d = sr[s] * a.dt;
    if (!(d > 1e-12 || d < -1e-12)) d = 0.0;
    if (srg) d += (double)a.srg_inc[(long long)s * g.N + n];
    inject(a, s, p, d);
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    double dB = br[b] * a.dt;
    if (!(dB > 1e-12 || dB < -1e-12)) dB = 0.0;
    if (srg) dB += (double)a.srg_inc[(long long)(S + b) * g.N + n];
    if (dB > 1e-12 || dB < -1e-12) {
      R* bp = &a.bmass[(long long)b * g.N + n];
      *bp = (R)((double)*bp + dB);
    }
  }
}

// Dirichlet / Neumann completion of the population entering through an x face
// (MISS = 1: +x population at x = 0; MISS = 2: -x population at x = nx-1).
template <typename R, int S, int MISS>
__device__ __forceinline__ void complete_face(const AdeArgs<R>& a, R* p, const R* rho_bc) {
  constexpr int side = (MISS == 1) ? 0 : 1;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (a.immobile[s]) continue;
    const R cb = (a.bc_type[2 * s + side] == CLB_BC_DIRICHLET) ? a.bc_value[2 * s + side] : rho_bc[s];
    R others = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i)
      if (i != MISS) others += p[s * 7 + i];
    p[s * 7 + MISS] = (cb - a.c_ref[s]) - others;
  }
}

// Collide (with boundary completion), biotic and abiotic kinetics, optional
// post-kinetics concentration output and biomass threshold rule for one fluid node.
// FACE: 0 interior node (1 <= x <= nx-2), 1 the inlet node x = 0, 2 the outlet node
// x = nx-1 (known at compile time, so every population index is a constant). rho_bc[s] is
// the Neumann density (interior neighbour's pre-collision density) for face nodes. The
// advecting velocity is a.u[ui + k us], k = 0, 1, 2 (dense: ui = n, us = N; compact layout:
// the node's slot and the capacity).
template <typename R, int S, bool KIN, bool CBUF, int FACE = 0>
__device__ __forceinline__ void ade_node(const AdeArgs<R>& a, int x, int y, int z, int n, uint8_t code, R* p,
                                         const R* rho_bc, long long ui, long long us) {
  const Geom& g = a.geo;
  if (FACE == 1) complete_face<R, S, 1>(a, p, rho_bc);
  if (FACE == 2) complete_face<R, S, 2>(a, p, rho_bc);
  const R ux = a.u[ui], uy = a.u[us + ui], uz = a.u[2 * us + ui];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (a.immobile[s]) continue;
    const R rs = pop_sum(&p[s * 7]);
    const R rho = rs + a.c_ref[s];
    const R om = a.omega[s * 256 + code];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const R cu = R(Q7::cx[i]) * ux + R(Q7::cy[i]) * uy + R(Q7::cz[i]) * uz;
      const R eq = R(Q7::w(i)) * rs + R(4.0 * Q7::w(i)) * rho * cu;
      p[s * 7 + i] += om * (eq - p[s * 7 + i]);
    }
  }
  const bool react = (FACE == 0 && code != a.solid_code && code != a.bb_code);
  if (KIN && react && (a.kinetics || a.srg_inc)) {
    switch (a.nb) {
      case 0: biotic_kinetics<R, S, 0>(a, n, code, p); break;
      case 1: biotic_kinetics<R, S, 1>(a, n, code, p); break;
      case 2: biotic_kinetics<R, S, 2>(a, n, code, p); break;
      case 3: biotic_kinetics<R, S, 3>(a, n, code, p); break;
      default: biotic_kinetics<R, S, CLB_MAX_BIOMASS>(a, n, code, p); break;
    }
  }
  if (react && a.water) {
    // the water a neighbouring mineral sees: after the biotic kinetics, before the abiotic
    // block (surfaceDissolutionKinetics3D shares the abiotic block's starting state)
#pragma unroll
    for (int s = 0; s < S; ++s) a.water[(long long)s * g.N + n] = (R)(pop_sum(&p[s * 7]) + a.c_ref[s]);
  }
  if (KIN && react && a.abiotic) {
    bool gate = true;
    if (a.abiotic == 2) {
      // surfaceAbioticKinetics3D: only pore voxels with a wall among the 6 face neighbours
      const Nbr nb(x, y, z, g);
      gate = false;
#pragma unroll
      for (int i = 1; i < 7; ++i) {
        const uint8_t m = a.mask[nb.at(Q7::cx[i], Q7::cy[i], Q7::cz[i])];
        if (m == a.solid_code || m == a.bb_code) gate = true;
      }
    }
    if (gate) {
      KVec C(S, 0.0), sr(S, 0.0);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        double c0 = (double)pop_sum(&p[s * 7]) + (double)a.c_ref[s];
        if (c0 < 1e-12) c0 = 0;
        C[s] = c0;
      }
      defineAbioticRxnKinetics(C, sr, (plb::plint)code);
#pragma unroll
      for (int s = 0; s < S; ++s) inject(a, s, p, sr[s] * a.dt);
    }
  }
  if (react && a.mask_update && a.nb > 0) {
    // updateLocalMaskNtotalLattices3D with one pore type (several: mask_rule_kernel after
    // the step): the fields above 1e-14 make the total and the biofilm label
    double btot = 0;
    int label = 0;
#pragma unroll
    for (int b = 0; b < CLB_MAX_BIOMASS; ++b)
      if (b < a.nb) {
        const double bv = (double)a.bmass[(long long)b * g.N + n];
        if (bv > 1e-14) {
          btot += bv;
          label += a.mrule.bio[b];
        }
      }
    const bool pore = mr_is_pore(a.mrule, code);
    if (!pore && btot < a.bthreshold) {
      a.mask[n] = (uint8_t)a.mrule.pore[0];
    } else if (pore && btot >= a.bthreshold) {
      const int lab = a.mrule.sum_labels ? label : a.mrule.bio[0];
      if (lab > 0) a.mask[n] = (uint8_t)lab;
    }
  }
  if (CBUF && react) {
#pragma unroll
    for (int s = 0; s < S; ++s) a.cbuf[(long long)s * g.N + n] = (double)pop_sum(&p[s * 7]) + (double)a.c_ref[s];
  }
}

}  // namespace
}  // namespace clb
