// clb_surrogate.hip -- the surrogate metabolic model: run_surrogate3D
// (src/complab3d_processors_surrogate.hh:74-295), with a MATLAB-exported feed-forward network
// per microbe in place of the compiled defineSurrogateModel (surrogateModel.hh:252-500).
//
// The reference runs the processor after the collision, on the same concentrations and
// biomass as run_kinetics, and adds its increments to the same dC / dB lattices, which
// update_rxnLattices then applies once (src/complab.cpp:1264-1313). The BGK collision keeps
// every solute's density, so those concentrations are the sums of the populations a node
// loads at the start of the step. This kernel therefore runs before the transport kernel:
// it sums each reactive voxel's incoming populations (the AA slots of the step's parity,
// plus the deferred speciation increments of the fused step), evaluates the processor and
// leaves dC (S planes) and dB (nb planes) in `inc`; the transport kernel adds them to the
// kinetics increments before the 1e-12 test, as update_rxnLattices applies the sum of both.
//
// Per reactive voxel (1 <= x <= nx-2, fluid, neither solid nor bounce-back code): the
// surrogate microbes holding more than 1e-12 are present. Each one's Michaelis-Menten
// uptake estimate Fin (positive = consumed; vmax-limited, or the supply-limited exhaustion
// flux, capped by the supply, by the <substrate_lower_bounds> floor and, with
// fix_lower_bounds, by the supply again) goes through its network for the growth rate mu
// (1/h); Fout = Fin unless a closure ties consumption to growth. dB = mu B dt / 3600, or
// first-order decay -kd B dt when |mu| < 1e-12, never below -B. Each substrate's draw
// Fout B MW dt / 3.6e6 is summed over the present microbes and clamped once to what is left
// (component totals and a per-species split with use_totals). All arithmetic is FP64.
#include "clb_ade_node.h"

namespace clb {
namespace {

constexpr double kThrd = 1e-12;  // thrd (src/complab3d_processors.hh:60)

// exhaustionFlux (src/complab3d_metabolic.hh:504-520): the uptake (negative, mmol/gDW/h)
// that would empty C in one step, capped at -1e4.
__device__ __forceinline__ double exhaustion_flux(double C, double B, double mw, double dt) {
  const double denom = B * mw * dt;
  if (!(denom > 1e-30)) return 0.0;
  double v = -C * 3.6e6 / denom;
  if (v < -1e4) v = -1e4;
  return v;
}

__device__ __forceinline__ double tansig(double n) { return 2.0 / (1.0 + exp(-2.0 * n)) - 1.0; }

// One microbe's network (the evaluation of surrogateModel.hh:396-427): inputs mapped onto
// [-1, 1] (mapminmax), tansig hidden layers, one linear output neuron, the reverse map; a
// growth rate below 1e-8 or not finite is 0. Each neuron sums W x from 0 in input order and
// then adds its bias, as matMultp / matSum do. The weights are read at uniform addresses
// (scalar loads); rows and 4-column blocks beyond a layer's width are skipped by uniform
// branches, so a layer costs its own width, and the zero padding adds exact zeros.
__device__ __forceinline__ double net_growth(const SrgNet& nt, const double* fin) {
  constexpr int W = CLB_SRG_MAX_WIDTH;
  double a[W];
#pragma unroll
  for (int j = 0; j < W; ++j) a[j] = 0.0;
#pragma unroll
  for (int j = 0; j < CLB_MAX_SOLUTES; ++j) {
    if (j < nt.n_in) {
      const int sub = nt.in_sub[j];
      double x = 0.0;
#pragma unroll
      for (int s = 0; s < CLB_MAX_SOLUTES; ++s)
        if (sub == s) x = fin[s];
      a[j] = (x - nt.x_offset[j]) * nt.x_gain[j] + nt.x_ymin;
    }
  }
  int wprev = nt.n_in;
  for (int l = 0; l < nt.n_layers; ++l) {
    const int wl = nt.width[l];
    const bool last = (l == nt.n_layers - 1);
    double z[W];
#pragma unroll
    for (int r = 0; r < W; ++r) {
      double acc = 0.0;
      if (r < wl) {
#pragma unroll
        for (int cb = 0; cb < W; cb += 4) {
          if (cb < wprev) {
#pragma unroll
            for (int c = cb; c < cb + 4; ++c) acc += nt.W[l][r][c] * a[c];
          }
        }
        acc = nt.b[l][r] + acc;
        if (!last) acc = tansig(acc);
      }
      z[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < W; ++r) a[r] = z[r];
    wprev = wl;
  }
  double g = (a[0] - nt.y_ymin) / nt.y_gain + nt.y_offset;
  if (nt.log_output) g = pow(10.0, g);
  if (g < 1e-8) g = 0.0;
  if (__builtin_isnan(g) || __builtin_isinf(g)) g = 0.0;
  return g;
}

// MetabolicTotals (src/complab3d_metabolic.hh:222-382): the signed component total of
// substrate s over the species carrying it, and whether it is usable.
template <int S>
__device__ __forceinline__ bool totals_usable(const SrgTables& t, const double* conc, int s, double& tot) {
  tot = 0.0;
  double scale = 0.0;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const double cf = t.contrib[s][i];
    tot += cf * conc[i];
    scale += fabs(cf * conc[i]);
  }
  if (!t.totals_active || !t.has_contrib[s]) return false;
  return fabs(tot) > 1e-6 * scale && fabs(tot) > 1e-30;
}

// drawDown: remove `amount` of the element substrate s carries, from each carrier in
// proportion to the element it holds, every species clamped against what is left of it.
template <int S>
__device__ __forceinline__ void draw_down(const SrgTables& t, const double* conc, int s, double amount, double* dC) {
  double tot;
  if (!totals_usable<S>(t, conc, s, tot)) {
    double take = amount;
    const double remaining = conc[s] + dC[s];
    if (take > remaining) take = (remaining > 0.0) ? remaining : 0.0;
    dC[s] -= take;
    return;
  }
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const double cf = t.contrib[s][i];
    if (cf == 0.0 || !(conc[i] > 0.0)) continue;
    const double share = (cf * conc[i]) / tot;
    double take = amount * share / cf;
    const double remaining = conc[i] + dC[i];
    if (take > remaining) take = (remaining > 0.0) ? remaining : 0.0;
    dC[i] -= take;
  }
}

template <typename R, int S, bool ODD>
__global__ void __launch_bounds__(256) srg_kernel(AdeArgs<R> a, const SrgTables* __restrict__ tp, R* inc) {
  const long long n64 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const Geom& g = a.geo;
  if (n64 >= g.N) return;
  const int n = (int)n64;
  const int x = n % g.nx;
  if (x < 1 || x > g.nx - 2) return;
  const uint8_t code = a.mask[n];
  if (a.kind[code] != CLB_KIND_FLUID || code == a.solid_code || code == a.bb_code) return;
  const SrgTables& t = *tp;
  const long long N = g.N;

  double bm[CLB_MAX_BIOMASS], dB[CLB_MAX_BIOMASS], dC[S];
  bool any = false;
#pragma unroll
  for (int k = 0; k < CLB_MAX_BIOMASS; ++k) {
    bm[k] = 0.0;
    dB[k] = 0.0;
    if (k < t.nm) {
      const double b = (double)a.bmass[(long long)t.id[k] * N + n];
      bm[k] = (b > kThrd) ? b : 0.0;
      any = any || bm[k] > kThrd;
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) dC[s] = 0.0;

  if (any) {
    // ---- the concentrations run_kinetics sees: sums of the node's incoming populations
    const int y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
    const Nbr nb(x, y, z, g);
    int src[7];
    bool wall[7];
    src[0] = n;
    wall[0] = false;
#pragma unroll
    for (int i = 1; i < 7; ++i) {
      const int m = nb.at(-Q7::cx[i], -Q7::cy[i], -Q7::cz[i]);
      wall[i] = a.kind[a.mask[m]] != CLB_KIND_FLUID;
      src[i] = wall[i] ? n : m;
    }
    double conc[S], avail[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      double r = 0.0;
      if (a.immobile[s]) {
        r = (double)a.g[(long long)(s * 7) * N + n];
        if (a.pend_r) r += (double)a.pend_r[(long long)s * N + n];
      } else {
#pragma unroll
        for (int i = 0; i < 7; ++i) {
          if (!ODD) r += (double)a.g[(long long)(s * 7 + i) * N + n];
          else r += (double)a.g[(long long)(s * 7 + opp_dir(i)) * N + nb.at(-Q7::cx[i], -Q7::cy[i], -Q7::cz[i])];
        }
        if (a.pend_r) {
          const R* P1 = a.pend_r + (long long)s * N;
          const R* P2 = a.pend_w + (long long)s * N;
#pragma unroll
          for (int i = 0; i < 7; ++i)
            r += Q7::w(i) * (double)(wall[i] ? P2[n] : P1[src[i]]);
        }
      }
      const double c = r + (double)a.c_ref[s];
      conc[s] = (c > kThrd) ? c : 0.0;
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      double tot;
      avail[s] = conc[s];
      if (t.use_totals && totals_usable<S>(t, conc, s, tot)) avail[s] = fabs(tot);
    }

    // ---- per present microbe: uptake estimate, network, closure, biomass increment. A
    // uniform loop that is not unrolled (the network is inlined once, its weights read with
    // the uniform index k); the per-microbe registers are picked and set by selects.
    double draw[S];
#pragma unroll
    for (int s = 0; s < S; ++s) draw[s] = 0.0;
#pragma unroll 1
    for (int k = 0; k < t.nm; ++k) {
      double B = 0.0;
#pragma unroll
      for (int q = 0; q < CLB_MAX_BIOMASS; ++q)
        if (q == k) B = bm[q];
      if (!(B > kThrd)) continue;
      const double mw = t.mw[k];
      double fin[CLB_MAX_SOLUTES], fout[S];
#pragma unroll
      for (int s = 0; s < CLB_MAX_SOLUTES; ++s) fin[s] = 0.0;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const double kc = t.Kc[k][s];
        const double mm = (kc + avail[s] > 0.0) ? avail[s] / (kc + avail[s]) : 0.0;
        const double cap = -exhaustion_flux(avail[s], B, mw, a.dt);  // >= 0
        double v = (t.vmax[k][s] > kThrd) ? t.vmax[k][s] * mm : cap * mm;
        if (cap > 0.0 && v > cap) v = cap;
        const double floor_mag = -t.maxup[k][s];
        if (t.maxup[k][s] > -1e29 && floor_mag >= 0.0 && v > floor_mag) v = floor_mag;
        if (t.fixlb[s] && cap >= 0.0 && v > cap) v = cap;
        fin[s] = v;
        fout[s] = v;
      }
      const SrgNet& nt = t.net[k];
      double mu = 0.0;
      if (nt.on) {
        const double gr = net_growth(nt, fin);
        mu = gr;
        if (nt.closure == 1) {
          // consume what the predicted growth needs, at most the estimate
#pragma unroll
          for (int s = 0; s < S; ++s)
            if (nt.yield[s] > 0.0) {
              double need = gr / nt.yield[s];
              if (need > fin[s]) need = fin[s];
              fout[s] = need;
            }
        } else if (nt.closure == 2) {
          // growth limited by the scarcest substrate, consumption to match
          double gl = gr;
#pragma unroll
          for (int s = 0; s < S; ++s)
            if (nt.yield[s] > 0.0) {
              const double gmax = fin[s] * nt.yield[s];
              if (gmax < gl) gl = gmax;
            }
          mu = gl;
#pragma unroll
          for (int s = 0; s < S; ++s)
            if (nt.yield[s] > 0.0) fout[s] = gl / nt.yield[s];
        }
      }
      double d = (fabs(mu) < kThrd) ? -t.decay[k] * B * a.dt  // decayToDeltaB, per second
                                    : mu * B * a.dt / 3600.0;  // growthToDeltaB, per hour
      if (d + B < 0.0) d = -B;
#pragma unroll
      for (int q = 0; q < CLB_MAX_BIOMASS; ++q)
        if (q == k) dB[q] = d;
#pragma unroll
      for (int s = 0; s < S; ++s) draw[s] += -(fout[s] * (B * mw) * a.dt / 3.6e6);  // fluxToDeltaC
    }

    // ---- each substrate's draw, summed over the microbes, clamped once
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (t.fixc[s]) continue;
      double d = draw[s];
      if (fabs(d) <= kThrd) continue;
      const double head = avail[s] + (t.use_totals ? 0.0 : dC[s]);
      if (d + head < 0.0) d = -head;
      if (d > 0.0) d = 0.0;
      if (t.use_totals) draw_down<S>(t, conc, s, -d, dC);
      else dC[s] += d;
    }
  }

  // ---- deposit: what exceeds 1e-12, for update_rxnLattices to add to the kinetics' share
#pragma unroll
  for (int s = 0; s < S; ++s) inc[(long long)s * N + n] = (R)((fabs(dC[s]) <= kThrd) ? 0.0 : dC[s]);
#pragma unroll
  for (int b = 0; b < CLB_MAX_BIOMASS; ++b) {
    if (b >= a.nb) continue;
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < CLB_MAX_BIOMASS; ++k)
      if (k < t.nm && t.id[k] == b && fabs(dB[k]) > kThrd) v = dB[k];
    inc[(long long)(S + b) * N + n] = (R)v;
  }
}

template <typename R, int S>
void launch_srg_s(const AdeArgs<R>& a, const SrgTables* t, R* inc, bool odd, hipStream_t s) {
  const dim3 grid((unsigned)((a.geo.N + 255) / 256));
  if (odd) hipLaunchKernelGGL((srg_kernel<R, S, true>), grid, dim3(256), 0, s, a, t, inc);
  else hipLaunchKernelGGL((srg_kernel<R, S, false>), grid, dim3(256), 0, s, a, t, inc);
}

}  // namespace

template <typename R>
void launch_surrogate(const AdeArgs<R>& a, const SrgTables* t, R* inc, bool odd, hipStream_t s) {
  switch (a.S) {
    case 1: launch_srg_s<R, 1>(a, t, inc, odd, s); break;
    case 2: launch_srg_s<R, 2>(a, t, inc, odd, s); break;
    case 3: launch_srg_s<R, 3>(a, t, inc, odd, s); break;
    case 4: launch_srg_s<R, 4>(a, t, inc, odd, s); break;
    case 5: launch_srg_s<R, 5>(a, t, inc, odd, s); break;
    case 6: launch_srg_s<R, 6>(a, t, inc, odd, s); break;
    case 7: launch_srg_s<R, 7>(a, t, inc, odd, s); break;
    default: launch_srg_s<R, 8>(a, t, inc, odd, s); break;
  }
}

template void launch_surrogate<float>(const AdeArgs<float>&, const SrgTables*, float*, bool, hipStream_t);
template void launch_surrogate<double>(const AdeArgs<double>&, const SrgTables*, double*, bool, hipStream_t);

}  // namespace clb
