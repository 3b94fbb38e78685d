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

// MATLAB's tansig, 2 / (1 + exp(-2n)) - 1, with the division done as a reciprocal
// estimate refined by two Newton steps (within an ulp or two of the divided value; a
// saturated exponent gives -1, as the division does).
__device__ __forceinline__ double tansig(double n) {
  const double d = 1.0 + exp(-2.0 * n);
  if (!(d < 1e300)) return (d == d) ? -1.0 : d;
  double r = __builtin_amdgcn_rcp(d);
  r = r * (2.0 - d * r);
  r = r * (2.0 - d * r);
  return 2.0 * r - 1.0;
}

// LDS image of one network layer: W (16 x 16, row per neuron, zero padded) then b (16).
constexpr int kLayerDoubles = CLB_SRG_MAX_WIDTH * CLB_SRG_MAX_WIDTH + CLB_SRG_MAX_WIDTH;

// One microbe's network (the evaluation of surrogateModel.hh:396-427): inputs mapped onto
// [-1, 1] (mapminmax), tansig hidden layers, one linear output neuron, the reverse map; a
// growth rate below 1e-8 or not finite is 0. Each neuron sums W x from 0 in input order and
// then adds its bias, as matMultp / matSum do. The weights come from the block's LDS copy
// (`w`, kLayerDoubles per layer; every lane reads the same address, a broadcast). Neurons
// go four at a time and inputs four at a time; groups beyond a layer's width are skipped
// by uniform branches, and the zero padding inside a group adds exact zeros (a padded
// neuron is tansig(0) = 0).
template <int S>
__device__ __forceinline__ double net_growth(const SrgNet& nt, const double* w, const double* fin) {
  constexpr int W = CLB_SRG_MAX_WIDTH;
  double a[W];
#pragma unroll
  for (int j = 0; j < W; ++j) a[j] = 0.0;
#pragma unroll
  for (int j = 0; j < S; ++j) {  // n_in <= S (an input names a substrate)
    if (j < nt.n_in) {
      const int sub = nt.in_sub[j];
      double x = 0.0;
#pragma unroll
      for (int s = 0; s < S; ++s)
        if (sub == s) x = fin[s];
      a[j] = (x - nt.x_offset[j]) * nt.x_gain[j] + nt.x_ymin;
    }
  }
  int wprev = nt.n_in;
  for (int l = 0; l < nt.n_layers; ++l) {
    const int wl = nt.width[l];
    const bool last = (l == nt.n_layers - 1);
    const double* Wl = w + l * kLayerDoubles;
    double z[W];
#pragma unroll
    for (int rb = 0; rb < W; rb += 4) {
      // four neurons at a time: four independent sums (each still in input order)
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      if (rb < wl) {
#pragma unroll
        for (int cb = 0; cb < W; cb += 4) {
          if (cb < wprev) {
#pragma unroll
            for (int c = cb; c < cb + 4; ++c)
#pragma unroll
              for (int j = 0; j < 4; ++j) acc[j] += Wl[(rb + j) * W + c] * a[c];
          }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[j] = Wl[W * W + rb + j] + acc[j];
          if (!last) acc[j] = tansig(acc[j]);
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) z[rb + j] = acc[j];
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

// Whether voxel n is reactive (1 <= x <= nx-2, fluid, neither solid nor bounce-back code)
// and holds a surrogate organism above 1e-12. Branch-free loads (a voxel past the end reads
// the last one and is rejected), so a tile's passes issue their loads together.
template <typename R>
__device__ __forceinline__ bool srg_present(const AdeArgs<R>& a, const SrgTables& t, long long n) {
  const Geom& g = a.geo;
  const long long q = (n < g.N) ? n : g.N - 1;
  const int x = (int)(q % g.nx);
  const uint8_t code = a.mask[q];
  bool any = false;
#pragma unroll
  for (int k = 0; k < CLB_MAX_BIOMASS; ++k)
    if (k < t.nm) any = any || (double)a.bmass[(long long)t.id[k] * g.N + q] > kThrd;
  return any && n < g.N && x >= 1 && x <= g.nx - 2 && a.kind[code] == CLB_KIND_FLUID && code != a.solid_code &&
         code != a.bb_code;
}

// The processor at one voxel holding a surrogate organism; writes all S + nb of its
// increments (the transport kernel reads them only at such voxels).
template <typename R, int S, bool ODD>
__device__ __forceinline__ void srg_voxel(const AdeArgs<R>& a, const SrgTables& t, const double* wlds, R* inc,
                                          int n) {
  const Geom& g = a.geo;
  const int x = n % g.nx;
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
      double fin[S];
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
      }
      const SrgNet& nt = t.net[k];
      double mu = 0.0;
      int closure = 0;
      if (nt.on) {
        const double gr = net_growth<S>(nt, wlds + k * t.lds_layers * kLayerDoubles, fin);
        mu = gr;
        closure = nt.closure;
        if (closure == 2) {
          // growth limited by the scarcest substrate
#pragma unroll
          for (int s = 0; s < S; ++s)
            if (nt.yield[s] > 0.0) {
              const double gmax = fin[s] * nt.yield[s];
              if (gmax < mu) mu = gmax;
            }
        }
      }
      double d = (fabs(mu) < kThrd) ? -t.decay[k] * B * a.dt  // decayToDeltaB, per second
                                    : mu * B * a.dt / 3600.0;  // growthToDeltaB, per hour
      if (d + B < 0.0) d = -B;
#pragma unroll
      for (int q = 0; q < CLB_MAX_BIOMASS; ++q)
        if (q == k) dB[q] = d;
      // Fout: the estimate, or (closures) what the growth needs -- at most the estimate
      // (closure 1), exactly growth / yield (closure 2)
#pragma unroll
      for (int s = 0; s < S; ++s) {
        double fout = fin[s];
        if (closure != 0 && nt.yield[s] > 0.0) {
          fout = mu / nt.yield[s];
          if (closure == 1 && fout > fin[s]) fout = fin[s];
        }
        draw[s] += -(fout * (B * mw) * a.dt / 3.6e6);  // fluxToDeltaC
      }
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

// One block per tile of kTile voxels: the block first gathers the tile's voxels holding a
// surrogate organism into LDS (one LDS atomic per wavefront and pass), then its threads
// evaluate them, all lanes busy. The organism usually sits in a small part of the domain
// (biofilm), so most wavefronts of a plain one-voxel-per-thread launch would wait on one
// or two lanes running the network; here a tile's few voxels share one wavefront, and no
// global atomics or second launch are needed.
constexpr int kTileThreads = 256;
constexpr int kTilePasses = 8;
constexpr int kTile = kTileThreads * kTilePasses;

template <typename R, int S, bool ODD>
__global__ void __launch_bounds__(kTileThreads) srg_tile_kernel(AdeArgs<R> a, const SrgTables* __restrict__ tp,
                                                                R* inc) {
  __shared__ int slot[kTile];
  __shared__ unsigned int count;
  extern __shared__ double wlds[];  // nm x lds_layers layers of the networks
  const SrgTables& t = *tp;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  if (tid == 0) count = 0u;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kTile;
  bool pres[kTilePasses];
#pragma unroll
  for (int v = 0; v < kTilePasses; ++v) pres[v] = srg_present(a, t, base + (long long)v * kTileThreads + tid);
#pragma unroll
  for (int v = 0; v < kTilePasses; ++v) {
    const long long n = base + (long long)v * kTileThreads + tid;
    const bool present = pres[v];
    const unsigned long long m = __ballot(present);
    if (m != 0ull) {
      const int leader = __ffsll((long long)m) - 1;
      unsigned int at = 0u;
      if (lane == leader) at = atomicAdd(&count, (unsigned int)__popcll(m));
      at = __shfl(at, leader);
      if (present) slot[at + __popcll(m & ((1ull << lane) - 1ull))] = (int)n;
    }
  }
  __syncthreads();
  const int total = (int)count;
  if (total == 0) return;
  // the networks' weights into LDS (only for tiles with work)
  const int per = t.lds_layers * kLayerDoubles;
  for (int i = tid; i < t.nm * per; i += kTileThreads) {
    const int k = i / per, r = i - k * per, l = r / kLayerDoubles, j = r - l * kLayerDoubles;
    const SrgNet& nt = t.net[k];
    wlds[i] = (j < CLB_SRG_MAX_WIDTH * CLB_SRG_MAX_WIDTH) ? (&nt.W[l][0][0])[j]
                                                          : nt.b[l][j - CLB_SRG_MAX_WIDTH * CLB_SRG_MAX_WIDTH];
  }
  __syncthreads();
  for (int i = tid; i < total; i += kTileThreads) srg_voxel<R, S, ODD>(a, t, wlds, inc, slot[i]);
}

template <typename R, int S>
void launch_srg_s(const AdeArgs<R>& a, const SrgTables* t, int lds_bytes, R* inc, bool odd, hipStream_t s) {
  const dim3 grid((unsigned)((a.geo.N + kTile - 1) / kTile));
  if (odd) hipLaunchKernelGGL((srg_tile_kernel<R, S, true>), grid, dim3(kTileThreads), lds_bytes, s, a, t, inc);
  else hipLaunchKernelGGL((srg_tile_kernel<R, S, false>), grid, dim3(kTileThreads), lds_bytes, s, a, t, inc);
}

}  // namespace

template <typename R>
void launch_surrogate(const AdeArgs<R>& a, const SrgTables* t, int nm, int lds_layers, R* inc, bool odd,
                      hipStream_t s) {
  const int lds_bytes = nm * lds_layers * kLayerDoubles * (int)sizeof(double);
  switch (a.S) {
    case 1: launch_srg_s<R, 1>(a, t, lds_bytes, inc, odd, s); break;
    case 2: launch_srg_s<R, 2>(a, t, lds_bytes, inc, odd, s); break;
    case 3: launch_srg_s<R, 3>(a, t, lds_bytes, inc, odd, s); break;
    case 4: launch_srg_s<R, 4>(a, t, lds_bytes, inc, odd, s); break;
    case 5: launch_srg_s<R, 5>(a, t, lds_bytes, inc, odd, s); break;
    case 6: launch_srg_s<R, 6>(a, t, lds_bytes, inc, odd, s); break;
    case 7: launch_srg_s<R, 7>(a, t, lds_bytes, inc, odd, s); break;
    default: launch_srg_s<R, 8>(a, t, lds_bytes, inc, odd, s); break;
  }
}

template void launch_surrogate<float>(const AdeArgs<float>&, const SrgTables*, int, int, float*, bool, hipStream_t);
template void launch_surrogate<double>(const AdeArgs<double>&, const SrgTables*, int, int, double*, bool,
                                       hipStream_t);

}  // namespace clb
