// clb_spec_kernel.h -- per-voxel equilibrium speciation, FP64 Newton, one persistent
// batched kernel with a work queue (instantiated by clb_spec_batch.hip and
// clb_spec_coupled_*.hip).
//
// The system is the reference's: species log10 c_i = logK_i + sum_j S_ij x_j (x = log10
// component activities, clamped to [-30, 1]), and the reference's own convergence judge,
// the PCF residual f_j = (1/mu0_j) (log10 S_P,j - log10 S_R,j) with its clamps and its
// zero-total guard (pcf_residual / calc_reactive_product_sums / calc_species,
// src/complab3d_processors_part4_eqsolver.hh:442-726). The Newton iteration solves
// f(x) = 0 directly:  J dx = -f with J_jk = df_j/dx_k
//   = (1/mu0_j) sum_i S_ik c_i a_ij,  a_ij = -S_ij / (S_ij < 0 ? S_P,j : S_R,j),
// LU with partial pivoting, step capped to 5 log units, Armijo backtracking on |f|^2/2,
// and the plain PCF update x += f (the reference's fixed-point step) when backtracking
// fails. A voxel is converged when |f|_2 < tol: the same test the reference applies
// (src/complab3d_processors_part4_eqsolver.hh:1033-1049); the Newton reserve solver of
// src/complab3d_processors_part4_eqsolver_NR_reserve.hh:512-615 is the specification of
// the Newton + backtracking structure.
//
// Work distribution: every lane owns one voxel at a time; a lane whose voxel has
// converged (or hit max_iter) takes the next voxel from a global counter, one atomic per
// wavefront per refill. Every loop trip is one residual+Jacobian evaluation on every
// busy lane, so lanes never wait on a slower voxel of their wavefront.
//
// Components are padded to the kernel's NC with zero columns and zero totals (the
// zero-total guard pins them: f_j = 0, J row j = e_j), so one instantiation serves every
// nc <= NC.
#pragma once

#include "clb_lattice.h"

namespace clb {

namespace {

constexpr double kMinConc = 1e-30, kMaxConc = 10.0, kMinLog = -30.0, kMaxLog = 1.0;
constexpr double kMaxStep = 5.0;
constexpr double kMinAlpha = 1.0 / 256.0;

__device__ __forceinline__ double clampd(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

// Residual (and optionally Jacobian) at x. Returns |f|_2^2.
template <int NC, int NSM>
__device__ __forceinline__ double evaluate(const Tableau& tb, const int* vsp, int nvar, const double* lkv,
                                           const double* x, const double* T, double* f, double* J, double* csp) {
  double sp[NC], sn[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) { sp[j] = 0.0; sn[j] = 0.0; }
#pragma unroll
  for (int i = 0; i < NSM; ++i) {
    csp[i] = 0.0;
    if (i < tb.ns && tb.is_eq[i]) {
      double lk = tb.logK[i];
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (v < nvar && vsp[v] == i) lk = lkv[v];
      double lc = lk;
#pragma unroll
      for (int j = 0; j < NC; ++j) lc += tb.S[i * CLB_MAX_COMPONENTS + j] * x[j];
      lc = clampd(lc, kMinLog, kMaxLog);
      const double c = exp10(lc);
      csp[i] = c;
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const double s = tb.S[i * CLB_MAX_COMPONENTS + j];
        if (s > 1e-15) sp[j] += s * c;
        else if (s < -1e-15) sn[j] -= s * c;
      }
    }
  }
  double rP[NC], rR[NC];
  double nrm2 = 0.0;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    double SR, SP;
    if (T[j] >= 0.0) { SR = sp[j]; SP = T[j] + sn[j]; }
    else { SR = -T[j] + sp[j]; SP = sn[j]; }
    SR = fmax(SR, kMinConc);
    SP = fmax(SP, kMinConc);
    double fj = tb.inv_mu0[j] * (log10(SP) - log10(SR));
    fj = clampd(fj, -10.0, 10.0);
    if (!isfinite(fj)) fj = 0.0;
    if (T[j] <= 1e-18) fj = 0.0;
    f[j] = fj;
    nrm2 += fj * fj;
    rP[j] = tb.inv_mu0[j] / SP;
    rR[j] = tb.inv_mu0[j] / SR;
  }
  if (J) {
#pragma unroll
    for (int q = 0; q < NC * NC; ++q) J[q] = 0.0;
#pragma unroll
    for (int i = 0; i < NSM; ++i) {
      if (i < tb.ns && tb.is_eq[i]) {
        const double c = csp[i];
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          const double s = tb.S[i * CLB_MAX_COMPONENTS + j];
          if (s == 0.0) continue;
          const double a = -s * c * ((s < 0.0) ? rP[j] : rR[j]);
#pragma unroll
          for (int k = 0; k < NC; ++k) J[j * NC + k] += a * tb.S[i * CLB_MAX_COMPONENTS + k];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NC; ++j)
      if (T[j] <= 1e-18) {
#pragma unroll
        for (int k = 0; k < NC; ++k) J[j * NC + k] = (k == j) ? 1.0 : 0.0;
      }
  }
  return nrm2;
}

// Solve J d = -f by LU with partial pivoting; row swaps are done with selects so that
// every array index stays a compile-time constant (registers, no scratch).
template <int NC>
__device__ __forceinline__ bool newton_direction(double* J, const double* f, double* d) {
  double b[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) b[j] = -f[j];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    int p = k;
    double pv = fabs(J[k * NC + k]);
#pragma unroll
    for (int r = k + 1; r < NC; ++r) {
      const double v = fabs(J[r * NC + k]);
      if (v > pv) { pv = v; p = r; }
    }
    if (!(pv > 1e-300)) return false;
#pragma unroll
    for (int r = k + 1; r < NC; ++r) {
      const bool sw = (p == r);
#pragma unroll
      for (int c = k; c < NC; ++c) {
        const double a = J[k * NC + c], bb = J[r * NC + c];
        J[k * NC + c] = sw ? bb : a;
        J[r * NC + c] = sw ? a : bb;
      }
      const double a = b[k], bb = b[r];
      b[k] = sw ? bb : a;
      b[r] = sw ? a : bb;
    }
    const double inv = 1.0 / J[k * NC + k];
#pragma unroll
    for (int r = k + 1; r < NC; ++r) {
      const double m = J[r * NC + k] * inv;
#pragma unroll
      for (int c = k + 1; c < NC; ++c) J[r * NC + c] -= m * J[k * NC + c];
      b[r] -= m * b[k];
    }
  }
#pragma unroll
  for (int k = NC - 1; k >= 0; --k) {
    double s = b[k];
#pragma unroll
    for (int c = k + 1; c < NC; ++c) s -= J[k * NC + c] * d[c];
    d[k] = s / J[k * NC + k];
    if (!isfinite(d[k])) return false;
  }
  return true;
}

struct CoupledIO {
  void* g;              // populations (R*)
  const uint8_t* mask;
  const int* list;      // reactive node indices
  const double* cbuf;   // S*N post-kinetics concentrations
  Geom geo;
  int odd_done;         // parity of the step whose post-collision values are in g
};

template <typename R>
__device__ __forceinline__ void add_increment(const CoupledIO& io, int s, int n, double dC) {
  R* g = (R*)io.g;
  const Geom& geo = io.geo;
  if (!io.odd_done) {
    // even step wrote f*_i of node n to slot (opp(i), n): every slot of node n
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      R* q = &g[(long long)(s * 7 + i) * geo.N + n];
      *q = (R)((double)*q + dC * Q7::w(i));
    }
  } else {
    const int x = n % geo.nx, y = (n / geo.nx) % geo.ny, z = n / (geo.nx * geo.ny);
    const Nbr nb(x, y, z, geo);
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      R* q = &g[(long long)(s * 7 + i) * geo.N + nb.at(Q7::cx[i], Q7::cy[i], Q7::cz[i])];
      *q = (R)((double)*q + dC * Q7::w(i));
    }
  }
}

// MODE 0: batch (totals in, log activities out). MODE 1: coupled step (concentrations in,
// limited increments added to the populations, run_equilibrium_biotic semantics).
template <typename R, int NC, int NSM, int MODE>
__global__ void __launch_bounds__(256) spec_kernel(Tableau tb, SpecArgs sa, int warm, CoupledIO io) {
  const int lane = threadIdx.x & 63;
  const long long n_items = sa.n;
  int vsp[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) vsp[v] = sa.var_species[v];
  const int nvar = sa.nvar;

  long long item = -1;
  bool have = false, exhausted = false;
  double x[NC], T[NC], fx[NC], d[NC], lkv[4];
  double phi = 0.0, alpha = 1.0;
  int evals = 0;
  bool trial = false;
  int node = 0;
  double conc[NSM];
  unsigned long long my_solves = 0, my_conv = 0, my_iters = 0;

  while (true) {
    // ---- refill idle lanes from the work queue (one atomic per wavefront)
    const bool need = !have && !exhausted;
    const unsigned long long want = __ballot(need);
    if (want) {
      const int leader = __ffsll((long long)want) - 1;
      unsigned int base = 0;
      if (lane == leader) base = atomicAdd(sa.counter, (unsigned int)__popcll(want));
      base = __shfl(base, leader);
      if (need) {
        const unsigned long long below = want & ((1ull << lane) - 1ull);
        const long long v = (long long)base + __popcll(below);
        if (v < n_items) {
          item = v;
          have = true;
          trial = false;
          evals = 0;
          alpha = 1.0;
#pragma unroll
          for (int q = 0; q < 4; ++q) lkv[q] = (q < nvar) ? sa.logk_var[v * nvar + q] : 0.0;
          if (MODE == 0) {
#pragma unroll
            for (int j = 0; j < NC; ++j) T[j] = (j < tb.nc) ? sa.totals[v * tb.nc + j] : 0.0;
            bool user = false;
            if (warm != CLB_WARM_REF) {
              user = true;
#pragma unroll
              for (int j = 0; j < NC; ++j)
                if (j < tb.nc && !isfinite(sa.logc[v * tb.nc + j])) user = false;
            }
#pragma unroll
            for (int j = 0; j < NC; ++j) {
              if (j >= tb.nc) x[j] = 0.0;
              else if (user) x[j] = clampd(sa.logc[v * tb.nc + j], kMinLog, kMaxLog);
              else x[j] = log10(clampd(T[j], kMinConc, kMaxConc));
            }
          } else {
            node = io.list[v];
#pragma unroll
            for (int i = 0; i < NSM; ++i)
              conc[i] = (i < tb.ns) ? fmax(io.cbuf[(long long)i * io.geo.N + node], kMinConc) : 0.0;
            // calc_component_totals (:520-538)
#pragma unroll
            for (int j = 0; j < NC; ++j) {
              double s = 0.0;
#pragma unroll
              for (int i = 0; i < NSM; ++i)
                if (i < tb.ns && tb.is_eq[i]) s += tb.S[i * CLB_MAX_COMPONENTS + j] * clampd(conc[i], kMinConc, kMaxConc);
              T[j] = (j < tb.nc) ? fmax(s, kMinConc) : 0.0;
            }
            // initial guess: log10 of the component species' concentration (:977-994)
#pragma unroll
            for (int j = 0; j < NC; ++j) {
              double cc = 1e-7;
              const int cs = (j < tb.nc) ? tb.comp_species[j] : -1;
#pragma unroll
              for (int i = 0; i < NSM; ++i)
                if (i == cs) cc = conc[i];
              x[j] = (j < tb.nc) ? log10(clampd(cc, kMinConc, kMaxConc)) : 0.0;
            }
          }
        } else {
          exhausted = true;
        }
      }
    }
    if (__ballot(have) == 0) break;
    if (!have) continue;

    // ---- one evaluation on every busy lane
    double p[NC], fp[NC], J[NC * NC], csp[NSM];
#pragma unroll
    for (int j = 0; j < NC; ++j) p[j] = trial ? clampd(x[j] + alpha * d[j], kMinLog, kMaxLog) : x[j];
    const double nrm2 = evaluate<NC, NSM>(tb, vsp, nvar, lkv, p, T, fp, J, csp);
    ++evals;
    const double phip = 0.5 * nrm2;
    bool accept = !trial || (phip <= (1.0 - 2e-4 * alpha) * phi);
    bool done = false, converged = false;
    if (accept) {
#pragma unroll
      for (int j = 0; j < NC; ++j) { x[j] = p[j]; fx[j] = fp[j]; }
      phi = phip;
      if (sqrt(nrm2) < tb.tol) {
        done = true;
        converged = true;
      } else if (evals >= tb.max_iter) {
        done = true;
      } else {
        bool ok = newton_direction<NC>(J, fp, d);
        double m = 0.0;
#pragma unroll
        for (int j = 0; j < NC; ++j) m = fmax(m, fabs(d[j]));
        if (ok && m > kMaxStep) {
          const double sc = kMaxStep / m;
#pragma unroll
          for (int j = 0; j < NC; ++j) d[j] *= sc;
        }
        if (ok) {
          trial = true;
          alpha = 1.0;
        } else {
          // singular Jacobian: the reference's fixed-point step
#pragma unroll
          for (int j = 0; j < NC; ++j) x[j] = clampd(x[j] + fx[j], kMinLog, kMaxLog);
          trial = false;
        }
      }
    } else {
      alpha *= 0.5;
      if (evals >= tb.max_iter) {
        done = true;
      } else if (alpha < kMinAlpha) {
#pragma unroll
        for (int j = 0; j < NC; ++j) x[j] = clampd(x[j] + fx[j], kMinLog, kMaxLog);
        trial = false;
      }
    }
    if (!done) continue;

    // ---- finished voxel: write results, release the lane
    have = false;
    my_solves += 1;
    my_conv += converged ? 1 : 0;
    my_iters += evals;
    if (MODE == 0) {
#pragma unroll
      for (int j = 0; j < NC; ++j)
        if (j < tb.nc) sa.logc[item * tb.nc + j] = x[j];
      if (sa.iters) sa.iters[item] = evals;
      if (sa.resid) sa.resid[item] = sqrt(2.0 * phi);
      if (sa.conv) sa.conv[item] = converged ? 1 : 0;
    } else {
      // species at the accepted point (calc_species), NaN guard and clamp (:1187-1214)
      double cfin[NSM];
      {
        double ftmp[NC];
        evaluate<NC, NSM>(tb, vsp, nvar, lkv, x, T, ftmp, nullptr, cfin);
      }
#pragma unroll
      for (int i = 0; i < NSM; ++i) {
        if (i >= tb.ns) continue;
        double eq;
        if (!tb.is_eq[i]) eq = (conc[i] > kMinConc) ? fmin(conc[i], kMaxConc) : kMinConc;
        else eq = cfin[i];
        if (!isfinite(eq)) eq = (conc[i] > kMinConc) ? fmin(conc[i], kMaxConc) : kMinConc;
        eq = clampd(eq, kMinConc, kMaxConc);
        // 10 % change limiter with a 1e-9 floor, concentration floor 1e-20 (:1434-1447)
        double dC = eq - conc[i];
        const double mx = fmax(1e-9, 0.1 * fabs(conc[i]));
        if (fabs(dC) > mx) dC = (dC > 0.0) ? mx : -mx;
        if (conc[i] + dC < 1e-20) dC = 1e-20 - conc[i];
        if (fabs(dC) > 1e-12) add_increment<R>(io, i, node, dC);
      }
    }
  }
  // wavefront reduction of the statistics, one atomic per wavefront
  for (int off = 32; off > 0; off >>= 1) {
    my_solves += __shfl_xor(my_solves, off);
    my_conv += __shfl_xor(my_conv, off);
    my_iters += __shfl_xor(my_iters, off);
  }
  if (lane == 0 && sa.stats) {
    atomicAdd(&sa.stats[0], my_solves);
    atomicAdd(&sa.stats[1], my_conv);
    atomicAdd(&sa.stats[2], my_iters);
  }
}

template <typename R, int MODE>
void dispatch_spec(const Tableau& tb, const SpecArgs& sa, int warm, const CoupledIO& io, int grid, hipStream_t s) {
  const int nc = tb.nc, ns = tb.ns;
  const dim3 blk(256);
  // (components, species) shapes compiled in; smaller tableaux are padded up
#define CLB_SPEC(NC_, NS_) hipLaunchKernelGGL((spec_kernel<R, NC_, NS_, MODE>), dim3(grid), blk, 0, s, tb, sa, warm, io)
  if (nc <= 2 && ns <= 4) CLB_SPEC(2, 4);
  else if (nc <= 3 && ns <= 6) CLB_SPEC(3, 6);
  else if (nc <= 4 && ns <= 8) CLB_SPEC(4, 8);
  else if (ns <= 16) CLB_SPEC(8, 16);
  else if (ns <= 24) CLB_SPEC(8, 24);
  else CLB_SPEC(8, 32);
#undef CLB_SPEC
}

}  // namespace
}  // namespace clb
