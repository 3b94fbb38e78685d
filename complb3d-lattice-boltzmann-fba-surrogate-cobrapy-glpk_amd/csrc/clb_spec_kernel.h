// clb_spec_kernel.h -- per-voxel equilibrium speciation, FP64 Newton-type solve, one
// persistent batched kernel with a work queue (instantiated by clb_spec_batch.hip and
// clb_spec_coupled_*.hip).
//
// The system is the reference's: species log10 c_i = logK_i + sum_j S_ij x_j with x the
// log10 component activities in [-30, 1], and the reference's own convergence judge, the
// PCF residual f_j = (1/mu0_j) (log10 S_P,j - log10 S_R,j) with its clamps and zero-total
// guard (calc_species / calc_reactive_product_sums / pcf_residual,
// src/complab3d_processors_part4_eqsolver.hh:442-726). A voxel is converged when
// |f|_2 < tol, the test of src/complab3d_processors_part4_eqsolver.hh:1033-1049.
//
// Iteration (the Newton + damping structure follows the reserve solver,
// src/complab3d_processors_part4_eqsolver_NR_reserve.hh:512-615, with the globalisation
// changed so that every voxel converges):
//   * Levenberg-Marquardt on the smooth log residual fs(x) (the judge without its clamps):
//     (J^T J + lambda I) d = -J^T fs, J_jk = dfs_j/dx_k = (1/mu0_j) sum_i S_ik c_i a_ij,
//     a_ij = -S_ij / (S_ij < 0 ? S_P,j : S_R,j); gain-ratio acceptance and Nielsen's
//     lambda update; steps capped at 5 log units.
//   * The log residual can have local minima away from the solution. When |fs| has not
//     dropped 5 % within 6 evaluations, the voxel switches to Newton on the strictly convex
//     potential G(x) = sum_i c_i / ln 10 - T.x (gradient S^T c - T, Hessian
//     ln 10 S^T diag(c) S, Armijo backtracking on G), which cannot stall; once
//     |f| < 1e-4 it returns to Levenberg-Marquardt for the fast local convergence.
//
// Work distribution: every lane owns one voxel at a time; a lane whose voxel has
// converged (or hit max_iter) takes the next voxel from a global counter, one atomic per
// wavefront per refill. Every loop trip is one residual+Jacobian evaluation on every
// busy lane, so lanes never wait on a slower voxel of their wavefront.
//
// Components are padded to the kernel's NC with zero columns and zero totals; the
// zero-total guard freezes them (zero residual, identity row), so one instantiation
// serves every nc <= NC.
#pragma once

#include "clb_fastmath.h"
#include "clb_lattice.h"

namespace clb {

namespace {

constexpr double kMinConc = 1e-30, kMaxConc = 10.0, kMinLog = -30.0, kMaxLog = 1.0;
constexpr double kLn10 = 2.302585092994045684;
constexpr double kMaxStep = 5.0;
constexpr int kStallEvals = 6;
constexpr double kBackToLM = 1e-4;

__device__ __forceinline__ double clampd(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

// packed upper triangle index of (j, k), j <= k
template <int NC>
__device__ __forceinline__ constexpr int up(int j, int k) {
  return j * NC - (j * (j - 1)) / 2 + (k - j);
}

template <int NC>
struct EvalOut {
  double judge2;         // |f_ref|^2 (the reference's judge)
  double merit;          // LM: |fs|^2 / 2;  G mode: G(x)
  double A[NC * (NC + 1) / 2];  // LM: J^T J;  G mode: Hessian of G (packed upper)
  double rhs[NC];        // LM: -J^T fs;  G mode: -(S^T c - T)
  double fsn;            // |fs| (LM) for the stall test
};

// One evaluation at x. gmode selects the system that is assembled. csp (optional)
// receives the species concentrations of calc_species (clamped log c).
template <int NC, int NSM>
__device__ __forceinline__ void evaluate(const Tableau& tb, const int* vsp, int nvar, const double* lkv,
                                         const double* x, const double* T, bool gmode, EvalOut<NC>& o,
                                         double* csp) {
  double craw[NSM];
  double spr[NC], snr[NC], sps[NC], sns[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) { spr[j] = 0.0; snr[j] = 0.0; sps[j] = 0.0; sns[j] = 0.0; }
  double gsum = 0.0;
#pragma unroll
  for (int i = 0; i < NSM; ++i) {
    craw[i] = 0.0;
    if (i < tb.ns && tb.is_eq[i]) {
      double lk = tb.logK[i];
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (v < nvar && vsp[v] == i) lk = lkv[v];
      double lc = lk;
#pragma unroll
      for (int j = 0; j < NC; ++j) lc += tb.S[i * CLB_MAX_COMPONENTS + j] * x[j];
      const double cr = fast_exp10(fmin(lc, 250.0));
      const double cc = (lc > kMaxLog) ? kMaxConc : ((lc < kMinLog) ? kMinConc : cr);
      craw[i] = cr;
      gsum += cr;
      if (csp) csp[i] = cc;
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const double s = tb.S[i * CLB_MAX_COMPONENTS + j];
        if (s > 1e-15) { spr[j] += s * cc; sps[j] += s * cr; }
        else if (s < -1e-15) { snr[j] -= s * cc; sns[j] -= s * cr; }
      }
    } else if (csp) {
      csp[i] = 0.0;
    }
  }
  // reference judge (clamped species, floors, +-10 clamp, zero-total guard)
  double judge2 = 0.0;
  double fs[NC], rP[NC], rR[NC];
  double fsn2 = 0.0, gtx = 0.0;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const bool live = T[j] > 1e-18;
    double SR, SP;
    if (T[j] >= 0.0) { SR = spr[j]; SP = T[j] + snr[j]; }
    else { SR = -T[j] + spr[j]; SP = snr[j]; }
    double fj = tb.inv_mu0[j] * fast_log10(fmax(SP, kMinConc) * fast_rcp(fmax(SR, kMinConc)));
    fj = clampd(fj, -10.0, 10.0);
    if (!isfinite(fj) || !live) fj = 0.0;
    judge2 += fj * fj;
    // smooth residual on the raw species
    double SRs, SPs;
    if (T[j] >= 0.0) { SRs = sps[j]; SPs = T[j] + sns[j]; }
    else { SRs = -T[j] + sps[j]; SPs = sns[j]; }
    SRs = fmax(SRs, 1e-300);
    SPs = fmax(SPs, 1e-300);
    const double fsj = live ? tb.inv_mu0[j] * (fast_log10(SPs) - fast_log10(SRs)) : 0.0;
    fs[j] = fsj;
    fsn2 += fsj * fsj;
    rP[j] = live ? tb.inv_mu0[j] / SPs : 0.0;
    rR[j] = live ? tb.inv_mu0[j] / SRs : 0.0;
    gtx += T[j] * x[j];
  }
  o.judge2 = judge2;
  o.fsn = sqrt(fsn2);
  // Jacobian of the chosen system, row-major NC x NC
  double J[NC * NC];
#pragma unroll
  for (int q = 0; q < NC * NC; ++q) J[q] = 0.0;
#pragma unroll
  for (int i = 0; i < NSM; ++i) {
    if (i < tb.ns && tb.is_eq[i]) {
      const double c = craw[i];
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const double s = tb.S[i * CLB_MAX_COMPONENTS + j];
        if (s == 0.0) continue;
        const double w = gmode ? kLn10 * s * c : -s * c * ((s < 0.0) ? rP[j] : rR[j]);
#pragma unroll
        for (int k = 0; k < NC; ++k) J[j * NC + k] += w * tb.S[i * CLB_MAX_COMPONENTS + k];
      }
    }
  }
  if (gmode) {
    o.merit = gsum / kLn10 - gtx;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const bool live = T[j] > 1e-18;
#pragma unroll
      for (int k = j; k < NC; ++k) {
        const bool livek = T[k] > 1e-18;
        o.A[up<NC>(j, k)] = (live && livek) ? J[j * NC + k] : ((j == k) ? 1.0 : 0.0);
      }
      o.rhs[j] = live ? -(sps[j] - sns[j] - T[j]) : 0.0;
    }
  } else {
    o.merit = 0.5 * fsn2;
    // a component under the zero-total guard keeps its starting activity (its residual
    // is zero, and the reference's update never moves it): drop its column
#pragma unroll
    for (int k = 0; k < NC; ++k)
      if (!(T[k] > 1e-18)) {
#pragma unroll
        for (int r = 0; r < NC; ++r) J[r * NC + k] = 0.0;
      }
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      double g = 0.0;
#pragma unroll
      for (int r = 0; r < NC; ++r) g += J[r * NC + j] * fs[r];
      o.rhs[j] = -g;
#pragma unroll
      for (int k = j; k < NC; ++k) {
        double a = 0.0;
#pragma unroll
        for (int r = 0; r < NC; ++r) a += J[r * NC + j] * J[r * NC + k];
        o.A[up<NC>(j, k)] = a;
      }
    }
  }
}

// Cholesky factor of A + lam I (A symmetric, packed upper): L packed lower, row-major
// (L(j,k), k <= j, at j(j+1)/2 + k), with reciprocal pivots inv, so a solve costs no
// division. Returns false when a pivot is not positive.
template <int NC>
__device__ __forceinline__ bool chol_factor(const double* A, double lam, double* L, double* inv) {
#pragma unroll
  for (int j = 0; j < NC; ++j) {
#pragma unroll
    for (int k = 0; k <= j; ++k) {
      double s = A[up<NC>(k, j)] + ((j == k) ? lam : 0.0);
#pragma unroll
      for (int m = 0; m < k; ++m) s -= L[j * (j + 1) / 2 + m] * L[k * (k + 1) / 2 + m];
      if (j == k) {
        if (!(s > 0.0) || !isfinite(s)) return false;
        inv[j] = fast_rsqrt(s);
        L[j * (j + 1) / 2 + j] = s * inv[j];
      } else {
        L[j * (j + 1) / 2 + k] = s * inv[k];
      }
    }
  }
  return true;
}

// d = (L L^T)^-1 b; false when the result is not finite.
template <int NC>
__device__ __forceinline__ bool chol_apply(const double* L, const double* inv, const double* b, double* d) {
  double y[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    double s = b[j];
#pragma unroll
    for (int m = 0; m < j; ++m) s -= L[j * (j + 1) / 2 + m] * y[m];
    y[j] = s * inv[j];
  }
  bool ok = true;
#pragma unroll
  for (int j = NC - 1; j >= 0; --j) {
    double s = y[j];
#pragma unroll
    for (int m = j + 1; m < NC; ++m) s -= L[m * (m + 1) / 2 + j] * d[m];
    d[j] = s * inv[j];
    ok = ok && isfinite(d[j]);
  }
  return ok;
}

// Solve (A + lam I) d = b by Cholesky.
template <int NC>
__device__ __forceinline__ bool chol_solve(const double* A, double lam, const double* b, double* d) {
  double L[NC * (NC + 1) / 2], inv[NC];
  return chol_factor<NC>(A, lam, L, inv) && chol_apply<NC>(L, inv, b, d);
}

template <int NC>
__device__ __forceinline__ double quad_form(const double* A, const double* d) {
  double q = 0.0;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    q += A[up<NC>(j, j)] * d[j] * d[j];
#pragma unroll
    for (int k = j + 1; k < NC; ++k) q += 2.0 * A[up<NC>(j, k)] * d[j] * d[k];
  }
  return q;
}

// ---- near-equilibrium solve inside the transport kernel (fused coupled step)
//
// In the coupled step a voxel's start is the previous step's equilibrium moved by one
// transport + kinetics update, so Newton converges in a few evaluations. This variant is
// Newton on the convex potential G (Hessian ln10 S^T diag(c) S, gradient S^T c - T) with
// backtracking: a trial point is taken when G satisfies the Armijo condition or when the
// reference judge decreases (near the solution G changes below its rounding while the
// judge still converges quadratically). A short Newton step gets the third-order
// (Chebyshev) correction: with c_i(x + d) = c_i(x) 10^(S_i.d) exactly, the gradient at
// x + d is g + H d + (ln10^2 / 2) sum_i c_i S_i (S_i.d)^2 + O(d^3), so
// d = d0 - H^-1 (ln10^2 / 2) sum_i c_i S_i (S_i.d0)^2 with d0 = -H^-1 g leaves an error of
// order |d0|^3 instead of |d0|^2, and a start one transport step away from equilibrium
// usually meets the judge at the first trial point. At most `max_evals` evaluations; a
// voxel that has not met the judge by then is handed, with its next trial point, to the
// persistent kernel above.
constexpr double kJudgeStep = 3e-4;   // judge a trial point reached by a step shorter than this
constexpr double kCubicStep = 0.05;   // correct Newton steps shorter than this (log10 units)

// Species concentration of calc_species (:442-496) at log activities x.
template <int NC>
__device__ __forceinline__ double species_conc(const Tableau& tb, int i, const double* x) {
  double lc = tb.logK[i];
#pragma unroll
  for (int j = 0; j < NC; ++j) lc += tb.S[i * CLB_MAX_COMPONENTS + j] * x[j];
  lc = clampd(lc, kMinLog, kMaxLog);
  return clampd(fast_exp10(lc), kMinConc, kMaxConc);
}

// Lean evaluation for the fused solve: G(x) = sum_i c_i / ln10 - T.x, the reference judge
// |f|^2 (clamped species, floors, +-10 clamp, zero-total guard, as in evaluate above), the
// packed Hessian ln10 S^T diag(c) S and rhs = -(S^T c - T); a component under the
// zero-total guard gets an identity row and a zero rhs, so it keeps its start.
//
// The weighted species concentrations are left in crs[i * BT] (LDS), for the correction
// of the Newton step.
// UNR: unroll count of the species loops (1 inside the transport kernel, whose registers
// are spoken for; a full unroll lets the species' exp10 chains overlap where registers
// are free).
template <int NC, int NSM, int BT, int UNR = 1>
__device__ __forceinline__ void eval_g(const Tableau& tb, const double* x, const double* T, bool JUDGE, double& G,
                                       double& judge2, double* A, double* rhs, double* crs) {
  constexpr int NP = NC * (NC + 1) / 2;
  // Branch-free: the sign split of each coefficient is fmax(+-s, 0) and a species outside
  // the equilibrium set has weight 0 (its row is zero as well), so every species and
  // component runs the same multiply-adds (uniform branches on tableau entries compile to
  // exec-mask juggling that costs more than the arithmetic).
  double pos[NC], neg[NC], raw[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) { pos[j] = 0.0; neg[j] = 0.0; raw[j] = 0.0; }
#pragma unroll
  for (int q = 0; q < NP; ++q) A[q] = 0.0;
  double gsum = 0.0;
  // one species at a time: the index is uniform, so the tableau entries come from scalar
  // loads and only one species' temporaries are live (the solve shares the kernel's
  // register budget with the transport phase)
#pragma unroll UNR
  for (int i = 0; i < NSM; ++i) {
    const double eqw = tb.eqw[i];
    double lc = tb.logK[i];
#pragma unroll
    for (int j = 0; j < NC; ++j) lc = fma(tb.S[i * CLB_MAX_COMPONENTS + j], x[j], lc);
    const double cr = eqw * fast_exp10(fmin(lc, 250.0));
    const double cc = eqw * ((lc > kMaxLog) ? kMaxConc : ((lc < kMinLog) ? kMinConc : cr));
    crs[i * BT] = cr;
    gsum += cr;
    const double lcr = kLn10 * cr;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const double s = tb.S[i * CLB_MAX_COMPONENTS + j];
      pos[j] = fma(fmax(s, 0.0), cc, pos[j]);
      neg[j] = fma(fmax(-s, 0.0), cc, neg[j]);
      raw[j] = fma(s, cr, raw[j]);
      const double w = lcr * s;
#pragma unroll
      for (int k = j; k < NC; ++k) A[up<NC>(j, k)] = fma(w, tb.S[i * CLB_MAX_COMPONENTS + k], A[up<NC>(j, k)]);
    }
  }
  double jg = 0.0, gtx = 0.0;
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const bool live = T[j] > 1e-18;
    if (JUDGE) {
      const double SR = (T[j] >= 0.0) ? pos[j] : -T[j] + pos[j];
      const double SP = (T[j] >= 0.0) ? T[j] + neg[j] : neg[j];
      double fj = tb.inv_mu0[j] * fast_log10(fmax(SP, kMinConc) * fast_rcp(fmax(SR, kMinConc)));
      fj = clampd(fj, -10.0, 10.0);
      if (!isfinite(fj) || !live) fj = 0.0;
      jg += fj * fj;
    }
    gtx = fma(T[j], x[j], gtx);
    rhs[j] = live ? -(raw[j] - T[j]) : 0.0;
#pragma unroll
    for (int k = j; k < NC; ++k)
      if (!(live && T[k] > 1e-18)) A[up<NC>(j, k)] = (j == k) ? 1.0 : 0.0;
  }
  judge2 = jg;
  G = gsum / kLn10 - gtx;
}

// Newton direction at a point whose Hessian A, right-hand side rhs (= -gradient) and
// species concentrations crs (LDS) eval_g has just produced: d = A^-1 rhs, with the
// third-order correction when the step is short, limited to kMaxStep. A component under
// the zero-total guard (identity row, zero rhs) keeps d_j = 0. Returns false when A is
// not positive definite.
template <int NC, int NSM, int BT, int UNR = 1>
__device__ __forceinline__ bool newton_dir(const Tableau& tb, const double* A, const double* rhs, const double* T,
                                           const double* crs, double* d, double& dmax) {
  double L[NC * (NC + 1) / 2], inv[NC];
  if (!chol_factor<NC>(A, 0.0, L, inv) || !chol_apply<NC>(L, inv, rhs, d)) return false;
  dmax = 0.0;
#pragma unroll
  for (int j = 0; j < NC; ++j) dmax = fmax(dmax, fabs(d[j]));
  if (dmax < kCubicStep) {
    double t3[NC], e[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) t3[j] = 0.0;
#pragma unroll UNR
    for (int i = 0; i < NSM; ++i) {
      double u = 0.0;
#pragma unroll
      for (int j = 0; j < NC; ++j) u = fma(tb.S[i * CLB_MAX_COMPONENTS + j], d[j], u);
      const double w = crs[i * BT] * u * u;
#pragma unroll
      for (int j = 0; j < NC; ++j) t3[j] = fma(w, tb.S[i * CLB_MAX_COMPONENTS + j], t3[j]);
    }
#pragma unroll
    for (int j = 0; j < NC; ++j) t3[j] = (T[j] > 1e-18) ? (-0.5 * kLn10 * kLn10) * t3[j] : 0.0;
    if (chol_apply<NC>(L, inv, t3, e)) {
      dmax = 0.0;
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        d[j] += e[j];
        dmax = fmax(dmax, fabs(d[j]));
      }
    }
  }
  if (dmax > kMaxStep) {
#pragma unroll
    for (int j = 0; j < NC; ++j) d[j] *= kMaxStep / dmax;
    dmax = kMaxStep;
  }
  return true;
}

// c: concentrations of the NSM species (tb.ns of them used), already floored at 1e-30 as
// run_equilibrium_biotic reads them (:1373-1400); crs: this thread's LDS column (stride
// BT) for the species concentrations. Returns true when the judge is met, with x the log
// activities; otherwise x is the next trial point, where a continuing solve starts.
// evals: the number of evaluations.
template <int NC, int NSM, int BT>
__device__ __forceinline__ bool spec_fused(const Tableau& tb, const double* c, double* x, int& evals, int max_evals,
                                           double* crs) {
  constexpr int NP = NC * (NC + 1) / 2;
  double T[NC], d[NC];
  // calc_component_totals (:520-538) and the component-slot guess (:977-994)
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NSM; ++i)
      s = fma(tb.eqw[i] * tb.S[i * CLB_MAX_COMPONENTS + j], clampd(c[i], kMinConc, kMaxConc), s);
    T[j] = (j < tb.nc) ? fmax(s, kMinConc) : 0.0;
    double cc = 1e-7;
    const int cs = (j < tb.nc) ? tb.comp_species[j] : -1;
#pragma unroll
    for (int i = 0; i < NSM; ++i)
      if (i == cs) cc = c[i];
    x[j] = (j < tb.nc) ? fast_log10(clampd(cc, kMinConc, kMaxConc)) : 0.0;
    d[j] = 0.0;
  }
  const double tol2 = tb.tol * tb.tol;
  // start: gradient and Hessian only. The solve always takes at least one step, and a
  // start that already met the judge stays within it after that (tiny) step, so the
  // judge is first evaluated at the first trial point.
  double G = 0.0, judge2 = 0.0, slope = 0.0, alpha = 1.0, dmax = 0.0;
  {
    double A[NP], rhs[NC], jdummy;
    eval_g<NC, NSM, BT>(tb, x, T, false, G, jdummy, A, rhs, crs);
    evals = 1;
    if (!newton_dir<NC, NSM, BT>(tb, A, rhs, T, crs, d, dmax)) return false;
#pragma unroll
    for (int j = 0; j < NC; ++j) slope -= rhs[j] * d[j];
  }
  // A trial point is taken on the Armijo condition for G, or when the judge improves on the
  // current point's (the first judged point: when it meets the tolerance). The judge is
  // evaluated at points reached by a short step (|d| < kJudgeStep) and at the last point the
  // budget allows; after a longer step the residual cannot meet 1e-10 yet, so its logs are
  // skipped. A voxel is declared converged only where the judge was evaluated and met.
  bool have_judge = false, conv = false;
#pragma unroll 1
  for (int it = 1; it < max_evals; ++it) {
    double p[NC], A[NP], rhs[NC], Gp, jp = 0.0;
#pragma unroll
    for (int j = 0; j < NC; ++j) p[j] = clampd(x[j] + alpha * d[j], kMinLog, kMaxLog);
    const bool judged = alpha * dmax < kJudgeStep || it + 1 == max_evals;
    eval_g<NC, NSM, BT>(tb, p, T, judged, Gp, jp, A, rhs, crs);
    ++evals;
    const bool armijo = Gp <= G + 1e-4 * alpha * slope;
    if (!(armijo || (judged && jp < (have_judge ? judge2 : tol2)))) {
      alpha *= 0.5;
      continue;
    }
#pragma unroll
    for (int j = 0; j < NC; ++j) x[j] = p[j];
    G = Gp;
    judge2 = jp;
    have_judge = judged;
    if (judged && judge2 < tol2) { conv = true; break; }
    if (!newton_dir<NC, NSM, BT>(tb, A, rhs, T, crs, d, dmax)) { alpha = 0.0; break; }
    slope = 0.0;
#pragma unroll
    for (int j = 0; j < NC; ++j) slope -= rhs[j] * d[j];
    alpha = 1.0;
  }
  if (!conv) {
#pragma unroll
    for (int j = 0; j < NC; ++j) x[j] = clampd(x[j] + alpha * d[j], kMinLog, kMaxLog);
  }
  return conv;
}

// Continuation of spec_fused from the trial point it handed over, for a voxel that had not
// met the judge within the fused step's budget: the same Newton iteration on G with the
// third-order correction of short steps, but judged from its first point, which may already
// meet the judge. T: the voxel's component totals as spec_fused formed them; x: the trial
// point in, and on failure the next trial point out. At most max_evals evaluations. Returns
// true when the judge is met, with the species of the judged point in crs.
template <int NC, int NSM, int BT, int UNR = 1>
__device__ __forceinline__ bool spec_fused_continue(const Tableau& tb, const double* T, double* x, int& evals,
                                                    int max_evals, double* crs) {
  constexpr int NP = NC * (NC + 1) / 2;
  const double tol2 = tb.tol * tb.tol;
  double G = 0.0, judge2 = 0.0, slope = 0.0, alpha = 1.0, dmax = 0.0, d[NC];
  {
    double A[NP], rhs[NC];
    eval_g<NC, NSM, BT, UNR>(tb, x, T, true, G, judge2, A, rhs, crs);
    evals = 1;
    if (judge2 < tol2) return true;
    if (!newton_dir<NC, NSM, BT, UNR>(tb, A, rhs, T, crs, d, dmax)) return false;
#pragma unroll
    for (int j = 0; j < NC; ++j) slope -= rhs[j] * d[j];
  }
  // trial points as in spec_fused: Armijo on G, or a judge below the current point's; the
  // judge at points reached by a short step and at the budget's last point
  bool conv = false;
#pragma unroll 1
  for (int it = 1; it < max_evals; ++it) {
    double p[NC], A[NP], rhs[NC], Gp, jp = 0.0;
#pragma unroll
    for (int j = 0; j < NC; ++j) p[j] = clampd(x[j] + alpha * d[j], kMinLog, kMaxLog);
    const bool judged = alpha * dmax < kJudgeStep || it + 1 == max_evals;
    eval_g<NC, NSM, BT, UNR>(tb, p, T, judged, Gp, jp, A, rhs, crs);
    ++evals;
    const bool armijo = Gp <= G + 1e-4 * alpha * slope;
    if (!(armijo || (judged && jp < judge2))) {
      alpha *= 0.5;
      continue;
    }
#pragma unroll
    for (int j = 0; j < NC; ++j) x[j] = p[j];
    G = Gp;
    if (judged) judge2 = jp;
    if (judged && judge2 < tol2) { conv = true; break; }
    if (!newton_dir<NC, NSM, BT, UNR>(tb, A, rhs, T, crs, d, dmax)) { alpha = 0.0; break; }
    slope = 0.0;
#pragma unroll
    for (int j = 0; j < NC; ++j) slope -= rhs[j] * d[j];
    alpha = 1.0;
  }
  if (!conv) {
#pragma unroll
    for (int j = 0; j < NC; ++j) x[j] = clampd(x[j] + alpha * d[j], kMinLog, kMaxLog);
  }
  return conv;
}

// Limited increment of a species whose equilibrium concentration is ceq (NaN guard and
// clamp of :1187-1214, then the 10 % limiter and 1e-20 floor of :1434-1447); c is its
// concentration before the solve. A species outside the equilibrium set keeps c.
__device__ __forceinline__ double spec_increment_eq(bool is_eq, double ceq, double c) {
  double eq = is_eq ? ceq : ((c > kMinConc) ? fmin(c, kMaxConc) : kMinConc);
  if (!isfinite(eq)) eq = (c > kMinConc) ? fmin(c, kMaxConc) : kMinConc;
  eq = clampd(eq, kMinConc, kMaxConc);
  double dd = eq - c;
  const double mx = fmax(1e-9, 0.1 * fabs(c));
  if (fabs(dd) > mx) dd = (dd > 0.0) ? mx : -mx;
  if (c + dd < 1e-20) dd = 1e-20 - c;
  return dd;
}

// Limited increment of species i after a converged solve at x (calc_species, then as above).
template <int NC>
__device__ __forceinline__ double spec_increment(const Tableau& tb, int i, const double* x, double c) {
  return spec_increment_eq(tb.is_eq[i], tb.is_eq[i] ? species_conc<NC>(tb, i, x) : 0.0, c);
}

struct CoupledIO {
  void* g;              // populations (R*)
  const uint8_t* mask;
  const int* list;      // reactive node indices
  const double* cbuf;   // S*N post-kinetics concentrations
  Geom geo;
  int odd_done;         // parity of the step whose post-collision values are in g
  int immobile[CLB_MAX_SOLUTES];
  void* pend_w;         // (R*) deferred increments: dC is written here instead of into g
  unsigned long long* blk_stats;  // the fused step's statistics pairs (solves, evaluations)
  long long blk_stats_n;          // and their number
};

template <typename R>
__device__ __forceinline__ void add_increment(const CoupledIO& io, int s, int n, double dC) {
  R* g = (R*)io.g;
  const Geom& geo = io.geo;
  if (io.pend_w) {
    // fused step: the increment is added when the next step loads the populations
    ((R*)io.pend_w)[(long long)s * geo.N + n] = (R)dC;
    return;
  }
  if (io.immobile[s]) {
    // immobile solute: the whole concentration sits in the node's rest slot
    R* q = &g[(long long)(s * 7) * geo.N + n];
    *q = (R)((double)*q + dC);
  } else if (!io.odd_done) {
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
  constexpr int NP = NC * (NC + 1) / 2;
  const int lane = threadIdx.x & 63;
  const long long n_items = sa.n_dev ? (long long)*sa.n_dev : sa.n;
  int vsp[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) vsp[v] = sa.var_species[v];
  const int nvar = sa.nvar;
  const double tol2 = tb.tol * tb.tol;

  long long item = -1;
  bool have = false, exhausted = false, first = true;
  double x[NC], T[NC], d[NC], lkv[4];
  double A[NP], rhs[NC];
  double merit = 0.0, lam = -1.0, nu = 2.0, alpha = 1.0, best = 0.0, judge2 = 0.0;
  int evals = 0, since = 0;
  bool trial = false, gmode = false, lm_final = false;
  int node = 0;
  double conc[NSM];
  unsigned long long my_solves = 0, my_conv = 0, my_iters = 0;

  while (true) {
    // ---- refill idle lanes. Every lane's first item is fixed (lane l of the grid takes item
    // l); the work queue hands out the items from the grid's lane count on, one atomic per
    // wavefront and refill, and is never touched when the grid covers every item. Atomics on
    // one address serialise across the device, so a grid launched for few items (the fused
    // step's stragglers) must not queue up one per wavefront.
    const bool need = !have && !exhausted;
    const unsigned long long want = __ballot(need);
    if (want) {
      const int leader = __ffsll((long long)want) - 1;
      const long long lanes = (long long)gridDim.x * blockDim.x;
      long long base;
      if (first) {
        base = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63);  // every lane wants
      } else if (lanes >= n_items) {
        base = n_items;
      } else {
        unsigned int q = 0;
        if (lane == leader) q = atomicAdd(sa.counter, (unsigned int)__popcll(want));
        base = lanes + (long long)__shfl(q, leader);
      }
      first = false;
      if (need) {
        const unsigned long long below = want & ((1ull << lane) - 1ull);
        const long long v = base + __popcll(below);
        if (v < n_items) {
          item = v;
          have = true;
          trial = false;
          gmode = false;
          lm_final = false;
          evals = 0;
          since = 0;
          lam = -1.0;
          nu = 2.0;
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
              else x[j] = fast_log10(clampd(T[j], kMinConc, kMaxConc));
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
            // initial guess: log10 of the component species' concentration (:977-994), or,
            // for a voxel handed over by the fused step, the point its solve reached
            // (sa.logc, stride 4)
            bool cont = sa.logc != nullptr && tb.nc <= 4;
#pragma unroll
            for (int j = 0; j < NC; ++j)
              if (j < tb.nc && j < 4 && cont && !isfinite(sa.logc[v * 4 + j])) cont = false;
#pragma unroll
            for (int j = 0; j < NC; ++j) {
              double cc = 1e-7;
              const int cs = (j < tb.nc) ? tb.comp_species[j] : -1;
#pragma unroll
              for (int i = 0; i < NSM; ++i)
                if (i == cs) cc = conc[i];
              x[j] = (j < tb.nc) ? fast_log10(clampd(cc, kMinConc, kMaxConc)) : 0.0;
              if (cont && j < tb.nc && j < 4) x[j] = clampd(sa.logc[v * 4 + j], kMinLog, kMaxLog);
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
    double p[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) p[j] = trial ? clampd(x[j] + alpha * d[j], kMinLog, kMaxLog) : x[j];
    EvalOut<NC> e;
    evaluate<NC, NSM>(tb, vsp, nvar, lkv, p, T, gmode, e, nullptr);
    ++evals;
    bool accept = true;
    if (trial) {
      if (!gmode) {
        // predicted decrease of |fs|^2/2 by the model: rhs.d - d.A.d / 2 (A = J^T J)
        const double pred = [&] {
          double r = 0.0;
#pragma unroll
          for (int j = 0; j < NC; ++j) r += rhs[j] * d[j];
          return r - 0.5 * quad_form<NC>(A, d);
        }();
        const double rho = (merit - e.merit) / fmax(pred, 1e-300);
        accept = rho > 0.0;
        if (accept) {
          const double t = 2.0 * rho - 1.0;
          lam *= fmax(1.0 / 3.0, 1.0 - t * t * t);
          nu = 2.0;
        } else {
          lam *= nu;
          nu *= 2.0;
        }
      } else {
        double slope = 0.0;
#pragma unroll
        for (int j = 0; j < NC; ++j) slope -= rhs[j] * d[j];
        accept = e.merit <= merit + 1e-4 * alpha * slope;
        if (!accept) alpha *= 0.5;
      }
    }
    bool done = false, converged = false;
    bool resolve = true;  // compute a new direction at the end of this trip
    if (accept) {
#pragma unroll
      for (int j = 0; j < NC; ++j) { x[j] = p[j]; rhs[j] = e.rhs[j]; }
#pragma unroll
      for (int q = 0; q < NP; ++q) A[q] = e.A[q];
      merit = e.merit;
      judge2 = e.judge2;
      if (judge2 < tol2) {
        done = true;
        converged = true;
      } else if (!gmode) {
        if (lam < 0.0) {
          // initial damping scaled by |fs|^2: a start near the solution (a good warm
          // start) takes plain Gauss-Newton steps, a distant one is damped
          double mx = 0.0;
#pragma unroll
          for (int j = 0; j < NC; ++j) mx = fmax(mx, A[up<NC>(j, j)]);
          lam = fmax(1e-3 * mx * fmin(1.0, e.fsn * e.fsn), 1e-300);
          best = e.fsn;
          since = 0;
        } else if (e.fsn < 0.95 * best) {
          best = e.fsn;
          since = 0;
        } else if (!lm_final && ++since >= kStallEvals) {
          gmode = true;      // re-evaluate x on the convex potential
          trial = false;
          resolve = false;
        }
      } else if (judge2 < kBackToLM * kBackToLM) {
        gmode = false;       // close to the solution: fast local convergence
        lm_final = true;
        lam = -1.0;
        nu = 2.0;
        trial = false;
        resolve = false;
      }
    } else if (!gmode) {
      since += 1;
      if (!lm_final && since >= kStallEvals) {
        gmode = true;
        trial = false;
        resolve = false;
      }
    } else if (alpha >= 1.0 / 1024.0) {
      resolve = false;       // G backtracking: same direction, smaller step
      trial = true;
    } else {
      // backtracking at the resolution of G: finish with Levenberg-Marquardt from x
      gmode = false;
      lm_final = true;
      lam = -1.0;
      nu = 2.0;
      trial = false;
      resolve = false;
    }
    if (!done && evals >= tb.max_iter) done = true;
    if (!done && resolve) {
      const bool ok = chol_solve<NC>(A, gmode ? 0.0 : lam, rhs, d);
      if (ok) {
        double m = 0.0;
#pragma unroll
        for (int j = 0; j < NC; ++j) m = fmax(m, fabs(d[j]));
        if (m > kMaxStep) {
          const double sc = kMaxStep / m;
#pragma unroll
          for (int j = 0; j < NC; ++j) d[j] *= sc;
        }
        trial = true;
        alpha = 1.0;
      } else {
        // no usable direction: take the convex route from here
        gmode = true;
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
      if (sa.resid) sa.resid[item] = sqrt(judge2);
      if (sa.conv) sa.conv[item] = converged ? 1 : 0;
    } else {
      // species at the accepted point (calc_species), NaN guard and clamp (:1187-1214)
      double cfin[NSM];
      {
        EvalOut<NC> e2;
        evaluate<NC, NSM>(tb, vsp, nvar, lkv, x, T, false, e2, cfin);
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
  if (lane == 0 && sa.stats && my_solves) {
    atomicAdd(&sa.stats[0], my_solves);
    atomicAdd(&sa.stats[1], my_conv);
    atomicAdd(&sa.stats[2], my_iters);
  }
}

// The fused step's stragglers, first pass. The transport kernel left them by segment
// (segment s: sa.seg_n[s] nodes from sa.seg_list[s * segcap] on, with the trial points its
// solve handed over in sa.seg_x, stride 4; their post-kinetics concentrations in io.cbuf).
// Each continues the fused solve (spec_fused_continue) for at most sa.straggler_evals
// evaluations and, when it meets the judge, writes its limited increments to pend_w as
// fused_solve_phase does. A voxel still short of the judge goes with its latest point to the
// failure list (sa.fail_list, sa.fail_x, sa.fail_n), which spec_kernel MODE 1 finishes.
// Wavefront w of the grid works on segment w mod kStragglerSegments, in turn with that
// segment's other wavefronts. The lean solve holds few registers (four wavefronts per SIMD)
// and the grid is kStragglerGridScale times the persistent one, so most stragglers have a
// lane of their own and a step's stragglers finish in a few evaluations' latency. Atomics:
// one per wavefront that leaves a voxel (rare), and the straggler count once per launch;
// each block adds its solves and evaluations to its own pair of io.blk_stats (the fused
// step's per-wavefront accumulators, summed when statistics are read; the fused step has
// finished with them).
constexpr int kStragglerGridScale = 4;

template <typename R, int NC, int NSM>
__global__ void __launch_bounds__(256) spec_straggler_kernel(Tableau tb, SpecArgs sa, CoupledIO io) {
  __shared__ double crs[NSM][256];
  const long long N = io.geo.N;
  R* pend = (R*)io.pend_w;
  const int lane = threadIdx.x & 63;
  const long long gw = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // wavefront of the grid
  const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
  const int seg = (int)(gw % kStragglerSegments);
  const long long wseg = nw / kStragglerSegments + ((long long)seg < nw % kStragglerSegments ? 1 : 0);
  const long long ns = (long long)sa.seg_n[seg];
  unsigned long long my_solves = 0, my_evals = 0;
  for (long long k = (gw / kStragglerSegments) * 64 + lane; k < ns; k += wseg * 64) {
    const long long v = seg * sa.segcap + k;
    const int node = sa.seg_list[v];
    double c[NSM], T[NC], x[NC];
#pragma unroll
    for (int i = 0; i < NSM; ++i) c[i] = (i < tb.ns) ? fmax(io.cbuf[(long long)i * N + node], kMinConc) : kMinConc;
    // calc_component_totals (:520-538), formed as spec_fused forms them
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < NSM; ++i)
        t = fma(tb.eqw[i] * tb.S[i * CLB_MAX_COMPONENTS + j], clampd(c[i], kMinConc, kMaxConc), t);
      T[j] = (j < tb.nc) ? fmax(t, kMinConc) : 0.0;
      x[j] = (j < tb.nc) ? clampd(sa.seg_x[v * 4 + j], kMinLog, kMaxLog) : 0.0;
    }
    int evals = 0;
    bool conv = false;
    if (sa.straggler_evals > 0)
      conv = spec_fused_continue<NC, NSM, 256>(tb, T, x, evals, sa.straggler_evals, &crs[0][threadIdx.x]);
    my_evals += (unsigned long long)evals;
    if (conv) {
      my_solves += 1;
#pragma unroll
      for (int s = 0; s < NSM; ++s) {
        if (s >= tb.ns) continue;
        double d = spec_increment_eq(tb.is_eq[s], crs[s][threadIdx.x], c[s]);
        if (!(d > 1e-12 || d < -1e-12)) d = 0.0;
        pend[(long long)s * N + node] = (R)d;
      }
    }
    // voxels left to the full iteration, with the point they reached
    const unsigned long long fb = __ballot(!conv);
    if (fb) {
      const int leader = __ffsll((long long)fb) - 1;
      unsigned int fbase = 0;
      if (lane == leader) fbase = atomicAdd(sa.fail_n, (unsigned int)__popcll(fb));
      fbase = __shfl(fbase, leader);
      if (!conv) {
        const long long f = (long long)fbase + __popcll(fb & ((1ull << lane) - 1ull));
        sa.fail_list[f] = node;
#pragma unroll
        for (int j = 0; j < 4; ++j) sa.fail_x[f * 4 + j] = (j < NC && j < tb.nc) ? x[j] : 0.0;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    my_solves += __shfl_xor(my_solves, off);
    my_evals += __shfl_xor(my_evals, off);
  }
  __shared__ unsigned long long red[2][4];
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = my_solves;
    red[1][threadIdx.x >> 6] = my_evals;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long bs = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    const unsigned long long be = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    if (io.blk_stats && (long long)blockIdx.x < io.blk_stats_n && (bs || be)) {
      io.blk_stats[2 * blockIdx.x] += bs;
      io.blk_stats[2 * blockIdx.x + 1] += be;
    }
    if (blockIdx.x == 0 && sa.stats) {
      unsigned long long total = 0;
      for (int q = 0; q < kStragglerSegments; ++q) total += sa.seg_n[q];
      if (total) atomicAdd(&sa.stats[3], total);
    }
  }
}

// The first straggler pass for the fused shapes (nc padded up to 2, species up to 4, 6 or 8:
// every tableau the fused step takes); false for a tableau outside them.
template <typename R>
bool dispatch_straggler(const Tableau& tb, const SpecArgs& sa, const CoupledIO& io, int grid, hipStream_t s) {
  const int nc = tb.nc, ns = tb.ns;
  if (nc < 1 || nc > 4 || ns > 8 || !sa.seg_n || !sa.fail_n) return false;
  const int NC = (nc <= 2) ? 2 : nc;
  const int NSM = (ns <= 4) ? 4 : ((ns <= 6) ? 6 : 8);
  const dim3 blk(256);
#define CLB_STRAG(NC_, NSM_) \
  case NC_ * 10 + NSM_:                                                                                       \
    hipLaunchKernelGGL((spec_straggler_kernel<R, NC_, NSM_>), dim3(kStragglerGridScale * grid), blk, 0, s, tb, sa, \
                       io);                                                                                           \
    return true;
  switch (NC * 10 + (NSM < NC ? NC : NSM)) {
    CLB_STRAG(2, 4) CLB_STRAG(2, 6) CLB_STRAG(2, 8)
    CLB_STRAG(3, 4) CLB_STRAG(3, 6) CLB_STRAG(3, 8)
    CLB_STRAG(4, 4) CLB_STRAG(4, 6) CLB_STRAG(4, 8)
    default: return false;
  }
#undef CLB_STRAG
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
