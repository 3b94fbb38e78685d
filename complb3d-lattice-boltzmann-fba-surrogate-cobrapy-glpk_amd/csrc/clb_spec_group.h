// clb_spec_group.h -- the batched speciation solve (spec_kernel MODE 0) for tableaux of more
// than four components, one voxel per group of eight lanes.
//
// spec_kernel keeps a voxel's whole solve in one lane. With eight components that is the
// 8 x 8 Jacobian, J^T J, its factor and the 24 species of an evaluation, well over 512
// registers per lane: the kernel ran one wavefront per SIMD with heavy spills and waited on
// memory most of its cycles. Here lane c of a group owns component c (its log activity,
// total, step, row c of the system and of the Jacobian) and species c, c + 8, c + 16, ...;
// the tableau sits in LDS, and the group exchanges species concentrations and Jacobian rows
// through a small LDS area of its own, written and read by the same wavefront (a wavefront's
// LDS operations complete in order; wave_sync keeps the compiler from reordering them). The
// Cholesky factor is computed by rows across the group (column k's pivot and multipliers
// broadcast in turn), the substitutions likewise. Sums over the group use the xor butterfly,
// whose every step adds the same two partial sums in every lane, so all eight lanes hold
// bit-identical values and take the same decisions.
//
// From the reference's start the iteration is spec_kernel's: Levenberg-Marquardt on the
// smooth log residual with Nielsen damping, Newton on the convex potential G when it stalls,
// Levenberg-Marquardt again near the solution. From a caller's or the network's start it is
// the polish below, which falls back to that iteration. The judge is the reference's PCF
// residual |f|_2 < tol (src/complab3d_processors_part4_eqsolver.hh:664-726, 1033-1049).
#pragma once

#include "clb_spec_kernel.h"

namespace clb {

namespace {

constexpr int kGrp = 8;  // lanes per voxel: one per component

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum / max over the eight lanes of a group, the same bits in every lane.
__device__ __forceinline__ double gsum(double v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  return v;
}
__device__ __forceinline__ double gmax(double v) {
  v = fmax(v, __shfl_xor(v, 1));
  v = fmax(v, __shfl_xor(v, 2));
  v = fmax(v, __shfl_xor(v, 4));
  return v;
}
// Whether pred holds in every / any lane of the group (base: its first lane).
__device__ __forceinline__ bool gall(bool pred, int base) {
  return ((__ballot(!pred) >> base) & 0xffull) == 0ull;
}

// The polish (the start of a solve from a caller's or the network's log activities, close to
// the solution): Newton on G with the third-order correction of steps shorter than
// kPolishCubic, the judge only at points reached by a step shorter than kJudgeStep, at most
// kPolishEvals evaluations; a polish that runs out of them, or backtracks below
// kPolishMinAlpha, continues as the full iteration from its current point, and so does one
// whose first Newton step is longer than kPolishFar (log10 units).
constexpr int kPolishEvals = 10;
constexpr double kPolishFar = 1.0;
constexpr double kPolishMinAlpha = 1.0 / 64.0;
constexpr double kPolishCubic = 0.05;

// Per-group exchange area, in doubles: species cr[NSM], cc[NSM], then an 8 x 8 matrix
// (Jacobian rows or Cholesky rows) and fs[8]; one spare double keeps the groups' areas on
// different LDS banks.
template <int NSM>
constexpr int grp_area() { return 2 * NSM + 64 + 8 + 1; }

// Held to three wavefronts per SIMD (168 VGPRs, a few spills): on the 128^3 field of the
// 8 / 24 tableau it ran 105.8 M solves/s at the compiler's two (198 VGPRs), 115.3 M at three
// and 105.4 M at four (128 VGPRs, 84 spilled).
// POLISH: whether the kernel carries the polish (a grid launched for the reference's starts
// alone is built without it, which leaves the full iteration its registers).
template <int NSM, bool POLISH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
spec_group_kernel(Tableau tb, SpecArgs sa, int warm, int method) {
  constexpr int NC = 8, NT = NSM / kGrp;
  constexpr int AREA = grp_area<NSM>();
  constexpr int OJ = 2 * NSM, OF = 2 * NSM + 64;
  __shared__ double sS[NSM][NC];
  __shared__ double slk[NSM], seqw[NSM], smu[NC];
  __shared__ double xch[256 / kGrp][AREA];
  for (int q = threadIdx.x; q < NSM * NC; q += blockDim.x)
    sS[q / NC][q % NC] = tb.S[(q / NC) * CLB_MAX_COMPONENTS + q % NC];
  for (int q = threadIdx.x; q < NSM; q += blockDim.x) {
    slk[q] = tb.logK[q];
    seqw[q] = tb.eqw[q];
  }
  if (threadIdx.x < NC) smu[threadIdx.x] = tb.inv_mu0[threadIdx.x];
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int c = lane & (kGrp - 1);      // this lane's component
  const int base = lane & ~(kGrp - 1);  // the group's first lane
  double* X = xch[threadIdx.x / kGrp];
  const long long n_items = sa.n_dev ? (long long)*sa.n_dev : sa.n;
  const int nvar = sa.nvar;
  int vsp[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) vsp[v] = sa.var_species[v];
  const double tol2 = tb.tol * tb.tol;

  long long item = -1;
  bool have = false, exhausted = false, first = true;
  // the wavefront's current range of the work queue [qcur, qend), taken kChunk items at a
  // time: with short solves (a warm start) one atomic per refill would make the queue head
  // the bottleneck
  constexpr unsigned int kChunk = 32;
  long long qcur = 0, qend = 0;
  bool drained = false;
  double xc = 0.0, Tc = 0.0, dc = 0.0, rhsc = 0.0, lkv[4] = {0.0, 0.0, 0.0, 0.0};
  double arow[NC];  // row c of the accepted system (J^T J, or the Hessian of G)
#pragma unroll
  for (int k = 0; k < NC; ++k) arow[k] = 0.0;
  unsigned int live = 0;  // bit k: component k has a positive total
  double merit = 0.0, lam = -1.0, nu = 2.0, alpha = 1.0, best = 0.0, judge2 = 0.0, dmax = 0.0;
  int evals = 0, since = 0;
  bool trial = false, gmode = false, lm_final = false, pol = false, have_judge = false;
  unsigned long long my_solves = 0, my_conv = 0, my_iters = 0;

  while (true) {
    // ---- refill idle groups from the wavefront's range, which takes kChunk more items from
    // the work queue when it is used up; a group that finds none this trip waits for the next
    const bool need = !have && !exhausted;
    const unsigned long long want = __ballot(need && c == 0);
    if (want) {
      if (qcur >= qend && !drained) {
        // a wavefront's first range is fixed (wavefront w of the grid takes items
        // [w kChunk, (w + 1) kChunk)); the queue hands out ranges after those, and is never
        // touched when the fixed ranges cover every item (atomics on one address serialise
        // across the device)
        const int leader = __ffsll((long long)want) - 1;
        const long long fixed = (long long)gridDim.x * (blockDim.x >> 6) * kChunk;
        long long qb;
        if (first) {
          qb = ((long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kChunk;
        } else if (fixed >= n_items) {
          qb = n_items;
        } else {
          unsigned int q = 0;
          if (lane == leader) q = atomicAdd(sa.counter, kChunk);
          qb = fixed + (long long)__shfl(q, leader);
        }
        first = false;
        if (qb >= n_items) {
          drained = true;
        } else {
          qcur = qb;
          qend = qb + kChunk < n_items ? qb + kChunk : n_items;
        }
      }
      const long long avail = qend - qcur;
      if (need) {
        const long long rank = __popcll(want & ((1ull << base) - 1ull));
        const long long v = qcur + rank;
        if (rank < avail) {
          item = v;
          have = true;
          trial = false;
          lm_final = false;
          have_judge = false;
          evals = 0;
          since = 0;
          lam = -1.0;
          nu = 2.0;
          alpha = 1.0;
          dc = 0.0;
          dmax = 0.0;
#pragma unroll
          for (int q = 0; q < 4; ++q) lkv[q] = (q < nvar) ? sa.logk_var[v * nvar + q] : 0.0;
          Tc = (c < tb.nc) ? sa.totals[v * tb.nc + c] : 0.0;
          live = (unsigned int)((__ballot(Tc > 1e-18) >> base) & 0xffull);
          bool user = warm != CLB_WARM_REF;
          if (user) user = gall(c >= tb.nc || isfinite(sa.logc[v * tb.nc + c]), base);
          if (c >= tb.nc) xc = 0.0;
          else if (user) xc = clampd(sa.logc[v * tb.nc + c], kMinLog, kMaxLog);
          else xc = fast_log10(clampd(Tc, kMinConc, kMaxConc));
          pol = POLISH && (method == CLB_BATCH_POLISH || (method == CLB_BATCH_AUTO && user));
          gmode = pol;
        } else if (drained) {
          exhausted = true;
        }
      }
      qcur += (avail < (long long)__popcll(want)) ? avail : (long long)__popcll(want);
    }
    if (__ballot(have) == 0) break;
    if (!have) continue;

    // ---- one evaluation per group: the species, the merit of the current system and, where
    // the point can end the solve, the reference judge. The system's matrix (J^T J, or the
    // Hessian of G) is assembled below only when a new direction is needed, so the
    // evaluation that meets the judge builds none.
    const double pc = trial ? clampd(xc + alpha * dc, kMinLog, kMaxLog) : xc;
    double pall[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) pall[k] = __shfl(pc, base + k);
    double gpart = 0.0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int i = c + kGrp * t;
      double lk = slk[i];
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (v < nvar && vsp[v] == i) lk = lkv[v];
      double lc = lk;
#pragma unroll
      for (int k = 0; k < NC; ++k) lc = fma(sS[i][k], pall[k], lc);
      const double w = seqw[i];
      const double cr = w * fast_exp10(fmin(lc, 250.0));
      const double cc = w * ((lc > kMaxLog) ? kMaxConc : ((lc < kMinLog) ? kMinConc : cr));
      X[i] = cr;
      X[NSM + i] = cc;
      gpart += cr;
    }
    wave_sync();
    double pos = 0.0, neg = 0.0, sps = 0.0, sns = 0.0;
#pragma unroll 4
    for (int i = 0; i < NSM; ++i) {
      const double s = sS[i][c];
      const double cr = X[i], cc = X[NSM + i];
      pos = fma(fmax(s, 0.0), cc, pos);
      neg = fma(fmax(-s, 0.0), cc, neg);
      sps = fma(fmax(s, 0.0), cr, sps);
      sns = fma(fmax(-s, 0.0), cr, sns);
    }
    const bool livec = Tc > 1e-18;
    const double imu = smu[c];
    // the polish judges a point reached by a short step, and the last point of its budget:
    // after a longer step the residual cannot meet the tolerance yet, so its logs are skipped
    const bool judged = !pol || (trial && alpha * dmax < kJudgeStep) || evals + 1 >= kPolishEvals;
    double e_judge2 = 0.0;
    if (judged) {
      // the reference judge (clamped species, floors, +-10 clamp, zero-total guard)
      const double SR = (Tc >= 0.0) ? pos : -Tc + pos;
      const double SP = (Tc >= 0.0) ? Tc + neg : neg;
      double fj = imu * fast_log10(fmax(SP, kMinConc) * fast_rcp(fmax(SR, kMinConc)));
      fj = clampd(fj, -10.0, 10.0);
      if (!isfinite(fj) || !livec) fj = 0.0;
      e_judge2 = gsum(fj * fj);
    }
    double fsc = 0.0, rP = 0.0, rR = 0.0, e_fsn = 0.0, e_merit;
    if (gmode) {
      e_merit = gsum(gpart) / kLn10 - gsum(Tc * pc);
    } else {
      // the smooth residual on the raw species, for component c
      double SRs = (Tc >= 0.0) ? sps : -Tc + sps;
      double SPs = (Tc >= 0.0) ? Tc + sns : sns;
      SRs = fmax(SRs, 1e-300);
      SPs = fmax(SPs, 1e-300);
      fsc = livec ? imu * (fast_log10(SPs) - fast_log10(SRs)) : 0.0;
      rP = livec ? imu / SPs : 0.0;
      rR = livec ? imu / SRs : 0.0;
      e_fsn = sqrt(gsum(fsc * fsc));
      e_merit = 0.5 * e_fsn * e_fsn;
    }
    ++evals;

    bool accept = true;
    if (trial) {
      if (!gmode) {
        // predicted decrease of |fs|^2/2 by the model: rhs.d - d.A.d / 2 (A = J^T J)
        double ad = 0.0;
#pragma unroll
        for (int k = 0; k < NC; ++k) ad = fma(arow[k], __shfl(dc, base + k), ad);
        const double pred = gsum(rhsc * dc) - 0.5 * gsum(dc * ad);
        const double rho = (merit - e_merit) / fmax(pred, 1e-300);
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
        const double slope = -gsum(rhsc * dc);
        accept = e_merit <= merit + 1e-4 * alpha * slope;
        // near the solution G changes below its rounding while the judge still falls
        if (!accept && pol && judged) accept = e_judge2 < (have_judge ? judge2 : tol2);
        if (!accept) alpha *= 0.5;
      }
    }
    bool done = false, converged = false, resolve = true, lam_init = false;
    // leave the polish for the full iteration, which starts over at the current point
    auto leave_polish = [&]() {
      pol = false;
      gmode = false;
      lm_final = false;
      lam = -1.0;
      nu = 2.0;
      since = 0;
      trial = false;
      resolve = false;
    };
    if (accept) {
      xc = pc;
      merit = e_merit;
      have_judge = judged;
      if (judged) judge2 = e_judge2;
      if (judged && judge2 < tol2) {
        done = true;
        converged = true;
      } else if (pol) {
        if (evals >= kPolishEvals) leave_polish();
      } else if (!gmode) {
        if (lam < 0.0) {
          lam_init = true;
          best = e_fsn;
          since = 0;
        } else if (e_fsn < 0.95 * best) {
          best = e_fsn;
          since = 0;
        } else if (!lm_final && ++since >= kStallEvals) {
          gmode = true;
          trial = false;
          resolve = false;
        }
      } else if (judge2 < kBackToLM * kBackToLM) {
        gmode = false;
        lm_final = true;
        lam = -1.0;
        nu = 2.0;
        trial = false;
        resolve = false;
      }
    } else if (pol) {
      if (alpha < kPolishMinAlpha || evals >= kPolishEvals) {
        leave_polish();
      } else {
        resolve = false;
        trial = true;
      }
    } else if (!gmode) {
      since += 1;
      if (!lm_final && since >= kStallEvals) {
        gmode = true;
        trial = false;
        resolve = false;
      }
    } else if (alpha >= 1.0 / 1024.0) {
      resolve = false;
      trial = true;
    } else {
      gmode = false;
      lm_final = true;
      lam = -1.0;
      nu = 2.0;
      trial = false;
      resolve = false;
    }
    if (!done && evals >= tb.max_iter) done = true;
    if (!done && resolve) {
      if (accept) {
        // row c of the system at the new point: the Hessian of G, or J^T J of the log residual
        double jrow[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) jrow[k] = 0.0;
#pragma unroll 2
        for (int i = 0; i < NSM; ++i) {
          const double s = sS[i][c];
          const double cr = X[i];
          const double w = gmode ? kLn10 * s * cr : -s * cr * ((s < 0.0) ? rP : rR);
#pragma unroll
          for (int k = 0; k < NC; ++k) jrow[k] = fma(w, sS[i][k], jrow[k]);
        }
        if (gmode) {
#pragma unroll
          for (int k = 0; k < NC; ++k)
            arow[k] = (livec && ((live >> k) & 1u)) ? jrow[k] : ((k == c) ? 1.0 : 0.0);
          rhsc = livec ? -(sps - sns - Tc) : 0.0;
        } else {
          // a component under the zero-total guard keeps its starting activity: drop its column
#pragma unroll
          for (int k = 0; k < NC; ++k) {
            if (!((live >> k) & 1u)) jrow[k] = 0.0;
            X[OJ + c * NC + k] = jrow[k];
          }
          X[OF + c] = fsc;
          wave_sync();
          // row c of J^T J and -(J^T fs)_c
          double g = 0.0;
#pragma unroll
          for (int k = 0; k < NC; ++k) arow[k] = 0.0;
#pragma unroll
          for (int r = 0; r < NC; ++r) {
            const double jrc = X[OJ + r * NC + c];
            g = fma(jrc, X[OF + r], g);
#pragma unroll
            for (int k = 0; k < NC; ++k) arow[k] = fma(jrc, X[OJ + r * NC + k], arow[k]);
          }
          rhsc = -g;
          if (lam_init) {
            double dg = 0.0;
#pragma unroll
            for (int k = 0; k < NC; ++k)
              if (k == c) dg = arow[k];
            lam = fmax(1e-3 * gmax(dg) * fmin(1.0, e_fsn * e_fsn), 1e-300);
          }
        }
      }
      // (A + lam I) d = rhs: Cholesky by rows across the group. Lane c holds row c; for
      // column k the pivot comes from lane k, then lane m's multiplier L_mk updates the
      // rows below.
      const double lm = gmode ? 0.0 : lam;
      double a[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) a[k] = arow[k] + ((k == c) ? lm : 0.0);
      bool ok = true;
      double invc = 0.0;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const double piv = __shfl(a[k], base + k);
        ok = ok && piv > 0.0 && isfinite(piv);
        const double ik = fast_rsqrt(piv);
        if (c == k) {
          invc = ik;
          a[k] = piv * ik;
        } else if (c > k) {
          a[k] *= ik;
        }
        const double lck = (c > k) ? a[k] : 0.0;
#pragma unroll
        for (int m = k + 1; m < NC; ++m) {
          const double lmk = __shfl(lck, base + m);
          if (c >= m) a[m] -= lck * lmk;
        }
      }
      // the group's rows of L in the exchange area, for the backward substitutions
      wave_sync();
#pragma unroll
      for (int k = 0; k < NC; ++k) X[OJ + c * NC + k] = (k <= c) ? a[k] : 0.0;
      wave_sync();
      // L L^T d = r for component c's entry r: forward L y = r with row c of L in registers,
      // backward L^T d = y with column c of L from the group's rows
      auto chol_solve = [&](double r) {
        double acc = r;
#pragma unroll
        for (int m = 0; m < NC; ++m) {
          const double ym = __shfl(acc * invc, base + m);
          if (c > m) acc -= a[m] * ym;
        }
        acc *= invc;
#pragma unroll
        for (int m = NC - 1; m >= 0; --m) {
          const double dm = __shfl(acc * invc, base + m);
          if (c < m) acc -= X[OJ + m * NC + c] * dm;
        }
        return acc * invc;
      };
      double dnew = chol_solve(rhsc);
      ok = gall(ok && isfinite(dnew), base);
      if (ok && pol && gmax(fabs(dnew)) < kPolishCubic) {
        // third-order (Chebyshev) correction of a short Newton step on G: c_i(x + d) =
        // c_i(x) 10^(S_i.d) exactly, so the gradient at x + d is g + H d +
        // (ln10^2 / 2) sum_i c_i S_i (S_i.d)^2 + O(d^3); solving H e = -(ln10^2 / 2)
        // sum_i c_i S_i (S_i.d0)^2 leaves d0 + e an error of order |d0|^3
        double dall[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) dall[k] = __shfl(dnew, base + k);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int i = c + kGrp * t;
          double u = 0.0;
#pragma unroll
          for (int k = 0; k < NC; ++k) u = fma(sS[i][k], dall[k], u);
          X[NSM + i] = X[i] * u * u;
        }
        wave_sync();
        double t3 = 0.0;
#pragma unroll 4
        for (int i = 0; i < NSM; ++i) t3 = fma(sS[i][c], X[NSM + i], t3);
        t3 = livec ? (-0.5 * kLn10 * kLn10) * t3 : 0.0;
        const double e = chol_solve(t3);
        if (gall(isfinite(dnew + e), base)) dnew += e;
      }
      if (ok) {
        dc = dnew;
        const double mx = gmax(fabs(dc));
        if (mx > kMaxStep) dc *= kMaxStep / mx;
        dmax = fmin(mx, kMaxStep);
        trial = true;
        alpha = 1.0;
        // a first Newton step this long: the start is far from the solution
        if (pol && evals == 1 && mx > kPolishFar) leave_polish();
      } else if (pol) {
        leave_polish();  // no usable Newton direction: the full iteration from here
      } else {
        gmode = true;  // no usable direction: take the convex route from here
        trial = false;
      }
    }
    wave_sync();  // the exchange area is written again by the next evaluation
    if (!done) continue;

    // ---- finished voxel: write results, release the group
    have = false;
    if (c == 0) {
      my_solves += 1;
      my_conv += converged ? 1 : 0;
      my_iters += evals;
      if (sa.iters) sa.iters[item] = evals;
      if (sa.resid) sa.resid[item] = sqrt(judge2);
      if (sa.conv) sa.conv[item] = converged ? 1 : 0;
    }
    if (c < tb.nc) sa.logc[item * tb.nc + c] = xc;
  }
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

}  // namespace
}  // namespace clb
