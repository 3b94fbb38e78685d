// clb_workloads.cpp -- deterministic synthetic inputs (host code, no GPU needed):
// sphere-pack pore masks, random speciation tableaux, spatially correlated component
// totals, biomass seeding. Every generator is driven by SplitMix64 so a seed gives the
// same bytes on every machine.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/clb.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
  double normal() {                                                               // Box-Muller
    double u1 = u01(), u2 = u01();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

const int D19[18][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
                        {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1},
                        {1, 0, -1}, {-1, 0, 1}, {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};

// Damped Newton on the PCF residual (the same equations the device solver uses), for the
// tableau acceptance check only. Returns false if it does not converge.
bool host_solve(int nc, int ns, const double* S, const double* logK, const double* T, double* x, double* maxlogc) {
  for (int j = 0; j < nc; ++j) x[j] = std::log10(std::min(std::max(T[j], 1e-30), 10.0));
  std::vector<double> f(nc), J(nc * nc), d(nc), c(ns), xt(nc), ft(nc);
  auto eval = [&](const double* xx, double* ff, bool wantJ) {
    std::vector<double> sp(nc, 0.0), sn(nc, 0.0);
    for (int i = 0; i < ns; ++i) {
      double lc = logK[i];
      for (int j = 0; j < nc; ++j) lc += S[i * nc + j] * xx[j];
      lc = std::min(std::max(lc, -30.0), 1.0);
      c[i] = std::pow(10.0, lc);
      for (int j = 0; j < nc; ++j) {
        double s = S[i * nc + j];
        if (s > 0) sp[j] += s * c[i];
        else if (s < 0) sn[j] -= s * c[i];
      }
    }
    double n2 = 0;
    std::vector<double> rP(nc), rR(nc);
    for (int j = 0; j < nc; ++j) {
      double SR = std::max(sp[j], 1e-30), SP = std::max(T[j] + sn[j], 1e-30);
      ff[j] = std::log10(SP) - std::log10(SR);
      n2 += ff[j] * ff[j];
      rP[j] = 1.0 / SP;
      rR[j] = 1.0 / SR;
    }
    if (wantJ) {
      std::fill(J.begin(), J.end(), 0.0);
      for (int i = 0; i < ns; ++i)
        for (int j = 0; j < nc; ++j) {
          double s = S[i * nc + j];
          if (s == 0) continue;
          double a = -s * c[i] * (s < 0 ? rP[j] : rR[j]);
          for (int k = 0; k < nc; ++k) J[j * nc + k] += a * S[i * nc + k];
        }
    }
    return n2;
  };
  double n2 = eval(x, f.data(), true);
  for (int it = 0; it < 200; ++it) {
    if (std::sqrt(n2) < 1e-11) {
      double m = -1e300;
      for (int i = 0; i < ns; ++i) {
        double lc = logK[i];
        for (int j = 0; j < nc; ++j) lc += S[i * nc + j] * x[j];
        m = std::max(m, lc);
      }
      *maxlogc = m;
      return true;
    }
    // LU with partial pivoting
    std::vector<double> A(J), b(nc);
    for (int j = 0; j < nc; ++j) b[j] = -f[j];
    bool ok = true;
    for (int k = 0; k < nc && ok; ++k) {
      int p = k;
      for (int r = k + 1; r < nc; ++r)
        if (std::fabs(A[r * nc + k]) > std::fabs(A[p * nc + k])) p = r;
      if (std::fabs(A[p * nc + k]) < 1e-300) { ok = false; break; }
      if (p != k) {
        for (int cc = 0; cc < nc; ++cc) std::swap(A[k * nc + cc], A[p * nc + cc]);
        std::swap(b[k], b[p]);
      }
      for (int r = k + 1; r < nc; ++r) {
        double m = A[r * nc + k] / A[k * nc + k];
        for (int cc = k; cc < nc; ++cc) A[r * nc + cc] -= m * A[k * nc + cc];
        b[r] -= m * b[k];
      }
    }
    if (ok)
      for (int k = nc - 1; k >= 0; --k) {
        double s = b[k];
        for (int cc = k + 1; cc < nc; ++cc) s -= A[k * nc + cc] * d[cc];
        d[k] = s / A[k * nc + k];
      }
    double alpha = 1.0;
    if (ok) {
      double m = 0;
      for (int j = 0; j < nc; ++j) m = std::max(m, std::fabs(d[j]));
      if (m > 5.0) for (int j = 0; j < nc; ++j) d[j] *= 5.0 / m;
      bool acc = false;
      while (alpha >= 1.0 / 256) {
        for (int j = 0; j < nc; ++j) xt[j] = std::min(std::max(x[j] + alpha * d[j], -30.0), 1.0);
        double nt = eval(xt.data(), ft.data(), false);
        if (nt <= (1.0 - 2e-4 * alpha) * n2) { acc = true; break; }
        alpha *= 0.5;
      }
      if (!acc) ok = false;
    }
    if (!ok)
      for (int j = 0; j < nc; ++j) xt[j] = std::min(std::max(x[j] + f[j], -30.0), 1.0);
    for (int j = 0; j < nc; ++j) x[j] = xt[j];
    n2 = eval(x, f.data(), true);
  }
  return false;
}

}  // namespace

extern "C" {

int clb_gen_sphere_pack(int nx, int ny, int nz, double rmin, double rmax, double porosity, uint64_t seed,
                        int pore_code, int solid_code, int bb_code, uint8_t* mask) {
  if (!mask || nx < 4 || ny < 1 || nz < 1 || rmin <= 0 || rmax < rmin || porosity <= 0 || porosity > 1)
    return CLB_ERR_ARG;
  const long long N = (long long)nx * ny * nz;
  std::vector<uint8_t> solid(N, 0);
  auto idx = [&](int x, int y, int z) { return x + (long long)nx * (y + (long long)ny * z); };
  const long long interior = (long long)(nx - 2) * ny * nz;
  long long nsolid = 0;
  SplitMix64 rng(seed);
  const long long max_spheres = 100000000;
  for (long long sph = 0; sph < max_spheres && (double)(interior - nsolid) / (double)interior > porosity; ++sph) {
    const double cx = 1.0 + rng.u01() * (double)(nx - 3);
    const double cy = rng.u01() * (double)ny;
    const double cz = rng.u01() * (double)nz;
    const double r = rmin + (rmax - rmin) * rng.u01();
    const int ir = (int)std::ceil(r);
    const int x0 = std::max(1, (int)std::floor(cx - ir)), x1 = std::min(nx - 2, (int)std::ceil(cx + ir));
    for (int dz = -ir - 1; dz <= ir + 1; ++dz)
      for (int dy = -ir - 1; dy <= ir + 1; ++dy) {
        const int zz = (int)std::floor(cz) + dz, yy = (int)std::floor(cy) + dy;
        const double ddz = (double)zz - cz, ddy = (double)yy - cy;
        if (ddz * ddz + ddy * ddy > r * r) continue;
        const int wz = ((zz % nz) + nz) % nz, wy = ((yy % ny) + ny) % ny;
        for (int x = x0; x <= x1; ++x) {
          const double ddx = (double)x - cx;
          if (ddx * ddx + ddy * ddy + ddz * ddz <= r * r) {
            const long long n = idx(x, wy, wz);
            if (!solid[n]) { solid[n] = 1; ++nsolid; }
          }
        }
      }
  }
  // inlet / outlet planes duplicate the first and last interior slices
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) {
      solid[idx(0, y, z)] = solid[idx(1, y, z)];
      solid[idx(nx - 1, y, z)] = solid[idx(nx - 2, y, z)];
    }
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const long long n = idx(x, y, z);
        if (!solid[n]) { mask[n] = (uint8_t)pore_code; continue; }
        bool adj = false;
        for (int q = 0; q < 18 && !adj; ++q) {
          const int xx = x + D19[q][0];
          if (xx < 0 || xx >= nx) continue;
          const int yy = (y + D19[q][1] + ny) % ny, zz = (z + D19[q][2] + nz) % nz;
          if (!solid[idx(xx, yy, zz)]) adj = true;
        }
        mask[n] = (uint8_t)(adj ? bb_code : solid_code);
      }
  return CLB_OK;
}

int clb_gen_tableau(int nc, int ns, double logk_min, double logk_max, int smin, int smax, uint64_t seed,
                    double log_t_min, double log_t_max, double* S, double* logK) {
  if (!S || !logK || nc < 1 || nc > CLB_MAX_COMPONENTS || ns < nc || ns > CLB_MAX_SPECIES || smax < smin)
    return CLB_ERR_ARG;
  SplitMix64 rng(seed);
  std::fill(S, S + ns * nc, 0.0);
  for (int j = 0; j < nc; ++j) { S[j * nc + j] = 1.0; logK[j] = 0.0; }
  // Species rows are drawn one at a time; a row is redrawn when it is zero, duplicates an
  // earlier row, or makes the tableau leave the clamp-free region: the solutions at the
  // corners of the totals box (plus random interior points) must converge and keep every
  // species below 1 M (log c <= 0; the solver clamps at 10 M).
  std::vector<double> T(nc), x(nc);
  const int ncorner = 1 << std::min(nc, 8);
  for (int i = nc; i < ns; ++i) {
    bool accepted = false;
    for (int attempt = 0; attempt < 5000 && !accepted; ++attempt) {
      bool nz = false;
      for (int j = 0; j < nc; ++j) {
        const int v = smin + (int)(rng.next() % (uint64_t)(smax - smin + 1));
        S[i * nc + j] = (double)v;
        if (v != 0) nz = true;
      }
      logK[i] = logk_min + (logk_max - logk_min) * rng.u01();
      bool dup = !nz;
      for (int k = 0; k < i && !dup; ++k) {
        bool same = true;
        for (int j = 0; j < nc; ++j)
          if (S[k * nc + j] != S[i * nc + j]) { same = false; break; }
        dup = same;
      }
      if (dup) continue;
      bool ok = true;
      for (int q = 0; q < ncorner + 16 && ok; ++q) {
        for (int j = 0; j < nc; ++j) {
          double lt;
          if (q < ncorner) lt = ((q >> j) & 1) ? log_t_max : log_t_min;
          else lt = log_t_min + (log_t_max - log_t_min) * rng.u01();
          T[j] = std::pow(10.0, lt);
        }
        double mx;
        if (!host_solve(nc, i + 1, S, logK, T.data(), x.data(), &mx) || mx > 0.0) ok = false;
      }
      accepted = ok;
    }
    if (!accepted) return CLB_ERR_UNSUP;
  }
  return CLB_OK;
}

int clb_gen_totals_field(int nx, int ny, int nz, int nc, double sigma, double log_t_min, double log_t_max,
                         uint64_t seed, double* totals) {
  if (!totals || nx < 1 || ny < 1 || nz < 1 || nc < 1 || sigma < 0) return CLB_ERR_ARG;
  const long long N = (long long)nx * ny * nz;
  SplitMix64 rng(seed);
  const int rad = (int)std::ceil(4.0 * sigma);
  std::vector<double> ker(2 * rad + 1);
  double ks = 0;
  for (int k = -rad; k <= rad; ++k) {
    ker[k + rad] = (sigma > 0) ? std::exp(-0.5 * (double)k * k / (sigma * sigma)) : (k == 0 ? 1.0 : 0.0);
    ks += ker[k + rad];
  }
  for (double& v : ker) v /= ks;
  std::vector<double> a(N), b(N);
  for (int j = 0; j < nc; ++j) {
    for (long long n = 0; n < N; ++n) a[n] = rng.normal();
    // periodic separable blur along x, y, z
    for (int dim = 0; dim < 3; ++dim) {
      const int L = dim == 0 ? nx : (dim == 1 ? ny : nz);
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x) {
            double s = 0;
            const int c0 = dim == 0 ? x : (dim == 1 ? y : z);
            for (int k = -rad; k <= rad; ++k) {
              const int c = ((c0 + k) % L + L) % L;
              const int xx = dim == 0 ? c : x, yy = dim == 1 ? c : y, zz = dim == 2 ? c : z;
              s += ker[k + rad] * a[xx + (long long)nx * (yy + (long long)ny * zz)];
            }
            b[x + (long long)nx * (y + (long long)ny * z)] = s;
          }
      std::swap(a, b);
    }
    double mean = 0, var = 0;
    for (long long n = 0; n < N; ++n) mean += a[n];
    mean /= (double)N;
    for (long long n = 0; n < N; ++n) var += (a[n] - mean) * (a[n] - mean);
    const double sd = std::sqrt(var / (double)N);
    for (long long n = 0; n < N; ++n) {
      const double zs = (sd > 0) ? (a[n] - mean) / sd : 0.0;
      const double u = 0.5 * std::erfc(-zs / std::sqrt(2.0));
      totals[n * nc + j] = std::pow(10.0, log_t_min + (log_t_max - log_t_min) * u);
    }
  }
  return CLB_OK;
}

int clb_gen_biomass_seed(long N, const uint8_t* mask, int pore_code, double fraction, double bmin, double bmax,
                         uint64_t seed_sites, uint64_t seed_density, double* b) {
  if (!mask || !b || N < 1 || fraction < 0 || fraction > 1) return CLB_ERR_ARG;
  std::vector<long long> pore;
  for (long long n = 0; n < N; ++n) {
    b[n] = 0.0;
    if (mask[n] == pore_code) pore.push_back(n);
  }
  const long long k = (long long)std::llround(fraction * (double)pore.size());
  SplitMix64 r1(seed_sites), r2(seed_density);
  for (long long i = 0; i < k; ++i) {  // partial Fisher-Yates
    const long long j = i + (long long)(r1.next() % (uint64_t)(pore.size() - i));
    std::swap(pore[i], pore[j]);
  }
  std::sort(pore.begin(), pore.begin() + k);
  for (long long i = 0; i < k; ++i) b[pore[i]] = bmin + (bmax - bmin) * r2.u01();
  return CLB_OK;
}

}  // extern "C"
