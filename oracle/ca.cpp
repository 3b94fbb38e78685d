// ca.cpp -- FP64 CPU restatement of the CA biomass redistribution (test infrastructure
// only: loaded by tests/ through oracle.py, never by the product).
//
// Follows the reference's processors with the order-independent sweep the GPU defines:
//   pushExcessBiomass3D / halfPushExcessBiomass3D  src/complab3d_processors_part1.hh:432-894
//   updateLocalMaskNtotalLattices3D                src/complab3d_processors_part2.hh:99-210
//   updateAgeDistance3D                            src/complab3d_processors_part2.hh:718-842
//   createAgeDomain3D                              src/complab3d_processors_part3.hh:347-367
//   the sweep loop with its stop conditions        src/complab.cpp:1680-1731
// Each sweep decides all pushes from the state at its start (the reference reads its total
// lattice as it was before the sweep too), then moves the biomass. Neighbour orders come
// from a splitmix64-based counter hash of (seed, sweep, voxel, draw). Built with
// -ffp-contract=off so each product and sum rounds as written.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "mask_rule.h"

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

const double kThrd = 1e-12;      // thrd, src/complab3d_processors.hh:60
const double kThrdMask = 1e-14;  // COMPLAB_THRD, src/complab_functions.hh:64

uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t draw(uint64_t seed, uint64_t sweep, long n, int k) {
  return mix64(seed ^ mix64(sweep ^ mix64((uint64_t)n * 8ull + (uint64_t)k)));
}

struct Box {
  int nx, ny, nz;
  long N;
  long at(int x, int y, int z) const { return x + (long)nx * (y + (long)ny * z); }
  // face neighbour d in the reference's order +x, -x, +y, -y, +z, -z; x limited to
  // [x0, x1], y and z to the box (no wrap); -1 outside
  long nbr(long n, int d, int x0, int x1) const {
    const int x = (int)(n % nx), y = (int)((n / nx) % ny), z = (int)(n / ((long)nx * ny));
    static const int D[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    const int xx = x + D[d][0], yy = y + D[d][1], zz = z + D[d][2];
    if (xx < x0 || xx > x1 || yy < 0 || yy >= ny || zz < 0 || zz >= nz) return -1;
    return at(xx, yy, zz);
  }
};

struct Codes {
  int solid, bb;
  OrcMaskRule rule;
  bool wall(int m) const { return m == solid || m == bb; }
  bool pore(int m) const { return rule.is_pore(m); }
};

// updateLocalMaskNtotalLattices3D; returns the max total over the planes 1 <= x <= nx-2,
// adds mask flips to *flips. With several pore types every voxel decides from the mask
// before the pass.
double mask_total(const Box& g, const Codes& k, int nb, const double* b, uint8_t* mask, double* btot, double bthr,
                  bool fresh, long long* flips) {
  double mx = 0.0;
  bool nan = false;
  std::vector<uint8_t> snapshot;
  if (k.rule.npore > 1) snapshot.assign(mask, mask + g.N);
  const uint8_t* prev = k.rule.npore > 1 ? snapshot.data() : mask;
  for (long n = 0; n < g.N; ++n) {
    const int m = prev[n];
    if (!k.wall(m)) {
      double bm = 0.0;
      int label = 0;
      for (int i = 0; i < nb; ++i) {
        const double v = b[(long)i * g.N + n];
        if (v > kThrdMask) {
          bm += v;
          label += (i < 4) ? k.rule.bio[i] : 0;
        }
      }
      if (fresh || std::fabs(bm - btot[n]) > kThrdMask) btot[n] = bm;
      const int x = (int)(n % g.nx);
      if (x >= 1 && x <= g.nx - 2) {
        const int y = (int)((n / g.nx) % g.ny), z = (int)(n / ((long)g.nx * g.ny));
        const int nm = orc_mask_rule(k.rule, prev, g.nx, g.ny, g.nz, x, y, z, m, bm, label, bthr, k.solid, k.bb);
        if (nm != m) {
          mask[n] = (uint8_t)nm;
          ++*flips;
        }
      }
    }
    const int xn = (int)(n % g.nx);
    if (xn < 1 || xn > g.nx - 2) continue;  // the inlet / outlet planes take no part in the CA
    const double t = btot[n];
    if (t != t) nan = true;
    if (t > mx) mx = t;
  }
  return nan ? NAN : mx;
}

// One sweep's decisions: out[d*N + n] = biomass voxel n sends across face d.
void push_decide(const Box& g, const Codes& k, const uint8_t* mask, const double* btot, const int* age, double bmax,
                 int half, uint64_t seed, uint64_t sweep, double* out) {
  std::memset(out, 0, sizeof(double) * 6 * (size_t)g.N);
  for (long n = 0; n < g.N; ++n) {
    const int x = (int)(n % g.nx);
    const int m = mask[n];
    if (x < 1 || x > g.nx - 2 || k.wall(m)) continue;
    if (half && k.pore(m)) continue;
    const double T = btot[n];
    if (!(T > bmax)) continue;
    double rest = half ? T * 0.5 : T - bmax;
    if (!half && !(rest > kThrd)) continue;
    int dirs[6];
    long nbrs[6];
    int nk = 0;
    for (int d = 0; d < 6; ++d) {
      const long q = g.nbr(n, d, 1, g.nx - 2);
      if (q >= 0 && !k.wall(mask[q])) {
        dirs[nk] = d;
        nbrs[nk] = q;
        ++nk;
      }
    }
    if (nk == 0) continue;  // encased: the excess stays
    int order[6] = {0, 1, 2, 3, 4, 5};
    for (int i = nk - 1; i >= 1; --i) {
      const int j = (int)(draw(seed, sweep, n, i) % (uint64_t)(i + 1));
      std::swap(order[i], order[j]);
    }
    double* o = out;
    bool placed = false;
    for (int t = 0; t < nk && !placed; ++t) {
      const int s = order[t];
      const double room_used = k.pore(mask[nbrs[s]]) ? 0.0 : btot[nbrs[s]];
      if (room_used < bmax) {
        const double room = bmax - room_used;
        double part;
        if (rest > room) {
          part = room;
          rest -= room;
        } else {
          part = rest;
          rest = 0.0;
          placed = true;
        }
        o[(long)dirs[s] * g.N + n] += part;
      }
    }
    if (!placed) {
      int cand[6], nc = 0;
      for (int s = 0; s < nk; ++s) {
        const bool younger = half ? (age[n] >= age[nbrs[s]]) : (age[n] > age[nbrs[s]]);
        if (younger) cand[nc++] = dirs[s];
      }
      int d;
      if (nc > 1) d = cand[(int)(draw(seed, sweep, n, 6) % (uint64_t)nc)];
      else if (nc == 1) d = cand[0];
      else d = dirs[nk > 1 ? (int)(draw(seed, sweep, n, 7) % (uint64_t)nk) : 0];
      o[(long)d * g.N + n] += rest;
    }
  }
}

// Move the decided amounts, species by species in proportion to the sender's biomass.
void push_apply(const Box& g, const Codes& k, int nb, const std::vector<double>& b0, double* b, const uint8_t* mask,
                const double* btot, const double* out) {
  for (long n = 0; n < g.N; ++n) {
    if (k.wall(mask[n])) continue;
    for (int i = 0; i < nb; ++i) {
      const double bi = b0[(size_t)i * g.N + n];
      double v = bi;
      bool touched = false;
      for (int d = 0; d < 6; ++d) {
        const double o = out[(long)d * g.N + n];
        if (o != 0.0) {
          touched = true;
          const double s = bi / btot[n] * o;
          if (s > kThrd) v -= s;
        }
      }
      for (int d = 0; d < 6; ++d) {
        const long q = g.nbr(n, d, 0, g.nx - 1);
        if (q < 0) continue;
        const double o = out[(long)(d ^ 1) * g.N + q];  // q's send across its face towards n
        if (o == 0.0) continue;
        touched = true;
        const double s = b0[(size_t)i * g.N + q] / btot[q] * o;
        if (s > kThrd) v += s;
      }
      if (touched) b[(long)i * g.N + n] = v;
    }
  }
}

// One Jacobi pass of updateAgeDistance3D; returns the number of ages that changed.
long age_pass(const Box& g, const int* in, int* outa, const uint8_t* dpos, const double* btot, double bmax) {
  long changed = 0;
  for (long n = 0; n < g.N; ++n) {
    const int a = in[n];
    int na = a;
    const double B = btot[n];
    if (B > kThrdMask) {
      long nb[6];
      int nk = 0;
      for (int d = 0; d < 6; ++d) {
        const long q = g.nbr(n, d, 0, g.nx - 1);
        if (q >= 0 && dpos[q]) nb[nk++] = q;
      }
      if (a == 0) {
        na = 1;
      } else if (a == 1 && (B - bmax) > -kThrdMask) {
        bool any0 = false;
        for (int s = 0; s < nk; ++s) any0 = any0 || in[nb[s]] == 0;
        if (!any0) na = 2;
      } else {
        int cnt = 0;
        for (int s = 0; s < nk; ++s)
          if (in[nb[s]] >= a && (btot[nb[s]] - bmax) > -kThrdMask) ++cnt;
        if (cnt == nk) na = a + 1;
      }
    }
    outa[n] = na;
    changed += (na != a);
  }
  return changed;
}

// Age passes until nothing changes or `cap` passes; returns passes, negative at the cap.
int age_relax(const Box& g, int* age, const uint8_t* dpos, const double* btot, double bmax, int cap) {
  std::vector<int> tmp((size_t)g.N);
  int passes = 0;
  while (true) {
    const long ch = age_pass(g, age, tmp.data(), dpos, btot, bmax);
    std::memcpy(age, tmp.data(), sizeof(int) * (size_t)g.N);
    ++passes;
    if (ch == 0) return passes;
    if (passes >= cap) return -passes;
  }
}

}  // namespace

// createAgeDomain3D and the distance field's sign, from the mask.
// rule: {npore, pore[8], bio[4], sum_labels} of several material types, or null (pore_code).
ORC_API void orc_ca_init(long N, const uint8_t* mask, int solid, int bb, int pore, int* age, uint8_t* dpos,
                         const int* rule) {
  const OrcMaskRule r = OrcMaskRule::from(rule, pore, pore);
  for (long n = 0; n < N; ++n) {
    const int m = mask[n];
    const bool wall = m == solid || m == bb;
    age[n] = wall ? -1 : (r.is_pore(m) ? 0 : 1);
    dpos[n] = wall ? 0 : 1;
  }
}

// The sweep loop of src/complab.cpp:1680-1731. b: nb*N biomass (in/out); mask, age in/out;
// btot: N total biomass lattice (in/out; recomputed first); sweep0: index of the first
// sweep (keys the draws). rep[0] sweeps, [1] age passes, [2] status (0 done, 1 stalled,
// 2 sweep cap, 3 age cap, 4 non-finite), [3] mask flips; mx[0] max before, mx[1] after.
ORC_API void orc_ca_redistribute(int nx, int ny, int nz, int nb, double* b, uint8_t* mask, int* age,
                                 const uint8_t* dpos, double* btot, int solid, int bb, int pore, int biofilm,
                                 double bthr, double bmax, int half, uint64_t seed, uint64_t sweep0, long long* rep,
                                 double* mx, const int* rule) {
  const Box g{nx, ny, nz, (long)nx * ny * nz};
  const Codes k{solid, bb, OrcMaskRule::from(rule, pore, biofilm)};
  for (int i = 0; i < 4; ++i) rep[i] = 0;
  long long flips = 0;
  double m = mask_total(g, k, nb, b, mask, btot, bthr, true, &flips);
  mx[0] = mx[1] = m;
  rep[3] = flips;
  if (!std::isfinite(m)) {
    rep[2] = 4;
    return;
  }
  if (!(m - bmax > kThrd)) return;
  std::vector<double> out((size_t)6 * g.N), b0;
  double prev = m;
  int stall = 0;
  uint64_t sweep = sweep0;
  for (int whilecount = 0; m - bmax > kThrd; ++whilecount) {
    b0.assign(b, b + (size_t)nb * g.N);
    push_decide(g, k, mask, btot, age, bmax, half, seed, sweep, out.data());
    push_apply(g, k, nb, b0, b, mask, btot, out.data());
    ++sweep;
    ++rep[0];
    m = mask_total(g, k, nb, b, mask, btot, bthr, false, &flips);
    rep[3] = flips;
    mx[1] = m;
    if (m > prev - kThrd) {
      if (++stall >= 100) {
        rep[2] = 1;
        break;
      }
    } else {
      stall = 0;
    }
    prev = m;
    if (whilecount % 50 == 0) {
      const int p = age_relax(g, age, dpos, btot, bmax, 1000);
      rep[1] += p < 0 ? -p : p;
      if (p < 0) {
        rep[2] = 3;
        break;
      }
    }
    if (whilecount > 1000) {
      rep[2] = 2;
      break;
    }
  }
}

// Age passes alone (src/complab.cpp:1739 runs one); returns passes, negative at the cap.
// fdDiffusion3D::process (src/complab3d_processors_part2.hh:276-423), bdryGap 1, over the
// copy `c` into `b` (nb fields). The neighbour reads past the last y or z plane (made only
// when a voxel there has a wall below it) wrap, as the lattice does in y and z.
void fd_step(const Box& g, int nb, const double* c, double* b, const uint8_t* mask, double nu) {
  const int gap = 1;
  auto M = [&](int x, int y, int z) { return (int)mask[g.at(x, (y + g.ny) % g.ny, (z + g.nz) % g.nz)]; };
  for (int z = 0; z < g.nz; ++z)
    for (int y = 0; y < g.ny; ++y)
      for (int x = gap; x <= g.nx - 1 - gap; ++x) {
        if (!(M(x, y, z) > 1)) continue;
        int mxp = 0, mxn = 0, myp = 0, myn = 0, mzp = 0, mzn = 0;
        if (x == gap) mxp = M(x + 1, y, z);
        else if (x == g.nx - 1 - gap) mxn = M(x - 1, y, z);
        else { mxp = M(x + 1, y, z); mxn = M(x - 1, y, z); }
        if (y == 0) myp = M(x, y + 1, z);
        else if (y == g.ny - 1) myn = M(x, y - 1, z);
        else { myp = M(x, y + 1, z); myn = M(x, y - 1, z); }
        if (z == 0) mzp = M(x, y, z + 1);
        else if (z == g.nz - 1) mzn = M(x, y, z - 1);
        else { mzp = M(x, y, z + 1); mzn = M(x, y, z - 1); }
        for (int k = 0; k < nb; ++k) {
          const double* ck = c + (long)k * g.N;
          auto B = [&](int xx, int yy, int zz) { return ck[g.at(xx, (yy + g.ny) % g.ny, (zz + g.nz) % g.nz)]; };
          double b0 = B(x, y, z), bxp, bxn, byp, byn, bzp, bzn;
          if (x == gap || mxn < 2) { bxp = B(x + 1, y, z); bxn = b0; }
          else if (x == g.nx - 1 - gap || mxp < 2) { bxp = b0; bxn = B(x - 1, y, z); }
          else { bxp = B(x + 1, y, z); bxn = B(x - 1, y, z); }
          if (y == 0 || myn < 2) { byp = B(x, y + 1, z); byn = b0; }
          else if (y == g.ny - 1 || myp < 2) { byp = b0; byn = B(x, y - 1, z); }
          else { byp = B(x, y + 1, z); byn = B(x, y - 1, z); }
          if (z == 0 || mzn < 2) { bzp = B(x, y, z + 1); bzn = b0; }
          else if (z == g.nz - 1 || mzp < 2) { bzp = b0; bzn = B(x, y, z - 1); }
          else { bzp = B(x, y, z + 1); bzn = B(x, y, z - 1); }
          double* v[7] = {&b0, &bxp, &bxn, &byp, &byn, &bzp, &bzn};
          for (double* p : v)
            if (*p < kThrdMask) *p = 0.0;
          const double nbm = b0 + nu * ((bxp - 2 * b0 + bxn) + (byp - 2 * b0 + byn) + (bzp - 2 * b0 + bzn));
          if (nbm > kThrdMask) b[(long)k * g.N + g.at(x, y, z)] = nbm;
        }
      }
}

ORC_API int orc_ca_update_age(int nx, int ny, int nz, int* age, const uint8_t* dpos, const double* btot, double bmax,
                              int max_passes) {
  const Box g{nx, ny, nz, (long)nx * ny * nz};
  return age_relax(g, age, dpos, btot, bmax, max_passes);
}

// The FD biofilm solver's block (src/complab.cpp:1716-1731): mask and totals, the copy, one
// fdDiffusion3D step, mask and totals. rep[0] mask flips; mx[0] max total before, mx[1] after.
ORC_API void orc_fd_biomass(int nx, int ny, int nz, int nb, double* b, uint8_t* mask, double* btot, int solid, int bb,
                            int pore, int biofilm, double bthr, double nu, long long* rep, double* mx,
                            const int* rule) {
  const Box g{nx, ny, nz, (long)nx * ny * nz};
  const Codes k{solid, bb, OrcMaskRule::from(rule, pore, biofilm)};
  long long flips = 0;
  mx[0] = mask_total(g, k, nb, b, mask, btot, bthr, false, &flips);
  std::vector<double> copy(b, b + (long)nb * g.N);
  fd_step(g, nb, copy.data(), b, mask, nu);
  mx[1] = mask_total(g, k, nb, b, mask, btot, bthr, false, &flips);
  rep[0] = flips;
}
