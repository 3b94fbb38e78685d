// surrogate.cpp -- FP64 CPU restatement of the surrogate metabolic processor.
//
// TEST INFRASTRUCTURE ONLY: part of the oracle library (liboracle.so), loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
//
// Restates, per voxel:
//   * run_surrogate3D::process (src/complab3d_processors_surrogate.hh:95-274): present
//     microbes, Michaelis-Menten uptake estimates with their caps, the per-organism model
//     call, growth / decay increments, summed and clamped substrate draws, the deposit test;
//   * MetabolicConfig's unit conversions (src/complab3d_metabolic.hh:478-520) and
//     MetabolicTotals build / total / drawDown (src/complab3d_metabolic.hh:222-382);
//   * the model: defineSurrogateModel's MATLAB feed-forward evaluation
//     (surrogateModel.hh:396-427: mapminmax, matMultp + matSum + tansig per hidden layer,
//     linear output, reverse mapminmax, the 1e-8 and non-finite guards), with the network
//     as data, and the two stoichiometric closures its comment block spells out
//     (surrogateModel.hh:443-488).
// Written in the reference's own shape (vectors per voxel, Fin / Fout indexed by the global
// microbe id, one model call per present organism); nothing here is shared with the HIP
// path. Pinned by closed forms in tests/test_oracle_surrogate.py.
#include <cmath>
#include <utility>
#include <vector>

namespace {

const double kThrd = 1e-12;  // thrd (src/complab3d_processors.hh:60)

}  // namespace

extern "C" {

// One microbe's network (row-major W per layer, then b; width[-1] = n_in). n_layers = 0:
// no network (the model leaves bioR at 0).
struct orc_srg_net {
  int n_in;
  const int* in_sub;
  int n_layers;
  const int* width;
  const double* params;
  const double* x_offset;
  const double* x_gain;
  double x_ymin, y_ymin, y_gain, y_offset;
  int log_output;
  int closure;           // 0 none, 1 consume by yield, 2 growth limited by yield
  const double* yield;   // [S] or null
};

struct orc_srg {
  int n;                 // surrogate microbes
  const int* id;         // [n] global microbe id = biomass field
  const double* Kc;      // [n*S]
  const double* decay;   // [n] 1/s
  const double* vmax;    // [n*S] 0 = unset
  const double* maxup;   // [n*S] floor, <= -1e29 unset
  const double* mw;      // [n] gDW/mol
  const int* fixc;       // [S]
  const int* fixlb;      // [S]
  int use_totals;
  int nc;                // tableau for the totals (0: none)
  const double* eqS;     // [S*nc] species x component
  const int* eqCompSpecies;  // [nc]
  const orc_srg_net* nets;   // [n]
};

}  // extern "C"

namespace {

std::vector<double> mat_multp(const std::vector<std::vector<double>>& m, const std::vector<double>& v) {
  std::vector<double> r(m.size(), 0.0);
  for (size_t i = 0; i < m.size(); ++i) {
    r[i] = 0.0;
    for (size_t j = 0; j < v.size(); ++j) r[i] += m[i][j] * v[j];
  }
  return r;
}

std::vector<double> mat_sum(const std::vector<double>& a, const std::vector<double>& b) {
  std::vector<double> r(a.size());
  for (size_t i = 0; i < a.size(); ++i) r[i] = a[i] + b[i];
  return r;
}

std::vector<double> tansig(const std::vector<double>& n) {
  std::vector<double> r(n.size());
  for (size_t i = 0; i < n.size(); ++i) r[i] = 2.0 / (1.0 + std::exp(-2.0 * n[i])) - 1.0;
  return r;
}

double net_eval(const orc_srg_net& nt, const std::vector<double>& fin) {
  std::vector<double> a(nt.n_in);
  for (int k = 0; k < nt.n_in; ++k) a[k] = (fin[nt.in_sub[k]] - nt.x_offset[k]) * nt.x_gain[k] + nt.x_ymin;
  size_t off = 0;
  int prev = nt.n_in;
  double out = 0.0;
  for (int l = 0; l < nt.n_layers; ++l) {
    const int w = nt.width[l];
    std::vector<std::vector<double>> W(w, std::vector<double>(prev));
    for (int r = 0; r < w; ++r)
      for (int c = 0; c < prev; ++c) W[r][c] = nt.params[off + (size_t)r * prev + c];
    off += (size_t)w * prev;
    std::vector<double> b(nt.params + off, nt.params + off + w);
    off += w;
    if (l < nt.n_layers - 1) {
      a = tansig(mat_sum(b, mat_multp(W, a)));
    } else {
      // linear output neuron: b5 + vecMultp(LW, a)
      double d = 0.0;
      for (int c = 0; c < prev; ++c) d += W[0][c] * a[c];
      out = b[0] + d;
    }
    prev = w;
  }
  double g = (out - nt.y_ymin) / nt.y_gain + nt.y_offset;
  if (nt.log_output) g = std::pow(10.0, g);
  if (g < 1e-8) g = 0.0;
  if (std::isnan(g) || std::isinf(g)) g = 0.0;
  return g;
}

// defineSurrogateModel for one organism: bioR[id] = growth; Fout row by the closure.
void model(const orc_srg_net& nt, int id, int S, const std::vector<std::vector<double>>& Fin,
           std::vector<std::vector<double>>& Fout, std::vector<double>& bioR) {
  if (nt.n_layers == 0) return;
  const double g = net_eval(nt, Fin[id]);
  bioR[id] = g;
  if (nt.closure == 1) {
    for (int s = 0; s < S; ++s) {
      if (!(nt.yield[s] > 0.0)) continue;
      double needed = g / nt.yield[s];
      if (needed > Fin[id][s]) needed = Fin[id][s];
      Fout[id][s] = needed;
    }
  } else if (nt.closure == 2) {
    double gl = g;
    for (int s = 0; s < S; ++s) {
      if (!(nt.yield[s] > 0.0)) continue;
      const double gmax = Fin[id][s] * nt.yield[s];
      if (gmax < gl) gl = gmax;
    }
    bioR[id] = gl;
    for (int s = 0; s < S; ++s)
      if (nt.yield[s] > 0.0) Fout[id][s] = gl / nt.yield[s];
  }
}

struct Totals {
  bool active = false;
  std::vector<std::vector<std::pair<int, double>>> contrib;

  void build(const orc_srg& P, int S) {
    contrib.assign(S, {});
    active = false;
    if (!P.use_totals || P.nc <= 0) return;
    int mapped = 0;
    for (int j = 0; j < P.nc; ++j) {
      const int sidx = P.eqCompSpecies[j];
      if (sidx < 0 || sidx >= S) continue;
      for (int i = 0; i < S; ++i) {
        const double cf = P.eqS[(size_t)i * P.nc + j];
        if (cf > 1e-12 || cf < -1e-12) contrib[sidx].push_back({i, cf});
      }
      if (!contrib[sidx].empty()) ++mapped;
    }
    active = mapped > 0;
  }
  double raw_total(const std::vector<double>& c, int s) const {
    double t = 0.0;
    for (auto& e : contrib[s]) t += e.second * c[e.first];
    return t;
  }
  double abs_scale(const std::vector<double>& c, int s) const {
    double t = 0.0;
    for (auto& e : contrib[s]) t += std::fabs(e.second * c[e.first]);
    return t;
  }
  bool usable(const std::vector<double>& c, int s) const {
    if (!active || contrib[s].empty()) return false;
    const double t = raw_total(c, s);
    return std::fabs(t) > 1e-6 * abs_scale(c, s) && std::fabs(t) > 1e-30;
  }
  double total(const std::vector<double>& c, int s) const {
    return usable(c, s) ? std::fabs(raw_total(c, s)) : c[s];
  }
  void draw_down(const std::vector<double>& c, int s, double amount, std::vector<double>& dC) const {
    if (!usable(c, s)) {
      double take = amount;
      const double remaining = c[s] + dC[s];
      if (take > remaining) take = (remaining > 0.0) ? remaining : 0.0;
      dC[s] -= take;
      return;
    }
    const double tot = raw_total(c, s);
    for (auto& e : contrib[s]) {
      const int i = e.first;
      const double cf = e.second;
      if (!(c[i] > 0.0)) continue;
      const double share = (cf * c[i]) / tot;
      double take = amount * share / cf;
      const double remaining = c[i] + dC[i];
      if (take > remaining) take = (remaining > 0.0) ? remaining : 0.0;
      dC[i] -= take;
    }
  }
};

double exhaustion_flux(const orc_srg& P, int k, double C, double B, double dt) {
  const double denom = B * P.mw[k] * dt;
  if (!(denom > 1e-30)) return 0.0;
  double v = -C * 3.6e6 / denom;
  if (v < -1e4) v = -1e4;
  return v;
}

}  // namespace

extern "C" {

// The network alone (for the oracle's own tests).
__attribute__((visibility("default"))) double orc_srg_net_eval(const orc_srg_net* nt, const double* fin, int S) {
  std::vector<double> f(fin, fin + S);
  return net_eval(*nt, f);
}

// One voxel of run_surrogate3D. conc[S] / bmass[nb]: the post-collision concentrations
// and biomass fields as the lattices hold them. dC[S], dB[nb]: the deposited increments
// (only those above 1e-12, as the processor writes them into the dC / dB lattices).
__attribute__((visibility("default"))) void orc_srg_voxel(const orc_srg* P, int S, int nb, const double* conc_in,
                                                          const double* bmass_in, double dt, double* dC_out,
                                                          double* dB_out) {
  for (int s = 0; s < S; ++s) dC_out[s] = 0.0;
  for (int b = 0; b < nb; ++b) dB_out[b] = 0.0;
  const int n = P->n;
  int total_microbes = nb;
  for (int k = 0; k < n; ++k)
    if (P->id[k] + 1 > total_microbes) total_microbes = P->id[k] + 1;
  Totals totals;
  totals.build(*P, S);

  std::vector<double> conc(S), avail(S), dC(S, 0.0), bmass(n), dB(n, 0.0);
  std::vector<int> bloc;
  for (int k = 0; k < n; ++k) {
    const double b = bmass_in[P->id[k]];
    bmass[k] = (b > kThrd) ? b : 0.0;
    if (bmass[k] > kThrd) bloc.push_back(k);
  }
  if (bloc.empty()) return;
  for (int s = 0; s < S; ++s) conc[s] = (conc_in[s] > kThrd) ? conc_in[s] : 0.0;
  for (int s = 0; s < S; ++s) avail[s] = P->use_totals ? totals.total(conc, s) : conc[s];

  std::vector<std::vector<double>> Fin(total_microbes, std::vector<double>(S, 0.0)), Fout;
  std::vector<double> bioR(total_microbes, 0.0);
  for (int k : bloc) {
    const int gm = P->id[k];
    for (int s = 0; s < S; ++s) {
      const double kc = P->Kc[k * S + s];
      const double mm = (kc + avail[s] > 0.0) ? avail[s] / (kc + avail[s]) : 0.0;
      double v;
      if (P->vmax[k * S + s] > kThrd) v = P->vmax[k * S + s] * mm;
      else v = -exhaustion_flux(*P, k, avail[s], bmass[k], dt) * mm;
      const double cap = -exhaustion_flux(*P, k, avail[s], bmass[k], dt);
      if (cap > 0.0 && v > cap) v = cap;
      const double floor_mag = -P->maxup[k * S + s];
      if (P->maxup[k * S + s] > -1e29 && floor_mag >= 0.0 && v > floor_mag) v = floor_mag;
      if (P->fixlb[s]) {
        const double av_cap = -exhaustion_flux(*P, k, avail[s], bmass[k], dt);
        if (av_cap >= 0.0 && v > av_cap) v = av_cap;
      }
      Fin[gm][s] = v;
    }
  }
  Fout = Fin;
  for (int k : bloc) model(P->nets[k], P->id[k], S, Fin, Fout, bioR);
  for (int k : bloc) {
    const double mu = bioR[P->id[k]];
    dB[k] = (std::fabs(mu) < kThrd) ? -P->decay[k] * bmass[k] * dt   // decayToDeltaB
                                    : mu * bmass[k] * dt / 3600.0;    // growthToDeltaB
    if (dB[k] + bmass[k] < 0.0) dB[k] = -bmass[k];
  }
  for (int s = 0; s < S; ++s) {
    if (P->fixc[s]) continue;
    double draw = 0.0;
    for (int k : bloc) draw += -(Fout[P->id[k]][s] * (bmass[k] * P->mw[k]) * dt / 3.6e6);  // fluxToDeltaC
    if (std::fabs(draw) <= kThrd) continue;
    const double headroom = avail[s] + (P->use_totals ? 0.0 : dC[s]);
    if (draw + headroom < 0.0) draw = -headroom;
    if (draw > 0.0) draw = 0.0;
    if (P->use_totals) totals.draw_down(conc, s, -draw, dC);
    else dC[s] += draw;
  }
  for (int s = 0; s < S; ++s)
    if (std::fabs(dC[s]) > kThrd) dC_out[s] = dC[s];
  for (int k = 0; k < n; ++k)
    if (std::fabs(dB[k]) > kThrd) dB_out[P->id[k]] = dB[k];
}

}  // extern "C"
