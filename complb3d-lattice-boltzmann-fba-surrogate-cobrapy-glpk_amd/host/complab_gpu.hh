// complab_gpu.hh -- C++ host interface of the MI355X path, over the C ABI in clb.h.
//
// Two pieces, mirroring what CompLaB3D's C++ driver holds today:
//
//  * complab_gpu::EquilibriumChemistry<T> has the public surface of the reference's
//    EquilibriumChemistry<T> (src/complab3d_processors_part4_eqsolver.hh:187-1320): the same
//    setters, getters, defaults (maxIterations 200, tolerance 1e-8) and the same
//    calculate_species_concentrations(std::vector<T>) contract -- totals from the clamped
//    input (calc_component_totals, :520-538), the component-slot initial guess (:977-994),
//    species from the mass-action law with the [-30, 1] log clamp (calc_species, :442-496),
//    NaN / inf replaced by the clamped input and a final clamp to [1e-30, 10] (:1187-1214).
//    The solve runs on the GPU's FP64 Newton kernel, judged by the reference's PCF residual
//    (2-norm < tolerance). A batched overload solves many voxels in one launch.
//    Like the reference, it never throws from a solve: a failed device call leaves the
//    clamped input in place, counts as not converged, and getLastError() says why.
//
//  * complab_gpu::Lattice owns one clb_ctx (one GPU's share of the domain) and replaces the
//    MultiBlockLattice3D objects and processor calls of the driver loop (src/complab.cpp:
//    592-1860) with calls on the device-resident state. Its methods throw complab_gpu::Error
//    on any CLB_ERR_* return, carrying clb_last_error().
//
//  * complab_gpu::readGeometry / writeVTI: the geometry file the driver reads and the VTK
//    ImageData files it writes (host only; they need no GPU).
//
// Header-only; link with -lclb. Requires a GPU at construction: there is no CPU fallback.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/clb.h"

namespace complab_gpu {

typedef long long plint;

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace detail {
inline void check(int rc, const clb_ctx* c, const char* what) {
  if (rc != CLB_OK)
    throw Error(rc, std::string(what) + ": " + (c ? clb_last_error(c) : "context creation failed") + " (code " +
                        std::to_string(rc) + ")");
}
inline clb_ctx* create(int device, int precision, int nx, int ny, int nz) {
  clb_config cfg{device, precision, nx, ny, nz};
  clb_ctx* c = nullptr;
  const int rc = clb_create(&cfg, &c);
  if (rc != CLB_OK) {
    std::string msg = c ? clb_last_error(c) : "no usable HIP device";
    if (c) clb_destroy(c);
    throw Error(rc, "clb_create: " + msg);
  }
  return c;
}
}  // namespace detail

// ------------------------------------------------------------------------------------------
// ValueTracer: Palabos' util::ValueTracer (Palabos v2.3.0, src/core/util.h; not in the
// reference tree), restated from its published behaviour, as the driver's flow loops use it
// (src/complab.cpp:593, 609-610, 1149, 1797-1798): ValueTracer(u, L, epsilon) keeps the last
// deltaT = (int)(L / u / 2) values; converged once the window is full and |stddev / mean| <
// epsilon (stddev with n - 1), or when that ratio is NaN (a diverged run). Host only.
class ValueTracer {
 public:
  ValueTracer(double u, double L, double epsilon) : deltaT_((int)(L / u / 2.)), epsilon_(epsilon) {}
  int deltaT() const { return deltaT_; }
  int size() const { return (int)values_.size(); }
  void takeValue(double v) {
    values_.push_back(v);
    if ((int)values_.size() > std::abs(deltaT_)) values_.erase(values_.begin());
    ++t_;
  }
  void resetValues() {
    values_.clear();
    t_ = 0;
  }
  bool hasConverged() const {
    if ((int)values_.size() < std::abs(deltaT_)) return false;
    const double avg = average();
    const double ratio = stdDev(avg) / avg;
    if (std::isnan(ratio)) return true;
    return std::fabs(ratio) < epsilon_;
  }
  double average() const {
    double s = 0.0;
    for (double v : values_) s += v;
    return s / (double)values_.size();
  }
  double stdDev(double avg) const {
    double s = 0.0;
    for (double v : values_) s += (v - avg) * (v - avg);
    return std::sqrt(s / (double)(values_.size() - 1));
  }

 private:
  int deltaT_;
  double epsilon_;
  std::vector<double> values_;
  long long t_ = 0;
};

// ------------------------------------------------------------------------------------------
// Lattice: the device-resident state of one coupled simulation.
class Lattice {
 public:
  Lattice(int nx, int ny, int nz, int precision = CLB_PREC_FP32, int device = 0)
      : c_(detail::create(device, precision, nx, ny, nz)), nx_(nx), ny_(ny), nz_(nz) {}
  ~Lattice() {
    if (c_) clb_destroy(c_);
  }
  Lattice(const Lattice&) = delete;
  Lattice& operator=(const Lattice&) = delete;
  Lattice(Lattice&& o) noexcept : c_(o.c_), nx_(o.nx_), ny_(o.ny_), nz_(o.nz_) { o.c_ = nullptr; }

  clb_ctx* ctx() const { return c_; }
  long long nodes() const { return (long long)nx_ * ny_ * nz_; }
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  int nz() const { return nz_; }

  // geometry and dynamics (readGeometry + NSdomainSetup / soluteDomainSetup material map)
  void uploadMask(const std::vector<uint8_t>& m) { need(m.size(), nodes(), "mask"); chk(clb_upload_mask(c_, m.data()), "clb_upload_mask"); }
  std::vector<uint8_t> mask() const {
    std::vector<uint8_t> m((size_t)nodes());
    chk(clb_download_mask(c_, m.data()), "clb_download_mask");
    return m;
  }
  void setDynamics(const std::array<int, 256>& kind, const std::array<double, 256>& flowOmega) {
    chk(clb_set_dynamics(c_, kind.data(), flowOmega.data()), "clb_set_dynamics");
  }
  // soluteDomainSetup's own map: every biofilm code is advection-diffusion for the solutes,
  // whether or not the flow sees it as bounce-back
  void setSoluteDynamics(const std::array<int, 256>& kind) {
    chk(clb_set_solute_dynamics(c_, kind.data()), "clb_set_solute_dynamics");
  }

  // D3Q19 flow
  void flowInit(double rhoIn, double rhoOut) { chk(clb_flow_init(c_, rhoIn, rhoOut), "clb_flow_init"); }
  void flowSteps(int n, std::vector<double>* energy = nullptr) {
    if (energy) energy->assign((size_t)n, 0.0);
    chk(clb_flow_steps(c_, n, energy ? energy->data() : nullptr), "clb_flow_steps");
  }
  std::vector<double> velocity() const {
    std::vector<double> u((size_t)(3 * nodes()));
    chk(clb_flow_download_velocity(c_, u.data(), nullptr), "clb_flow_download_velocity");
    return u;
  }
  std::array<double, 4> flowStats() const {
    std::array<double, 4> s{};
    chk(clb_flow_stats(c_, s.data()), "clb_flow_stats");
    return s;
  }
  // A flow solve of the driver's loops (src/complab.cpp:607-620, 1794-1800, 1836-1840):
  // collideAndStream, the stored average energy into the tracer, stop once it has converged,
  // at most maxSteps; the tracer keeps its window across calls, as ns_convg2 does. Steps
  // that cannot end the solve (window not yet full) run as one batch. Returns the steps run;
  // *converged tells whether the tracer stopped it.
  long long flowResolve(ValueTracer& tracer, long long maxSteps, bool* converged = nullptr) {
    long long done = 0;
    std::vector<double> e;
    while (done < maxSteps) {
      const long long k = std::min<long long>(std::max(1, tracer.deltaT() - tracer.size()), maxSteps - done);
      flowSteps((int)k, &e);
      for (double v : e) tracer.takeValue(v);
      done += k;
      if (tracer.hasConverged()) {
        if (converged) *converged = true;
        return done;
      }
    }
    if (converged) *converged = false;
    return done;
  }
  // The same solves with the tracer on the device (clb_flow_tracer / clb_flow_resolve): no
  // host round trip per step. deviceTracer starts a tracer with an empty window; each
  // flowResolveDevice feeds it (the window persists across calls, as ns_convg2's does).
  void deviceTracer(double u, double L, double epsilon) {
    detail::check(clb_flow_tracer(c_, u, L, epsilon), c_, "clb_flow_tracer");
  }
  long long flowResolveDevice(int maxSteps, bool* converged = nullptr) {
    int steps = 0, conv = 0;
    detail::check(clb_flow_resolve(c_, maxSteps, &steps, &conv, nullptr), c_, "clb_flow_resolve");
    if (converged) *converged = conv != 0;
    return steps;
  }
  // The percolation test after a re-solve that used all its steps (src/complab.cpp:1801-1805):
  // the mean x velocity over the plane x = nx-2 above thrd.
  bool percolates(double thrd = 1e-12) const { return flowStats()[2] > thrd; }

  // D3Q7 solutes and reactions
  void setSolutes(int S, const std::vector<double>& omega, const std::vector<int>& bcType,
                  const std::vector<double>& bcValue) {
    need(omega.size(), (long long)S * 256, "omega");
    need(bcType.size(), 2LL * S, "bc_type");
    need(bcValue.size(), 2LL * S, "bc_value");
    chk(clb_set_solutes(c_, S, omega.data(), bcType.data(), bcValue.data()), "clb_set_solutes");
    S_ = S;
  }
  void setImmobile(const std::vector<int>& flags) { chk(clb_set_immobile(c_, flags.data()), "clb_set_immobile"); }
  void solutesInit(const std::vector<double>& c0) { chk(clb_solutes_init(c_, c0.data()), "clb_solutes_init"); }
  void solutesReset(const std::vector<double>& c0) { chk(clb_solutes_reset(c_, c0.data()), "clb_solutes_reset"); }
  void coupleVelocity() { chk(clb_couple_velocity(c_), "clb_couple_velocity"); }
  void setVelocity(const std::vector<double>& u) { need(u.size(), 3 * nodes(), "u"); chk(clb_set_velocity(c_, u.data()), "clb_set_velocity"); }
  void setBiomass(int nb, const std::vector<double>& b) {
    need(b.size(), nb * nodes(), "biomass");
    chk(clb_set_biomass(c_, nb, b.data()), "clb_set_biomass");
    nb_ = nb;
  }
  void setKinetics(int kind, double dt, int solidCode, int bbCode) {
    chk(clb_set_kinetics(c_, kind, dt, solidCode, bbCode), "clb_set_kinetics");
  }
  void setAbiotic(int mode) { chk(clb_set_abiotic(c_, mode), "clb_set_abiotic"); }
  void setTableau(int nc, int ns, const std::vector<double>& S, const std::vector<double>& logK,
                  const std::vector<int>& compSpecies, double tol, int maxIter) {
    chk(clb_set_tableau(c_, nc, ns, S.data(), logK.data(), compSpecies.data(), tol, maxIter), "clb_set_tableau");
  }
  void setMaskUpdate(bool enable, double threshold, int poreCode, int biofilmCode) {
    chk(clb_set_mask_update(c_, enable ? 1 : 0, threshold, poreCode, biofilmCode), "clb_set_mask_update");
  }
  // updateLocalMaskNtotalLattices3D with several material types: pore_dynamics and
  // bio_dynamics[i][0] of each microbe (after setMaskUpdate)
  void setMaskCodes(const std::vector<int>& poreCodes, const std::vector<int>& bioCodes) {
    chk(clb_set_mask_codes(c_, (int)poreCodes.size(), poreCodes.data(), (int)bioCodes.size(),
                           bioCodes.empty() ? nullptr : bioCodes.data()),
        "clb_set_mask_codes");
  }
  void setPhase(const std::vector<uint8_t>& ph) { need(ph.size(), nodes(), "phase"); chk(clb_set_phase(c_, ph.data()), "clb_set_phase"); }
  long long nodeConversion(int mineralSolute, int poreCode, int precipCode, int phaseId, double precipMax,
                           double reopen) {
    long long changed = 0;
    chk(clb_node_conversion(c_, mineralSolute, poreCode, precipCode, phaseId, precipMax, reopen, &changed),
        "clb_node_conversion");
    return changed;
  }
  // Surrogate metabolic processor (run_surrogate3D, src/complab3d_processors_surrogate.hh:
  // 74-295): the microbes it grows (by biomass field) with their uptake settings, then each
  // one's network. Per-microbe tables are n x S, row per microbe; empty vectors take the
  // defaults (no Vmax, no floor, molar mass 24.6, no held substrates).
  struct SurrogateNet {
    int nIn = 0;
    std::vector<int> inSub;         // input k = Fin[id][inSub[k]]
    std::vector<int> width;         // neurons per weight layer, the last 1
    std::vector<double> params;     // per layer W (row per neuron) then b
    std::vector<double> xOffset, xGain;
    double xYmin = -1.0, yYmin = -1.0, yGain = 1.0, yOffset = 0.0;
    bool logOutput = false;
  };
  void setSurrogate(const std::vector<int>& ids, const std::vector<double>& Kc, const std::vector<double>& decay,
                    const std::vector<double>& vmax = {}, const std::vector<double>& maxUptake = {},
                    const std::vector<double>& molarMass = {}, const std::vector<int>& fixConc = {},
                    const std::vector<int>& fixLb = {}, bool useTotals = false) {
    const long long n = (long long)ids.size();
    need(Kc.size(), n * S_, "Kc");
    need(decay.size(), n, "decay");
    if (!vmax.empty()) need(vmax.size(), n * S_, "vmax");
    if (!maxUptake.empty()) need(maxUptake.size(), n * S_, "maxUptake");
    if (!molarMass.empty()) need(molarMass.size(), n, "molarMass");
    if (!fixConc.empty()) need(fixConc.size(), S_, "fixConc");
    if (!fixLb.empty()) need(fixLb.size(), S_, "fixLb");
    auto p = [](const auto& v) { return v.empty() ? nullptr : v.data(); };
    chk(clb_set_surrogate(c_, (int)n, p(ids), p(Kc), p(decay), p(vmax), p(maxUptake), p(molarMass), p(fixConc),
                          p(fixLb), useTotals ? 1 : 0),
        "clb_set_surrogate");
  }
  void setSurrogateNet(int id, const SurrogateNet& net, int closure = 0, const std::vector<double>& yield = {}) {
    if (!yield.empty()) need(yield.size(), S_, "yield");
    chk(clb_set_surrogate_net(c_, id, net.nIn, net.inSub.data(), (int)net.width.size(), net.width.data(),
                              net.params.data(), net.xOffset.data(), net.xGain.data(), net.xYmin, net.yYmin,
                              net.yGain, net.yOffset, net.logOutput ? 1 : 0, closure,
                              yield.empty() ? nullptr : yield.data()),
        "clb_set_surrogate_net");
  }
  void clearSurrogateNet(int id) {
    chk(clb_set_surrogate_net(c_, id, 0, nullptr, 0, nullptr, nullptr, nullptr, nullptr, 0, 0, 1, 0, 0, 0, nullptr),
        "clb_set_surrogate_net");
  }
  void setGraph(bool enable, int stepsPerGraph) { chk(clb_set_graph(c_, enable ? 1 : 0, stepsPerGraph), "clb_set_graph"); }
  void step(int n) { chk(clb_step(c_, n), "clb_step"); }
  void saveState(const std::string& path) { chk(clb_save_state(c_, path.c_str()), "clb_save_state"); }
  void loadState(const std::string& path) { chk(clb_load_state(c_, path.c_str()), "clb_load_state"); }
  // Fastest launch shapes (and fused budget) for the current state; the state is kept.
  clb_tune_report autotune(int steps = 10) {
    clb_tune_report r{};
    chk(clb_autotune(c_, steps, &r), "clb_autotune");
    return r;
  }
  std::vector<double> concentrations() const {
    std::vector<double> c((size_t)(S_ * nodes()));
    chk(clb_download_conc(c_, c.data()), "clb_download_conc");
    return c;
  }
  std::vector<double> biomass() const {
    std::vector<double> b((size_t)(nb_ * nodes()));
    chk(clb_download_biomass(c_, b.data()), "clb_download_biomass");
    return b;
  }
  // Execution choices with no reference counterpart (clb.h documents each).
  void setSpeciationMode(int mode) { chk(clb_set_speciation_mode(c_, mode), "clb_set_speciation_mode"); }
  void setFusedBudget(int evals) { chk(clb_set_fused_budget(c_, evals), "clb_set_fused_budget"); }
  void setStragglerBudget(int evals) { chk(clb_set_straggler_budget(c_, evals), "clb_set_straggler_budget"); }
  void setLayout(int layout) { chk(clb_set_layout(c_, layout), "clb_set_layout"); }
  void setCompactOrder(int bx, int by, int bz) { chk(clb_set_compact_order(c_, bx, by, bz), "clb_set_compact_order"); }
  void setBlockShape(int kernel, int bx, int by) { chk(clb_set_block_shape(c_, kernel, bx, by), "clb_set_block_shape"); }

  // Mineral dissolution (surfaceDissolutionKinetics3D) and one solute's field at rest.
  void setDissolution(bool enable, const std::vector<int>& phaseIds, const std::vector<int>& substrates,
                      bool surfaceOnly = true) {
    need(substrates.size(), (long long)phaseIds.size(), "substrates");
    chk(clb_set_dissolution(c_, enable ? 1 : 0, (int)phaseIds.size(), phaseIds.data(), substrates.data(),
                            surfaceOnly ? 1 : 0),
        "clb_set_dissolution");
  }
  void setSoluteField(int solute, const std::vector<double>& c) {
    need(c.size(), nodes(), "field");
    chk(clb_set_solute_field(c_, solute, c.data()), "clb_set_solute_field");
  }

  // Biofilm: CA redistribution (the CA solver) and the FD step (the FD solver).
  void caInit() { chk(clb_ca_init(c_), "clb_ca_init"); }
  clb_ca_report caRedistribute(double bmax, bool half = false, unsigned long long seed = 1) {
    clb_ca_report r{};
    chk(clb_ca_redistribute(c_, bmax, half ? 1 : 0, seed, &r), "clb_ca_redistribute");
    return r;
  }
  int caUpdateAge(double bmax, int maxPasses = 1) {
    int passes = 0;
    chk(clb_ca_update_age(c_, bmax, maxPasses, &passes), "clb_ca_update_age");
    return passes;
  }
  std::vector<int> caAge() const {
    std::vector<int> a((size_t)nodes());
    chk(clb_ca_download_age(c_, a.data()), "clb_ca_download_age");
    return a;
  }
  void caSetAge(const std::vector<int>& age) { need(age.size(), nodes(), "age"); chk(clb_ca_set_age(c_, age.data()), "clb_ca_set_age"); }
  clb_ca_report fdBiomass(double nu) {
    clb_ca_report r{};
    chk(clb_fd_biomass(c_, nu, &r), "clb_fd_biomass");
    return r;
  }

  std::array<long long, 3> speciationStats() const {
    std::array<long long, 3> s{};
    chk(clb_speciation_stats(c_, s.data()), "clb_speciation_stats");
    return s;
  }
  void sync() { chk(clb_sync(c_), "clb_sync"); }

 private:
  void chk(int rc, const char* what) const { detail::check(rc, c_, what); }
  static void need(size_t got, long long want, const char* what) {
    if ((long long)got != want)
      throw Error(CLB_ERR_ARG, std::string(what) + ": expected " + std::to_string(want) + " values, got " +
                                   std::to_string(got));
  }
  clb_ctx* c_;
  int nx_, ny_, nz_;
  int S_ = 0, nb_ = 0;
};

// ------------------------------------------------------------------------------------------
// EquilibriumChemistry<T>: the reference's per-voxel speciation object, solved on the GPU.
template <typename T>
class EquilibriumChemistry {
 public:
  static constexpr T MIN_CONC = (T)1e-30;
  static constexpr T MAX_CONC = (T)10.0;
  static constexpr T MIN_LOG_C = (T)-30.0;
  static constexpr T MAX_LOG_C = (T)1.0;

  explicit EquilibriumChemistry(int device = 0) : c_(detail::create(device, CLB_PREC_FP64, 4, 1, 1)) {}
  ~EquilibriumChemistry() {
    if (c_) clb_destroy(c_);
  }
  EquilibriumChemistry(const EquilibriumChemistry&) = delete;
  EquilibriumChemistry& operator=(const EquilibriumChemistry&) = delete;

  // ---- setters (same names and meaning as the reference)
  void setSpeciesNames(const std::vector<std::string>& names) {
    species_.clear();
    for (const auto& n : names)
      if (n != "H2O" && n != "h2o") species_.push_back(n);
    dirty_ = true;
  }
  void setComponentNames(const std::vector<std::string>& names) { components_ = names; dirty_ = true; }
  void setLogK(const std::vector<T>& logK) { logK_ = logK; dirty_ = true; }
  void setStoichiometryMatrix(const std::vector<std::vector<T>>& S) { S_ = S; dirty_ = true; }
  void setStoichiometryRow(size_t i, const std::vector<T>& row) {
    if (i < S_.size() && row.size() == components_.size()) { S_[i] = row; dirty_ = true; }
  }
  void setMaxIterations(plint n) { maxIterations_ = n; dirty_ = true; }
  void setTolerance(T tol) { tolerance_ = tol; dirty_ = true; }
  void setVerbose(bool v) { verbose_ = v; }
  // Anderson-acceleration knobs of the reference's PCF solver; kept for source compatibility,
  // they do not change the GPU Newton solve.
  void setAndersonDepth(plint d) { andersonDepth_ = std::max(plint(1), d); }
  void setConditionTolerance(T t) { conditionTol_ = t; }
  void setBeta(T b) { beta_ = b; }

  // ---- getters
  plint getMaxIterations() const { return maxIterations_; }
  T getTolerance() const { return tolerance_; }
  std::vector<T> getLogK() const { return logK_; }
  bool didConverge() const { return lastConverged_; }
  std::vector<std::string> getSpeciesNames() const { return species_; }
  std::vector<std::string> getComponentNames() const { return components_; }
  plint getNumSpecies() const { return (plint)species_.size(); }
  plint getNumComponents() const { return (plint)components_.size(); }
  plint getLastIterations() const { return lastIterations_; }
  T getLastResidual() const { return lastResidual_; }
  plint getAndersonDepth() const { return andersonDepth_; }
  plint getTotalSolves() const { return totalSolves_; }
  plint getTotalConverged() const { return totalConverged_; }
  plint getTotalDiverged() const { return totalDiverged_; }
  void resetStatistics() { totalSolves_ = totalConverged_ = totalDiverged_ = 0; }
  const std::string& getLastError() const { return lastError_; }

  bool isEquilibriumSpecies(size_t i) const {
    if (i >= S_.size()) return false;
    for (T v : S_[i])
      if (std::abs(v) > (T)1e-10) return true;
    return false;
  }

  std::vector<T> calc_component_totals(const std::vector<T>& conc) const {
    const size_t nc = components_.size(), ns = species_.size();
    std::vector<T> tot(nc, (T)0);
    for (size_t j = 0; j < nc; ++j) {
      for (size_t i = 0; i < ns && i < conc.size(); ++i)
        if (isEquilibriumSpecies(i) && j < S_[i].size())
          tot[j] += S_[i][j] * std::max(std::min(conc[i], MAX_CONC), MIN_CONC);
      tot[j] = std::max(tot[j], MIN_CONC);
    }
    return tot;
  }

  std::vector<T> calc_species(const std::vector<T>& logC, const std::vector<T>& init = std::vector<T>()) const {
    const size_t ns = species_.size(), nc = components_.size();
    std::vector<T> c(ns);
    for (size_t i = 0; i < ns; ++i) {
      const T keep = (i < init.size() && init[i] > MIN_CONC) ? std::min(init[i], MAX_CONC) : MIN_CONC;
      if (!isEquilibriumSpecies(i)) { c[i] = keep; continue; }
      T lc = i < logK_.size() ? logK_[i] : (T)0;
      for (size_t j = 0; j < nc && j < S_[i].size(); ++j) lc += S_[i][j] * logC[j];
      lc = std::max(std::min(lc, MAX_LOG_C), MIN_LOG_C);
      c[i] = std::pow((T)10, lc);
      if (std::isnan(c[i]) || std::isinf(c[i])) c[i] = (i < init.size() && init[i] > MIN_CONC) ? init[i] : MIN_CONC;
      c[i] = std::max(std::min(c[i], MAX_CONC), MIN_CONC);
    }
    return c;
  }

  // One voxel (the reference's per-voxel call).
  std::vector<T> calculate_species_concentrations(const std::vector<T>& initial) {
    const size_t ns = species_.size();
    std::vector<T> in(initial), out(ns);
    in.resize(ns, (T)0);
    if (components_.empty() || species_.empty()) return initial;
    solve(1, in.data(), out.data());
    return out;
  }

  // Many voxels in one launch: conc_in / conc_out hold n rows of getNumSpecies() values.
  // The last* getters then describe the batch: converged only if every voxel converged,
  // the largest iteration count and residual.
  void calculate_species_concentrations(long n, const T* conc_in, T* conc_out) {
    if (components_.empty() || species_.empty()) {
      std::copy(conc_in, conc_in + n * (long)species_.size(), conc_out);
      return;
    }
    solve(n, conc_in, conc_out);
  }

 private:
  bool upload() {
    const int nc = (int)components_.size(), ns = (int)species_.size();
    if (nc > CLB_MAX_COMPONENTS || ns > CLB_MAX_SPECIES || (int)S_.size() < ns || (int)logK_.size() < ns) {
      lastError_ = "tableau outside the device limits or incomplete (nc <= " + std::to_string(CLB_MAX_COMPONENTS) +
                   ", ns <= " + std::to_string(CLB_MAX_SPECIES) + ")";
      return false;
    }
    std::vector<double> S((size_t)ns * nc, 0.0), lk((size_t)ns);
    for (int i = 0; i < ns; ++i) {
      lk[i] = (double)logK_[i];
      for (int j = 0; j < nc && j < (int)S_[i].size(); ++j) S[(size_t)i * nc + j] = (double)S_[i][j];
    }
    compSpecies_.assign((size_t)nc, -1);
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < ns; ++i)
        if (species_[i] == components_[j]) { compSpecies_[j] = i; break; }
    const int rc = clb_set_tableau(c_, nc, ns, S.data(), lk.data(), compSpecies_.data(), (double)tolerance_,
                                   (int)maxIterations_);
    if (rc != CLB_OK) { lastError_ = clb_last_error(c_); return false; }
    dirty_ = false;
    return true;
  }

  void solve(long n, const T* in, T* out) {
    const size_t ns = species_.size(), nc = components_.size();
    auto fail = [&](const std::string& why) {
      if (!why.empty()) lastError_ = why;
      for (long v = 0; v < n; ++v)
        for (size_t i = 0; i < ns; ++i) {
          const T x = in[v * ns + i];
          out[v * ns + i] = (x > MIN_CONC) ? std::min(x, MAX_CONC) : MIN_CONC;
        }
      lastConverged_ = false;
      totalSolves_ += n;
      totalDiverged_ += n;
    };
    if (dirty_ && !upload()) return fail("");
    std::vector<double> tot((size_t)n * nc), guess((size_t)n * nc);
    std::vector<T> row(ns);
    for (long v = 0; v < n; ++v) {
      row.assign(in + v * ns, in + (v + 1) * ns);
      const std::vector<T> t = calc_component_totals(row);
      for (size_t j = 0; j < nc; ++j) {
        tot[v * nc + j] = (double)t[j];
        T cc = (T)1e-7;
        if (compSpecies_[j] >= 0) cc = row[(size_t)compSpecies_[j]];
        guess[v * nc + j] = (double)std::log10(std::max(std::min(cc, MAX_CONC), MIN_CONC));
      }
    }
    std::vector<double> logc((size_t)n * nc), resid((size_t)n);
    std::vector<int> iters((size_t)n);
    std::vector<uint8_t> conv((size_t)n);
    int rc = clb_batch_load(c_, n, tot.data(), 0, nullptr, nullptr);
    if (rc == CLB_OK) rc = clb_batch_solve(c_, CLB_WARM_USER, guess.data());
    if (rc == CLB_OK) rc = clb_batch_fetch(c_, logc.data(), iters.data(), resid.data(), conv.data());
    if (rc != CLB_OK) return fail(clb_last_error(c_));
    bool all = true;
    plint itmax = 0;
    double rmax = 0.0;
    std::vector<T> lc(nc);
    for (long v = 0; v < n; ++v) {
      for (size_t j = 0; j < nc; ++j) lc[j] = (T)logc[v * nc + j];
      row.assign(in + v * ns, in + (v + 1) * ns);
      const std::vector<T> c = calc_species(lc, row);
      for (size_t i = 0; i < ns; ++i) {
        T x = c[i];
        if (std::isnan(x) || std::isinf(x)) x = (row[i] > MIN_CONC) ? std::min(row[i], MAX_CONC) : MIN_CONC;
        out[v * ns + i] = std::max(std::min(x, MAX_CONC), MIN_CONC);
      }
      all = all && conv[v];
      itmax = std::max(itmax, (plint)iters[v]);
      rmax = std::max(rmax, resid[v]);
      totalSolves_ += 1;
      if (conv[v]) totalConverged_ += 1;
      else totalDiverged_ += 1;
    }
    lastConverged_ = all;
    lastIterations_ = itmax;
    lastResidual_ = (T)rmax;
  }

  clb_ctx* c_;
  std::vector<std::string> species_, components_;
  std::vector<T> logK_;
  std::vector<std::vector<T>> S_;
  std::vector<int> compSpecies_;
  plint maxIterations_ = 200;
  T tolerance_ = (T)1e-8;
  bool verbose_ = false;
  plint andersonDepth_ = 4;
  T conditionTol_ = (T)1e10;
  T beta_ = (T)1;
  bool dirty_ = true;
  bool lastConverged_ = true;
  plint lastIterations_ = 0;
  T lastResidual_ = (T)0;
  plint totalSolves_ = 0, totalConverged_ = 0, totalDiverged_ = 0;
  std::string lastError_;
};


// ------------------------------------------------------------------ files (host only)
enum class GeometryOrder {
  XFastest,  // one integer per line, x fastest, then y, then z (the reference's makeGeometry.py)
  XSlices    // x slowest, each x slice z fastest (readGeometry's slice-by-slice reads)
};

// readGeometry (src/complab_functions.hh:127-167): nx*ny*nz material numbers from a text
// file, returned as the (nx + 2)*ny*nz lattice mask (x fastest) whose planes x = 0 and
// x = nx + 1 copy x = 1 and x = nx. The reference's reader consumes the file in slices
// along x while its makeGeometry.py writes x fastest, so the caller says which order the
// file has. Throws Error(CLB_ERR_ARG, ...) on a missing file, a wrong count or a code
// outside 0..255.
inline std::vector<uint8_t> readGeometry(const std::string& path, int nx, int ny, int nz,
                                         GeometryOrder order = GeometryOrder::XFastest) {
  std::ifstream in(path);
  if (!in) throw Error(CLB_ERR_ARG, "cannot open geometry file " + path);
  const long long n = (long long)nx * ny * nz, NX = nx + 2;
  std::vector<uint8_t> m((size_t)(NX * ny * nz));
  long long k = 0, v = 0;
  while (k < n && (in >> v)) {
    if (v < 0 || v > 255) throw Error(CLB_ERR_ARG, path + ": material number outside 0..255");
    long long x, y, z;
    if (order == GeometryOrder::XFastest) {
      x = k % nx;
      y = (k / nx) % ny;
      z = k / ((long long)nx * ny);
    } else {
      z = k % nz;
      y = (k / nz) % ny;
      x = k / ((long long)nz * ny);
    }
    m[(size_t)(x + 1 + NX * (y + (long long)ny * z))] = (uint8_t)v;
    ++k;
  }
  if (k != n || (in >> v)) throw Error(CLB_ERR_ARG, path + ": expected " + std::to_string(n) + " values");
  for (long long z = 0; z < nz; ++z)
    for (long long y = 0; y < ny; ++y) {
      const long long row = NX * (y + (long long)ny * z);
      m[(size_t)row] = m[(size_t)(row + 1)];
      m[(size_t)(row + NX - 1)] = m[(size_t)(row + NX - 2)];
    }
  return m;
}

namespace detail {
inline std::string base64(const std::vector<uint8_t>& b) {
  static const char* t = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string o;
  o.reserve((b.size() + 2) / 3 * 4);
  for (size_t i = 0; i < b.size(); i += 3) {
    const uint32_t v = (uint32_t)b[i] << 16 | (i + 1 < b.size() ? (uint32_t)b[i + 1] << 8 : 0u) |
                       (i + 2 < b.size() ? (uint32_t)b[i + 2] : 0u);
    o += t[(v >> 18) & 63];
    o += t[(v >> 12) & 63];
    o += (i + 1 < b.size()) ? t[(v >> 6) & 63] : '=';
    o += (i + 2 < b.size()) ? t[v & 63] : '=';
  }
  return o;
}
}  // namespace detail

// One field of a VTK file: `comps` values per voxel, x fastest.
struct VtiField {
  std::string name;
  const double* data;
  int comps;
};

// VTK XML ImageData (Float64, base64 binary, little endian) of lattice fields over the
// planes 1 <= x <= nx-2, the box the reference's writers output (writeAdvVTI / writeNsVTI /
// saveGeometry, src/complab_functions.hh:86-125). Throws Error(CLB_ERR_ARG, ...) when the
// file cannot be written.
inline void writeVTI(const std::string& path, int nx, int ny, int nz, const std::vector<VtiField>& fields) {
  std::ofstream out(path);
  if (!out) throw Error(CLB_ERR_ARG, "cannot write " + path);
  const std::string ext = "0 " + std::to_string(nx - 3) + " 0 " + std::to_string(ny - 1) + " 0 " + std::to_string(nz - 1);
  out << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n"
      << "  <ImageData WholeExtent=\"" << ext << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
      << "    <Piece Extent=\"" << ext << "\">\n      <PointData>\n";
  for (const VtiField& f : fields) {
    std::vector<uint8_t> raw(4);
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 1; x <= nx - 2; ++x)
          for (int c = 0; c < f.comps; ++c) {
            const double v = f.data[((size_t)x + (size_t)nx * (y + (size_t)ny * z)) * f.comps + c];
            const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
            raw.insert(raw.end(), p, p + sizeof(double));
          }
    const uint32_t nbytes = (uint32_t)(raw.size() - 4);
    for (int i = 0; i < 4; ++i) raw[i] = (uint8_t)(nbytes >> (8 * i));
    out << "        <DataArray type=\"Float64\" Name=\"" << f.name << "\" NumberOfComponents=\"" << f.comps
        << "\" format=\"binary\">\n" << detail::base64(raw) << "\n        </DataArray>\n";
  }
  out << "      </PointData>\n    </Piece>\n  </ImageData>\n</VTKFile>\n";
  if (!out) throw Error(CLB_ERR_ARG, "short write to " + path);
}

}  // namespace complab_gpu
