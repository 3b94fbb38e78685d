// host_mirror_check.cpp -- drives the C++ host interface (host/complab_gpu.hh) the way the
// reference's C++ driver would, and writes raw float64 results for
// tests/test_gpu_host_cpp.py to compare with the FP64 oracle.
//
// usage: host_mirror_check <dir>
//   reads  <dir>/eq_in.bin   n x 4 carbonate species vectors (HCO3, H, CO3, H2CO3)
//   writes <dir>/eq_out.bin  batched calculate_species_concentrations
//          <dir>/eq_one.bin  the single-voxel overload on row 0
//          <dir>/u.bin       flow velocity (3N) after 200 FP64 steps on the 32x16x16 pack
//          <dir>/conc.bin    4 solutes (4N) after 30 coupled steps with dual-Monod kinetics
//          <dir>/bmass.bin   biomass (N)
//   prints one line "ok <converged> <iterations> <fluid nodes>" on success.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "complab_gpu.hh"

static std::vector<double> read_bin(const std::string& p) {
  std::ifstream f(p, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot read " + p);
  const std::streamsize n = f.tellg();
  f.seekg(0);
  std::vector<double> v((size_t)n / sizeof(double));
  f.read(reinterpret_cast<char*>(v.data()), n);
  return v;
}

static void write_bin(const std::string& p, const std::vector<double>& v) {
  std::ofstream f(p, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
  if (!f) throw std::runtime_error("cannot write " + p);
}

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  try {
    // ---- speciation: the carbonate tableau of the reference's example 04
    complab_gpu::EquilibriumChemistry<double> eq;
    eq.setSpeciesNames({"HCO3", "H", "CO3", "H2CO3", "H2O"});
    eq.setComponentNames({"HCO3", "H"});
    eq.setLogK({0.0, 0.0, -10.33, 6.35});
    eq.setStoichiometryMatrix({{1, 0}, {0, 1}, {1, -1}, {1, 1}});
    eq.setTolerance(1e-12);
    eq.setMaxIterations(100);
    const std::vector<double> in = read_bin(dir + "/eq_in.bin");
    const long n = (long)in.size() / 4;
    std::vector<double> out(in.size());
    eq.calculate_species_concentrations(n, in.data(), out.data());
    const bool conv_batch = eq.didConverge();
    write_bin(dir + "/eq_out.bin", out);
    const std::vector<double> one = eq.calculate_species_concentrations(std::vector<double>(in.begin(), in.begin() + 4));
    write_bin(dir + "/eq_one.bin", one);
    if (eq.getTotalSolves() != n + 1 || eq.getNumSpecies() != 4) throw std::runtime_error("solve bookkeeping");

    // ---- lattice: flow, then transport + kinetics on the same pack
    const int nx = 32, ny = 16, nz = 16;
    const long long N = (long long)nx * ny * nz;
    std::vector<uint8_t> mask((size_t)N);
    if (clb_gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.45, 0, 2, 0, 1, mask.data()) != CLB_OK)
      throw std::runtime_error("sphere pack");
    std::array<int, 256> kind;
    std::array<double, 256> om;
    kind.fill(CLB_KIND_FLUID);
    kind[0] = CLB_KIND_NONE;
    kind[1] = CLB_KIND_BB;
    om.fill(1.0 / 0.8);
    complab_gpu::Lattice lat(nx, ny, nz, CLB_PREC_FP64);
    lat.setDynamics(kind, om);
    lat.uploadMask(mask);
    lat.flowInit(1.0, 0.999);
    std::vector<double> energy;
    lat.flowSteps(200, &energy);
    write_bin(dir + "/u.bin", lat.velocity());
    const std::array<double, 4> st = lat.flowStats();

    std::vector<double> b((size_t)N);
    if (clb_gen_biomass_seed(N, mask.data(), 2, 0.02, 10.0, 100.0, 1, 2, b.data()) != CLB_OK)
      throw std::runtime_error("biomass seed");
    lat.setSolutes(4, std::vector<double>(4 * 256, 1.0 / 0.8), {0, 1, 0, 1, 1, 1, 1, 1},
                   {1.0, 0, 0.5, 0, 0, 0, 0, 0});
    lat.coupleVelocity();
    lat.setBiomass(1, b);
    lat.setKinetics(CLB_KINETICS_PLUGIN, 1.0, 0, 1);
    lat.solutesInit({0.2, 0.1, 0.0, 0.0});
    lat.step(10);
    // the autotuner keeps the state: the 30 steps still match the oracle's
    const clb_tune_report tr = lat.autotune(4);
    // flow and transport on the compact layouts: 3 block sizes each
    if (tr.trials != 6 || tr.ade_bx * tr.ade_by < 64) throw std::runtime_error("autotune report");
    lat.step(20);
    write_bin(dir + "/conc.bin", lat.concentrations());
    write_bin(dir + "/bmass.bin", lat.biomass());

    // ---- the flow solve with the tracer on the host and on the device: same steps, same state
    {
      complab_gpu::Lattice fa(nx, ny, nz, CLB_PREC_FP64), fb(nx, ny, nz, CLB_PREC_FP64);
      for (complab_gpu::Lattice* f : {&fa, &fb}) {
        f->setDynamics(kind, om);
        f->uploadMask(mask);
        f->flowInit(1.0, 0.999);
      }
      complab_gpu::ValueTracer tracer(1.0, 1000.0, 1e-4);
      fb.deviceTracer(1.0, 1000.0, 1e-4);
      for (int cap : {3000, 1}) {
        bool ca = false, cb = false;
        const long long na = fa.flowResolve(tracer, cap, &ca);
        const long long nb = fb.flowResolveDevice(cap, &cb);
        if (na != nb || ca != cb || na < 1) throw std::runtime_error("device tracer steps");
        if (fa.velocity() != fb.velocity()) throw std::runtime_error("device tracer state");
      }
    }

    // ---- surrogate processor: uniform field at rest, constant growth 0.4 / h, no Vmax.
    // One step: every reactive voxel's biomass is B (1 + 0.4 dt / 3600); a voxel whose
    // neighbours all reacted holds C - C^2 / (Kc + C) (the supply-limited draw).
    {
      const int sx = 10, sy = 4, sz = 4;
      const long long sn = (long long)sx * sy * sz;
      const double B = 0.8, dt = 50.0, Kc = 0.3, g = 0.4;
      complab_gpu::Lattice sl(sx, sy, sz, CLB_PREC_FP64);
      sl.setDynamics(kind, om);
      sl.uploadMask(std::vector<uint8_t>((size_t)sn, 2));
      sl.setSolutes(2, std::vector<double>(2 * 256, 1.0 / 0.8), {1, 1, 1, 1}, {0, 0, 0, 0});
      sl.setVelocity(std::vector<double>((size_t)(3 * sn), 0.0));
      sl.setBiomass(1, std::vector<double>((size_t)sn, B));
      sl.setKinetics(CLB_KINETICS_NONE, dt, 0, 1);
      sl.solutesInit({0.5, 0.2});
      sl.setSurrogate({0}, {Kc, Kc}, {1e-6});
      complab_gpu::Lattice::SurrogateNet net;
      net.nIn = 2;
      net.inSub = {0, 1};
      net.width = {1};
      net.params = {0.0, 0.0, 2.0 * g - 1.0};
      net.xOffset = {0.0, 0.0};
      net.xGain = {1.0, 1.0};
      net.yGain = 2.0;
      sl.setSurrogateNet(0, net);
      sl.step(1);
      const std::vector<double> sb = sl.biomass(), sc = sl.concentrations();
      const double c0[2] = {0.5, 0.2};
      for (long long q = 0; q < sn; ++q) {
        const int x = (int)(q % sx);
        const double wantB = (x >= 1 && x <= sx - 2) ? B + g * B * dt / 3600.0 : B;
        if (std::fabs(sb[(size_t)q] - wantB) > 1e-14 * wantB) throw std::runtime_error("surrogate biomass");
        if (x < 2 || x > sx - 3) continue;
        for (int s = 0; s < 2; ++s) {
          const double want = c0[s] - c0[s] * c0[s] / (Kc + c0[s]);
          if (std::fabs(sc[(size_t)(s * sn + q)] - want) > 1e-13 * want) throw std::runtime_error("surrogate draw");
        }
      }
    }

    // ---- FD biofilm step through the mirror: a spike in an open box keeps its mass and
    // sends nu b to each face neighbour
    {
      const int fx = 9, fy = 7, fz = 7;
      const long long fn = (long long)fx * fy * fz;
      complab_gpu::Lattice fl(fx, fy, fz, CLB_PREC_FP64);
      fl.setDynamics(kind, om);
      fl.uploadMask(std::vector<uint8_t>((size_t)fn, 2));
      std::vector<double> fb((size_t)fn, 0.0);
      const long long c = 4 + fx * (3 + fy * 3);
      fb[(size_t)c] = 10.0;
      fl.setBiomass(1, fb);
      fl.setKinetics(CLB_KINETICS_NONE, 1.0, 0, 1);
      fl.setMaskUpdate(true, 1e9, 2, 3);
      fl.caInit();
      const clb_ca_report r = fl.fdBiomass(0.1);
      const std::vector<double> out = fl.biomass();
      double total = 0.0;
      for (double v : out) total += v;
      if (r.sweeps != 1 || out[(size_t)c] != 4.0 || out[(size_t)(c + 1)] != 1.0 || std::fabs(total - 10.0) > 1e-12)
        throw std::runtime_error("fd biomass");
    }
    std::printf("ok %d %lld %.0f %.6e\n", conv_batch ? 1 : 0, (long long)eq.getLastIterations(), st[3],
                energy.back());
    return 0;
  } catch (const complab_gpu::Error& e) {
    std::fprintf(stderr, "complab_gpu error %d: %s\n", e.code(), e.what());
    return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
