// io_check.cpp -- host-only check of complab_gpu::readGeometry and writeVTI (no GPU, no
// library calls). Usage: io_check <geometry> <nx> <ny> <nz> <order 0|1> <mask.bin> <out.vti>
// Writes the lattice mask it read as raw bytes, and a VTI of two fields of that lattice:
// "tag" (the mask as doubles) and "xyz" (the coordinates, 3 components).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "complab_gpu.hh"

int main(int argc, char** argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: io_check geometry nx ny nz order mask.bin out.vti\n");
    return 2;
  }
  const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]);
  const auto order = std::atoi(argv[5]) ? complab_gpu::GeometryOrder::XSlices : complab_gpu::GeometryOrder::XFastest;
  try {
    const std::vector<uint8_t> m = complab_gpu::readGeometry(argv[1], nx, ny, nz, order);
    FILE* f = std::fopen(argv[6], "wb");
    if (!f || std::fwrite(m.data(), 1, m.size(), f) != m.size()) return 1;
    std::fclose(f);
    const int NX = nx + 2;
    std::vector<double> tag(m.begin(), m.end()), xyz(3 * m.size());
    for (size_t n = 0; n < m.size(); ++n) {
      xyz[3 * n] = (double)(n % NX);
      xyz[3 * n + 1] = (double)((n / NX) % ny);
      xyz[3 * n + 2] = (double)(n / ((size_t)NX * ny));
    }
    complab_gpu::writeVTI(argv[7], NX, ny, nz, {{"tag", tag.data(), 1}, {"xyz", xyz.data(), 3}});
  } catch (const complab_gpu::Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
    return 3;
  }
  std::printf("ok\n");
  return 0;
}
