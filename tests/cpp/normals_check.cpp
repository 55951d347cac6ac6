// Prints the discrete normal of every cell of material 2 or 3 at least two
// cells inside the padded block (overlap 3), as the C++ host mirror computes
// it: "x y z type nx ny nz" per line, type as DiscreteNormalType (0 flat,
// 1 external corner, 2 internal corner, 3 external edge, 4 internal edge).
//
//   normals_check cavity N    the cavity3dBenchmark geometry at RESOLUTION N
//   normals_check file PATH   int32 nx, ny, nz, then uint8 core materials
//                             (x slowest, z fastest)
//
// Built and run by tests/test_host_mirror.py; needs no GPU.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "olb_gfx950.h"

int main(int argc, char** argv)
{
  using namespace olb;
  if (argc != 3) {
    std::fprintf(stderr, "usage: normals_check cavity N | file PATH\n");
    return 2;
  }
  const std::string mode = argv[1];
  std::int32_t e[3] = {0, 0, 0};
  std::vector<std::uint8_t> core;
  if (mode == "cavity") {
    const int n = std::atoi(argv[2]);
    e[0] = e[1] = e[2] = n + 1;
  } else {
    std::ifstream f(argv[2], std::ios::binary);
    f.read(reinterpret_cast<char*>(e), sizeof(e));
    core.resize((std::size_t)e[0] * e[1] * e[2]);
    f.read(reinterpret_cast<char*>(core.data()), (std::streamsize)core.size());
    if (!f) {
      std::fprintf(stderr, "cannot read %s\n", argv[2]);
      return 1;
    }
  }
  SuperGeometry<float> geometry({e[0], e[1], e[2]}, 3);
  if (mode == "cavity") {
    const int n = e[0] - 1;
    geometry.rename(0, 2);
    geometry.rename(2, 1, {1, 1, 1});
    geometry.rename(2, 3, [n](int x, int y, int z) { return y == n && x > 0 && x < n && z > 0 && z < n; });
  } else {
    std::size_t k = 0;
    geometry.forCore([&](int x, int y, int z) { geometry.set(x, y, z, core[k++]); });
  }
  geometry.forSpatialLocations(2, [&](int x, int y, int z) {
    const int m = geometry.get(x, y, z);
    if (m != 2 && m != 3) return;
    const auto [type, nv] = discreteNormal(geometry, x, y, z);
    std::printf("%d %d %d %d %d %d %d\n", x, y, z, (int)type, nv[0], nv[1], nv[2]);
  });
  return 0;
}
