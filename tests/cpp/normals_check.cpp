// Prints the discrete normal of every cell of material 2 or 3 of the
// cavity3dBenchmark geometry at the given resolution, as the C++ host
// mirror computes it: "x y z type nx ny nz" per line (type 0 flat,
// 1 external corner, 3 external edge, -2 unclassified).  Built and run by
// tests/test_host_mirror.py; needs no GPU.
#include <cstdio>
#include <cstdlib>

#include "olb_gfx950.h"

int main(int argc, char** argv)
{
  using namespace olb;
  const int n = argc > 1 ? std::atoi(argv[1]) : 12;
  SuperGeometry<float> geometry({n + 1, n + 1, n + 1}, 3);
  geometry.rename(0, 2);
  geometry.rename(2, 1, {1, 1, 1});
  geometry.rename(2, 3, [n](int x, int y, int z) { return y == n && x > 0 && x < n && z > 0 && z < n; });
  geometry.forCore([&](int x, int y, int z) {
    const int m = geometry.get(x, y, z);
    if (m != 2 && m != 3) return;
    const auto [type, nv] = discreteNormal(geometry, x, y, z);
    const int t = type == NormalType::Flat ? 0
                : type == NormalType::ExternalCorner ? 1
                : type == NormalType::ExternalEdge ? 3 : -2;
    std::printf("%d %d %d %d %d %d %d\n", x, y, z, t, nv[0], nv[1], nv[2]);
  });
  return 0;
}
