// Prints the BOUZIDI_DISTANCE field the C++ host mirror computes for the
// forced Bouzidi Poiseuille pipe of the reference's test: one line
// "cell direction q" per entry.  The materials of the padded block come from
// a raw uint8 file (CellID order).  Built and run by
// tests/test_host_mirror.py; needs no GPU.
//   bouzidi_check MATERIALS NX NY NZ PADDING OX OY OZ DELTA
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <vector>

#include "olb_gfx950.h"

int main(int argc, char** argv)
{
  using namespace olb;
  if (argc != 10) return 2;
  const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]), p = std::atoi(argv[5]);
  SuperGeometry<double> geometry({nx, ny, nz}, p, {true, false, false});
  geometry.setPhysical({std::atof(argv[6]), std::atof(argv[7]), std::atof(argv[8])}, std::atof(argv[9]));
  std::ifstream in(argv[1], std::ios::binary);
  const std::vector<char> m((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  if (m.size() != geometry.getNcells()) return 3;
  for (int x = -p; x < nx + p; ++x)
    for (int y = -p; y < ny + p; ++y)
      for (int z = -p; z < nz + p; ++z) geometry.set(x, y, z, (unsigned char)m[geometry.getCellId(x, y, z)]);
  // testPoiseuille3DBGK.cpp:130-151: the pipe stretched by dx/2 + 3 dx
  const double dx = geometry.getDeltaR(), r = 0.5, lx = 2.0;
  std::array<double, 3> c0{0.0, r, r}, c1{lx, r, r};
  c0[0] -= 0.5 * dx;
  c1[0] += 0.5 * dx;
  c0[0] -= 3. * dx;
  c1[0] += 3. * dx;
  IndicatorCylinder3D<double> wall(c0, c1, r);
  const std::vector<double> q = bouzidiDistances(geometry, 2, wall);
  const std::size_t n = geometry.getNcells();
  for (int i = 0; i < 19; ++i)
    for (std::size_t c = 0; c < n; ++c)
      if (q[(std::size_t)i * n + c] >= 0) std::printf("%zu %d %.17g\n", c, i, q[(std::size_t)i * n + c]);
  return 0;
}
