/*
 * Channel with a velocity inlet and a pressure outlet on the gfx950
 * platform, written the way an OpenLB case is: BounceBack walls,
 * boundary::set<boundary::LocalVelocity> on the inflow face and
 * boundary::set<boundary::LocalPressure> on the outflow face (the boundary
 * pair of the reference's non-forced Poiseuille test), BGK fluid at rest.
 *
 *   channel3d [--NX n] [--NY n] [--NZ n] [--TIME_STEPS k] [--U_IN u] [--TAU t]
 *             [--DUMP file] [--BOUNDARY local|interpolated]
 *
 * --BOUNDARY interpolated uses boundary::InterpolatedVelocity and
 * boundary::InterpolatedPressure instead (the pair of the test's
 * interpolated rows).
 *
 * --DUMP writes the materials and the populations of the padded block after
 * the run (logical order, float32) for the tests.
 */
#include <fstream>
#include <iostream>
#include <string>

#include "olb_gfx950.h"

using namespace olb;
using T = float;
using DESCRIPTOR = descriptors::D3Q19;

int main(int argc, char** argv)
{
  try {
    int nx = 48, ny = 16, nz = 16, steps = 100;
    T uIn = T(0.05), tau = T(0.8);
    std::string dump;
    bool interpolated = false;
    for (int i = 1; i + 1 < argc; i += 2) {
      const std::string k = argv[i], v = argv[i + 1];
      if (k == "--NX") nx = std::stoi(v);
      else if (k == "--NY") ny = std::stoi(v);
      else if (k == "--NZ") nz = std::stoi(v);
      else if (k == "--TIME_STEPS") steps = std::stoi(v);
      else if (k == "--U_IN") uIn = std::stof(v);
      else if (k == "--TAU") tau = std::stof(v);
      else if (k == "--DUMP") dump = v;
      else if (k == "--BOUNDARY") {
        if (v != "local" && v != "interpolated") throw std::invalid_argument("--BOUNDARY is local or interpolated");
        interpolated = v == "interpolated";
      }
      else throw std::invalid_argument("unknown option " + k);
    }
    SuperGeometry<T> geometry({nx, ny, nz}, 3);
    geometry.rename(0, 2);
    geometry.rename(2, 1, [&](int, int y, int z) { return y > 0 && y < ny - 1 && z > 0 && z < nz - 1; });
    geometry.rename(1, 3, [](int x, int, int) { return x == 0; });
    geometry.rename(1, 4, [&](int x, int, int) { return x == nx - 1; });

    SuperLattice<T, DESCRIPTOR> lattice(geometry);
    dynamics::set<BGKdynamics>(lattice, geometry, 1);
    boundary::set<boundary::BounceBack>(lattice, geometry, 2);
    if (interpolated) {
      boundary::set<boundary::InterpolatedVelocity>(lattice, geometry, 3);
      boundary::set<boundary::InterpolatedPressure>(lattice, geometry, 4);
    } else {
      boundary::set<boundary::LocalVelocity>(lattice, geometry, 3);
      boundary::set<boundary::LocalPressure>(lattice, geometry, 4);
    }
    lattice.defineU(geometry, 3, {uIn, T(0), T(0)});
    lattice.defineRho(geometry, 4, T(1));
    lattice.setParameter<descriptors::OMEGA>(T(1) / tau);
    lattice.setStatisticsOff();
    lattice.initialize();
    for (int iT = 0; iT < steps; ++iT) {
      lattice.collideAndStream();
    }
    lattice.setProcessingContext(ProcessingContext::Evaluation);
    T rho, u[3];
    lattice.computeRhoU(geometry.getCellId(nx / 2, ny / 2, nz / 2), rho, u);
    std::cout << "centre u_x " << u[0] << " rho " << rho << std::endl;
    if (!dump.empty()) {
      std::ofstream out(dump, std::ios::binary);
      const auto& m = geometry.materials();
      out.write(reinterpret_cast<const char*>(m.data()), m.size());
      for (int i = 0; i < DESCRIPTOR::q; ++i) {
        for (CellID c = 0; c < (CellID)lattice.getNcells(); ++c) {
          const T f = lattice.population(c, i);
          out.write(reinterpret_cast<const char*>(&f), sizeof(T));
        }
      }
    }
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
