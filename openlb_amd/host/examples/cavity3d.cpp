/*
 * Lid-driven cavity on the gfx950 platform, written the way an OpenLB case
 * is (examples/laminar/cavity3dBenchmark/cavity3d.cpp with
 * LID_DRIVEN_BOUNCE_BACK): geometry, lattice, initial values, then timed
 * collideAndStream steps after 10 warm-up steps, printing
 * "RESOLUTION, TIME_STEPS, CUBOIDS, ranks, threads, MLUPs" as the reference's
 * benchmark does (cavity3d.cpp:225-234).
 *
 *   cavity3d [--RESOLUTION n] [--TIME_STEPS k] [--LID_VELOCITY u] [--REYNOLDS re]
 *            [--STORAGE f32|f16|bf16] [--DUMP file] [--TUNE_CACHE file]
 *            [--BOUNDARY local_velocity|bounce_back] [--SAVE file] [--LOAD file]
 *            [--VTK dir] [--STEP s]
 *
 * --DUMP writes the materials and the populations of the padded block after
 * the run (logical order, float32) for the tests.  --VTK writes velocity,
 * density, physVelocity and physPressure after the run with SuperVTMwriter3D
 * into dir (cuboid of length 1, spacing 1/RESOLUTION, lattice-unit factors),
 * as the reference's getResults does with its own writer.  --STEP s makes
 * the core cells with x, y, z <= s a solid cube (wall material 2 next to the
 * fluid, outside elsewhere), so the LocalVelocity walls take every boundary
 * shape; wall cells without a discrete normal stay BounceBack.
 */
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "olb_gfx950.h"
#include "olb_gfx950_vtk.h"

using namespace olb;
using T = float;
using DESCRIPTOR = descriptors::D3Q19;

struct Parameters {
  int resolution = 64;
  int timeSteps = 100;
  T lidVelocity = T(0.1);
  T reynolds = T(100);
  Storage storage = Storage::F32;
  std::string dump;
  std::string tuneCache;  // "" = no autotuning
  bool localVelocity = true;  // cavity3d.cpp:111-124: LID_DRIVEN without LID_DRIVEN_BOUNCE_BACK
  std::string save;           // checkpoint after the timed steps
  std::string load;           // checkpoint to restart from before the warm-up
  std::string vtk;            // VTK output directory
  int step = -1;              // solid corner cube x, y, z <= step (-1: none)
};

static Parameters fromCLI(int argc, char** argv)
{
  Parameters p;
  for (int i = 1; i + 1 < argc; i += 2) {
    const std::string k = argv[i], v = argv[i + 1];
    if (k == "--RESOLUTION") p.resolution = std::stoi(v);
    else if (k == "--TIME_STEPS") p.timeSteps = std::stoi(v);
    else if (k == "--LID_VELOCITY") p.lidVelocity = std::stof(v);
    else if (k == "--REYNOLDS") p.reynolds = std::stof(v);
    else if (k == "--DUMP") p.dump = v;
    else if (k == "--TUNE_CACHE") p.tuneCache = v;
    else if (k == "--SAVE") p.save = v;
    else if (k == "--LOAD") p.load = v;
    else if (k == "--VTK") p.vtk = v;
    else if (k == "--STEP") p.step = std::stoi(v);
    else if (k == "--BOUNDARY") {
      if (v != "local_velocity" && v != "bounce_back") {
        throw std::invalid_argument("--BOUNDARY is local_velocity or bounce_back");
      }
      p.localVelocity = v == "local_velocity";
    }
    else if (k == "--STORAGE") {
      p.storage = v == "f16" ? Storage::F16 : (v == "bf16" ? Storage::BF16 : Storage::F32);
    } else {
      throw std::invalid_argument("unknown option " + k);
    }
  }
  return p;
}

// the cavity3dBenchmark geometry at RESOLUTION n: (n+1)^3 nodes, walls
// (material 2), the lid face y = n without its rim (material 3), fluid (1)
static void prepareGeometry(SuperGeometry<T>& geometry, int n, int step)
{
  geometry.rename(0, 2);
  geometry.rename(2, 1, {1, 1, 1});
  geometry.rename(2, 3, [n](int x, int y, int z) {
    return y == n && x > 0 && x < n && z > 0 && z < n;
  });
  if (step < 0) {
    return;
  }
  if (step < 1 || step > n - 3) {
    throw std::invalid_argument("--STEP must leave fluid on every side of the cube");
  }
  auto inCube = [step](int x, int y, int z) { return x <= step && y <= step && z <= step; };
  geometry.rename(1, 2, [&](int x, int y, int z) { return inCube(x, y, z); });
  // cube cells with no fluid among their 26 neighbours are outside
  std::vector<std::array<int, 3>> outside;
  geometry.forCore([&](int x, int y, int z) {
    if (!inCube(x, y, z) || x == 0 || y == 0 || z == 0) return;
    for (int i = -1; i <= 1; ++i)
      for (int j = -1; j <= 1; ++j)
        for (int k = -1; k <= 1; ++k)
          if (geometry.getOrOutside(x + i, y + j, z + k) == 1) return;
    outside.push_back({x, y, z});
  });
  for (const auto& c : outside) geometry.set(c[0], c[1], c[2], 0);
}

static void prepareLattice(SuperLattice<T, DESCRIPTOR>& lattice, SuperGeometry<T>& geometry,
                           const Parameters& p)
{
  dynamics::set<BGKdynamics>(lattice, geometry, 1);
  if (p.localVelocity) {
    if (p.step >= 0) {
      boundary::set<boundary::BounceBack>(lattice, geometry, 2);
      boundary::set<boundary::BounceBack>(lattice, geometry, 3);
    }
    boundary::set<boundary::LocalVelocity>(lattice, geometry, 2);
    boundary::set<boundary::LocalVelocity>(lattice, geometry, 3);
  } else {
    boundary::set<boundary::BounceBack>(lattice, geometry, 2);
    dynamics::set<BounceBackVelocity>(lattice, geometry, 3);
  }
  lattice.setStatisticsOff();
}

static void setInitialValues(SuperLattice<T, DESCRIPTOR>& lattice, SuperGeometry<T>& geometry,
                             const Parameters& p)
{
  lattice.defineU(geometry, 3, {p.lidVelocity, T(0), T(0)});
  const T nu = p.lidVelocity * T(p.resolution) / p.reynolds;
  lattice.setParameter<descriptors::OMEGA>(T(1) / (T(3) * nu + T(0.5)));
  lattice.initialize();
}

int main(int argc, char** argv)
{
  try {
    const Parameters p = fromCLI(argc, argv);
    const int n = p.resolution;
    SuperGeometry<T> geometry({n + 1, n + 1, n + 1}, 3);
    prepareGeometry(geometry, n, p.step);
    SuperLattice<T, DESCRIPTOR> lattice(geometry, p.storage);
    prepareLattice(lattice, geometry, p);
    setInitialValues(lattice, geometry, p);
    if (!p.load.empty()) {
      lattice.load(p.load);
    }
    if (!p.tuneCache.empty()) {
      const olbgfx_tune_trial t = lattice.autotune(p.tuneCache);
      std::cerr << "autotune: block " << t.block_threads << ", waves/SIMD " << t.waves_per_simd
                << ", order " << t.block_order << ", " << t.ms_per_step << " ms/step" << std::endl;
    }

    for (int iT = 0; iT < 10; ++iT) {
      lattice.collideAndStream();
    }
    check(olbgfx_synchronize(lattice.handle()));
    const auto t0 = std::chrono::steady_clock::now();
    for (int iT = 0; iT < p.timeSteps; ++iT) {
      lattice.collideAndStream();
    }
    check(olbgfx_synchronize(lattice.handle()));
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double mlups = double(geometry.getNvoxel()) * p.timeSteps / s / 1e6;

    if (!p.save.empty()) {
      lattice.save(p.save);
    }
    if (!p.vtk.empty()) {
      geometry.setPhysical({T(0), T(0), T(0)}, T(1) / T(n));
      SuperVTMwriter3D<T> vtm(p.localVelocity ? "cavity_lv" : "cavity", p.vtk);
      vtm.addFunctor(VtkFunctor::Velocity);
      vtm.addFunctor(VtkFunctor::Density);
      vtm.addFunctor(VtkFunctor::PhysVelocity);
      vtm.addFunctor(VtkFunctor::PhysPressure);
      vtm.createMasterFile();
      vtm.write(lattice, 10 + p.timeSteps);
    }
    lattice.setProcessingContext(ProcessingContext::Evaluation);
    std::cout << p.resolution << ", " << p.timeSteps << ", 1, 1, 1, " << mlups << std::endl;
    if (!p.dump.empty()) {
      std::ofstream out(p.dump, std::ios::binary);
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
