/*
 * ref_driver.cpp -- runs OpenLB's own CPU path on the workloads of this
 * repository and writes binary dumps used as golden fixtures.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the OpenLB
 * sources where they lie (never copied); the binary lands in oracle/_ref/ and
 * is a test oracle: tests compare the gfx950 platform and the C restatement with
 * what it writes.  It is our own driver code calling the public OpenLB API
 * (SuperGeometry, SuperLattice, dynamics/boundary setters), the way an OpenLB
 * case does.
 *
 * Usage (REF_VALUE = float or double decides the value type):
 *   olb_ref cavity   RES STEPS ULID RE OUT        cavity3dBenchmark geometry, BGK +
 *                                                BounceBack walls + BounceBackVelocity lid
 *   olb_ref cavity_lv RES STEPS ULID RE OUT       the same with LocalVelocity walls and lid
 *   olb_ref tgv      N STEPS U0 TAU OUT           periodic Taylor-Green, BGK
 *   olb_ref vtk_tgv  N STEPS OUT DIR              the same (u0 0.04, tau 0.56), then SuperVTMwriter3D
 *                                                output into DIR/vtkData (cavity modes: optional DIR)
 *   olb_ref tgv_smagorinsky N STEPS U0 TAU CS OUT  the same with SmagorinskyBGKdynamics
 *   olb_ref porous   NX NY NZ SPHERES R TAU MAGIC FX STEPS OUT
 *                                                periodic sphere pack, ForcedTRT + BounceBack
 *   olb_ref poiseuille N OUT                      forced BGK + BounceBack pipe, run to
 *                                                convergence (the reference's KAT)
 *   olb_ref poiseuille_bouzidi N OUT              the same with the Bouzidi wall (its bouzidi row)
 *   olb_ref poiseuille_local N OUT [STEPS]        its forced local row (LocalVelocity wall); STEPS
 *                                                stops after that many steps instead
 *   olb_ref poiseuille_local_nonforced N OUT      its non-forced local row (LocalVelocity wall and
 *                                                inflow, LocalPressure outflow)
 *   olb_ref poiseuille_interpolated N OUT [STEPS]  its forced interpolated row (InterpolatedVelocity
 *                                                wall); _nonforced: InterpolatedVelocity wall and
 *                                                inflow, InterpolatedPressure outflow
 *   olb_ref poiseuille_bouzidi_nonforced N OUT [STEPS]  its non-forced bouzidi rows (Bouzidi wall,
 *                                                velocity Bouzidi inflow, InterpolatedPressure outflow)
 *   olb_ref poiseuille_bounceback_nonforced N OUT [STEPS]  its non-forced bounceBack row (BounceBack
 *                                                wall, interpolated inflow and outflow)
 *   olb_ref normals IN OUT                       discrete normals (boundary::set's classifier)
 *                                                of materials 2 and 3 on the geometry in IN
 *   olb_ref lv_geometry IN STEPS ULID RE OUT     LocalVelocity walls (2) and lid (3) on the
 *                                                geometry in IN, BounceBack where no normal
 *   olb_ref bench_tgv N STEPS                    CPU throughput of the TGV step (MLUPS)
 *   olb_ref checkpoint N STEPS OUT NAME           TGV after STEPS steps: populations dump and
 *                                                OpenLB's own checkpoint (SuperLattice::save)
 *   olb_ref restore N OUT NAME                    TGV lattice loaded from checkpoint NAME
 *                                                (SuperLattice::load), populations dumped
 *
 * Dump format (little endian): text header line "OLBDUMP v1 <key=value ...>\n",
 * then named arrays, each "ARRAY <name> <dtype> <count>\n" followed by raw data.
 */
#include <olb.h>

#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

using namespace olb;
using T = REF_VALUE;

namespace {

struct Dump {
  std::ofstream out;
  explicit Dump(const std::string& path, const std::string& header)
    : out(path, std::ios::binary)
  {
    if (!out) {
      throw std::runtime_error("cannot open " + path);
    }
    out << "OLBDUMP v1 " << header << "\n";
  }
  template <typename V>
  void array(const std::string& name, const std::string& dtype, const std::vector<V>& v)
  {
    out << "ARRAY " << name << " " << dtype << " " << v.size() << "\n";
    out.write(reinterpret_cast<const char*>(v.data()), sizeof(V) * v.size());
  }
};

const char* tname() { return sizeof(T) == 4 ? "f4" : "f8"; }

// materials of the padded block in CellID order
std::vector<int32_t> materials(SuperGeometry<T,3>& G, int padding)
{
  auto& bg = G.getBlockGeometry(0);
  std::vector<int32_t> m;
  const int nx = bg.getNx(), ny = bg.getNy(), nz = bg.getNz();
  for (int x = -padding; x < nx + padding; ++x) {
    for (int y = -padding; y < ny + padding; ++y) {
      for (int z = -padding; z < nz + padding; ++z) {
        m.push_back(bg.get({x, y, z}));
      }
    }
  }
  return m;
}

// populations of the padded block in logical order (evaluation view)
template <typename DESCRIPTOR>
std::vector<T> populations(SuperLattice<T,DESCRIPTOR>& L)
{
  L.setProcessingContext(ProcessingContext::Evaluation);
  auto& b = L.getBlock(0);
  const std::size_t n = b.getNcells();
  std::vector<T> f(DESCRIPTOR::q * n);
  for (std::size_t iCell = 0; iCell < n; ++iCell) {
    auto cell = b.get(iCell);
    for (int i = 0; i < DESCRIPTOR::q; ++i) {
      f[i * n + iCell] = cell[i];
    }
  }
  return f;
}

template <typename DESCRIPTOR>
std::string block_header(SuperLattice<T,DESCRIPTOR>& L)
{
  auto& b = L.getBlock(0);
  return "nx=" + std::to_string(b.getNx()) + " ny=" + std::to_string(b.getNy()) +
         " nz=" + std::to_string(b.getNz()) + " padding=" + std::to_string(b.getPadding()) +
         " value=" + tname();
}

template <typename DESCRIPTOR>
std::vector<int64_t> distances(SuperLattice<T,DESCRIPTOR>& L)
{
  std::vector<int64_t> d;
  for (int i = 0; i < DESCRIPTOR::q; ++i) {
    d.push_back(L.getBlock(0).getNeighborDistance(descriptors::c<DESCRIPTOR>(i)));
  }
  return d;
}

template <typename DESCRIPTOR>
UnitConverter<T,DESCRIPTOR> lattice_units(T nu)
{
  // physical units equal lattice units: dx = dt = 1
  return UnitConverter<T,DESCRIPTOR>(T(1), T(1), T(1), T(1), nu, T(1));
}

// ---- lid-driven cavity: the geometry of examples/laminar/cavity3dBenchmark ----
// discrete normals of the boundary materials as boundary::set computes them
// (boundary/setBoundary.h:76-100, geometry/discreteNormals.hh:176): type per
// padded cell (-1 where not a boundary cell) and the normal, CellID order
void dump_normals(Dump& d, SuperGeometry<T,3>& G, int padding, std::initializer_list<int> boundary)
{
  auto fluid = G.getMaterialIndicator(1);
  auto outside = G.getMaterialIndicator(0);
  auto& fl = fluid->getBlockIndicatorF(0);
  auto& os = outside->getBlockIndicatorF(0);
  auto& bg = G.getBlockGeometry(0);
  std::vector<int32_t> type, normal;
  const int nx = bg.getNx(), ny = bg.getNy(), nz = bg.getNz();
  for (int x = -padding; x < nx + padding; ++x) {
    for (int y = -padding; y < ny + padding; ++y) {
      for (int z = -padding; z < nz + padding; ++z) {
        const int m = bg.get({x, y, z});
        bool on = false;
        for (int b : boundary) {
          on = on || m == b;
        }
        if (on && bg.getNeighborhoodRadius({x, y, z}) >= 2) {
          const auto [t, n] = computeBoundaryTypeAndNormal<T>(fl, os, {x, y, z});
          type.push_back((int32_t)t);
          normal.insert(normal.end(), {n[0], n[1], n[2]});
        } else {
          type.push_back(-1);
          normal.insert(normal.end(), {0, 0, 0});
        }
      }
    }
  }
  d.array("normal_type", "i4", type);
  d.array("normals", "i4", normal);
}

// SuperVTMwriter3D output (io/superVtmWriter3D.hh) of the lattice at iT
// into DIR/vtkData/: velocity, density, physVelocity and physPressure, once
// compressed (the default) as NAME and once as plain binary as NAME_raw
// relative L2 error of the physical strain rate over the fluid (material 1),
// as testPoiseuille3DBGK.cpp:277-287 computes it
template <typename DESCRIPTOR>
double strain_error_l2(SuperLattice<T,DESCRIPTOR>& L, SuperGeometry<T,3>& G,
                       const UnitConverter<T,DESCRIPTOR>& converter)
{
  int tmp[] = {int()};
  T result[2] = {T(), T()}, resultSol[2] = {T(), T()};
  CirclePoiseuilleStrainRate3D<T,DESCRIPTOR> sSol(converter, T(0.5));
  SuperLatticePhysStrainRate3D<T,DESCRIPTOR> s(L, converter);
  SuperLatticeFfromAnalyticalF3D<T,DESCRIPTOR> sSolLattice(sSol, L);
  SuperL2Norm3D<T> sL2Norm(sSolLattice - s, G, 1);
  SuperL2Norm3D<T> sSolL2Norm(sSolLattice, G, 1);
  sL2Norm(result, tmp);
  sSolL2Norm(resultSol, tmp);
  return result[0] / resultSol[0];
}

template <typename DESCRIPTOR>
void write_vtk(SuperLattice<T,DESCRIPTOR>& L, const UnitConverter<T,DESCRIPTOR>& converter,
               const std::string& dir, const std::string& name, int iT)
{
  singleton::directories().setOutputDir(dir + "/");
  SuperLatticeVelocity3D<T,DESCRIPTOR> velocity(L);
  SuperLatticeDensity3D<T,DESCRIPTOR> density(L);
  SuperLatticePhysVelocity3D<T,DESCRIPTOR> physVelocity(L, converter);
  SuperLatticePhysPressure3D<T,DESCRIPTOR> physPressure(L, converter);
  SuperVTMwriter3D<T> packed(name);
  SuperVTMwriter3D<T> raw(name + "_raw", 1, true, false);
  for (SuperVTMwriter3D<T>* w : {&packed, &raw}) {
    w->addFunctor(velocity);
    w->addFunctor(density);
    w->addFunctor(physVelocity);
    w->addFunctor(physPressure);
    w->createMasterFile();
    w->write(iT);
  }
}

int run_cavity(int res, int steps, double ulid, double re, const std::string& out, bool local_velocity,
               const std::string& vtk = "")
{
  using DESCRIPTOR = descriptors::D3Q19<>;
  const T length = 1;
  const T dx = length / res;
  Vector<T,3> origin((T)0);
  IndicatorCuboid3D<T> meshCuboid(Vector<T,3>(length + T(0.5) * dx), origin);
  CuboidDecomposition3D<T> cuboids(meshCuboid, dx, 1);
  HeuristicLoadBalancer<T> balancer(cuboids);
  SuperGeometry<T,3> G(cuboids, balancer, 3);

  IndicatorCuboid3D<T> domain(Vector<T,3>(length + T(0.25) * dx), origin);
  G.rename(0, 2, domain);
  G.rename(2, 1, {1, 1, 1});
  const T eps = dx;
  IndicatorCuboid3D<T> lid(Vector<T,3>(length + 2 * eps, 2 * eps, length + 2 * eps),
                           Vector<T,3>(-eps, length - eps, -eps));
  G.rename(2, 3, 1, lid);
  G.clean(false);
  G.innerClean(false);
  G.checkForErrors(false);

  // Re = U N / nu with N the resolution (RES = 64 gives nu = 0.064 for U = 0.1)
  const T nu = T(ulid * res / re);
  SuperLattice<T,DESCRIPTOR> L(lattice_units<DESCRIPTOR>(nu), G);
  L.setStatisticsOff();
  dynamics::set<BGKdynamics>(L, G, 1);
  if (local_velocity) {
    // the default of cavity3dBenchmark (LID_DRIVEN without LID_DRIVEN_BOUNCE_BACK)
    boundary::set<boundary::LocalVelocity<T,DESCRIPTOR,BGKdynamics<T,DESCRIPTOR>>>(L, G, 2);
    boundary::set<boundary::LocalVelocity<T,DESCRIPTOR,BGKdynamics<T,DESCRIPTOR>>>(L, G, 3);
  } else {
    boundary::set<boundary::BounceBack>(L, G, 2);
    dynamics::set<BounceBackVelocity>(L, G, 3);
  }
  AnalyticalConst3D<T,T> uLid(T(ulid), T(0), T(0));
  L.defineU(G.getMaterialIndicator(3), uLid);
  const T omega = T(1) / (T(3) * nu + T(0.5));
  L.setParameter<descriptors::OMEGA>(omega);
  L.initialize();

  Dump d(out, block_header(L) + (local_velocity ? " case=cavity_lv" : " case=cavity") + " res=" +
                  std::to_string(res) + " steps=" + std::to_string(steps) + " omega=" +
                  std::to_string((double)omega) + " ulid=" + std::to_string(ulid) + " nu=" +
                  std::to_string((double)nu));
  d.array("materials", "i4", materials(G, 3));
  if (local_velocity) {
    dump_normals(d, G, 3, {2, 3});
  }
  d.array("distances", "i8", distances(L));
  d.array("f0", tname(), populations(L));
  for (int t = 0; t < steps; ++t) {
    L.collideAndStream();
  }
  d.array("f1", tname(), populations(L));
  if (!vtk.empty()) {
    write_vtk(L, lattice_units<DESCRIPTOR>(nu), vtk, local_velocity ? "cavity_lv" : "cavity", steps);
  }
  return 0;
}

// ---- Taylor-Green vortex ----
struct TgvField : public AnalyticalF3D<T,T> {
  double u0, k;
  bool density;
  TgvField(double u0_, int n, bool density_)
    : AnalyticalF3D<T,T>(density_ ? 1 : 3), u0(u0_), k(2 * M_PI / n), density(density_) { }
  bool operator()(T output[], const T input[]) override
  {
    const double x = k * input[0], y = k * input[1], z = k * input[2];
    if (density) {
      const double p = u0 * u0 / 16.0 * (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2.0);
      output[0] = T(1.0 + 3.0 * p);
    } else {
      output[0] = T(u0 * std::sin(x) * std::cos(y) * std::cos(z));
      output[1] = T(-u0 * std::cos(x) * std::sin(y) * std::cos(z));
      output[2] = T(0);
    }
    return true;
  }
};

template <typename DESCRIPTOR>
void build_periodic_box(int nx, int ny, int nz, CuboidDecomposition3D<T>*& cuboids,
                        HeuristicLoadBalancer<T>*& balancer, SuperGeometry<T,3>*& G)
{
  cuboids = new CuboidDecomposition3D<T>(Vector<T,3>(T(0)), T(1), Vector<int,3>(nx, ny, nz), 1);
  cuboids->setPeriodicity({true, true, true});
  balancer = new HeuristicLoadBalancer<T>(*cuboids);
  G = new SuperGeometry<T,3>(*cuboids, *balancer, 3);
  G->rename(0, 1);
  G->communicate();
}

int run_tgv(int n, int steps, double u0, double tau, const std::string& out, bool bench, double smagorinsky = -1,
            const std::string& checkpoint = "")
{
  using DESCRIPTOR = descriptors::D3Q19<>;
  CuboidDecomposition3D<T>* cuboids;
  HeuristicLoadBalancer<T>* balancer;
  SuperGeometry<T,3>* G;
  build_periodic_box<DESCRIPTOR>(n, n, n, cuboids, balancer, G);
  const T nu = T((tau - 0.5) / 3.0);
  SuperLattice<T,DESCRIPTOR> L(lattice_units<DESCRIPTOR>(nu), *G);
  L.setStatisticsOff();
  if (smagorinsky >= 0) {
    // the bulk dynamics of examples/turbulence/tgv3d (tgv3d.cpp:143, 165)
    dynamics::set<SmagorinskyBGKdynamics>(L, *G, 1);
    L.setParameter<collision::LES::SMAGORINSKY>(T(smagorinsky));
  } else {
    dynamics::set<BGKdynamics>(L, *G, 1);
  }
  L.setParameter<descriptors::OMEGA>(T(1.0 / tau));
  TgvField rho(u0, n, true), u(u0, n, false);
  L.iniEquilibrium(*G, 1, rho, u);
  L.initialize();
  if (bench) {
    for (int t = 0; t < 2; ++t) {
      L.collideAndStream();
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int t = 0; t < steps; ++t) {
      L.collideAndStream();
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("MLUPS %.3f\nseconds %.3f\n", double(n) * n * n * steps / s / 1e6, s);
    return 0;
  }
  if (!checkpoint.empty() && steps < 0) {
    // SuperLattice::load: a checkpoint written elsewhere, then the dump of
    // what the lattice holds
    L.load(checkpoint);
    Dump d(out, block_header(L) + " case=tgv_restore n=" + std::to_string(n));
    d.array("f1", tname(), populations(L));
    return 0;
  }
  if (!checkpoint.empty()) {
    // SuperLattice::save (core/serializer.h:181): OpenLB's own checkpoint of
    // the lattice after the steps, next to the dump of its populations
    for (int t = 0; t < steps; ++t) {
      L.collideAndStream();
    }
    L.save(checkpoint);
    Dump d(out, block_header(L) + " case=tgv_checkpoint n=" + std::to_string(n) + " steps=" + std::to_string(steps));
    d.array("materials", "i4", materials(*G, 3));
    d.array("distances", "i8", distances(L));
    d.array("f1", tname(), populations(L));
    return 0;
  }
  Dump d(out, block_header(L) + " case=tgv n=" + std::to_string(n) + " steps=" + std::to_string(steps) +
                  " u0=" + std::to_string(u0) + " tau=" + std::to_string(tau) +
                  (smagorinsky >= 0 ? " smagorinsky=" + std::to_string(smagorinsky) : std::string()));
  d.array("materials", "i4", materials(*G, 3));
  d.array("distances", "i8", distances(L));
  d.array("f0", tname(), populations(L));
  for (int t = 0; t < steps; ++t) {
    L.collideAndStream();
  }
  d.array("f1", tname(), populations(L));
  return 0;
}

// ---- periodic sphere pack, ForcedTRT + BounceBack ----
int run_porous(int nx, int ny, int nz, const std::string& spheres, double radius, double tau,
               double magic, double fx, int steps, const std::string& out)
{
  using DESCRIPTOR = descriptors::D3Q19<descriptors::FORCE>;
  CuboidDecomposition3D<T>* cuboids;
  HeuristicLoadBalancer<T>* balancer;
  SuperGeometry<T,3>* G;
  build_periodic_box<DESCRIPTOR>(nx, ny, nz, cuboids, balancer, G);
  std::ifstream in(spheres);
  double cx, cy, cz;
  const int ext[3] = {nx, ny, nz};
  while (in >> cx >> cy >> cz) {
    // the sphere and its periodic images
    for (int ix = -1; ix <= 1; ++ix) {
      for (int iy = -1; iy <= 1; ++iy) {
        for (int iz = -1; iz <= 1; ++iz) {
          const double img[3] = {cx + ix * ext[0], cy + iy * ext[1], cz + iz * ext[2]};
          bool touches = true;  // images clear of the block and its overlap rename nothing
          for (int a = 0; a < 3; ++a) {
            touches = touches && img[a] + radius >= -4.0 && img[a] - radius <= ext[a] + 3.0;
          }
          if (!touches) {
            continue;
          }
          const Vector<T,3> c{T(img[0]), T(img[1]), T(img[2])};
          IndicatorSphere3D<T> s(c, T(radius));
          G->rename(1, 2, s);
        }
      }
    }
  }
  G->communicate();
  const T nu = T((tau - 0.5) / 3.0);
  SuperLattice<T,DESCRIPTOR> L(lattice_units<DESCRIPTOR>(nu), *G);
  L.setStatisticsOff();
  dynamics::set<ForcedTRTdynamics>(L, *G, 1);
  dynamics::set<BounceBack>(L, *G, 2);
  L.setParameter<descriptors::OMEGA>(T(1.0 / tau));
  L.setParameter<collision::TRT::MAGIC>(T(magic));
  AnalyticalConst3D<T,T> force(T(fx), T(0), T(0));
  L.defineField<descriptors::FORCE>(*G, 1, force);
  L.initialize();
  Dump d(out, block_header(L) + " case=porous steps=" + std::to_string(steps) + " tau=" +
                  std::to_string(tau) + " magic=" + std::to_string(magic) + " fx=" + std::to_string(fx));
  d.array("materials", "i4", materials(*G, 3));
  d.array("distances", "i8", distances(L));
  d.array("f0", tname(), populations(L));
  for (int t = 0; t < steps; ++t) {
    L.collideAndStream();
  }
  d.array("f1", tname(), populations(L));
  return 0;
}


// ---- Poiseuille pipe, forced BGK + BounceBack, run to convergence: the set-up
// of the reference's known-answer test tests/benchmarks/testPoiseuille3DBGK.cpp
// (forced, bounceBack row), driven through the same public API ----
// Taylor-Green vortex (u0 = 0.04, tau = 0.56) after STEPS steps, dumped and
// written as VTK through a converter with non-unit factors: dx = 1, dt = 0.5,
// physical density 1.2 (velocity factor 2, pressure factor 4.8)
// discrete normals on a given geometry: IN holds int32 nx, ny, nz and then
// the uint8 materials of the core cells (x slowest, z fastest); the block
// is non-periodic with overlap 3, so the overlap cells stay material 0.
// OUT gets the padded materials and the normals of materials 2 and 3 as
// boundary::set computes them (fluid = 1, outside = 0)
struct GeometryFile {
  int32_t n[3];
  std::vector<uint8_t> core;
};

GeometryFile read_geometry(const std::string& in)
{
  GeometryFile g;
  std::ifstream f(in, std::ios::binary);
  f.read(reinterpret_cast<char*>(g.n), sizeof(g.n));
  if (!f || g.n[0] < 1 || g.n[1] < 1 || g.n[2] < 1) {
    throw std::runtime_error("bad geometry file " + in);
  }
  g.core.resize((size_t)g.n[0] * g.n[1] * g.n[2]);
  f.read(reinterpret_cast<char*>(g.core.data()), g.core.size());
  if (!f) {
    throw std::runtime_error("short geometry file " + in);
  }
  return g;
}

// a single non-periodic cuboid of the file's extent (dx = 1, overlap 3)
// holding the file's core materials
struct LoadedGeometry {
  std::unique_ptr<CuboidDecomposition3D<T>> cuboids;
  std::unique_ptr<HeuristicLoadBalancer<T>> balancer;
  std::unique_ptr<SuperGeometry<T,3>> G;
};

LoadedGeometry load_geometry(const GeometryFile& g)
{
  LoadedGeometry L;
  IndicatorCuboid3D<T> mesh(Vector<T,3>(g.n[0] - T(0.5), g.n[1] - T(0.5), g.n[2] - T(0.5)), Vector<T,3>((T)0));
  L.cuboids = std::make_unique<CuboidDecomposition3D<T>>(mesh, T(1), 1);
  L.balancer = std::make_unique<HeuristicLoadBalancer<T>>(*L.cuboids);
  L.G = std::make_unique<SuperGeometry<T,3>>(*L.cuboids, *L.balancer, 3);
  auto& bg = L.G->getBlockGeometry(0);
  if (bg.getNx() != g.n[0] || bg.getNy() != g.n[1] || bg.getNz() != g.n[2]) {
    throw std::runtime_error("cuboid extent does not match the geometry file");
  }
  size_t k = 0;
  for (int x = 0; x < g.n[0]; ++x) {
    for (int y = 0; y < g.n[1]; ++y) {
      for (int z = 0; z < g.n[2]; ++z) {
        bg.set({x, y, z}, g.core[k++]);
      }
    }
  }
  return L;
}

// LocalVelocity cavity on the geometry in IN: BGK fluid (1), walls (2) at
// rest and lid (3) moving at (ULID, 0, 0), both BounceBack where
// boundary::set<LocalVelocity> finds no discrete normal; nu = ULID (nx-1) /
// RE; populations, normals and materials after STEPS steps into OUT
int run_lv_geometry(const std::string& in, int steps, double ulid, double re, const std::string& out)
{
  using DESCRIPTOR = descriptors::D3Q19<>;
  const GeometryFile g = read_geometry(in);
  LoadedGeometry geo = load_geometry(g);
  SuperGeometry<T,3>& G = *geo.G;
  const int res = g.n[0] - 1;
  const T nu = T(ulid * res / re);
  SuperLattice<T,DESCRIPTOR> L(lattice_units<DESCRIPTOR>(nu), G);
  L.setStatisticsOff();
  dynamics::set<BGKdynamics>(L, G, 1);
  boundary::set<boundary::BounceBack>(L, G, 2);
  boundary::set<boundary::BounceBack>(L, G, 3);
  boundary::set<boundary::LocalVelocity<T,DESCRIPTOR,BGKdynamics<T,DESCRIPTOR>>>(L, G, 2);
  boundary::set<boundary::LocalVelocity<T,DESCRIPTOR,BGKdynamics<T,DESCRIPTOR>>>(L, G, 3);
  AnalyticalConst3D<T,T> uLid(T(ulid), T(0), T(0));
  L.defineU(G.getMaterialIndicator(3), uLid);
  const T omega = T(1) / (T(3) * nu + T(0.5));
  L.setParameter<descriptors::OMEGA>(omega);
  L.initialize();

  Dump d(out, block_header(L) + " case=lv_geometry res=" + std::to_string(res) + " steps=" +
                  std::to_string(steps) + " omega=" + std::to_string((double)omega) + " ulid=" +
                  std::to_string(ulid) + " nu=" + std::to_string((double)nu));
  d.array("materials", "i4", materials(G, 3));
  dump_normals(d, G, 3, {2, 3});
  d.array("distances", "i8", distances(L));
  d.array("f0", tname(), populations(L));
  for (int t = 0; t < steps; ++t) {
    L.collideAndStream();
  }
  d.array("f1", tname(), populations(L));
  return 0;
}

int run_normals(const std::string& in, const std::string& out)
{
  const GeometryFile g = read_geometry(in);
  LoadedGeometry geo = load_geometry(g);
  Dump d(out, "nx=" + std::to_string(g.n[0]) + " ny=" + std::to_string(g.n[1]) + " nz=" +
                  std::to_string(g.n[2]) + " padding=3 case=normals");
  d.array("materials", "i4", materials(*geo.G, 3));
  dump_normals(d, *geo.G, 3, {2, 3});
  return 0;
}

int run_vtk_tgv(int n, int steps, const std::string& out, const std::string& dir)
{
  using DESCRIPTOR = descriptors::D3Q19<>;
  CuboidDecomposition3D<T>* cuboids;
  HeuristicLoadBalancer<T>* balancer;
  SuperGeometry<T,3>* G;
  build_periodic_box<DESCRIPTOR>(n, n, n, cuboids, balancer, G);
  const double u0 = 0.04, tau = 0.56;
  const T nu = T((tau - 0.5) / 3.0);
  UnitConverter<T,DESCRIPTOR> converter(T(1), T(0.5), T(1), T(2 * u0), T(2) * nu, T(1.2));
  SuperLattice<T,DESCRIPTOR> L(converter, *G);
  L.setStatisticsOff();
  dynamics::set<BGKdynamics>(L, *G, 1);
  L.setParameter<descriptors::OMEGA>(T(1.0 / tau));
  TgvField rho(u0, n, true), u(u0, n, false);
  L.iniEquilibrium(*G, 1, rho, u);
  L.initialize();
  Dump d(out, block_header(L) + " case=tgv n=" + std::to_string(n) + " steps=" + std::to_string(steps) +
                  " u0=" + std::to_string(u0) + " tau=" + std::to_string(tau) + " velocity_factor=" +
                  std::to_string((double)converter.getConversionFactorVelocity()) + " pressure_factor=" +
                  std::to_string((double)converter.getConversionFactorPressure()));
  d.array("materials", "i4", materials(*G, 3));
  d.array("distances", "i8", distances(L));
  d.array("f0", tname(), populations(L));
  for (int t = 0; t < steps; ++t) {
    L.collideAndStream();
  }
  d.array("f1", tname(), populations(L));
  write_vtk(L, converter, dir, "tgv", steps);
  return 0;
}

// the "local" rows of testPoiseuille3DBGK (lines 55-60): LocalVelocity on
// the pipe wall (and the inflow disc), LocalPressure on the outflow disc when
// not forced; geometry, dynamics and initial values as the test's
// prepareGeometry/prepareLattice (lines 72-242).  Dumps, besides populations
// and energies, the discrete normals of materials 2-4, every cell's
// computeRhoU after initialize (the velocity and density the boundary momenta
// hold) and the analytic velocity and pressure of the fluid.
int run_poiseuille_local(int n, const std::string& out, bool forced, int fixedSteps = -1,
                         bool interpolated = false, bool bouzidi = false, bool bounceBack = false)
{
  using DESCRIPTOR = descriptors::D3Q19<descriptors::FORCE>;
  const T lx = 2, diameter = 1, r = 0.5, Re = 10, maxPhysT = 20, physInterval = 0.25, residuum = 1e-5;
  UnitConverterFromResolutionAndRelaxationTime<T,DESCRIPTOR> const converter(n, T(0.8), T(1), T(1), T(1) / Re, T(1));
  const T dx = converter.getPhysDeltaX();
  Vector<T,3> a0(0, r, r), a1(lx, r, r);
  IndicatorCylinder3D<T> pipe(a0, a1, r);
  IndicatorLayer3D<T> extended(pipe, dx);
  CuboidDecomposition3D<T> cuboids(extended, dx, 1);
  if (forced) {
    cuboids.setPeriodicity({true, false, false});
  }
  HeuristicLoadBalancer<T> balancer(cuboids);
  SuperGeometry<T,3> G(cuboids, balancer, 2);

  Vector<T,3> c0(-dx * 0.2, r, r), c1(lx, r, r);
  if (forced) {
    c0[0] -= 3. * dx;
    c1[0] += 3. * dx;
  }
  IndicatorCylinder3D<T> pipeLong(c0, c1, r);
  G.rename(0, 2);
  G.rename(2, 1, pipeLong);
  if (!forced) {
    Vector<T,3> o(0, r, r), e = o;
    o[0] = -dx * 2;
    e[0] = dx * 2;
    IndicatorCylinder3D<T> inflow(o, e, r);
    G.rename(2, 3, 1, inflow);
    o[0] = lx - 2 * dx;
    e[0] = lx + 2 * dx;
    IndicatorCylinder3D<T> outflow(e, o, r);
    G.rename(2, 4, 1, outflow);
  }
  G.clean(false);
  G.innerClean(false);
  G.checkForErrors(false);

  SuperLattice<T,DESCRIPTOR> L(converter, G);
  L.template defineDynamics<NoDynamics>(G, 0);
  if (forced) {
    L.template defineDynamics<ForcedBGKdynamics>(G, 1);
    L.template defineDynamics<ForcedBGKdynamics>(G, 2);
  } else {
    L.template defineDynamics<BGKdynamics>(G, 1);
    L.template defineDynamics<BGKdynamics>(G, 2);
  }
  // the interpolated row (testPoiseuille3DBGK.cpp:158-160, 183-197) and the
  // non-forced bouzidi rows (lines 141-151, 168-172, 195-197, 239-241)
  // the Bouzidi pipe of prepareLattice (testPoiseuille3DBGK.cpp:130-131,
  // 145-146): a fresh cylinder from x = -dx/2 to lx + dx/2, not the
  // geometry's pipe (which starts at -0.2 dx)
  Vector<T,3> b0(0, r, r), b1(lx, r, r);
  b0[0] -= 0.5 * dx;
  b1[0] += 0.5 * dx;
  IndicatorCylinder3D<T> bouzidiPipe(b0, b1, r);
  if (bouzidi) {
    L.template defineDynamics<NoDynamics>(G, 2);
    setBouzidiBoundary<T, DESCRIPTOR, BouzidiPostProcessor>(L, G, 2, bouzidiPipe);
    L.template defineDynamics<NoDynamics>(G, 3);
    setBouzidiBoundary<T, DESCRIPTOR, BouzidiVelocityPostProcessor>(L, G, 3, bouzidiPipe);
    L.template defineDynamics<BGKdynamics>(G, 4);
    boundary::set<boundary::InterpolatedPressure>(L, G, 4);
  } else if (bounceBack) {
    // the non-forced bounceBack row: BounceBack wall, interpolated inflow
    // and outflow (lines 137-139, 172-197)
    L.template defineDynamics<BounceBack>(G, 2);
  } else if (interpolated) {
    boundary::set<boundary::InterpolatedVelocity>(L, G, 2);
  } else {
    boundary::set<boundary::LocalVelocity>(L, G, 2);
  }
  if (!forced && !bouzidi) {
    L.template defineDynamics<BGKdynamics>(G, 3);
    L.template defineDynamics<BGKdynamics>(G, 4);
    if (interpolated || bounceBack) {
      boundary::set<boundary::InterpolatedVelocity>(L, G, 3);
      boundary::set<boundary::InterpolatedPressure>(L, G, 4);
    } else {
      boundary::set<boundary::LocalVelocity>(L, G, 3);
      boundary::set<boundary::LocalPressure>(L, G, 4);
    }
  }
  std::vector<T> axisPoint = {lx, r, r};
  std::vector<T> axisDirection = {1, 0, 0};
  CirclePoiseuille3D<T> poiseuilleU(axisPoint, axisDirection, converter.getCharLatticeVelocity(), r);
  T force0 = 0;
  if (forced) {
    const T D = converter.getLatticeLength(diameter);
    std::vector<T> f(3, T());
    f[0] = 4. * converter.getLatticeViscosity() * converter.getCharLatticeVelocity() / (D * D / 4.);
    force0 = f[0];
    AnalyticalConst3D<T,T> force(f);
    L.template defineField<descriptors::FORCE>(G, 1, force);
    L.template defineField<descriptors::FORCE>(G, 2, force);
    AnalyticalConst3D<T,T> rhoF(1);
    L.defineRhoU(G, 1, rhoF, poiseuilleU);
    L.iniEquilibrium(G, 1, rhoF, poiseuilleU);
    L.defineRhoU(G, 2, rhoF, poiseuilleU);
    L.iniEquilibrium(G, 2, rhoF, poiseuilleU);
  } else {
    T p0 = 4. * converter.getPhysViscosity() * converter.getCharPhysVelocity() * lx / (r * r);
    p0 = converter.getLatticePressure(p0);
    AnalyticalLinear3D<T,T> rho(-p0 / lx * descriptors::invCs2<T,DESCRIPTOR>(), 0, 0,
                                p0 * descriptors::invCs2<T,DESCRIPTOR>() + 1);
    std::vector<T> velocity(3, T());
    AnalyticalConst3D<T,T> uF(velocity);
    L.defineRhoU(G, 0, rho, uF);
    L.iniEquilibrium(G, 0, rho, uF);
    for (int m = 1; m <= 4; ++m) {
      L.defineRhoU(G, m, rho, poiseuilleU);
      L.iniEquilibrium(G, m, rho, poiseuilleU);
      if (m == 3 && bouzidi) {
        setBouzidiVelocity(L, G, 3, poiseuilleU);
      }
    }
  }
  const T omega = converter.getLatticeRelaxationFrequency();
  L.template setParameter<descriptors::OMEGA>(omega);
  L.initialize();

  const int padding = G.getOverlap();
  std::vector<T> f0 = populations(L);
  // what the boundary momenta hold: computeRhoU of every cell at t = 0
  std::vector<T> rho0, u0;
  {
    auto& b = L.getBlock(0);
    for (std::size_t c = 0; c < b.getNcells(); ++c) {
      T rr, uu[3];
      b.get(c).computeRhoU(rr, uu);
      rho0.push_back(rr);
      u0.insert(u0.end(), uu, uu + 3);
    }
  }
  util::ValueTracer<T> converge(converter.getLatticeTime(physInterval), residuum);
  std::vector<double> energy;
  int iT;
  for (iT = 0; iT < (int)converter.getLatticeTime(maxPhysT); ++iT) {
    if (fixedSteps >= 0 ? iT == fixedSteps : converge.hasConverged()) {
      break;
    }
    L.collideAndStream();
    energy.push_back(L.getStatistics().getAverageEnergy());
    converge.takeValue(L.getStatistics().getAverageEnergy(), false);
  }

  int tmp[] = {int()};
  T result[2] = {T(), T()}, resultSol[2] = {T(), T()};
  const T maxVelocity = converter.getCharPhysVelocity();
  CirclePoiseuille3D<T> uSol(axisPoint, axisDirection, maxVelocity, r);
  SuperLatticePhysVelocity3D<T,DESCRIPTOR> u(L, converter);
  SuperLatticeFfromAnalyticalF3D<T,DESCRIPTOR> uSolLattice(uSol, L);
  SuperL2Norm3D<T> uL2Norm(uSolLattice - u, G, 1);
  SuperL2Norm3D<T> uSolL2Norm(uSolLattice, G, 1);
  uL2Norm(result, tmp);
  uSolL2Norm(resultSol, tmp);
  const double velocityErrorL2 = result[0] / resultSol[0];
  const double strainErrorL2 = strain_error_l2(L, G, converter);
  double pressureErrorL2 = 0;
  const T p0phys = 4. * converter.getPhysViscosity() * maxVelocity * lx / (r * r);
  AnalyticalLinear3D<T,T> pressureSol(-p0phys / lx, 0, 0, p0phys);
  if (!forced) {
    SuperLatticePhysPressure3D<T,DESCRIPTOR> pressure(L, converter);
    SuperLatticeFfromAnalyticalF3D<T,DESCRIPTOR> pressureSolLattice(pressureSol, L);
    SuperL2Norm3D<T> pL2Norm(pressureSolLattice - pressure, G, 1);
    SuperL2Norm3D<T> pSolL2Norm(pressureSolLattice, G, 1);
    pL2Norm(result, tmp);
    pSolL2Norm(resultSol, tmp);
    pressureErrorL2 = result[0] / resultSol[0];
  }

  auto& bg = G.getBlockGeometry(0);
  std::vector<T> usol, psol;
  CirclePoiseuilleStrainRate3D<T,DESCRIPTOR> sSol(converter, r);
  std::vector<T> ssol;
  for (int x = -padding; x < bg.getNx() + padding; ++x) {
    for (int y = -padding; y < bg.getNy() + padding; ++y) {
      for (int z = -padding; z < bg.getNz() + padding; ++z) {
        T v[3] = {0, 0, 0}, pv[1] = {0}, sv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (bg.get({x, y, z}) == 1) {
          Vector<T,3> physR;
          bg.getPhysR(physR, {x, y, z});
          uSol(v, physR.data());
          pressureSol(pv, physR.data());
          sSol(sv, physR.data());
        }
        usol.insert(usol.end(), v, v + 3);
        psol.push_back(pv[0]);
        ssol.insert(ssol.end(), sv, sv + 9);
      }
    }
  }
  const auto& cuboid0 = G.getCuboidDecomposition().get(0);
  const Vector<T,3> origin = cuboid0.getOrigin();
  char buf[896];
  std::snprintf(buf, sizeof(buf),
                " case=poiseuille boundary=%s forced=%d n=%d iT=%d omega=%.17g force=%.17g u_factor=%.17g "
                "p_factor=%.17g deltaT=%zu residuum=%.17g velocityErrorL2=%.17g pressureErrorL2=%.17g max_steps=%zu "
                "strainErrorL2=%.17g dt=%.17g tau=%.17g origin=%.17g,%.17g,%.17g delta=%.17g",
                bouzidi ? "bouzidi" : (bounceBack ? "bounceBack" : (interpolated ? "interpolated" : "local")),
                forced ? 1 : 0, n, iT,
                (double)omega, (double)force0,
                (double)converter.getConversionFactorVelocity(), (double)converter.getConversionFactorPressure(),
                converter.getLatticeTime(physInterval), (double)residuum, velocityErrorL2, pressureErrorL2,
                converter.getLatticeTime(maxPhysT), strainErrorL2, (double)converter.getConversionFactorTime(),
                (double)converter.getLatticeRelaxationTime(), (double)origin[0], (double)origin[1],
                (double)origin[2], (double)cuboid0.getDeltaR());
  Dump d(out, block_header(L) + buf);
  d.array("materials", "i4", materials(G, padding));
  dump_normals(d, G, padding, {2, 3, 4});
  d.array("distances", "i8", distances(L));
  d.array("f0", tname(), f0);
  d.array("f1", tname(), populations(L));
  d.array("rho0", tname(), rho0);
  d.array("u0", tname(), u0);
  d.array("energy", "f8", energy);
  d.array("usol", tname(), usol);
  d.array("psol", tname(), psol);
  d.array("ssol", tname(), ssol);
  if (bouzidi) {
    // BOUZIDI_DISTANCE (-1: none) and BOUZIDI_VELOCITY of every cell,
    // population-major
    auto& b = L.getBlock(0);
    const std::size_t nc = b.getNcells();
    std::vector<T> q(DESCRIPTOR::q * nc), v(DESCRIPTOR::q * nc);
    for (std::size_t c = 0; c < nc; ++c) {
      const auto qc = b.get(c).template getField<descriptors::BOUZIDI_DISTANCE>();
      const auto vc = b.get(c).template getField<descriptors::BOUZIDI_VELOCITY>();
      for (int i = 0; i < DESCRIPTOR::q; ++i) {
        q[i * nc + c] = qc[i];
        v[i * nc + c] = vc[i];
      }
    }
    d.array("bouzidi", tname(), q);
    d.array("bouzidi_velocity", tname(), v);
  }
  return 0;
}

int run_poiseuille(int n, const std::string& out, bool bouzidi = false)
{
  using DESCRIPTOR = descriptors::D3Q19<descriptors::FORCE>;
  const T lx = 2, diameter = 1, r = 0.5, Re = 10, maxPhysT = 20, physInterval = 0.25, residuum = 1e-5;
  UnitConverterFromResolutionAndRelaxationTime<T,DESCRIPTOR> const converter(n, T(0.8), T(1), T(1), T(1) / Re, T(1));
  const T dx = converter.getPhysDeltaX();
  Vector<T,3> a0(0, r, r), a1(lx, r, r);
  IndicatorCylinder3D<T> pipe(a0, a1, r);
  IndicatorLayer3D<T> extended(pipe, dx);
  CuboidDecomposition3D<T> cuboids(extended, dx, 1);
  cuboids.setPeriodicity({true, false, false});
  HeuristicLoadBalancer<T> balancer(cuboids);
  SuperGeometry<T,3> G(cuboids, balancer, 2);

  Vector<T,3> c0(-dx * T(0.2) - 3 * dx, r, r), c1(lx + 3 * dx, r, r);
  IndicatorCylinder3D<T> pipeLong(c0, c1, r);
  G.rename(0, 2);
  G.rename(2, 1, pipeLong);
  G.clean(false);
  G.innerClean(false);
  G.checkForErrors(false);

  SuperLattice<T,DESCRIPTOR> L(converter, G);
  L.template defineDynamics<NoDynamics>(G, 0);
  L.template defineDynamics<ForcedBGKdynamics>(G, 1);
  if (bouzidi) {
    // the bouzidi row of the test (testPoiseuille3DBGK.cpp:130-151, forced)
    L.template defineDynamics<NoDynamics>(G, 2);
    Vector<T,3> b0(0, r, r), b1(lx, r, r);
    b0[0] -= 0.5 * dx;
    b1[0] += 0.5 * dx;
    b0[0] -= 3. * dx;
    b1[0] += 3. * dx;
    IndicatorCylinder3D<T> wall(b0, b1, r);
    setBouzidiBoundary<T, DESCRIPTOR, BouzidiPostProcessor>(L, G, 2, wall);
  } else {
    L.template defineDynamics<BounceBack>(G, 2);
  }
  const T D = converter.getLatticeLength(diameter);
  std::vector<T> f(3, T());
  f[0] = 4. * converter.getLatticeViscosity() * converter.getCharLatticeVelocity() / (D * D / 4.);
  AnalyticalConst3D<T,T> force(f);
  L.template defineField<descriptors::FORCE>(G, 1, force);
  L.template defineField<descriptors::FORCE>(G, 2, force);
  std::vector<T> axisPoint = {lx, r, r};
  std::vector<T> axisDirection = {1, 0, 0};
  CirclePoiseuille3D<T> poiseuilleU(axisPoint, axisDirection, converter.getCharLatticeVelocity(), r);
  AnalyticalConst3D<T,T> rhoF(1);
  L.defineRhoU(G, 1, rhoF, poiseuilleU);
  L.iniEquilibrium(G, 1, rhoF, poiseuilleU);
  L.defineRhoU(G, 2, rhoF, poiseuilleU);
  L.iniEquilibrium(G, 2, rhoF, poiseuilleU);
  const T omega = converter.getLatticeRelaxationFrequency();
  L.template setParameter<descriptors::OMEGA>(omega);
  L.initialize();

  const int padding = G.getOverlap();
  std::vector<T> f0 = populations(L);
  util::ValueTracer<T> converge(converter.getLatticeTime(physInterval), residuum);
  std::vector<double> energy;
  int iT;
  for (iT = 0; iT < (int)converter.getLatticeTime(maxPhysT); ++iT) {
    if (converge.hasConverged()) {
      break;
    }
    L.collideAndStream();
    energy.push_back(L.getStatistics().getAverageEnergy());
    converge.takeValue(L.getStatistics().getAverageEnergy(), false);
  }

  // relative L2 velocity error over the fluid (material 1) as the test computes it
  int tmp[] = {int()};
  T result[2] = {T(), T()}, resultSol[2] = {T(), T()};
  CirclePoiseuille3D<T> uSol(axisPoint, axisDirection, converter.getCharPhysVelocity(), r);
  SuperLatticePhysVelocity3D<T,DESCRIPTOR> u(L, converter);
  SuperLatticeFfromAnalyticalF3D<T,DESCRIPTOR> uSolLattice(uSol, L);
  SuperL2Norm3D<T> uL2Norm(uSolLattice - u, G, 1);
  SuperL2Norm3D<T> uSolL2Norm(uSolLattice, G, 1);
  uL2Norm(result, tmp);
  uSolL2Norm(resultSol, tmp);
  const double velocityErrorL2 = result[0] / resultSol[0];
  const double strainErrorL2 = strain_error_l2(L, G, converter);

  // analytic velocity and strain rate at every padded cell (zero outside material 1)
  auto& bg = G.getBlockGeometry(0);
  std::vector<T> usol, ssol;
  CirclePoiseuilleStrainRate3D<T,DESCRIPTOR> sSol(converter, r);
  for (int x = -padding; x < bg.getNx() + padding; ++x) {
    for (int y = -padding; y < bg.getNy() + padding; ++y) {
      for (int z = -padding; z < bg.getNz() + padding; ++z) {
        T v[3] = {0, 0, 0}, sv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (bg.get({x, y, z}) == 1) {
          Vector<T,3> physR;
          bg.getPhysR(physR, {x, y, z});
          uSol(v, physR.data());
          sSol(sv, physR.data());
        }
        usol.insert(usol.end(), v, v + 3);
        ssol.insert(ssol.end(), sv, sv + 9);
      }
    }
  }
  const auto& cuboid = G.getCuboidDecomposition().get(0);
  const Vector<T,3> origin = cuboid.getOrigin();
  const T delta = cuboid.getDeltaR();
  char buf[768];
  std::snprintf(buf, sizeof(buf),
                " case=poiseuille boundary=%s n=%d iT=%d omega=%.17g force=%.17g u_factor=%.17g deltaT=%zu "
                "residuum=%.17g velocityErrorL2=%.17g max_steps=%zu origin=%.17g,%.17g,%.17g delta=%.17g "
                "strainErrorL2=%.17g dt=%.17g tau=%.17g",
                bouzidi ? "bouzidi" : "bounceBack", n, iT, (double)omega, (double)f[0],
                (double)converter.getConversionFactorVelocity(),
                converter.getLatticeTime(physInterval), (double)residuum, velocityErrorL2,
                converter.getLatticeTime(maxPhysT), (double)origin[0], (double)origin[1],
                (double)origin[2], (double)delta, strainErrorL2, (double)converter.getConversionFactorTime(),
                (double)converter.getLatticeRelaxationTime());
  Dump d(out, block_header(L) + buf);
  d.array("materials", "i4", materials(G, padding));
  d.array("distances", "i8", distances(L));
  d.array("f0", tname(), f0);
  d.array("f1", tname(), populations(L));
  d.array("energy", "f8", energy);
  d.array("usol", tname(), usol);
  d.array("ssol", tname(), ssol);
  if (bouzidi) {
    // the Bouzidi distances of every cell, population-major (-1: none)
    auto& b = L.getBlock(0);
    const std::size_t nc = b.getNcells();
    std::vector<T> q(DESCRIPTOR::q * nc);
    for (std::size_t c = 0; c < nc; ++c) {
      const auto v = b.get(c).template getField<descriptors::BOUZIDI_DISTANCE>();
      for (int i = 0; i < DESCRIPTOR::q; ++i) q[i * nc + c] = v[i];
    }
    d.array("bouzidi", tname(), q);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv)
{
  olb::initialize(&argc, &argv, false, false);
  if (argc < 2) {
    std::cerr << "usage: see header of ref_driver.cpp\n";
    return 2;
  }
  const std::string mode = argv[1];
  try {
    if ((mode == "cavity" || mode == "cavity_lv") && (argc == 7 || argc == 8)) {
      return run_cavity(std::stoi(argv[2]), std::stoi(argv[3]), std::stod(argv[4]), std::stod(argv[5]), argv[6],
                        mode == "cavity_lv", argc == 8 ? argv[7] : "");
    }
    if (mode == "vtk_tgv" && argc == 6) {
      return run_vtk_tgv(std::stoi(argv[2]), std::stoi(argv[3]), argv[4], argv[5]);
    }
    if (mode == "tgv" && argc == 7) {
      return run_tgv(std::stoi(argv[2]), std::stoi(argv[3]), std::stod(argv[4]), std::stod(argv[5]), argv[6], false);
    }
    if (mode == "tgv_smagorinsky" && argc == 8) {
      return run_tgv(std::stoi(argv[2]), std::stoi(argv[3]), std::stod(argv[4]), std::stod(argv[5]), argv[7], false,
                     std::stod(argv[6]));
    }
    if (mode == "restore" && argc == 5) {
      return run_tgv(std::stoi(argv[2]), -1, 0.04, 0.56, argv[3], false, -1, argv[4]);
    }
    if (mode == "checkpoint" && argc == 6) {
      return run_tgv(std::stoi(argv[2]), std::stoi(argv[3]), 0.04, 0.56, argv[4], false, -1, argv[5]);
    }
    if (mode == "bench_tgv" && argc == 4) {
      return run_tgv(std::stoi(argv[2]), std::stoi(argv[3]), 0.04, 0.56, "", true);
    }
    if (mode == "poiseuille" && argc == 4) {
      return run_poiseuille(std::stoi(argv[2]), argv[3]);
    }
    if ((mode == "poiseuille_local" || mode == "poiseuille_local_nonforced") && (argc == 4 || argc == 5)) {
      return run_poiseuille_local(std::stoi(argv[2]), argv[3], mode == "poiseuille_local",
                                  argc == 5 ? std::stoi(argv[4]) : -1);
    }
    if (mode == "poiseuille_bounceback_nonforced" && (argc == 4 || argc == 5)) {
      return run_poiseuille_local(std::stoi(argv[2]), argv[3], false, argc == 5 ? std::stoi(argv[4]) : -1, false,
                                  false, true);
    }
    if (mode == "poiseuille_bouzidi_nonforced" && (argc == 4 || argc == 5)) {
      return run_poiseuille_local(std::stoi(argv[2]), argv[3], false, argc == 5 ? std::stoi(argv[4]) : -1, false,
                                  true);
    }
    if ((mode == "poiseuille_interpolated" || mode == "poiseuille_interpolated_nonforced") &&
        (argc == 4 || argc == 5)) {
      return run_poiseuille_local(std::stoi(argv[2]), argv[3], mode == "poiseuille_interpolated",
                                  argc == 5 ? std::stoi(argv[4]) : -1, true);
    }
    if (mode == "lv_geometry" && argc == 7) {
      return run_lv_geometry(argv[2], std::stoi(argv[3]), std::stod(argv[4]), std::stod(argv[5]), argv[6]);
    }
    if (mode == "normals" && argc == 4) {
      return run_normals(argv[2], argv[3]);
    }
    if (mode == "poiseuille_bouzidi" && argc == 4) {
      return run_poiseuille(std::stoi(argv[2]), argv[3], true);
    }
    if (mode == "porous" && argc == 12) {
      return run_porous(std::stoi(argv[2]), std::stoi(argv[3]), std::stoi(argv[4]), argv[5], std::stod(argv[6]),
                        std::stod(argv[7]), std::stod(argv[8]), std::stod(argv[9]), std::stoi(argv[10]), argv[11]);
    }
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  std::cerr << "bad arguments\n";
  return 2;
}
