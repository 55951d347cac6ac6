/*
 * olb_gfx950.h -- C++ host side of the gfx950 platform, mirroring the OpenLB
 * calls an application makes on the D3Q19 hot path:
 *
 *   SuperGeometry::rename                      geometry/superGeometry.h, blockGeometry.hh:401-466
 *   SuperLattice::defineDynamics<DYNAMICS>     core/superLattice.h:228-250
 *   dynamics::set / boundary::set              case/setters.h:303-345, boundary/setBoundary.h:150-157
 *   boundary::set<boundary::LocalVelocity>     boundary/localVelocity3D.h:36-120
 *   SuperLattice::setParameter<OMEGA / MAGIC / LES::SMAGORINSKY>  core/superLattice.h (setParameter)
 *   SuperLattice::defineField<FORCE>, defineU  core/superLattice.h:255-330
 *   SuperLattice::iniEquilibrium               core/blockLattice.hh:202-236
 *   SuperLattice::initialize, collideAndStream core/superLattice.hh:248-266, 663-725
 *   setStatisticsOn/Off, getStatistics         core/superLattice.h, core/latticeStatistics.h:43-127
 *   setProcessingContext(Evaluation)           core/superLattice.hh:184-195
 *   save / load (OpenLB checkpoint format)     core/serializer.h:181-189, io/serializerIO.hh:40-64
 *   setBouzidiBoundary<..., BouzidiPostProcessor> boundary/setBouzidiBoundary.h:193-330
 *   IndicatorCylinder3D, IndicatorF3D::distance functors/analytical/indicator/indicatorF3D.hh:312-418
 *   autotune (an addition of this platform)    no counterpart; OpenLB fixes launch shapes
 *
 * Everything below calls the C ABI of include/olbgfx.h; no HIP header is
 * included, so a host translation unit builds with any C++17 compiler and
 * links libolbgfx.so.  The dynamics are the closed set the platform provides;
 * asking for another one fails at compile time (there is no tag for it), and
 * an unknown kind passed through the ABI fails with std::domain_error, as the
 * reference's platform dispatch does for unsupported operators.
 *
 * One SuperLattice is one block (one cuboid) on one GPU, in lattice units.
 */
#ifndef OLB_GFX950_H
#define OLB_GFX950_H

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <iterator>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "olbgfx.h"

namespace olb {

using CellID = std::uint32_t;

/// error reporting as the reference does it: exceptions, by kind
inline void check(olbgfx_status status)
{
  if (status == OLBGFX_OK) {
    return;
  }
  const std::string msg = olbgfx_last_error();
  switch (status) {
  case OLBGFX_ERR_INVALID:
    throw std::invalid_argument(msg);
  case OLBGFX_ERR_UNSUPPORTED:
    throw std::domain_error(msg);
  default:
    throw std::runtime_error("gfx950 platform: " + msg);
  }
}

enum class ProcessingContext { Evaluation, Simulation };

namespace descriptors {

/// D3Q19 (descriptor/definition/common.h:152-200)
struct D3Q19 {
  static constexpr int d = 3;
  static constexpr int q = 19;
  static constexpr int c[19][3] = {
    {0, 0, 0},
    {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1},
    {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1},
    {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};
  static constexpr int tden[19] = {3, 18, 18, 18, 36, 36, 36, 36, 36, 36,
                                   18, 18, 18, 36, 36, 36, 36, 36, 36};
  static constexpr int opposite(int i) { return i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9); }
};

struct OMEGA { };
struct FORCE { static constexpr int32_t field = OLBGFX_FIELD_FORCE; };
struct VELOCITY { static constexpr int32_t field = OLBGFX_FIELD_VELOCITY; };

}  // namespace descriptors

namespace collision {
struct TRT {
  struct MAGIC { };
};
namespace LES {
struct SMAGORINSKY { };  // dynamics/collisionLES.h:37
}  // namespace LES
}  // namespace collision

// ---- dynamics (dynamics/dynamics.h) --------------------------------------
template <typename T, typename DESCRIPTOR> struct NoDynamics {
  static constexpr int32_t kind = OLBGFX_NO_DYNAMICS;
};
template <typename T, typename DESCRIPTOR> struct BGKdynamics {
  static constexpr int32_t kind = OLBGFX_BGK;
};
template <typename T, typename DESCRIPTOR> struct TRTdynamics {
  static constexpr int32_t kind = OLBGFX_TRT;
};
template <typename T, typename DESCRIPTOR> struct ForcedBGKdynamics {
  static constexpr int32_t kind = OLBGFX_FORCED_BGK;
};
template <typename T, typename DESCRIPTOR> struct ForcedTRTdynamics {
  static constexpr int32_t kind = OLBGFX_FORCED_TRT;
};
template <typename T, typename DESCRIPTOR> struct BounceBack {
  static constexpr int32_t kind = OLBGFX_BOUNCE_BACK;
};
template <typename T, typename DESCRIPTOR> struct BounceBackVelocity {
  static constexpr int32_t kind = OLBGFX_BOUNCE_BACK_VELOCITY;
};
template <typename T, typename DESCRIPTOR> struct SmagorinskyBGKdynamics {
  static constexpr int32_t kind = OLBGFX_SMAGORINSKY_BGK;  // dynamics/smagorinskyBGKdynamics.h:37-44
};

/// population storage of the platform
enum class Storage : int32_t {
  F32 = OLBGFX_STORE_F32,
  F16 = OLBGFX_STORE_F16,
  BF16 = OLBGFX_STORE_BF16,
  F64 = OLBGFX_STORE_F64
};

// ---- geometry ------------------------------------------------------------

/// Material numbers of one cuboid with its overlap, in CellID order
/// (core/blockStructure.h:150-160); 0 marks cells outside the domain.
template <typename T>
class SuperGeometry {
public:
  SuperGeometry(std::array<int, 3> extent, int overlap = 3,
                std::array<bool, 3> periodic = {false, false, false})
    : _extent(extent), _overlap(overlap), _periodic(periodic)
  {
    if (extent[0] <= 0 || extent[1] <= 0 || extent[2] <= 0 || overlap < 1) {
      throw std::invalid_argument("extent must be positive and the overlap at least 1");
    }
    for (int a = 0; a < 3; ++a) {
      _size[a] = extent[a] + 2 * overlap;
    }
    _material.assign((std::size_t)_size[0] * _size[1] * _size[2], 0);
  }

  std::array<int, 3> getExtent() const { return _extent; }
  int getOverlap() const { return _overlap; }
  /// physical placement of the cuboid: origin of core cell (0, 0, 0) and the
  /// spacing (geometry/cuboid.h:90-92); lattice units by default
  void setPhysical(std::array<T, 3> origin, T deltaR)
  {
    _origin = origin;
    _deltaR = deltaR;
  }
  T getDeltaR() const { return _deltaR; }
  std::array<T, 3> getPhysR(int x, int y, int z) const
  {
    const int r[3] = {x, y, z};
    std::array<T, 3> p;
    for (int a = 0; a < 3; ++a) p[a] = _origin[a] + T(r[a]) * _deltaR;
    return p;
  }
  std::array<bool, 3> getPeriodicity() const { return _periodic; }
  std::size_t getNcells() const { return _material.size(); }

  CellID getCellId(int x, int y, int z) const
  {
    return (CellID)(((std::size_t)(x + _overlap) * _size[1] + (y + _overlap)) * _size[2] + (z + _overlap));
  }
  int get(int x, int y, int z) const { return _material[getCellId(x, y, z)]; }
  /// material, or 0 beyond the padded block
  int getOrOutside(int x, int y, int z) const { return getClamped(x, y, z); }
  void set(int x, int y, int z, int m) { _material[getCellId(x, y, z)] = (std::uint8_t)m; }

  /// every core cell of material fromM becomes toM (blockGeometry.hh:401-408)
  void rename(int fromM, int toM)
  {
    forCore([&](int x, int y, int z) {
      if (get(x, y, z) == fromM) set(x, y, z, toM);
    });
  }
  /// core cells of material fromM inside `indicator(x, y, z)` become toM
  /// (blockGeometry.hh:410-434, lattice coordinates instead of physical ones)
  void rename(int fromM, int toM, const std::function<bool(int, int, int)>& indicator)
  {
    forCore([&](int x, int y, int z) {
      if (get(x, y, z) == fromM && indicator(x, y, z)) set(x, y, z, toM);
    });
  }
  /// core cells of material fromM whose whole offset neighbourhood is fromM
  /// become toM (blockGeometry.hh:435-466)
  void rename(int fromM, int toM, std::array<int, 3> offset)
  {
    std::vector<CellID> hit;
    forCore([&](int x, int y, int z) {
      if (get(x, y, z) != fromM) return;
      for (int i = -offset[0]; i <= offset[0]; ++i)
        for (int j = -offset[1]; j <= offset[1]; ++j)
          for (int k = -offset[2]; k <= offset[2]; ++k)
            if (getClamped(x + i, y + j, z + k) != fromM) return;
      hit.push_back(getCellId(x, y, z));
    });
    for (CellID c : hit) _material[c] = (std::uint8_t)toM;
  }

  /// cells with material != 0 (SuperGeometryStatistics::getNvoxel)
  std::size_t getNvoxel() const
  {
    std::size_t n = 0;
    forCore([&](int x, int y, int z) { n += get(x, y, z) != 0; });
    return n;
  }

  const std::vector<std::uint8_t>& materials() const { return _material; }

  template <typename F>
  void forCore(F f) const
  {
    for (int x = 0; x < _extent[0]; ++x)
      for (int y = 0; y < _extent[1]; ++y)
        for (int z = 0; z < _extent[2]; ++z) f(x, y, z);
  }

  /// every cell of the padded block at least `radius` cells from its edge
  /// (forSpatialLocations with getNeighborhoodRadius,
  /// core/blockStructure.h:223-250)
  template <typename F>
  void forSpatialLocations(int radius, F f) const
  {
    const int lo = radius - _overlap;
    for (int x = lo; x < _extent[0] + _overlap - radius; ++x)
      for (int y = lo; y < _extent[1] + _overlap - radius; ++y)
        for (int z = lo; z < _extent[2] + _overlap - radius; ++z) f(x, y, z);
  }

private:
  int getClamped(int x, int y, int z) const
  {
    if (x < -_overlap || y < -_overlap || z < -_overlap || x >= _extent[0] + _overlap ||
        y >= _extent[1] + _overlap || z >= _extent[2] + _overlap) {
      return 0;
    }
    return get(x, y, z);
  }

  std::array<int, 3> _extent;
  int _overlap;
  std::array<bool, 3> _periodic;
  std::array<T, 3> _origin{T(0), T(0), T(0)};
  T _deltaR = T(1);
  std::array<int, 3> _size{};
  std::vector<std::uint8_t> _material;
};

// ---- statistics (core/latticeStatistics.h:43-127) -------------------------
template <typename T>
struct LatticeStatistics {
  std::size_t numCells = 0;
  T averageRho = 0;
  T averageEnergy = 0;  // 0.5 u^2 (core/latticeStatistics.hh:89)
  T maxU = 0;
  std::size_t getNumCells() const { return numCells; }
  T getAverageRho() const { return averageRho; }
  T getAverageEnergy() const { return averageEnergy; }
  T getMaxU() const { return maxU; }
};

// ---- base64 (RFC 4648 alphabet with padding, as OpenLB's Base64Encoder) ----
inline std::string base64(const std::uint8_t* data, std::size_t n)
{
  static const char* A = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  for (std::size_t i = 0; i < n; i += 3) {
    const std::uint32_t v = (std::uint32_t)data[i] << 16 | (i + 1 < n ? (std::uint32_t)data[i + 1] << 8 : 0) |
                            (i + 2 < n ? (std::uint32_t)data[i + 2] : 0);
    out += A[(v >> 18) & 63];
    out += A[(v >> 12) & 63];
    out += i + 1 < n ? A[(v >> 6) & 63] : '=';
    out += i + 2 < n ? A[v & 63] : '=';
  }
  return out;
}

inline std::vector<std::uint8_t> unbase64(const std::string& text)
{
  auto val = [](char c) -> int {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+') return 62;
    if (c == '/') return 63;
    return -1;
  };
  std::vector<std::uint8_t> out;
  out.reserve(text.size() / 4 * 3);
  std::uint32_t acc = 0;
  int bits = 0;
  for (char c : text) {
    const int v = val(c);
    if (v < 0) continue;  // padding and line breaks
    acc = acc << 6 | (std::uint32_t)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out.push_back((std::uint8_t)(acc >> bits));
    }
  }
  return out;
}

// ---- discrete normals (geometry/discreteNormals.h:33-57) ------------------
/// DiscreteNormalType, with the reference's values
enum class NormalType : int { Flat = 0, ExternalCorner = 1, InternalCorner = 2, ExternalEdge = 3, InternalEdge = 4 };

struct NormalPattern {
  NormalType type;
  std::array<int, 3> normal;
  const char* tests;  // "Kxyz" words: K in F (fluid), B (neither fluid nor outside), N (not fluid)
};

/// the neighbourhood patterns of computeBoundaryTypeAndNormal
/// (geometry/discreteNormals.hh:176-1190) in the reference's order; every
/// pattern also needs the cell itself to be B.  The first flat, corner or
/// external-edge pattern that matches decides; an internal-edge pattern that
/// matches afterwards replaces it (the last one wins).
inline const std::vector<NormalPattern>& normalPatterns()
{
  static const std::vector<NormalPattern> patterns = {
      {NormalType::Flat, {-1, 0, 0}, "F+00 B0+0 B0-0 B00+ B00-"},
      {NormalType::Flat, {1, 0, 0}, "F-00 B0+0 B0-0 B00+ B00-"},
      {NormalType::Flat, {0, -1, 0}, "F0+0 B+00 B-00 B00+ B00-"},
      {NormalType::Flat, {0, 1, 0}, "F0-0 B+00 B-00 B00+ B00-"},
      {NormalType::Flat, {0, 0, -1}, "F00+ B+00 B-00 B0+0 B0-0"},
      {NormalType::Flat, {0, 0, 1}, "F00- B+00 B-00 B0+0 B0-0"},
      {NormalType::ExternalCorner, {-1, -1, -1}, "F+++ B++0 B+0+ B+00 B0++ B0+0 B00+"},
      {NormalType::ExternalCorner, {-1, 1, -1}, "F+-+ B+-0 B+0+ B+00 B0-+ B0-0 B00+"},
      {NormalType::ExternalCorner, {-1, 1, 1}, "F+-- B+-0 B+0- B+00 B0-- B0-0 B00-"},
      {NormalType::ExternalCorner, {-1, -1, 1}, "F++- B++0 B+0- B+00 B0+- B0+0 B00-"},
      {NormalType::ExternalCorner, {1, 1, 1}, "F--- B--0 B-0- B-00 B0-- B0-0 B00-"},
      {NormalType::ExternalCorner, {1, -1, 1}, "F-+- B-+0 B-0- B-00 B0+- B0+0 B00-"},
      {NormalType::ExternalCorner, {1, -1, -1}, "F-++ B-+0 B-0+ B-00 B0++ B0+0 B00+"},
      {NormalType::ExternalCorner, {1, 1, -1}, "F--+ B--0 B-0+ B-00 B0-+ B0-0 B00+"},
      {NormalType::InternalCorner, {1, 1, 1}, "F-00 F0-0 F00- B+00 B0+0 B00+"},
      {NormalType::InternalCorner, {1, -1, 1}, "F-00 F0+0 F00- B+00 B0-0 B00+"},
      {NormalType::InternalCorner, {1, -1, -1}, "F-00 F0+0 F00+ B+00 B0-0 B00-"},
      {NormalType::InternalCorner, {1, 1, -1}, "F-00 F0-0 F00+ B+00 B0+0 B00-"},
      {NormalType::InternalCorner, {-1, 1, 1}, "F+00 F0-0 F00- B-00 B0+0 B00+"},
      {NormalType::InternalCorner, {-1, -1, 1}, "F+00 F0+0 F00- B-00 B0-0 B00+"},
      {NormalType::InternalCorner, {-1, 1, -1}, "F+00 F0-0 F00+ B-00 B0+0 B00-"},
      {NormalType::InternalCorner, {-1, -1, -1}, "F+00 F0+0 F00+ B-00 B0-0 B00-"},
      {NormalType::ExternalEdge, {0, 1, -1}, "F0-+ B+00 B-00 B0-0 B00+ N+0+ N-0+"},
      {NormalType::ExternalEdge, {0, -1, -1}, "F0++ B+00 B-00 B0+0 B00+ N+0+ N-0+"},
      {NormalType::ExternalEdge, {0, -1, 1}, "F0+- B+00 B-00 B00- N+0- N-0- N0+0"},
      {NormalType::ExternalEdge, {0, 1, 1}, "F0-- B+00 B-00 B00- N+0- N-0- N0-0"},
      {NormalType::ExternalEdge, {-1, 0, -1}, "F+0+ B+00 B0+0 B0-0 B00+ N-00 N00-"},
      {NormalType::ExternalEdge, {1, 0, -1}, "F-0+ B-00 B0+0 B0-0 B00+ N+00 N00-"},
      {NormalType::ExternalEdge, {-1, 0, 1}, "F+0- B+00 B0+0 B0-0 B00- N-00 N00+"},
      {NormalType::ExternalEdge, {1, 0, 1}, "F-0- B-00 B0+0 B0-0 B00- N+00 N00+"},
      {NormalType::ExternalEdge, {-1, -1, 0}, "F++0 B0+0 B00+ B00- N+00"},
      {NormalType::ExternalEdge, {1, -1, 0}, "F-+0 B0+0 B00+ B00- N-00"},
      {NormalType::ExternalEdge, {1, 1, 0}, "F--0 B0-0 B00+ B00- N-00"},
      {NormalType::ExternalEdge, {-1, 1, 0}, "F+-0 B0-0 B00+ B00- N+00"},
      {NormalType::InternalEdge, {0, -1, -1}, "F0+0 F00+ B+-0 B+00 B--0 B-00 B0-0 B00-"},
      {NormalType::InternalEdge, {0, 1, -1}, "F0-0 F00+ B++0 B+00 B-+0 B-00 B0+0 B00-"},
      {NormalType::InternalEdge, {0, -1, 1}, "F0+0 F00- B+-0 B+00 B--0 B-00 B0-0 B00+"},
      {NormalType::InternalEdge, {0, 1, 1}, "F0-0 F00- B++0 B+00 B-+0 B-00 B0+0 B00+"},
      {NormalType::InternalEdge, {1, 0, 1}, "F-00 F00- B+00 B0++ B0+0 B0-+ B0-0 B00+"},
      {NormalType::InternalEdge, {1, 0, -1}, "F-00 F00+ B+00 B0+- B0+0 B0-- B0-0 B00-"},
      {NormalType::InternalEdge, {-1, 0, 1}, "F+00 F00- B-00 B0++ B0+0 B0-+ B0-0 B00+"},
      {NormalType::InternalEdge, {-1, 0, -1}, "F+00 F00+ B-00 B0+- B0+0 B0-- B0-0 B00-"},
      {NormalType::InternalEdge, {1, 1, 0}, "F--0 F-00 F0-0 B0+0 B00+ B00-"},
      {NormalType::InternalEdge, {-1, 1, 0}, "F+-0 F+00 F0-0 B0+0 B00+ B00-"},
      {NormalType::InternalEdge, {1, -1, 0}, "F-+0 F-00 F0+0 B0-0 B00+ B00-"},
      {NormalType::InternalEdge, {-1, -1, 0}, "F++0 F+00 F0+0 B0-0 B00+ B00-"},
  };
  return patterns;
}

/// type and outward normal of the cell at (x, y, z) (fluid = material 1,
/// outside = material 0); a cell no pattern matches gives Flat with the
/// zero normal, as the reference returns it
template <typename T>
std::pair<NormalType, std::array<int, 3>> discreteNormal(const SuperGeometry<T>& g, int x, int y, int z)
{
  auto holds = [&](char kind, int dx, int dy, int dz) {
    const int m = g.getOrOutside(x + dx, y + dy, z + dz);
    switch (kind) {
    case 'F': return m == 1;
    case 'N': return m != 1;
    default: return m != 1 && m != 0;
    }
  };
  std::pair<NormalType, std::array<int, 3>> out{NormalType::Flat, {0, 0, 0}};
  if (!holds('B', 0, 0, 0)) {
    return out;
  }
  auto step = [](char c) { return c == '+' ? 1 : (c == '-' ? -1 : 0); };
  bool decided = false;
  for (const NormalPattern& p : normalPatterns()) {
    if (decided && p.type != NormalType::InternalEdge) {
      continue;
    }
    bool match = true;
    for (const char* w = p.tests; *w && match; w += (w[4] ? 5 : 4)) {
      match = holds(w[0], step(w[1]), step(w[2]), step(w[3]));
    }
    if (match) {
      out = {p.type, p.normal};
      decided = true;
    }
  }
  return out;
}


// ---- the lattice ---------------------------------------------------------

template <typename T, typename DESCRIPTOR = descriptors::D3Q19>
class SuperLattice {
  static_assert(std::is_same_v<DESCRIPTOR, descriptors::D3Q19>, "the gfx950 platform provides D3Q19");
  static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "T is float or double");

public:
  explicit SuperLattice(SuperGeometry<T>& geometry,
                        Storage storage = std::is_same_v<T, double> ? Storage::F64 : Storage::F32)
    : _geometry(geometry), _storage(storage)
  {
    if ((storage == Storage::F64) != std::is_same_v<T, double>) {
      throw std::invalid_argument("FP64 storage goes with T = double, the others with T = float");
    }
    const auto e = geometry.getExtent();
    const auto p = geometry.getPeriodicity();
    olbgfx_lattice_desc desc{};
    desc.nx = e[0];
    desc.ny = e[1];
    desc.nz = e[2];
    desc.padding = geometry.getOverlap();
    for (int a = 0; a < 3; ++a) desc.periodic[a] = p[a] ? 1 : 0;
    desc.storage = (int32_t)storage;
    olbgfx_lattice* h = nullptr;
    check(olbgfx_lattice_create(&desc, &h));
    _handle.reset(h);
    _ncells = geometry.getNcells();
    _slotOfCell.assign(_ncells, 0);
    _slots.push_back(Slot{OLBGFX_NO_DYNAMICS});
  }

  SuperLattice(const SuperLattice&) = delete;
  SuperLattice& operator=(const SuperLattice&) = delete;

  olbgfx_lattice* handle() { return _handle.get(); }
  std::size_t getNcells() const { return _ncells; }

  /// cells of `material` get DYNAMICS (core/superLattice.h:246-250)
  template <template <typename, typename> class DYNAMICS>
  void defineDynamics(SuperGeometry<T>& geometry, int material)
  {
    const int slot = slotFor(DYNAMICS<T, DESCRIPTOR>::kind);
    forMaterial(geometry, material, [&](CellID c) { _slotOfCell[c] = (std::uint8_t)slot; });
    _dirty = true;
  }

  /// LocalVelocity boundary on the cells of `material`
  /// (boundary::set<boundary::LocalVelocity<T,D,BGKdynamics>>,
  /// boundary/setBoundary.h:76-100, boundary/localVelocity3D.h:46-115): each
  /// cell at least two cells inside the padded block gets its discrete
  /// normal (fluid = material 1, outside = material 0); flat faces become
  /// regularised velocity cells, external edges and corners fixed-velocity
  /// BGK cells with their post-processors, internal edges and corners the
  /// inner-edge and inner-corner velocity cells.  A cell without a discrete
  /// normal keeps the dynamics it had, as in the reference.  With
  /// flatKind = OLBGFX_INTERPOLATED_VELOCITY the flat faces are those of
  /// boundary::InterpolatedVelocity instead (BGK faces with the plane
  /// finite-difference processor, boundary/interpolatedVelocity3D.h:45-100).
  void setLocalVelocity(SuperGeometry<T>& geometry, int material, int flatKind = OLBGFX_LOCAL_VELOCITY)
  {
    if (_normals.empty()) {
      _normals.assign(3 * _ncells, 0);
    }
    geometry.forSpatialLocations(2, [&](int x, int y, int z) {
      if (geometry.get(x, y, z) != material) return;
      const auto [type, n] = discreteNormal(geometry, x, y, z);
      if (n[0] == 0 && n[1] == 0 && n[2] == 0) return;
      int kind = flatKind;
      switch (type) {
      case NormalType::Flat: kind = flatKind; break;
      case NormalType::ExternalEdge:
      case NormalType::ExternalCorner: kind = OLBGFX_FIXED_VELOCITY; break;
      case NormalType::InternalEdge: kind = OLBGFX_INNER_EDGE_VELOCITY; break;
      case NormalType::InternalCorner: kind = OLBGFX_INNER_CORNER_VELOCITY; break;
      }
      const CellID c = geometry.getCellId(x, y, z);
      _slotOfCell[c] = (std::uint8_t)slotFor(kind);
      for (int a = 0; a < 3; ++a) _normals[3 * (std::size_t)c + a] = (std::int8_t)n[a];
    });
    _dirty = true;
  }

  /// the cells that run BouzidiPostProcessor and their 19 distances each
  /// (setBouzidiBoundary below computes them); replaces any earlier list
  void setBouzidi(std::vector<std::int64_t> cells, std::vector<double> q)
  {
    if (q.size() != (std::size_t)DESCRIPTOR::q * cells.size()) {
      throw std::invalid_argument("19 Bouzidi distances per cell");
    }
    _bouzidiCells = std::move(cells);
    _bouzidiQ = std::move(q);
    _dirty = true;
  }

  /// the cells that run BouzidiVelocityPostProcessor with their 19
  /// distances and 19 BOUZIDI_VELOCITY coefficients (c_i . u_wall at the wall
  /// intersection, as setBouzidiVelocity writes them,
  /// boundary/setBouzidiBoundary.h:340-420); applied after the setBouzidi
  /// list; replaces any earlier velocity list
  void setBouzidiVelocity(std::vector<std::int64_t> cells, std::vector<double> q, std::vector<double> velocity)
  {
    if (q.size() != (std::size_t)DESCRIPTOR::q * cells.size() || velocity.size() != q.size()) {
      throw std::invalid_argument("19 Bouzidi distances and 19 velocity coefficients per cell");
    }
    _bouzidiVelocityCells = std::move(cells);
    _bouzidiVelocityQ = std::move(q);
    _bouzidiVelocityV = std::move(velocity);
    _dirty = true;
  }

  /// LocalPressure on the flat faces of `material` (boundary::set<boundary::
  /// LocalPressure>, boundary/localPressure3D.h:47-58): regularised faces
  /// with the density the RHO field (defineRho) or 1 gives and the velocity's
  /// normal component from the known populations; any other boundary shape
  /// throws std::runtime_error, as the reference does
  void setLocalPressure(SuperGeometry<T>& geometry, int material)
  {
    const int face = slotFor(OLBGFX_LOCAL_PRESSURE);
    if (_normals.empty()) {
      _normals.assign(3 * _ncells, 0);
    }
    geometry.forSpatialLocations(2, [&](int x, int y, int z) {
      if (geometry.get(x, y, z) != material) return;
      const auto [type, n] = discreteNormal(geometry, x, y, z);
      if (type != NormalType::Flat) {
        throw std::runtime_error("No valid discrete normal found. This BC is not suited for curved walls.");
      }
      if (n[0] == 0 && n[1] == 0 && n[2] == 0) {
        // boundary/setBoundary3D.h:204-226
        throw std::runtime_error("Could not set Boundary.");
      }
      const CellID c = geometry.getCellId(x, y, z);
      _slotOfCell[c] = (std::uint8_t)face;
      for (int a = 0; a < 3; ++a) _normals[3 * (std::size_t)c + a] = (std::int8_t)n[a];
    });
    _dirty = true;
  }

  /// InterpolatedPressure on the flat faces of `material`
  /// (boundary::set<boundary::InterpolatedPressure>,
  /// boundary/interpolatedPressure3D.h:45-85): BGK faces with fixed density
  /// (defineRho, or 1) and the normal velocity from the known populations,
  /// run through the plane finite-difference processor; every other cell
  /// keeps its dynamics
  void setInterpolatedPressure(SuperGeometry<T>& geometry, int material)
  {
    if (_normals.empty()) {
      _normals.assign(3 * _ncells, 0);
    }
    geometry.forSpatialLocations(2, [&](int x, int y, int z) {
      if (geometry.get(x, y, z) != material) return;
      const auto [type, n] = discreteNormal(geometry, x, y, z);
      if (type != NormalType::Flat || (n[0] == 0 && n[1] == 0 && n[2] == 0)) return;
      const CellID c = geometry.getCellId(x, y, z);
      _slotOfCell[c] = (std::uint8_t)slotFor(OLBGFX_INTERPOLATED_PRESSURE);
      for (int a = 0; a < 3; ++a) _normals[3 * (std::size_t)c + a] = (std::int8_t)n[a];
    });
    _dirty = true;
  }

  /// the RHO field (FixedDensity momenta) on the cells of `material`
  void defineRho(SuperGeometry<T>& geometry, int material, T rho)
  {
    auto& field = fieldOf(OLBGFX_FIELD_RHO);
    forMaterial(geometry, material, [&](CellID c) {
      field.values[c] = rho;
      field.set[c] = 1;
    });
    _dirty = true;
  }

  /// set a dynamics parameter on every dynamics that has it
  template <typename PARAMETER>
  void setParameter(T value)
  {
    for (Slot& s : _slots) {
      if constexpr (std::is_same_v<PARAMETER, descriptors::OMEGA>) {
        if (s.kind != OLBGFX_NO_DYNAMICS && s.kind != OLBGFX_BOUNCE_BACK &&
            s.kind != OLBGFX_BOUNCE_BACK_VELOCITY) {
          s.omega = value;
        }
      } else if constexpr (std::is_same_v<PARAMETER, collision::TRT::MAGIC>) {
        if (s.kind == OLBGFX_TRT || s.kind == OLBGFX_FORCED_TRT) {
          s.magic = value;
        }
      } else if constexpr (std::is_same_v<PARAMETER, collision::LES::SMAGORINSKY>) {
        if (s.kind == OLBGFX_SMAGORINSKY_BGK) {
          s.smagorinsky = value;
        }
      } else {
        static_assert(sizeof(PARAMETER) == 0, "parameter not provided by the gfx950 platform");
      }
    }
    _dirty = true;
  }

  /// per-cell FORCE or VELOCITY on the cells of `material`
  template <typename FIELD>
  void defineField(SuperGeometry<T>& geometry, int material, std::array<T, 3> value)
  {
    auto& field = fieldOf(FIELD::field);
    forMaterial(geometry, material, [&](CellID c) {
      for (int a = 0; a < 3; ++a) field.values[3 * (std::size_t)c + a] = value[a];
      field.set[c] = 1;
    });
    _dirty = true;
  }

  /// u of cells with externally stored velocity (momenta::setVelocity on
  /// BounceBackVelocity cells writes their VELOCITY field)
  void defineU(SuperGeometry<T>& geometry, int material, std::array<T, 3> u)
  {
    defineField<descriptors::VELOCITY>(geometry, material, u);
  }

  /// f_i = feq_i(rho(x), u(x)) on the cells of `material`
  void iniEquilibrium(SuperGeometry<T>& geometry, int material,
                      const std::function<T(int, int, int)>& rho,
                      const std::function<std::array<T, 3>(int, int, int)>& u)
  {
    std::vector<std::uint8_t> sel(_ncells, 0);
    std::vector<T> r(_ncells, T(1)), v(3 * _ncells, T(0));
    geometry.forCore([&](int x, int y, int z) {
      if (geometry.get(x, y, z) != material) return;
      const CellID c = geometry.getCellId(x, y, z);
      sel[c] = 1;
      r[c] = rho(x, y, z);
      const auto uu = u(x, y, z);
      for (int a = 0; a < 3; ++a) v[3 * (std::size_t)c + a] = uu[a];
    });
    check(olbgfx_ini_equilibrium(handle(), sel.data(), valueType(), r.data(), v.data()));
  }

  /// upload dynamics, parameters and fields (core/superLattice.hh:248-266)
  void initialize()
  {
    if (_slots.size() > OLBGFX_MAX_DYNAMICS) {
      throw std::invalid_argument("too many distinct dynamics for one lattice");
    }
    std::vector<olbgfx_dynamics_desc> descs;
    for (const Slot& s : _slots) descs.push_back(s.desc());
    resolveField(OLBGFX_FIELD_FORCE, descs);
    resolveField(OLBGFX_FIELD_VELOCITY, descs);
    for (std::size_t k = 0; k < descs.size(); ++k) {
      check(olbgfx_set_dynamics(handle(), (uint32_t)k, &descs[k]));
    }
    check(olbgfx_set_dynamics_mask(handle(), _slotOfCell.data()));
    check(olbgfx_set_cell_normals(handle(), _normals.empty() ? nullptr : _normals.data()));
    check(olbgfx_set_bouzidi(handle(), (int64_t)_bouzidiCells.size(), _bouzidiCells.data(), OLBGFX_VALUE_F64,
                             _bouzidiQ.data()));
    check(olbgfx_set_bouzidi_velocity(handle(), (int64_t)_bouzidiVelocityCells.size(),
                                      _bouzidiVelocityCells.data(), OLBGFX_VALUE_F64, _bouzidiVelocityQ.data(),
                                      _bouzidiVelocityV.data()));
    for (auto& kv : _fields) {
      check(olbgfx_set_cell_field(handle(), kv.first, valueType(),
                                  kv.second.uniform ? nullptr : kv.second.values.data()));
    }
    check(olbgfx_set_statistics(handle(), _statistics ? 1 : 0));
    if (!_initialized) {
      // BlockLattice::initialize runs the PostStream processors once
      // (core/blockLattice.hh:56-67)
      check(olbgfx_post_process(handle(), nullptr));
    }
    _dirty = false;
    _initialized = true;
  }

  /// one step on the default stream (core/superLattice.hh:663-725)
  void collideAndStream()
  {
    if (!_initialized || _dirty) {
      initialize();
    }
    check(olbgfx_collide_and_stream(handle(), 1, nullptr));
    _hostValid = false;
  }

  /// Launch autotuner of the platform (no reference counterpart): times the
  /// candidate launch shapes on this lattice, applies the fastest and puts
  /// the populations back, so the simulation does not advance.  An empty
  /// cachePath tunes without a cache file.
  olbgfx_tune_trial autotune(const std::string& cachePath = "", std::uint32_t steps = 4)
  {
    if (!_initialized || _dirty) {
      initialize();
    }
    olbgfx_tune_result r;
    check(olbgfx_autotune_lattice(handle(), steps, nullptr, cachePath.empty() ? nullptr : cachePath.c_str(),
                                  nullptr, &r));
    return r.best;
  }

  void setStatisticsOn() { setStatistics(true); }
  void setStatisticsOff() { setStatistics(false); }

  LatticeStatistics<T> getStatistics()
  {
    double s[4];
    check(olbgfx_get_statistics(handle(), s));
    LatticeStatistics<T> st;
    st.numCells = (std::size_t)s[0];
    if (s[0] > 0) {
      st.averageRho = T(s[1] / s[0]);
      st.averageEnergy = T(0.5 * s[2] / s[0]);
      st.maxU = T(std::sqrt(s[3]));
    }
    return st;
  }

  /// Evaluation copies the populations into a host mirror in logical order
  /// (core/platform/gpu/hip/column.hh:289-300)
  void setProcessingContext(ProcessingContext context)
  {
    if (context == ProcessingContext::Evaluation) {
      _host.resize((std::size_t)DESCRIPTOR::q * _ncells);
      check(olbgfx_get_populations(handle(), valueType(), _host.data()));
      _hostValid = true;
    } else if (_hostValid) {
      check(olbgfx_set_populations(handle(), valueType(), _host.data()));
    }
  }

  /// Checkpoint in OpenLB's format (SuperLattice::save, core/serializer.h:181,
  /// io/serializerIO.hh:40-64): base64 of the 8-byte binary size, then
  /// base64 of the 19 population column blocks [shift][count][values]
  /// (core/platform/cpu/simd/column.hh:65-87 for CpuSimd, two copies of a
  /// page-aligned ring; cpu/sisd/column.hh:67-89 for CpuSisd, one ring of
  /// the padded cell count).  Unlike OpenLB, `fileName` is the whole path.
  enum class CheckpointLayout { CpuSimd, CpuSisd };

  void save(const std::string& fileName, CheckpointLayout layout = CheckpointLayout::CpuSimd)
  {
    setProcessingContext(ProcessingContext::Evaluation);
    const std::size_t n = _ncells;
    const std::size_t count = layout == CheckpointLayout::CpuSimd ? ((n * sizeof(T) - 1) / 4096 + 1) * 4096 / sizeof(T) : n;
    const std::size_t copies = layout == CheckpointLayout::CpuSimd ? 2 : 1;
    std::vector<std::uint8_t> body;
    body.reserve(DESCRIPTOR::q * (16 + copies * count * sizeof(T)));
    auto put = [&](const void* p, std::size_t bytes) {
      const auto* b = static_cast<const std::uint8_t*>(p);
      body.insert(body.end(), b, b + bytes);
    };
    std::vector<T> ring(count);
    for (int i = 0; i < DESCRIPTOR::q; ++i) {
      const std::int64_t shift = 0;
      const std::uint64_t c = count;
      put(&shift, 8);
      put(&c, 8);
      std::fill(ring.begin(), ring.end(), T(0));
      std::copy(_host.begin() + (std::ptrdiff_t)(i * n), _host.begin() + (std::ptrdiff_t)((i + 1) * n), ring.begin());
      for (std::size_t k = 0; k < copies; ++k) put(ring.data(), count * sizeof(T));
    }
    const std::uint64_t size = body.size();
    std::ofstream out(fileName, std::ios::binary);
    if (!out) throw std::runtime_error("cannot write " + fileName);
    out << base64(reinterpret_cast<const std::uint8_t*>(&size), 8) << base64(body.data(), body.size());
  }

  /// restart from a checkpoint of either layout written by SuperLattice::save
  /// of a lattice of this block's size and value type
  void load(const std::string& fileName)
  {
    std::ifstream in(fileName, std::ios::binary);
    if (!in) throw std::runtime_error("cannot read " + fileName);
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    if (text.size() < 12) throw std::invalid_argument("not an OpenLB checkpoint");
    const std::vector<std::uint8_t> head = unbase64(text.substr(0, 12));
    std::uint64_t size = 0;
    std::memcpy(&size, head.data(), 8);
    const std::vector<std::uint8_t> body = unbase64(text.substr(12));
    if (body.size() != size || size % DESCRIPTOR::q != 0) throw std::invalid_argument("checkpoint size mismatch");
    const std::size_t block = size / DESCRIPTOR::q, n = _ncells;
    _host.resize((std::size_t)DESCRIPTOR::q * n);
    for (int i = 0; i < DESCRIPTOR::q; ++i) {
      const std::uint8_t* b = body.data() + i * block;
      std::int64_t shift = 0;
      std::uint64_t count = 0;
      std::memcpy(&shift, b, 8);
      std::memcpy(&count, b + 8, 8);
      if (count < n || (16 + count * sizeof(T) != block && 16 + 2 * count * sizeof(T) != block)) {
        throw std::invalid_argument("checkpoint column does not fit this lattice");
      }
      const T* data = reinterpret_cast<const T*>(b + 16);
      for (std::size_t c = 0; c < n; ++c) {
        const std::int64_t j = ((std::int64_t)c + shift) % (std::int64_t)count;
        std::memcpy(&_host[(std::size_t)i * n + c], data + (j < 0 ? j + (std::int64_t)count : j), sizeof(T));
      }
    }
    _hostValid = true;
    setProcessingContext(ProcessingContext::Simulation);
  }

  /// density and velocity of every padded cell through the momenta of its
  /// dynamics (Cell::computeRhoU), computed on the device
  void computeRhoU(std::vector<T>& rho, std::vector<T>& u)
  {
    if (!_initialized || _dirty) {
      initialize();
    }
    rho.resize(_ncells);
    u.resize(3 * _ncells);
    check(olbgfx_compute_rho_u(handle(), valueType(), rho.data(), u.data()));
  }

  SuperGeometry<T>& getGeometry() { return _geometry; }

  /// population iPop of cell iCell, shifted by -t_i (evaluation mirror)
  T& population(CellID iCell, int iPop)
  {
    if (!_hostValid) {
      throw std::logic_error("call setProcessingContext(Evaluation) first");
    }
    return _host[(std::size_t)iPop * _ncells + iCell];
  }

  /// density and velocity of one cell from the evaluation mirror
  void computeRhoU(CellID iCell, T& rho, T u[3])
  {
    T r = 0;
    T j[3] = {0, 0, 0};
    for (int i = 0; i < DESCRIPTOR::q; ++i) {
      const T f = population(iCell, i);
      r += f;
      for (int a = 0; a < 3; ++a) j[a] += f * T(DESCRIPTOR::c[i][a]);
    }
    rho = r + T(1);
    for (int a = 0; a < 3; ++a) u[a] = j[a] / rho;
  }

private:
  struct Slot {
    int32_t kind;
    T omega = 0;
    T magic = 0;
    T smagorinsky = 0;
    olbgfx_dynamics_desc desc() const
    {
      olbgfx_dynamics_desc d{};
      d.kind = kind;
      d.omega = (double)omega;
      d.magic = (double)magic;
      d.smagorinsky = (double)smagorinsky;
      d.wall_rho = 1.0;
      return d;
    }
  };
  struct Field {
    std::vector<T> values;
    std::vector<std::uint8_t> set;
    bool uniform = false;
  };
  struct Deleter {
    void operator()(olbgfx_lattice* h) const { olbgfx_lattice_destroy(h); }
  };

  int32_t valueType() const { return std::is_same_v<T, double> ? OLBGFX_VALUE_F64 : OLBGFX_VALUE_F32; }

  int slotFor(int32_t kind)
  {
    for (std::size_t k = 0; k < _slots.size(); ++k) {
      if (_slots[k].kind == kind) return (int)k;
    }
    _slots.push_back(Slot{kind});
    return (int)_slots.size() - 1;
  }

  // RHO holds one value per cell (1 where undefined, FixedDensity's
  // initial value), the vector fields three
  Field& fieldOf(int32_t id)
  {
    Field& f = _fields[id];
    if (f.values.empty()) {
      const bool scalar = id == OLBGFX_FIELD_RHO;
      f.values.assign((scalar ? 1 : 3) * _ncells, T(scalar ? 1 : 0));
      f.set.assign(_ncells, 0);
    }
    return f;
  }

  template <typename F>
  void forMaterial(SuperGeometry<T>& geometry, int material, F f)
  {
    for (CellID c = 0; c < (CellID)_ncells; ++c) {
      if (geometry.materials()[c] == material) f(c);
    }
  }

  // A field whose value is the same on every cell of each dynamics that has
  // it (and that covers all of that dynamics' cells) goes into the dynamics
  // table and costs no per-cell reads; otherwise it is uploaded per cell.
  void resolveField(int32_t id, std::vector<olbgfx_dynamics_desc>& descs)
  {
    auto it = _fields.find(id);
    if (it == _fields.end()) {
      return;
    }
    Field& f = it->second;
    const std::size_t ns = _slots.size();
    std::vector<std::array<T, 3>> value(ns);
    std::vector<char> seen(ns, 0), missing(ns, 0);
    bool table = true;
    for (CellID c = 0; c < (CellID)_ncells && table; ++c) {
      const std::size_t k = _slotOfCell[c];
      if (!f.set[c]) {
        missing[k] = 1;
        continue;
      }
      const std::array<T, 3> w{f.values[3 * (std::size_t)c], f.values[3 * (std::size_t)c + 1],
                               f.values[3 * (std::size_t)c + 2]};
      if (!seen[k]) {
        value[k] = w;
        seen[k] = 1;
      } else if (w != value[k]) {
        table = false;
      }
    }
    for (std::size_t k = 0; k < ns && table; ++k) {
      if (seen[k] && missing[k]) table = false;
    }
    f.uniform = table;
    if (!table) {
      return;
    }
    for (std::size_t k = 0; k < ns; ++k) {
      if (!seen[k]) continue;
      for (int a = 0; a < 3; ++a) {
        if (id == OLBGFX_FIELD_FORCE) {
          descs[k].force[a] = (double)value[k][a];
        } else {
          descs[k].wall_velocity[a] = (double)value[k][a];
        }
      }
    }
  }

  void setStatistics(bool on)
  {
    _statistics = on;
    check(olbgfx_set_statistics(handle(), on ? 1 : 0));
  }

  SuperGeometry<T>& _geometry;
  Storage _storage;
  std::unique_ptr<olbgfx_lattice, Deleter> _handle;
  std::size_t _ncells = 0;
  std::vector<Slot> _slots;
  std::vector<std::uint8_t> _slotOfCell;
  std::vector<std::int8_t> _normals;  // discrete normal per cell for LocalVelocity
  std::vector<std::int64_t> _bouzidiCells;
  std::vector<double> _bouzidiQ;
  std::vector<std::int64_t> _bouzidiVelocityCells;
  std::vector<double> _bouzidiVelocityQ;
  std::vector<double> _bouzidiVelocityV;
  std::map<int32_t, Field> _fields;
  std::vector<T> _host;
  bool _hostValid = false;
  bool _statistics = true;  // on by default (core/blockLattice.hh:42)
  bool _dirty = true;
  bool _initialized = false;
};

// ---- setters (case/setters.h:303-345, boundary/setBoundary.h:150-157) -----
namespace dynamics {
template <template <typename, typename> class DYNAMICS, typename T, typename DESCRIPTOR>
void set(SuperLattice<T, DESCRIPTOR>& lattice, SuperGeometry<T>& geometry, int material)
{
  lattice.template defineDynamics<DYNAMICS>(geometry, material);
}
}  // namespace dynamics

namespace boundary {
struct BounceBack {
  template <typename T, typename DESCRIPTOR>
  using dynamics = olb::BounceBack<T, DESCRIPTOR>;
};
/// boundary::LocalVelocity<T,D,BGKdynamics> (boundary/localVelocity3D.h:36-120)
struct LocalVelocity {
};
/// boundary::LocalPressure<T,D,BGKdynamics> (boundary/localPressure3D.h:36-62)
struct LocalPressure {
};
/// boundary::InterpolatedVelocity<T,D,BGKdynamics> (boundary/interpolatedVelocity3D.h:36-100)
struct InterpolatedVelocity {
};
/// boundary::InterpolatedPressure<T,D,BGKdynamics> (boundary/interpolatedPressure3D.h:36-85)
struct InterpolatedPressure {
};
template <typename BC, typename T, typename DESCRIPTOR>
void set(SuperLattice<T, DESCRIPTOR>& lattice, SuperGeometry<T>& geometry, int material)
{
  if constexpr (std::is_same_v<BC, LocalVelocity>) {
    lattice.setLocalVelocity(geometry, material);
  } else if constexpr (std::is_same_v<BC, LocalPressure>) {
    lattice.setLocalPressure(geometry, material);
  } else if constexpr (std::is_same_v<BC, InterpolatedVelocity>) {
    lattice.setLocalVelocity(geometry, material, OLBGFX_INTERPOLATED_VELOCITY);
  } else if constexpr (std::is_same_v<BC, InterpolatedPressure>) {
    lattice.setInterpolatedPressure(geometry, material);
  } else {
    lattice.template defineDynamics<BC::template dynamics>(geometry, material);
  }
}
}  // namespace boundary

// ---- Bouzidi boundaries (boundary/setBouzidiBoundary.h:193-330) ------------

/// analytical indicator: inside test, bounding box and the ray distance of
/// IndicatorF3D (functors/analytical/indicator/indicatorBaseF3D.hh:57-87,
/// the bisection of indicatorBase.h:35-93 with precision 1e-4, pitch 1/2)
template <typename S>
class IndicatorF3D {
public:
  virtual ~IndicatorF3D() = default;
  virtual bool operator()(bool output[1], const S input[3]) = 0;
  const std::array<S, 3>& getMin() const { return _myMin; }
  const std::array<S, 3>& getMax() const { return _myMax; }
  bool isInsideBox(const std::array<S, 3>& p) const
  {
    bool in = true;
    for (int a = 0; a < 3; ++a) in = in && _myMin[a] <= p[a] && p[a] <= _myMax[a];
    return in;
  }
  /// distance from origin along direction to the first change of inside /
  /// outside; false if the ray leaves the bounding box from outside first
  virtual bool distance(S& distance, const std::array<S, 3>& origin, const std::array<S, 3>& direction,
                        int iC = -1)
  {
    (void)iC;
    const S precision = S(.0001);
    S pitch = S(0.5);
    auto at = [](const std::array<S, 3>& o, S t, const std::array<S, 3>& d) {
      return std::array<S, 3>{o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
    };
    auto inside = [this](const std::array<S, 3>& p) {
      bool v = false;
      (*this)(&v, p.data());
      return v;
    };
    const bool nearIn = inside(at(origin, 10 * precision, direction));
    const bool nearOut = inside(at(origin, -(10 * precision), direction));
    if (nearIn == !nearOut) {
      distance = 0;
      return true;
    }
    const bool originValue = inside(origin);
    std::array<S, 3> cur = origin;
    bool value = originValue;
    while (value == originValue && isInsideBox(cur)) {
      for (int a = 0; a < 3; ++a) cur[a] += direction[a];
      value = inside(cur);
    }
    if (!isInsideBox(cur) && !originValue) {
      return false;
    }
    while (pitch >= precision) {
      if (!isInsideBox(cur) && originValue) {
        for (int a = 0; a < 3; ++a) cur[a] -= pitch * direction[a];
      } else {
        value = inside(cur);
        const S sign = value == originValue ? S(1) : S(-1);
        for (int a = 0; a < 3; ++a) cur[a] += sign * pitch * direction[a];
      }
      pitch *= S(0.5);
    }
    S n2 = 0;
    for (int a = 0; a < 3; ++a) n2 += (cur[a] - origin[a]) * (cur[a] - origin[a]);
    distance = std::sqrt(n2);
    return true;
  }

protected:
  std::array<S, 3> _myMin{}, _myMax{};
};

/// util::epsRelative (utilities/floatingPointRelationalOperators.h:278-296)
template <typename S>
S epsRelative(S a)
{
  const S e = std::numeric_limits<S>::epsilon();
  if (a == S() || (a > -e && a < e)) {
    return std::numeric_limits<S>::min() / e;
  }
  return S(3) * e * std::fabs(a);
}

/// IndicatorCylinder3D(center1, center2, radius)
/// (functors/analytical/indicator/indicatorF3D.hh:312-418)
template <typename S>
class IndicatorCylinder3D : public IndicatorF3D<S> {
public:
  IndicatorCylinder3D(std::array<S, 3> center1, std::array<S, 3> center2, S radius)
    : _c1(center1), _c2(center2), _radius2(radius * radius)
  {
    S baba = 0;
    for (int a = 0; a < 3; ++a) baba += (_c2[a] - _c1[a]) * (_c2[a] - _c1[a]);
    const S length = std::sqrt(baba);
    std::array<S, 3> K, I, J;
    for (int a = 0; a < 3; ++a) K[a] = (_c2[a] - _c1[a]) / length;
    if (_c2[1] == _c1[1] && _c2[0] == _c1[0]) {
      I = {1, 0, 0};
      J = {0, 1, 0};
    } else {
      const S normi = std::sqrt(K[1] * K[1] + K[0] * K[0]);
      I = {-K[1] / normi, K[0] / normi, 0};
      J = {K[1] * I[2] - K[2] * I[1], K[2] * I[0] - K[0] * I[2], K[0] * I[1] - K[1] * I[0]};
    }
    const S r = std::sqrt(_radius2);
    for (int a = 0; a < 3; ++a) {
      const S w = std::sqrt(I[a] * I[a] + J[a] * J[a]) * r;
      this->_myMax[a] = _c1[a] + w + std::max(K[a] * length, S(0));
      this->_myMin[a] = _c1[a] - w + std::min(K[a] * length, S(0));
    }
    _epsMax = std::max(epsRelative(this->_myMax[0]),
                       std::max(epsRelative(this->_myMax[1]), epsRelative(this->_myMax[2])));
    _epsRad = epsRelative(std::sqrt(_radius2));
  }

  bool operator()(bool output[1], const S input[3]) override
  {
    S pa[3], d[3];
    for (int a = 0; a < 3; ++a) {
      pa[a] = input[a] - _c1[a];
      d[a] = _c2[a] - _c1[a];
    }
    const S h = (pa[0] * d[0] + pa[1] * d[1] + pa[2] * d[2]) / (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (h < S(0) - _epsMax || h > S(1) + _epsMax) {
      output[0] = false;
      return true;
    }
    S dist2 = 0;
    for (int a = 0; a < 3; ++a) {
      const S r = input[a] - (_c1[a] + h * d[a]);
      dist2 += r * r;
    }
    output[0] = dist2 < _radius2 + _epsRad;
    return true;
  }

private:
  std::array<S, 3> _c1, _c2;
  S _radius2;
  S _epsMax = 0, _epsRad = 0;
};

/// BOUZIDI_DISTANCE field (19 per padded cell, -1 where unset) that
/// setBouzidiBoundary (setBouzidiBoundary.h:229-330) writes for the solid
/// material and the bulk materials: for every solid cell of the padded block
/// and every direction i with x_b = x_s + c_i inside the block, q[opp i] of
/// x_b is the ray distance along -c_i normed by |c_i| dx (0.5 if not found or
/// outside [0, 1], or if x_b + c_i is not bulk) when x_b is bulk, and 0.5
/// when x_b is neither bulk nor material 0
template <typename T>
std::vector<double> bouzidiDistances(const SuperGeometry<T>& geometry, int material, IndicatorF3D<T>& indicator,
                                     const std::vector<int>& bulkMaterials = {1})
{
  using D = descriptors::D3Q19;
  const auto e = geometry.getExtent();
  const int p = geometry.getOverlap();
  std::vector<double> q((std::size_t)D::q * geometry.getNcells(), -1.0);
  auto bulk = [&](int x, int y, int z) {
    const int m = geometry.getOrOutside(x, y, z);
    return std::find(bulkMaterials.begin(), bulkMaterials.end(), m) != bulkMaterials.end();
  };
  auto inside = [&](int x, int y, int z) {
    return x >= -p && y >= -p && z >= -p && x < e[0] + p && y < e[1] + p && z < e[2] + p;
  };
  const T dx = geometry.getDeltaR();
  for (int x = -p; x < e[0] + p; ++x) {
    for (int y = -p; y < e[1] + p; ++y) {
      for (int z = -p; z < e[2] + p; ++z) {
        if (geometry.get(x, y, z) != material) continue;
        for (int i = 1; i < D::q; ++i) {
          const int bx = x + D::c[i][0], by = y + D::c[i][1], bz = z + D::c[i][2];
          if (!inside(bx, by, bz)) continue;
          const std::size_t at = (std::size_t)D::opposite(i) * geometry.getNcells() + geometry.getCellId(bx, by, bz);
          if (bulk(bx, by, bz)) {
            const std::array<T, 3> origin = geometry.getPhysR(bx, by, bz);
            const std::array<T, 3> direction = {-dx * T(D::c[i][0]), -dx * T(D::c[i][1]), -dx * T(D::c[i][2])};
            const T norm = dx * std::sqrt(T(D::c[i][0] * D::c[i][0] + D::c[i][1] * D::c[i][1] +
                                            D::c[i][2] * D::c[i][2]));
            T dist = -1;
            T qi = 0.5;
            if (indicator.distance(dist, origin, direction)) {
              qi = dist / norm;
              if (qi < 0 || qi > 1) qi = 0.5;
            }
            q[at] = bulk(bx + D::c[i][0], by + D::c[i][1], bz + D::c[i][2]) ? (double)qi : 0.5;
          } else if (geometry.getOrOutside(bx, by, bz) != 0) {
            q[at] = 0.5;
          }
        }
      }
    }
  }
  return q;
}

/// setBouzidiBoundary<T, DESCRIPTOR, BouzidiPostProcessor>(lattice, geometry,
/// material, indicator, bulkMaterials): the non-padding cells with an entry
/// run the post-processor after each stream.  The reference applies them in
/// ascending cell id; the platform updates them concurrently, which differs
/// only for q = 1/2 copies between two solid cells that read each other's
/// output -- there the lower id takes the other's value and the higher id's
/// copy would write back what it holds, so it is dropped here.
template <typename T, typename DESCRIPTOR>
void setBouzidiBoundary(SuperLattice<T, DESCRIPTOR>& lattice, SuperGeometry<T>& geometry, int material,
                        IndicatorF3D<T>& indicator, std::vector<int> bulkMaterials = {1})
{
  using D = descriptors::D3Q19;
  const std::vector<double> field = bouzidiDistances(geometry, material, indicator, bulkMaterials);
  const std::size_t n = geometry.getNcells();
  const auto e = geometry.getExtent();
  std::vector<std::int64_t> cells;
  std::vector<std::int64_t> listed(n, -1);
  geometry.forCore([&](int x, int y, int z) {
    const CellID c = geometry.getCellId(x, y, z);
    for (int i = 1; i < D::q; ++i) {
      if (field[(std::size_t)i * n + c] >= 0) {
        listed[c] = (std::int64_t)cells.size();
        cells.push_back(c);
        return;
      }
    }
  });
  (void)e;
  std::sort(cells.begin(), cells.end());
  for (std::size_t k = 0; k < cells.size(); ++k) listed[(std::size_t)cells[k]] = (std::int64_t)k;
  std::vector<double> q(D::q * cells.size());
  for (std::size_t k = 0; k < cells.size(); ++k) {
    for (int i = 0; i < D::q; ++i) q[D::q * k + i] = field[(std::size_t)i * n + (std::size_t)cells[k]];
  }
  const std::int64_t step[3] = {(std::int64_t)(e[1] + 2 * geometry.getOverlap()) * (e[2] + 2 * geometry.getOverlap()),
                                (std::int64_t)(e[2] + 2 * geometry.getOverlap()), 1};
  for (std::size_t k = 0; k < cells.size(); ++k) {
    for (int j = 1; j < D::q; ++j) {
      if (q[D::q * k + j] < 0) continue;
      const std::int64_t dj = D::c[j][0] * step[0] + D::c[j][1] * step[1] + D::c[j][2] * step[2];
      const std::int64_t y = listed[(std::size_t)(cells[k] + dj)];
      const int o = D::opposite(j);
      if (y >= 0 && q[D::q * (std::size_t)y + o] >= 0) {
        if (q[D::q * k + j] != 0.5 || q[D::q * (std::size_t)y + o] != 0.5) {
          throw std::invalid_argument("two Bouzidi cells read each other's updates");
        }
        if ((std::size_t)y < k) q[D::q * k + j] = -1.0;
      }
    }
  }
  std::vector<std::int64_t> keptCells;
  std::vector<double> keptQ;
  for (std::size_t k = 0; k < cells.size(); ++k) {
    bool any = false;
    for (int i = 1; i < D::q; ++i) any = any || q[D::q * k + i] >= 0;
    if (!any) continue;
    keptCells.push_back(cells[k]);
    keptQ.insert(keptQ.end(), q.begin() + (std::ptrdiff_t)(D::q * k), q.begin() + (std::ptrdiff_t)(D::q * (k + 1)));
  }
  lattice.setBouzidi(std::move(keptCells), std::move(keptQ));
}

}  // namespace olb

#endif  // OLB_GFX950_H
