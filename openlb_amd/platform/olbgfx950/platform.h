/*
 * olbgfx950/platform.h -- Platform::GPU_GFX950, a platform of OpenLB whose
 * blocks step on an MI355X through libolbgfx (include/olbgfx.h).
 *
 * Included by OpenLB's core/blockLattice.h when PLATFORM_GPU_GFX950 is
 * defined.  It supplies what core/platform/gpu/hip supplies for GPU_HIP:
 *
 *   ConcreteBlockMask<T,GPU_GFX950>                     (gpu/hip/mask.h:31-85)
 *   ConcreteBlockCollisionO<T,D,GPU_GFX950,DYNAMICS>    (gpu/hip/operator.h:47-124)
 *   ConcreteBlockO<T,D,GPU_GFX950,OPERATOR,SCOPE>       (gpu/hip/operator.h:128-245)
 *   ConcreteBlockCommunicator<ConcreteBlockLattice<T,D,GPU_GFX950>>
 *                                                       (gpu/hip/communicator.h)
 *   StatisticsPostProcessor::type<ConcreteBlockLattice<T,D,GPU_GFX950>>
 *                                                       (gpu/hip/statistics.hh:79-115)
 *   checkPlatform<GPU_GFX950>                           (gpu/hip/communicator.hh:45-81)
 *
 * How a block runs.  The host side of a GPU_GFX950 block is OpenLB's own
 * (cell data, setters, functors, serializer all work on host columns, the
 * population columns being a mirror of the device rings, olbgfx950/column.h).
 * Each block has a device twin, an olbgfx lattice of the same extent,
 * padding and periodicity:
 *
 * - collide() of the first collision operator with cells runs one fused
 *   olbgfx step: collision of every core cell by the kind its dynamics maps
 *   to, the crossing populations written into the periodic overlap (the
 *   PostCollide exchange of a single block), the ring rotation and the
 *   PostStream boundary processors the kinds imply.  stream() then rotates
 *   the host mirror (no device work), the PostCollide and PostStream
 *   exchanges and the known post-processors are no-ops, the statistics come
 *   from the step;
 * - setProcessingContext(Evaluation) copies the populations to the mirror
 *   (and fills its periodic overlap as the reference's communicate() does);
 *   setProcessingContext(Simulation) after host changes uploads mirror,
 *   dynamics, parameters and fields before the next step.
 *
 * Dynamics without a kernel kind, post-processors the kinds do not cover and
 * more than one block per lattice throw std::domain_error at set-up, as the
 * reference throws for what a platform does not provide.
 */
#ifndef OLBGFX950_PLATFORM_H
#define OLBGFX950_PLATFORM_H

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <typeindex>
#include <vector>

#include "olbgfx.h"

namespace olb {

template<typename T, typename DESCRIPTOR, Platform PLATFORM> class ConcreteBlockLattice;
template<typename T, Platform PLATFORM> class ConcreteBlockMask;

namespace gfx950 {

/// Throw the library's last error as the reference's exception kinds
inline void check(olbgfx_status status, const char* what)
{
  if (status == OLBGFX_OK) {
    return;
  }
  const std::string msg = std::string(what) + ": " + olbgfx_last_error();
  switch (status) {
  case OLBGFX_ERR_INVALID:
    throw std::invalid_argument(msg);
  case OLBGFX_ERR_UNSUPPORTED:
    throw std::domain_error(msg);
  default:
    throw std::runtime_error(msg);
  }
}

/// A collision operator of a block, as one entry of the device's dynamics table
struct CollisionSource {
  virtual ~CollisionSource() = default;
  /// kind and the parameters it reads now
  virtual olbgfx_dynamics_desc describe() const = 0;
  /// 1 for every padded cell (core or overlap) that has this dynamics
  virtual const std::vector<std::uint8_t>& cells() const = 0;
  /// discrete normal the kind implies (LocalVelocity faces), else 0
  virtual int normal(int axis) const = 0;
  virtual bool hasNormal() const = 0;
};

/// A post-processor of a block that carries a discrete normal (the
/// LocalVelocity edge and corner processors): the normal of its cells
struct NormalSource {
  virtual ~NormalSource() = default;
  virtual const std::vector<std::uint8_t>& cells() const = 0;
  virtual int normal(int axis) const = 0;
};

/// Storage of the device rings: FP32 for float blocks, FP64 for double
/// ones; OLBGFX_STORAGE=f16|bf16 stores float blocks as shifted FP16/BF16
inline int storageFor(bool isDouble)
{
  if (isDouble) {
    return OLBGFX_STORE_F64;
  }
  if (const char* s = std::getenv("OLBGFX_STORAGE")) {
    if (std::strcmp(s, "f16") == 0) {
      return OLBGFX_STORE_F16;
    }
    if (std::strcmp(s, "bf16") == 0) {
      return OLBGFX_STORE_BF16;
    }
  }
  return OLBGFX_STORE_F32;
}

/// The device copy of one block lattice
class BlockTwin {
public:
  // typed access to the block, set by twinOf()
  std::function<void(void*)> gatherPopulations;      // mirror -> logical-order buffer
  std::function<void(const void*)> scatterPopulations; // logical-order buffer -> mirror
  /// per-cell fields the kinds read; given the cells' table ids and the table
  std::function<void(olbgfx_lattice*, const std::vector<std::uint8_t>&,
                     const std::vector<olbgfx_dynamics_desc>&)> uploadFields;
  std::function<bool()> statisticsEnabled;
  std::function<void()> fillPeriodicOverlap;           // mirror's overlap from its images

  int nx = 0, ny = 0, nz = 0, padding = 0;
  std::size_t ncells = 0;
  int valueType = OLBGFX_VALUE_F32;
  std::size_t valueSize = 4;
  int storage = OLBGFX_STORE_F32;
  bool periodic[3] = {false, false, false};

  BlockTwin() = default;
  BlockTwin(const BlockTwin&) = delete;
  BlockTwin& operator=(const BlockTwin&) = delete;

  ~BlockTwin()
  {
    if (_lattice) {
      olbgfx_lattice_destroy(_lattice);
    }
  }

  void addSource(CollisionSource* s)
  {
    _sources.push_back(s);
    _maskStale = true;
  }

  void removeSource(CollisionSource* s)
  {
    _sources.erase(std::remove(_sources.begin(), _sources.end(), s), _sources.end());
    _maskStale = true;
  }

  void addNormalSource(NormalSource* s)
  {
    _normalSources.push_back(s);
    _maskStale = true;
  }

  void removeNormalSource(NormalSource* s)
  {
    _normalSources.erase(std::remove(_normalSources.begin(), _normalSources.end(), s), _normalSources.end());
    _maskStale = true;
  }

  void cellsChanged()
  {
    _maskStale = true;
  }

  void setPeriodic(int axis)
  {
    if (!periodic[axis]) {
      periodic[axis] = true;
      _layoutStale = true;
    }
  }

  bool deviceOwnsPopulations() const
  {
    return _deviceValid && !_hostValid;
  }

  /// collide() of the block: one fused step per stream() cycle
  void collide()
  {
    if (_stepped) {
      return;
    }
    prepare();
    check(olbgfx_set_statistics(_lattice, statisticsEnabled() ? 1 : 0), "olbgfx_set_statistics");
    check(olbgfx_collide_and_stream(_lattice, 1, nullptr), "olbgfx_collide_and_stream");
    _hostValid = false;
    _stepped = true;
    _processed = true;
    _statisticsFresh = true;
  }

  /// the block's PostStream processors: a step runs them itself after its
  /// stream, so they run here only on populations no step or earlier call
  /// has processed (BlockLattice::initialize, core/blockLattice.hh:56-67);
  /// every processor type of the block calls this, the first does the work
  void postProcess()
  {
    if (_processed && _deviceValid) {
      return;
    }
    prepare();
    check(olbgfx_post_process(_lattice, nullptr), "olbgfx_post_process");
    _hostValid = false;
    _processed = true;
  }

  /// {nCells, sum rho, sum uSqr, max uSqr} of the last step, once per step
  bool takeStatistics(double out[4])
  {
    if (!_lattice || !_statisticsFresh) {
      return false;
    }
    _statisticsFresh = false;
    check(olbgfx_get_statistics(_lattice, out), "olbgfx_get_statistics");
    return true;
  }

  void streamed()
  {
    _stepped = false;
  }

  void contextSwitch(ProcessingContext context)
  {
    if (context == ProcessingContext::Evaluation) {
      hostNeeded();
    } else if (_hostValid) {
      // the mirror may have changed: the device takes it before the next step
      _deviceValid = false;
    }
  }

  void hostNeeded()
  {
    if (_lattice && _deviceValid && !_hostValid) {
      std::vector<char> buf(19 * ncells * valueSize);
      check(olbgfx_get_populations(_lattice, valueType, buf.data()), "olbgfx_get_populations");
      scatterPopulations(buf.data());
      // the device's overlap holds only the populations that cross into the
      // core; the reference's communicate() before evaluation fills all 19
      fillPeriodicOverlap();
      _hostValid = true;
    }
  }

  void hostLoaded()
  {
    _hostValid = true;
    _deviceValid = false;
  }

  olbgfx_lattice* lattice()
  {
    return _lattice;
  }

private:
  olbgfx_lattice* _lattice = nullptr;
  std::vector<CollisionSource*> _sources;
  std::vector<NormalSource*> _normalSources;
  std::vector<olbgfx_dynamics_desc> _table;
  std::vector<std::uint8_t> _ids;
  bool _hostValid = true;     // the mirror holds the latest populations
  bool _deviceValid = false;  // the device holds them
  bool _stepped = false;      // a step ran since the last stream()
  bool _processed = false;    // the PostStream processors ran on the current populations
  bool _statisticsFresh = false;
  bool _maskStale = true;
  bool _layoutStale = true;
  std::uint64_t _fieldEpoch = 0;

  /// field by field: the struct's padding bytes carry no value
  static bool sameDesc(const olbgfx_dynamics_desc& a, const olbgfx_dynamics_desc& b)
  {
    bool same = a.kind == b.kind && a.omega == b.omega && a.magic == b.magic && a.wall_rho == b.wall_rho
             && a.smagorinsky == b.smagorinsky;
    for (int k = 0; k < 3; ++k) {
      same = same && a.force[k] == b.force[k] && a.wall_velocity[k] == b.wall_velocity[k];
    }
    return same;
  }

  void createLattice()
  {
    if (_lattice) {
      check(olbgfx_lattice_destroy(_lattice), "olbgfx_lattice_destroy");
      _lattice = nullptr;
    }
    olbgfx_lattice_desc d{};
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.padding = padding;
    for (int a = 0; a < 3; ++a) {
      d.periodic[a] = periodic[a] ? 1 : 0;
    }
    d.storage = storage;
    check(olbgfx_lattice_create(&d, &_lattice), "olbgfx_lattice_create");
    // a host timer around a run of steps misses at most two of them
    check(olbgfx_set_inflight_limit(_lattice, 2), "olbgfx_set_inflight_limit");
    _table.clear();
    _maskStale = true;
    _layoutStale = false;
    _deviceValid = false;
    _fieldEpoch = 0;
  }

  /// dynamics table and per-cell ids (and normals) from the operators
  void uploadDynamics()
  {
    std::vector<olbgfx_dynamics_desc> table;
    std::vector<std::uint8_t> ids(ncells, 0);
    std::vector<std::int8_t> normals;
    bool anyNormal = !_normalSources.empty();
    for (CollisionSource* s : _sources) {
      anyNormal = anyNormal || s->hasNormal();
    }
    if (anyNormal) {
      normals.assign(3 * ncells, 0);
    }
    olbgfx_dynamics_desc none{};
    none.kind = OLBGFX_NO_DYNAMICS;
    none.wall_rho = 1;
    table.push_back(none);
    for (CollisionSource* s : _sources) {
      const olbgfx_dynamics_desc d = s->describe();
      std::size_t id = 0;
      while (id < table.size() && !sameDesc(table[id], d)) {
        ++id;
      }
      if (id == table.size()) {
        if (table.size() == OLBGFX_MAX_DYNAMICS) {
          throw std::domain_error("GPU_GFX950: more than 16 distinct dynamics in one block");
        }
        table.push_back(d);
      }
      const std::vector<std::uint8_t>& cells = s->cells();
      for (std::size_t n = 0; n < ncells; ++n) {
        if (cells[n]) {
          ids[n] = static_cast<std::uint8_t>(id);
          if (s->hasNormal()) {
            for (int a = 0; a < 3; ++a) {
              normals[3 * n + a] = static_cast<std::int8_t>(s->normal(a));
            }
          }
        }
      }
    }
    for (NormalSource* s : _normalSources) {
      const std::vector<std::uint8_t>& cells = s->cells();
      for (std::size_t n = 0; n < ncells; ++n) {
        if (cells[n]) {
          for (int a = 0; a < 3; ++a) {
            normals[3 * n + a] = static_cast<std::int8_t>(s->normal(a));
          }
        }
      }
    }
    for (std::uint32_t k = 0; k < table.size(); ++k) {
      check(olbgfx_set_dynamics(_lattice, k, &table[k]), "olbgfx_set_dynamics");
    }
    check(olbgfx_set_cell_normals(_lattice, anyNormal ? normals.data() : nullptr), "olbgfx_set_cell_normals");
    check(olbgfx_set_dynamics_mask(_lattice, ids.data()), "olbgfx_set_dynamics_mask");
    _table = table;
    _ids = std::move(ids);
    _maskStale = false;
  }

  /// parameters changed by setParameter between steps
  void refreshParameters()
  {
    std::size_t id = 1;
    std::vector<olbgfx_dynamics_desc> table(1, _table.empty() ? olbgfx_dynamics_desc{} : _table[0]);
    for (CollisionSource* s : _sources) {
      const olbgfx_dynamics_desc d = s->describe();
      std::size_t k = 0;
      while (k < table.size() && !sameDesc(table[k], d)) {
        ++k;
      }
      if (k == table.size()) {
        table.push_back(d);
      }
    }
    if (table.size() != _table.size()) {
      _maskStale = true;
      return;
    }
    for (id = 0; id < table.size(); ++id) {
      if (!sameDesc(table[id], _table[id])) {
        _maskStale = true;
        return;
      }
    }
  }

  void prepare()
  {
    if (!_lattice || _layoutStale) {
      if (_lattice && !_hostValid) {
        hostNeeded();
      }
      createLattice();
    }
    if (!_maskStale) {
      refreshParameters();
    }
    bool dynamicsChanged = false;
    if (_maskStale) {
      uploadDynamics();
      dynamicsChanged = true;
    }
    if (dynamicsChanged || _fieldEpoch != fieldEpoch() || !_deviceValid) {
      uploadFields(_lattice, _ids, _table);
      _fieldEpoch = fieldEpoch();
    }
    if (!_deviceValid) {
      std::vector<char> buf(19 * ncells * valueSize);
      gatherPopulations(buf.data());
      check(olbgfx_set_populations(_lattice, valueType, buf.data()), "olbgfx_set_populations");
      _deviceValid = true;
      _processed = false;
    }
  }
};

inline void twinContextSwitch(BlockTwin& twin, ProcessingContext context)
{
  twin.contextSwitch(context);
}

inline void twinStreamed(BlockTwin& twin)
{
  twin.streamed();
}

inline void twinHostNeeded(BlockTwin& twin)
{
  twin.hostNeeded();
}

inline void twinHostLoaded(BlockTwin& twin)
{
  twin.hostLoaded();
}

/// LocalVelocity external edges and corners: BGK with fixed-velocity momenta;
/// their normal comes with their post-processor
template <typename T, typename DESCRIPTOR>
struct FixedVelocityKind {
  using type = typename BGKdynamics<T,DESCRIPTOR>::template exchange_momenta<momenta::FixedVelocityBoundaryTuple>;
};

/// The dynamics the device kernels provide, and their kind
template <typename T, typename DESCRIPTOR, typename DYNAMICS>
struct KindOf {
  static constexpr int value =
    std::is_same_v<DYNAMICS, NoDynamics<T,DESCRIPTOR>>          ? OLBGFX_NO_DYNAMICS
  : std::is_same_v<DYNAMICS, BGKdynamics<T,DESCRIPTOR>>         ? OLBGFX_BGK
  : std::is_same_v<DYNAMICS, TRTdynamics<T,DESCRIPTOR>>         ? OLBGFX_TRT
  : std::is_same_v<DYNAMICS, ForcedBGKdynamics<T,DESCRIPTOR>>   ? OLBGFX_FORCED_BGK
  : std::is_same_v<DYNAMICS, ForcedTRTdynamics<T,DESCRIPTOR>>   ? OLBGFX_FORCED_TRT
  : std::is_same_v<DYNAMICS, BounceBack<T,DESCRIPTOR>>          ? OLBGFX_BOUNCE_BACK
  : std::is_same_v<DYNAMICS, BounceBackVelocity<T,DESCRIPTOR>>  ? OLBGFX_BOUNCE_BACK_VELOCITY
  : std::is_same_v<DYNAMICS, SmagorinskyBGKdynamics<T,DESCRIPTOR>> ? OLBGFX_SMAGORINSKY_BGK
  : std::is_same_v<DYNAMICS, typename FixedVelocityKind<T,DESCRIPTOR>::type> ? OLBGFX_FIXED_VELOCITY
  : -1;
  static constexpr int normal[3] = {0, 0, 0};
  static constexpr bool hasNormal = false;
};

/// Normal of the edge templates <plane, normal1, normal2>: zero along the
/// edge, normal1 and normal2 on the next two axes in cyclic order
/// (boundaryhelper::promiseForNormalSpecial, boundary/setBoundary3D.h)
constexpr int edgeNormal(int plane, int n1, int n2, int axis)
{
  return axis == plane ? 0 : (axis == (plane + 1) % 3 ? n1 : n2);
}

/// LocalVelocity faces: CombinedRLBdynamics around BGK with the regularised
/// velocity boundary momenta of one face (boundary/localVelocity3D.h)
template <typename T, typename DESCRIPTOR, int DIRECTION, int ORIENTATION>
struct KindOf<T, DESCRIPTOR, CombinedRLBdynamics<T, DESCRIPTOR, BGKdynamics<T,DESCRIPTOR>,
                                                 momenta::RegularizedVelocityBoundaryTuple<DIRECTION,ORIENTATION>>> {
  static constexpr int value = OLBGFX_LOCAL_VELOCITY;
  static constexpr int normal[3] = {DIRECTION == 0 ? ORIENTATION : 0,
                                    DIRECTION == 1 ? ORIENTATION : 0,
                                    DIRECTION == 2 ? ORIENTATION : 0};
  static constexpr bool hasNormal = true;
};

/// LocalPressure faces (boundary/localPressure3D.h)
template <typename T, typename DESCRIPTOR, int DIRECTION, int ORIENTATION>
struct KindOf<T, DESCRIPTOR, CombinedRLBdynamics<T, DESCRIPTOR, BGKdynamics<T,DESCRIPTOR>,
                                                 momenta::RegularizedPressureBoundaryTuple<DIRECTION,ORIENTATION>>> {
  static constexpr int value = OLBGFX_LOCAL_PRESSURE;
  static constexpr int normal[3] = {DIRECTION == 0 ? ORIENTATION : 0,
                                    DIRECTION == 1 ? ORIENTATION : 0,
                                    DIRECTION == 2 ? ORIENTATION : 0};
  static constexpr bool hasNormal = true;
};

/// LocalVelocity internal edges (boundary/localVelocity3D.h, InternalEdge)
template <typename T, typename DESCRIPTOR, int PLANE, int NORMAL1, int NORMAL2>
struct KindOf<T, DESCRIPTOR, CombinedRLBdynamics<T, DESCRIPTOR, BGKdynamics<T,DESCRIPTOR>,
                                                 momenta::InnerEdgeVelocityTuple3D<PLANE,NORMAL1,NORMAL2>>> {
  static constexpr int value = OLBGFX_INNER_EDGE_VELOCITY;
  static constexpr int normal[3] = {edgeNormal(PLANE, NORMAL1, NORMAL2, 0),
                                    edgeNormal(PLANE, NORMAL1, NORMAL2, 1),
                                    edgeNormal(PLANE, NORMAL1, NORMAL2, 2)};
  static constexpr bool hasNormal = true;
};

/// LocalVelocity internal corners (InternalCorner)
template <typename T, typename DESCRIPTOR, int X, int Y, int Z>
struct KindOf<T, DESCRIPTOR, CombinedRLBdynamics<T, DESCRIPTOR, BGKdynamics<T,DESCRIPTOR>,
                                                 momenta::InnerCornerVelocityTuple3D<X,Y,Z>>> {
  static constexpr int value = OLBGFX_INNER_CORNER_VELOCITY;
  static constexpr int normal[3] = {X, Y, Z};
  static constexpr bool hasNormal = true;
};

/// The dynamics a (CSE-wrapped) collision operator stands for
template <typename DYNAMICS>
struct Unwrapped {
  using type = DYNAMICS;
};

template <typename DYNAMICS>
struct Unwrapped<CSE<DYNAMICS>> {
  using type = DYNAMICS;
};

/// Post-processors whose work the fused step does for the kinds above
template <typename OPERATOR>
struct ProcessorTraits {
  static constexpr bool covered = false;
  static constexpr bool hasNormal = false;
  static constexpr int normal[3] = {0, 0, 0};
};

template <>
struct ProcessorTraits<StatisticsPostProcessor> {
  static constexpr bool covered = true;
  static constexpr bool hasNormal = false;
  static constexpr int normal[3] = {0, 0, 0};
};

/// LocalVelocity external edges (boundary/boundaryPostProcessors3D.h:88)
template <typename T, typename DESCRIPTOR, int PLANE, int NORMAL1, int NORMAL2>
struct ProcessorTraits<OuterVelocityEdgeProcessor3D<T,DESCRIPTOR,PLANE,NORMAL1,NORMAL2>> {
  static constexpr bool covered = true;
  static constexpr bool hasNormal = true;
  static constexpr int normal[3] = {edgeNormal(PLANE, NORMAL1, NORMAL2, 0),
                                    edgeNormal(PLANE, NORMAL1, NORMAL2, 1),
                                    edgeNormal(PLANE, NORMAL1, NORMAL2, 2)};
};

/// LocalVelocity external corners (boundary/boundaryPostProcessors3D.h:111)
template <typename T, typename DESCRIPTOR, int X, int Y, int Z>
struct ProcessorTraits<OuterVelocityCornerProcessor3D<T,DESCRIPTOR,X,Y,Z>> {
  static constexpr bool covered = true;
  static constexpr bool hasNormal = true;
  static constexpr int normal[3] = {X, Y, Z};
};

/// the generated (CSE) form of a processor does the same work
template <typename OPERATOR, typename DESCRIPTOR>
struct ProcessorTraits<CSE_O<OPERATOR,DESCRIPTOR>> : ProcessorTraits<OPERATOR> { };

}

/// Mask of the cells a dynamics collides (core cells only), host array
template <typename T>
class ConcreteBlockMask<T,Platform::GPU_GFX950> final : public Serializable {
private:
  cpu::sisd::Column<bool> _mask;
  std::size_t _weight;

public:
  ConcreteBlockMask(std::size_t size):
    _mask(size),
    _weight(0)
  { }

  AbstractColumn<bool>& asAbstract()
  {
    return _mask;
  }

  bool operator[](std::size_t i) const
  {
    return _mask[i];
  }

  void set(std::size_t i, bool active)
  {
    if (_mask[i] && !active) {
      _weight -= 1;
    } else if (!_mask[i] && active) {
      _weight += 1;
    }
    _mask[i] = active;
  }

  std::size_t weight() const
  {
    return _weight;
  }

  void setProcessingContext(ProcessingContext) { }

  std::size_t getNblock() const override
  {
    return 1 + _mask.getNblock();
  }

  std::size_t getSerializableSize() const override
  {
    return sizeof(_weight) + _mask.getSerializableSize();
  }

  bool* getBlock(std::size_t iBlock, std::size_t& sizeBlock, bool loadingMode) override
  {
    std::size_t currentBlock = 0;
    bool* dataPtr = nullptr;
    registerVar(iBlock, sizeBlock, currentBlock, dataPtr, _weight);
    registerSerializableOfConstSize(iBlock, sizeBlock, currentBlock, dataPtr, _mask, loadingMode);
    return dataPtr;
  }

};

namespace gfx950 {

/// The twin of a block, created on first use and owned by the block's
/// population column 0
template <typename T, typename DESCRIPTOR>
BlockTwin& twinOf(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>& block)
{
  static_assert(DESCRIPTOR::d == 3 && DESCRIPTOR::q == 19,
                "GPU_GFX950 provides D3Q19 lattices");
  static_assert(std::is_same_v<T,float> || std::is_same_v<T,double>,
                "GPU_GFX950 blocks use float or double");
  auto& pops = block.template getField<descriptors::POPULATION>();
  if (BlockTwin* twin = pops[0].getTwin()) {
    return *twin;
  }
  auto twin = std::make_shared<BlockTwin>();
  twin->nx = block.getNx();
  twin->ny = block.getNy();
  twin->nz = block.getNz();
  twin->padding = block.getPadding();
  twin->ncells = block.getNcells();
  const bool isDouble = std::is_same_v<T,double>;
  twin->valueType = isDouble ? OLBGFX_VALUE_F64 : OLBGFX_VALUE_F32;
  twin->valueSize = sizeof(T);
  twin->storage = storageFor(isDouble);
  auto* b = &block;
  const std::size_t n = twin->ncells;
  twin->gatherPopulations = [b, n](void* out) {
    auto& p = b->template getField<descriptors::POPULATION>();
    T* o = static_cast<T*>(out);
    for (unsigned i = 0; i < 19; ++i) {
      for (std::size_t c = 0; c < n; ++c) {
        o[i * n + c] = p[i][c];
      }
    }
  };
  twin->scatterPopulations = [b, n](const void* in) {
    auto& p = b->template getField<descriptors::POPULATION>();
    const T* s = static_cast<const T*>(in);
    for (unsigned i = 0; i < 19; ++i) {
      for (std::size_t c = 0; c < n; ++c) {
        p[i][c] = s[i * n + c];
      }
    }
  };
  twin->uploadFields = [b, n](olbgfx_lattice* L, const std::vector<std::uint8_t>& ids,
                              const std::vector<olbgfx_dynamics_desc>& table) {
    const int vt = std::is_same_v<T,double> ? OLBGFX_VALUE_F64 : OLBGFX_VALUE_F32;
    if constexpr (DESCRIPTOR::template provides<descriptors::FORCE>()) {
      auto& f = b->template getField<descriptors::FORCE>();
      std::vector<T> v(3 * n);
      for (std::size_t c = 0; c < n; ++c) {
        for (unsigned a = 0; a < 3; ++a) {
          v[3 * c + a] = f[a][c];
        }
      }
      check(olbgfx_set_cell_field(L, OLBGFX_FIELD_FORCE, vt, v.data()), "olbgfx_set_cell_field(FORCE)");
    }
    // one VELOCITY array on the device: BounceBackVelocity cells read
    // descriptors::VELOCITY (momenta::FixedVelocityMomentum), the
    // LocalVelocity kinds FixedVelocityMomentumGeneric::VELOCITY
    using GENERIC = momenta::FixedVelocityMomentumGeneric::VELOCITY;
    const bool external = b->template hasData<Array<descriptors::VELOCITY>>();
    const bool generic = b->template hasData<Array<GENERIC>>();
    if (external || generic) {
      auto* ext = external ? &b->template getField<descriptors::VELOCITY>() : nullptr;
      auto* gen = generic ? &b->template getField<GENERIC>() : nullptr;
      std::vector<T> v(3 * n, T(0));
      for (std::size_t c = 0; c < n; ++c) {
        const int kind = ids.empty() ? OLBGFX_NO_DYNAMICS : table[ids[c]].kind;
        if (kind == OLBGFX_BOUNCE_BACK_VELOCITY) {
          if (ext) {
            for (unsigned a = 0; a < 3; ++a) {
              v[3 * c + a] = (*ext)[a][c];
            }
          }
        } else if (gen) {
          for (unsigned a = 0; a < 3; ++a) {
            v[3 * c + a] = (*gen)[a][c];
          }
        }
      }
      check(olbgfx_set_cell_field(L, OLBGFX_FIELD_VELOCITY, vt, v.data()), "olbgfx_set_cell_field(VELOCITY)");
    }
    using RHO = momenta::FixedDensity::RHO;
    if (b->template hasData<Array<RHO>>()) {
      auto& r = b->template getField<RHO>();
      std::vector<T> v(n);
      for (std::size_t c = 0; c < n; ++c) {
        v[c] = r[0][c];
      }
      check(olbgfx_set_cell_field(L, OLBGFX_FIELD_RHO, vt, v.data()), "olbgfx_set_cell_field(RHO)");
    }
  };
  twin->statisticsEnabled = [b]() {
    return b->statisticsEnabled();
  };
  BlockTwin* tw = twin.get();
  twin->fillPeriodicOverlap = [b, tw]() {
    const int ext[3] = {tw->nx, tw->ny, tw->nz};
    const int p = tw->padding;
    if (!tw->periodic[0] && !tw->periodic[1] && !tw->periodic[2]) {
      return;
    }
    auto& pops = b->template getField<descriptors::POPULATION>();
    for (int x = -p; x < ext[0] + p; ++x) {
      for (int y = -p; y < ext[1] + p; ++y) {
        for (int z = -p; z < ext[2] + p; ++z) {
          int c[3] = {x, y, z};
          bool outside = false, ok = true;
          for (int a = 0; a < 3; ++a) {
            if (c[a] < 0 || c[a] >= ext[a]) {
              outside = true;
              if (!tw->periodic[a]) {
                ok = false;
              } else {
                c[a] = c[a] < 0 ? c[a] + ext[a] : c[a] - ext[a];
              }
            }
          }
          if (!outside || !ok) {
            continue;
          }
          const CellID to = b->getCellId(x, y, z);
          const CellID from = b->getCellId(c[0], c[1], c[2]);
          for (unsigned i = 0; i < 19; ++i) {
            pops[i][to] = pops[i][from];
          }
        }
      }
    }
  };
  BlockTwin& ref = *twin;
  pops[0].attachTwin(std::move(twin));
  return ref;
}

}

/// Collision of one DYNAMICS on a GPU_GFX950 block: an entry of the device
/// twin's dynamics table; apply() runs the fused step of the whole block
template <typename T, typename DESCRIPTOR, typename DYNAMICS>
class ConcreteBlockCollisionO<T,DESCRIPTOR,Platform::GPU_GFX950,DYNAMICS> final
  : public BlockCollisionO<T,DESCRIPTOR,Platform::GPU_GFX950>
  , public gfx950::CollisionSource {
private:
  using PLAIN = typename gfx950::Unwrapped<DYNAMICS>::type;
  using KIND = gfx950::KindOf<T,DESCRIPTOR,PLAIN>;

  std::unique_ptr<DYNAMICS> _dynamics;
  ConcreteParametersD<T,DESCRIPTOR,Platform::GPU_GFX950,typename DYNAMICS::parameters>* _parameters;
  std::vector<std::uint8_t> _cells;
  std::size_t _weight;
  gfx950::BlockTwin* _twin;

public:
  ConcreteBlockCollisionO():
    _dynamics(new DYNAMICS()),
    _parameters(nullptr),
    _weight(0),
    _twin(nullptr)
  { }

  ~ConcreteBlockCollisionO()
  {
    if (_twin) {
      _twin->removeSource(this);
    }
  }

  std::type_index id() const override
  {
    return typeid(DYNAMICS);
  }

  std::size_t weight() const override
  {
    return _weight;
  }

  void set(CellID iCell, bool state, bool overlap) override
  {
    if (iCell >= _cells.size()) {
      throw std::out_of_range("GPU_GFX950: cell index beyond the block");
    }
    const bool was = _cells[iCell] != 0;
    _cells[iCell] = state ? 1 : 0;
    if (!overlap && was != state) {
      _weight = state ? _weight + 1 : _weight - 1;
    }
    if (_twin) {
      _twin->cellsChanged();
    }
  }

  Dynamics<T,DESCRIPTOR>* getDynamics() override
  {
    return _dynamics.get();
  }

  void setup(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>& block) override
  {
    if constexpr (KIND::value < 0) {
      throw std::domain_error("GPU_GFX950 provides no kernel for dynamics " + _dynamics->getName());
    } else {
      _parameters = &block.template getData<OperatorParameters<DYNAMICS>>();
      _cells.assign(block.getNcells(), 0);
      _twin = &gfx950::twinOf(block);
      _twin->addSource(this);
    }
  }

  void apply(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>& block,
             ConcreteBlockMask<T,Platform::GPU_GFX950>&,
             CollisionDispatchStrategy) override
  {
    gfx950::twinOf(block).collide();
  }

  olbgfx_dynamics_desc describe() const override
  {
    olbgfx_dynamics_desc d{};
    d.kind = KIND::value;
    d.wall_rho = 1;
    using PARAMETERS = typename DYNAMICS::parameters;
    if constexpr (PARAMETERS::template contains<descriptors::OMEGA>()) {
      d.omega = _parameters->parameters.template get<descriptors::OMEGA>();
    }
    if constexpr (PARAMETERS::template contains<collision::TRT::MAGIC>()) {
      d.magic = _parameters->parameters.template get<collision::TRT::MAGIC>();
    }
    if constexpr (PARAMETERS::template contains<collision::LES::SMAGORINSKY>()) {
      d.smagorinsky = _parameters->parameters.template get<collision::LES::SMAGORINSKY>();
    }
    return d;
  }

  const std::vector<std::uint8_t>& cells() const override
  {
    return _cells;
  }

  int normal(int axis) const override
  {
    return KIND::normal[axis];
  }

  bool hasNormal() const override
  {
    return KIND::hasNormal;
  }

};

/// Post-processors of GPU_GFX950 blocks, per cell: the fused step applies
/// those its kinds imply; any other is refused at set-up
template <typename T, typename DESCRIPTOR, typename OPERATOR>
class ConcreteBlockO<T,DESCRIPTOR,Platform::GPU_GFX950,OPERATOR,OperatorScope::PerCell> final
  : public BlockO<T,DESCRIPTOR,Platform::GPU_GFX950>
  , public gfx950::NormalSource {
private:
  using TRAITS = gfx950::ProcessorTraits<OPERATOR>;

  std::vector<std::uint8_t> _cells;
  std::size_t _weight = 0;
  gfx950::BlockTwin* _twin = nullptr;

public:
  ~ConcreteBlockO()
  {
    if (_twin) {
      _twin->removeNormalSource(this);
    }
  }

  std::type_index id() const override
  {
    return typeid(OPERATOR);
  }

  std::size_t weight() const override
  {
    return _weight;
  }

  void set(CellID iCell, bool state) override
  {
    if (iCell >= _cells.size()) {
      throw std::out_of_range("GPU_GFX950: cell index beyond the block");
    }
    const bool was = _cells[iCell] != 0;
    _cells[iCell] = state ? 1 : 0;
    if (was != state) {
      _weight = state ? _weight + 1 : _weight - 1;
    }
    if (_twin) {
      _twin->cellsChanged();
    }
  }

  void setup(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>& block) override
  {
    if constexpr (!TRAITS::covered) {
      throw std::domain_error(std::string("GPU_GFX950 provides no kernel for post-processor ")
                              + typeid(OPERATOR).name());
    } else {
      _cells.assign(block.getNcells(), 0);
      if constexpr (TRAITS::hasNormal) {
        _twin = &gfx950::twinOf(block);
        _twin->addNormalSource(this);
      }
    }
  }

  const std::vector<std::uint8_t>& cells() const override
  {
    return _cells;
  }

  int normal(int axis) const override
  {
    return TRAITS::normal[axis];
  }

  void apply(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>& block) override
  {
    gfx950::twinOf(block).postProcess();
  }

};

template <typename T, typename DESCRIPTOR, typename OPERATOR>
class ConcreteBlockO<T,DESCRIPTOR,Platform::GPU_GFX950,OPERATOR,OperatorScope::PerCellWithParameters> final
  : public BlockO<T,DESCRIPTOR,Platform::GPU_GFX950> {
private:
  std::size_t _weight = 0;

public:
  std::type_index id() const override
  {
    return typeid(OPERATOR);
  }

  std::size_t weight() const override
  {
    return _weight;
  }

  void set(CellID, bool state) override
  {
    _weight += state ? 1 : 0;
  }

  void setup(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>&) override
  {
    if constexpr (!gfx950::ProcessorTraits<OPERATOR>::covered) {
      throw std::domain_error(std::string("GPU_GFX950 provides no kernel for post-processor ")
                              + typeid(OPERATOR).name());
    }
  }

  void apply(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>& block) override
  {
    gfx950::twinOf(block).postProcess();
  }

};

/// Block-wide operators (the statistics) run on the host with the device's results
template <typename T, typename DESCRIPTOR, typename OPERATOR>
class ConcreteBlockO<T,DESCRIPTOR,Platform::GPU_GFX950,OPERATOR,OperatorScope::PerBlock> final
  : public BlockO<T,DESCRIPTOR,Platform::GPU_GFX950> {
public:
  std::type_index id() const override
  {
    return typeid(OPERATOR);
  }

  std::size_t weight() const override
  {
    return 0;
  }

  void set(CellID, bool) override
  {
    throw std::logic_error("BlockO::set not supported for OperatorScope::PerBlock");
  }

  void setup(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>& block) override
  {
    if constexpr (!gfx950::ProcessorTraits<OPERATOR>::covered) {
      throw std::domain_error(std::string("GPU_GFX950 provides no kernel for block operator ")
                              + typeid(OPERATOR).name());
    }
    OPERATOR().setup(block);
  }

  void apply(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>& block) override
  {
    OPERATOR().apply(block);
  }

};

/// Statistics of a GPU_GFX950 block: the FP64 sums of the last step
template <typename T, typename DESCRIPTOR>
struct StatisticsPostProcessor::type<ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>> {

  void setup(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>&) { }

  void apply(ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>& blockLattice)
  {
    if (!blockLattice.statisticsEnabled()) {
      return;
    }
    double s[4];
    if (gfx950::twinOf(blockLattice).takeStatistics(s)) {
      typename LatticeStatistics<T>::Aggregatable statistics{
        .nCells = static_cast<std::size_t>(s[0]),
        .avRho = static_cast<T>(s[1]),
        .avEnergy = static_cast<T>(s[2]),
        .maxU = static_cast<T>(s[3])
      };
      blockLattice.getStatistics().incrementStats(statistics);
    }
    blockLattice.getStatistics().reset();
  }

};

/// Exchanges of a GPU_GFX950 block.  One block per lattice: its only
/// neighbour is itself across periodic faces, which tells the twin which
/// axes are periodic.  While the device owns the populations the fused step
/// has filled what an exchange would (the crossing populations of the
/// overlap, and the processors' neighbourhoods); in the Evaluation context
/// the exchange copies host data as the reference's CPU copy task does
/// (core/blockLattice.hh:763-799)
template <typename T, typename DESCRIPTOR>
class ConcreteBlockCommunicator<ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>>
  final : public BlockCommunicator {
private:
  using BLOCK = ConcreteBlockLattice<T,DESCRIPTOR,Platform::GPU_GFX950>;

  BLOCK& _block;
  std::vector<CellID> _targets;
  std::vector<CellID> _sources;
  std::unique_ptr<MultiConcreteCommunicatable<BLOCK>> _target;
  std::unique_ptr<MultiConcreteCommunicatable<BLOCK>> _source;
  std::vector<std::uint8_t> _buffer;

public:
  ConcreteBlockCommunicator(SuperLattice<T,DESCRIPTOR>& super,
                            LoadBalancer<T>& loadBalancer,
#ifdef PARALLEL_MODE_MPI
                            SuperCommunicationTagCoordinator<T>&,
                            MPI_Comm,
#endif
                            int iC,
                            const BlockCommunicationNeighborhood<T,DESCRIPTOR::d>& neighborhood):
    _block(super.template getBlock<BLOCK>(iC))
  {
    neighborhood.forNeighbors([&](int remoteC) {
      if (!loadBalancer.isLocal(remoteC) || loadBalancer.loc(remoteC) != iC) {
        throw std::domain_error("GPU_GFX950 runs one block per lattice (a single cuboid)");
      }
      if (!neighborhood.getCellsInboundFrom(remoteC).empty()) {
        _targets = neighborhood.getCellsInboundFrom(remoteC);
        _sources = neighborhood.getCellsRequestedFrom(remoteC);
        _target = std::make_unique<MultiConcreteCommunicatable<BLOCK>>(
          _block, neighborhood.getFieldsCommonWith(remoteC));
        _source = std::make_unique<MultiConcreteCommunicatable<BLOCK>>(
          _block, neighborhood.getFieldsCommonWith(remoteC));
        _buffer.resize(_source->size(_sources));
      }
    });
    // overlap cells filled from the block itself lie across periodic faces
    gfx950::BlockTwin& twin = gfx950::twinOf(_block);
    for (CellID c : _targets) {
      const auto loc = _block.getLatticeR(c);
      if (loc[0] < 0 || loc[0] >= _block.getNx()) twin.setPeriodic(0);
      if (loc[1] < 0 || loc[1] >= _block.getNy()) twin.setPeriodic(1);
      if (loc[2] < 0 || loc[2] >= _block.getNz()) twin.setPeriodic(2);
    }
  }

#ifdef PARALLEL_MODE_MPI
  void receive() override { }
  void send() override { }
  void unpack() override
  {
    copy();
  }
  void wait() override { }
#endif

  void copy()
#ifndef PARALLEL_MODE_MPI
  override
#endif
  {
    if (!_target || gfx950::twinOf(_block).deviceOwnsPopulations()) {
      return;
    }
    _source->serialize(_sources, _buffer.data());
    _target->deserialize(_targets, _buffer.data());
  }

};

/// GPU_GFX950 needs an MI355X (or another HIP device) visible
template <>
inline void checkPlatform<Platform::GPU_GFX950>()
{
  std::int32_t count = 0;
  gfx950::check(olbgfx_device_count(&count), "olbgfx_device_count");
  OstreamManager clout(std::cout, "GPU_GFX950");
  clout << "Found " << count << " HIP device(s); the lattice steps on device 0" << std::endl;
  if (count == 0) {
    throw std::runtime_error("GPU_GFX950 requested but no HIP device is visible");
  }
}

}

#endif
