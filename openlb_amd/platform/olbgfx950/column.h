/*
 * olbgfx950/column.h -- columns of the GPU_GFX950 platform of OpenLB.
 *
 * Included by OpenLB's core/platform/column.h when PLATFORM_GPU_GFX950 is
 * defined.  Plain field columns are host arrays: the device keeps its own
 * copy of the fields a step reads, uploaded again after a context switch to
 * Simulation.  Population columns are the host mirror of
 * the device rings of libolbgfx: the periodic-shift ring of
 * core/platform/cpu/sisd/column.h:129-204 (element i at (i + shift) mod n,
 * rotate(d) is shift -= d), which OpenLB's setters, functors and serializer
 * read and write in the Evaluation context.  Column 0 of a block's
 * POPULATION array owns that block's device twin (olbgfx950/platform.h),
 * which decides when the mirror and the device exchange data.
 */
#ifndef OLBGFX950_COLUMN_H
#define OLBGFX950_COLUMN_H

#include <array>
#include <cstddef>
#include <memory>
#include <stdexcept>

#include "core/platform/platform.h"
#include "core/serializer.h"
#include "communication/communicatable.h"

namespace olb {

namespace gfx950 {

/// Device copy of one block lattice (olbgfx950/platform.h)
class BlockTwin;

/// Hooks of the twin the population columns call; defined with BlockTwin
inline void twinContextSwitch(BlockTwin& twin, ProcessingContext context);
inline void twinStreamed(BlockTwin& twin);
inline void twinHostNeeded(BlockTwin& twin);
inline void twinHostLoaded(BlockTwin& twin);

/// Epoch of host changes to plain field columns (any block): bumped by a
/// field column's switch to Simulation, read by the twins before a step
inline std::uint64_t& fieldEpoch()
{
  static std::uint64_t epoch = 1;
  return epoch;
}

/// Plain field column of GPU_GFX950 blocks: a host array whose switch to
/// Simulation tells the twins to upload fields again
template <typename T>
class Column final : public AbstractColumn<T>
                   , public Serializable {
private:
  std::size_t _count;
  std::unique_ptr<T[]> _data;

public:
  using value_t = T;

  Column(std::size_t count):
    _count(count),
    _data(new T[count] {})
  { }

  Column():
    Column(0)
  { }

  Column(Column<T>&& rhs):
    _count(rhs._count),
    _data(rhs._data.release())
  { }

  Column(const Column<T>& rhs):
    _count(rhs._count),
    _data(new T[rhs._count] {})
  {
    std::copy(rhs._data.get(), rhs._data.get() + _count, _data.get());
  }

  void resize(std::size_t count)
  {
    std::unique_ptr<T[]> data(new T[count] {});
    std::copy(_data.get(), _data.get() + std::min(_count, count), data.get());
    _data.swap(data);
    _count = count;
  }

  const T& operator[](std::size_t i) const override
  {
    return _data[i];
  }

  T& operator[](std::size_t i) override
  {
    return _data[i];
  }

  std::size_t size() const override
  {
    return _count;
  }

  const T* data() const
  {
    return _data.get();
  }

  T* data()
  {
    return _data.get();
  }

  void setProcessingContext(ProcessingContext context)
  {
    if (context == ProcessingContext::Simulation) {
      ++fieldEpoch();
    }
  }

  std::size_t getNblock() const override
  {
    return 2;
  }

  std::size_t getSerializableSize() const override
  {
    return _count * sizeof(T) + sizeof(std::size_t);
  }

  bool* getBlock(std::size_t iBlock, std::size_t& sizeBlock, bool loadingMode) override
  {
    std::size_t currentBlock = 0;
    bool* dataPtr = nullptr;
    registerVar(iBlock, sizeBlock, currentBlock, dataPtr, _count);
    if (loadingMode && iBlock == 1) {
      resize(_count);
    }
    registerVar(iBlock, sizeBlock, currentBlock, dataPtr, *data(), _count);
    return dataPtr;
  }

};

/// Host mirror of one population ring
template <typename T>
class CyclicColumn final : public AbstractCyclicColumn<T>
                         , public Serializable {
private:
  std::size_t _count;
  std::unique_ptr<T[]> _data;
  /// logical element i lives at (i + _shift) mod _count, 0 <= _shift < _count
  std::ptrdiff_t _shift;
  /// set on column 0 of a POPULATION array once the block has a device twin
  std::shared_ptr<BlockTwin> _twin;

public:
  using value_t = T;

  CyclicColumn(std::size_t count):
    _count(count),
    _data(new T[count] {}),
    _shift(0)
  { }

  CyclicColumn(CyclicColumn<T>&& rhs):
    _count(rhs._count),
    _data(rhs._data.release()),
    _shift(rhs._shift),
    _twin(std::move(rhs._twin))
  { }

  ~CyclicColumn() = default;

  const T& operator[](std::size_t i) const override
  {
    std::size_t j = i + static_cast<std::size_t>(_shift);
    return _data[j >= _count ? j - _count : j];
  }

  T& operator[](std::size_t i) override
  {
    std::size_t j = i + static_cast<std::size_t>(_shift);
    return _data[j >= _count ? j - _count : j];
  }

  std::size_t size() const override
  {
    return _count;
  }

  T* data()
  {
    return _data.get();
  }

  const T* data() const
  {
    return _data.get();
  }

  std::ptrdiff_t getShift() const
  {
    return _shift;
  }

  void refresh() { }

  /// stream(): shift -= offset (mod count), as the reference's PS columns
  void rotate(std::ptrdiff_t offset)
  {
    const std::ptrdiff_t n = static_cast<std::ptrdiff_t>(_count);
    std::ptrdiff_t s = (_shift - offset) % n;
    _shift = s < 0 ? s + n : s;
    if (_twin) {
      twinStreamed(*_twin);
    }
  }

  void resize(std::size_t)
  {
    throw std::logic_error("Cyclic column can not be resized");
  }

  /// Column 0 of a block's POPULATION array takes ownership of its twin
  void attachTwin(std::shared_ptr<BlockTwin> twin)
  {
    _twin = std::move(twin);
  }

  BlockTwin* getTwin() const
  {
    return _twin.get();
  }

  void setProcessingContext(ProcessingContext context)
  {
    if (_twin) {
      twinContextSwitch(*_twin, context);
    }
  }

  std::size_t getNblock() const override
  {
    return 3;
  }

  std::size_t getSerializableSize() const override
  {
    return _count * sizeof(T) + sizeof(std::ptrdiff_t) + sizeof(std::size_t);
  }

  /// the CPU_SISD block layout (shift, count, values), so checkpoints move
  /// between the platforms; saving first brings the mirror up to date
  bool* getBlock(std::size_t iBlock, std::size_t& sizeBlock, bool loadingMode) override
  {
    if (_twin && !loadingMode && iBlock == 0) {
      twinHostNeeded(*_twin);
    }
    std::size_t currentBlock = 0;
    bool* dataPtr = nullptr;
    registerVar(iBlock, sizeBlock, currentBlock, dataPtr, _shift);
    registerVar(iBlock, sizeBlock, currentBlock, dataPtr, _count);
    registerVar(iBlock, sizeBlock, currentBlock, dataPtr, *_data.get(), _count);
    return dataPtr;
  }

  void postLoad() override
  {
    const std::ptrdiff_t n = static_cast<std::ptrdiff_t>(_count);
    _shift %= n;
    if (_shift < 0) {
      _shift += n;
    }
    if (_twin) {
      twinHostLoaded(*_twin);
    }
  }

};

}

/// Plain columns of GPU_GFX950 blocks are host arrays
template <typename T>
struct ImplementationOf<AbstractColumn<T>,Platform::GPU_GFX950> {
  using type = gfx950::Column<T>;
};

/// Defined in olbgfx950/platform.h; declared here because initialize()
/// (core/initialize.h) calls it before that header is read
template <>
inline void checkPlatform<Platform::GPU_GFX950>();

/// Population columns of GPU_GFX950 blocks mirror the device rings
template <typename T>
struct ImplementationOf<AbstractCyclicColumn<T>,Platform::GPU_GFX950> {
  using type = gfx950::CyclicColumn<T>;
};

}

#endif
