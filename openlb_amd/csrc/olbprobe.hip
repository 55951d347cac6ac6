// olbprobe.hip -- bandwidth probes of the step kernel's access patterns
// (declared in include/olbgfx_probe.h): a plain 16-byte copy, the 19 PS
// rings read, written, and read-modified-written in place, and read from one
// buffer and written to another, with the step kernel's cell indexing, block
// shape and block order.  They give the
// ceilings against which the step kernel's HBM fraction is read: the
// in-place pattern moves the same bytes as a step, without its arithmetic.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/olbgfx_probe.h"

namespace {

constexpr int Q = 19;
constexpr uint32_t kThreads = 256;

constexpr int kCx[Q] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
constexpr int kCy[Q] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
constexpr int kCz[Q] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};

struct RingArgs {
  void* pop;
  void* pop2;           // second buffer (two-buffer pattern)
  uint64_t stride;      // elements between rings
  uint32_t count;       // ring length = padded cells
  uint32_t shift[Q];
  uint32_t nyz, nz, pad, py, pz;
  uint32_t ncore;
  uint32_t nblocks;
  uint32_t blocks_per_xcd;
  int32_t run_log2;
};

__device__ __forceinline__ uint32_t probe_logical_block(uint32_t b, uint32_t blocks_per_xcd, int32_t run_log2)
{
  if (run_log2 < 0) {
    return (b & 7u) * blocks_per_xcd + (b >> 3);
  }
  const uint32_t k = (uint32_t)run_log2, g = b >> 3;
  return ((g >> k) << (k + 3)) + ((b & 7u) << k) + (g & ((1u << k) - 1u));
}

template <typename E>
__device__ __forceinline__ float to_f(E v) { return (float)v; }
template <typename E>
__device__ __forceinline__ E from_f(float v) { return (E)v; }

// PATTERN: 1 read, 2 write, 3 in place, 4 read pop and write pop2 at slot n
template <typename E, int PATTERN>
__global__ void __launch_bounds__(kThreads) probe_rings(const RingArgs a, float* sink)
{
  const uint32_t lb = probe_logical_block(blockIdx.x, a.blocks_per_xcd, a.run_log2);
  const uint32_t lin = lb * kThreads + threadIdx.x;
  if (lb >= a.nblocks || lin >= a.ncore) {
    return;
  }
  const uint32_t x = lin / a.nyz;
  const uint32_t r = lin - x * a.nyz;
  const uint32_t y = r / a.nz;
  const uint32_t z = r - y * a.nz;
  const uint32_t n = ((x + a.pad) * a.py + (y + a.pad)) * a.pz + (z + a.pad);
  char* base = (char*)a.pop;
  float acc = 0.f;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    uint32_t s = n + a.shift[i];
    s = s >= a.count ? s - a.count : s;
    E* p = (E*)(base + ((uint64_t)i * a.stride) * sizeof(E)) + s;
    if constexpr (PATTERN == 1) {
      acc += to_f(*p);
    } else if constexpr (PATTERN == 2) {
      *p = from_f<E>((float)(lin & 7u) * 1e-3f);
    } else if constexpr (PATTERN == 4) {
      E* q = (E*)((char*)a.pop2 + ((uint64_t)i * a.stride) * sizeof(E)) + n;
      *q = from_f<E>(to_f(*p) * 0.999f);
    } else {
      *p = from_f<E>(to_f(*p) * 0.999f);
    }
  }
  if constexpr (PATTERN == 1) {
    // never true for the values the probe holds; keeps the loads alive
    if (acc == -1.2345e30f) {
      sink[0] = acc;
    }
  }
}

// 16 bytes as a native vector (what the nontemporal builtins take)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ u32x4 copy_load(const u32x4* p)
{
  if constexpr (NT) {
    return __builtin_nontemporal_load(p);
  } else {
    return *p;
  }
}

template <bool NT>
__device__ __forceinline__ void copy_store(u32x4* p, u32x4 v)
{
  if constexpr (NT) {
    __builtin_nontemporal_store(v, p);
  } else {
    *p = v;
  }
}

// grid-stride 16-byte copy, four loads in flight per lane; NT: nontemporal
// loads and stores
template <bool NT>
__global__ void __launch_bounds__(kThreads) probe_copy(const u32x4* src, u32x4* dst, uint64_t n)
{
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  for (; k + 3 * stride < n; k += 4 * stride) {
    const u32x4 a = copy_load<NT>(src + k), b = copy_load<NT>(src + k + stride);
    const u32x4 c = copy_load<NT>(src + k + 2 * stride), d = copy_load<NT>(src + k + 3 * stride);
    copy_store<NT>(dst + k, a);
    copy_store<NT>(dst + k + stride, b);
    copy_store<NT>(dst + k + 2 * stride, c);
    copy_store<NT>(dst + k + 3 * stride, d);
  }
  for (; k < n; k += stride) {
    copy_store<NT>(dst + k, copy_load<NT>(src + k));
  }
}

// one 16-byte element per lane, one workgroup per 256 elements (the step
// kernel's launch shape); `first` offsets the grid for copies above 2^32 lanes
template <bool NT>
__global__ void __launch_bounds__(kThreads) probe_copy_flat(const u32x4* src, u32x4* dst, uint64_t first,
                                                            uint64_t n)
{
  const uint64_t k = first + (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (k < n) {
    copy_store<NT>(dst + k, copy_load<NT>(src + k));
  }
}

constexpr uint32_t kLdsPerCu = 160 * 1024;

// dynamic LDS per 256-thread block that caps the waves per SIMD (0: no cap),
// as the step kernel's launch configuration does
uint32_t lds_for_waves(uint32_t waves)
{
  if (waves == 0) {
    return 0;
  }
  const uint32_t blocks_per_cu = waves;  // 4 waves per block, 4 SIMDs per CU
  uint32_t lds = kLdsPerCu / blocks_per_cu / 1024 * 1024;
  if (lds * (blocks_per_cu + 1) <= kLdsPerCu) {
    lds += 1024;
  }
  return lds < kLdsPerCu - 4096 ? lds : kLdsPerCu - 4096;
}

}  // namespace

extern "C" {

olbgfx_status olbgfx_probe_bandwidth(const olbgfx_lattice_desc* desc, int32_t pattern, uint32_t shift_steps,
                                     int32_t block_order, uint32_t reps, void* stream, double* gbps)
{
  if (!desc || !gbps || reps == 0 || pattern < OLBGFX_PROBE_COPY || pattern > OLBGFX_PROBE_RING_TWO_BUFFER ||
      block_order < -1 || block_order > 12) {
    return OLBGFX_ERR_INVALID;
  }
  const olbgfx_lattice_desc& D = *desc;
  if (D.nx <= 0 || D.ny <= 0 || D.nz <= 0 || D.padding < 1) {
    return OLBGFX_ERR_INVALID;
  }
  const uint64_t elem = D.storage == OLBGFX_STORE_F32 ? 4 : (D.storage == OLBGFX_STORE_F64 ? 8 : 2);
  const uint64_t px = (uint64_t)D.nx + 2 * (uint64_t)D.padding;
  const uint64_t py = (uint64_t)D.ny + 2 * (uint64_t)D.padding;
  const uint64_t pz = (uint64_t)D.nz + 2 * (uint64_t)D.padding;
  const uint64_t count = px * py * pz;
  const uint64_t ncore = (uint64_t)D.nx * D.ny * D.nz;
  if (count >= (1ull << 31) || count * elem >= (1ull << 32)) {
    return OLBGFX_ERR_INVALID;
  }
  const uint64_t stride = (count + 256 / elem - 1) / (256 / elem) * (256 / elem);
  const uint64_t bytes = (uint64_t)Q * stride * elem;
  hipStream_t s = (hipStream_t)stream;

  void* a = nullptr;
  void* b = nullptr;
  float* sink = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) {
    return OLBGFX_ERR_NOMEM;
  }
  if (((pattern == OLBGFX_PROBE_COPY || pattern == OLBGFX_PROBE_RING_TWO_BUFFER) && hipMalloc(&b, bytes) != hipSuccess) ||
      hipMalloc(&sink, sizeof(float)) != hipSuccess) {
    (void)hipFree(a);
    (void)hipFree(b);
    return OLBGFX_ERR_NOMEM;
  }
  olbgfx_status st = OLBGFX_OK;
  if (hipMemsetAsync(a, 0, bytes, s) != hipSuccess || (b && hipMemsetAsync(b, 0, bytes, s) != hipSuccess)) {
    st = OLBGFX_ERR_HIP;
  }

  RingArgs r{};
  r.pop = a;
  r.pop2 = b;
  r.stride = stride;
  r.count = (uint32_t)count;
  const int64_t proj[3] = {(int64_t)(py * pz), (int64_t)pz, 1};
  for (int i = 0; i < Q; ++i) {
    const int64_t d = kCx[i] * proj[0] + kCy[i] * proj[1] + kCz[i] * proj[2];
    int64_t m = (-d * (int64_t)(shift_steps % count)) % (int64_t)count;
    if (m < 0) {
      m += (int64_t)count;
    }
    r.shift[i] = (uint32_t)m;
  }
  r.nyz = (uint32_t)(D.ny * D.nz);
  r.nz = (uint32_t)D.nz;
  r.pad = (uint32_t)D.padding;
  r.py = (uint32_t)py;
  r.pz = (uint32_t)pz;
  r.ncore = (uint32_t)ncore;
  r.nblocks = (uint32_t)((ncore + kThreads - 1) / kThreads);
  r.blocks_per_xcd = (r.nblocks + 7) / 8;
  r.run_log2 = block_order;
  if (block_order >= 0) {
    const uint32_t run = 1u << block_order;
    r.blocks_per_xcd = (r.blocks_per_xcd + run - 1) / run * run;
  }
  const dim3 grid(8 * r.blocks_per_xcd), block(kThreads);

  // the best of a small sweep: the copy over 256..16384 looping workgroups
  // and one workgroup per 256 elements, each with plain and nontemporal
  // accesses; the ring patterns over the occupancy caps the step kernel's
  // tuner tries
  const uint32_t copy_grids[] = {256, 512, 1024, 2048, 4096, 8192, 16384, 0};
  const uint32_t wave_caps[] = {0, 2, 3, 4, 5, 6};
  const uint32_t nconf = pattern == OLBGFX_PROBE_COPY ? 16 : 6;
  const uint64_t n16 = bytes / 16;
  auto launch = [&](uint32_t conf) {
    if (pattern == OLBGFX_PROBE_COPY) {
      const bool nt = conf >= 8;
      const uint32_t g = copy_grids[conf % 8];
      if (g != 0) {
        if (nt) probe_copy<true><<<g, kThreads, 0, s>>>((const u32x4*)a, (u32x4*)b, n16);
        else probe_copy<false><<<g, kThreads, 0, s>>>((const u32x4*)a, (u32x4*)b, n16);
        return;
      }
      // the work-items of one launch must fit in 32 bits: 2^23 workgroups
      // (2^31 lanes) per launch
      const uint64_t per_launch = (uint64_t)(1u << 23) * kThreads;
      for (uint64_t first = 0; first < n16; first += per_launch) {
        const uint64_t m = n16 - first < per_launch ? n16 - first : per_launch;
        const uint32_t blocks = (uint32_t)((m + kThreads - 1) / kThreads);
        if (nt) probe_copy_flat<true><<<blocks, kThreads, 0, s>>>((const u32x4*)a, (u32x4*)b, first, n16);
        else probe_copy_flat<false><<<blocks, kThreads, 0, s>>>((const u32x4*)a, (u32x4*)b, first, n16);
      }
      return;
    }
    const uint32_t lds = lds_for_waves(wave_caps[conf]);
    if (elem == 4) {
      if (pattern == OLBGFX_PROBE_RING_READ) probe_rings<float, 1><<<grid, block, lds, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_WRITE) probe_rings<float, 2><<<grid, block, lds, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_INPLACE) probe_rings<float, 3><<<grid, block, lds, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_TWO_BUFFER) probe_rings<float, 4><<<grid, block, lds, s>>>(r, sink);
    } else if (elem == 2) {
      if (pattern == OLBGFX_PROBE_RING_READ) probe_rings<_Float16, 1><<<grid, block, lds, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_WRITE) probe_rings<_Float16, 2><<<grid, block, lds, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_INPLACE) probe_rings<_Float16, 3><<<grid, block, lds, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_TWO_BUFFER) probe_rings<_Float16, 4><<<grid, block, lds, s>>>(r, sink);
    } else {
      if (pattern == OLBGFX_PROBE_RING_READ) probe_rings<double, 1><<<grid, block, lds, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_WRITE) probe_rings<double, 2><<<grid, block, lds, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_INPLACE) probe_rings<double, 3><<<grid, block, lds, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_TWO_BUFFER) probe_rings<double, 4><<<grid, block, lds, s>>>(r, sink);
    }
  };

  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (st == OLBGFX_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
    st = OLBGFX_ERR_HIP;
  }
  float ms = 0.f;
  for (uint32_t conf = 0; conf < nconf && st == OLBGFX_OK; ++conf) {
    float t = 0.f;
    launch(conf);  // untimed
    if (hipEventRecord(e0, s) != hipSuccess) {
      st = OLBGFX_ERR_HIP;
    }
    for (uint32_t k = 0; k < reps && st == OLBGFX_OK; ++k) {
      launch(conf);
    }
    if (st == OLBGFX_OK && (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
                            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)) {
      st = OLBGFX_ERR_HIP;
    }
    if (st == OLBGFX_OK && (conf == 0 || t < ms)) {
      ms = t;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(sink);
  if (st != OLBGFX_OK) {
    return st;
  }
  double moved = 0;
  if (pattern == OLBGFX_PROBE_COPY) {
    moved = 2.0 * (double)bytes;
  } else if (pattern == OLBGFX_PROBE_RING_INPLACE || pattern == OLBGFX_PROBE_RING_TWO_BUFFER) {
    moved = 2.0 * Q * (double)ncore * (double)elem;
  } else {
    moved = (double)Q * (double)ncore * (double)elem;
  }
  *gbps = moved * reps / (ms * 1e-3) / 1e9;
  return OLBGFX_OK;
}

}  // extern "C"
