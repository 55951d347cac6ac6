// olbprobe.hip -- bandwidth probes of the step kernel's access patterns
// (declared in include/olbgfx_probe.h): a plain 16-byte copy, and the 19 PS
// rings read, written, and read-modified-written in place, with the step
// kernel's cell indexing, block shape and block order.  They give the
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

// PATTERN: 1 read, 2 write, 3 in place
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

__global__ void __launch_bounds__(kThreads) probe_copy(const uint4* src, uint4* dst, uint64_t n)
{
  for (uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kThreads) {
    dst[k] = src[k];
  }
}

}  // namespace

extern "C" {

olbgfx_status olbgfx_probe_bandwidth(const olbgfx_lattice_desc* desc, int32_t pattern, uint32_t shift_steps,
                                     int32_t block_order, uint32_t reps, void* stream, double* gbps)
{
  if (!desc || !gbps || reps == 0 || pattern < OLBGFX_PROBE_COPY || pattern > OLBGFX_PROBE_RING_INPLACE ||
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
  if ((pattern == OLBGFX_PROBE_COPY && hipMalloc(&b, bytes) != hipSuccess) ||
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

  auto launch = [&]() {
    if (pattern == OLBGFX_PROBE_COPY) {
      probe_copy<<<4096, kThreads, 0, s>>>((const uint4*)a, (uint4*)b, bytes / 16);
      return;
    }
    if (elem == 4) {
      if (pattern == OLBGFX_PROBE_RING_READ) probe_rings<float, 1><<<grid, block, 0, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_WRITE) probe_rings<float, 2><<<grid, block, 0, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_INPLACE) probe_rings<float, 3><<<grid, block, 0, s>>>(r, sink);
    } else if (elem == 2) {
      if (pattern == OLBGFX_PROBE_RING_READ) probe_rings<_Float16, 1><<<grid, block, 0, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_WRITE) probe_rings<_Float16, 2><<<grid, block, 0, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_INPLACE) probe_rings<_Float16, 3><<<grid, block, 0, s>>>(r, sink);
    } else {
      if (pattern == OLBGFX_PROBE_RING_READ) probe_rings<double, 1><<<grid, block, 0, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_WRITE) probe_rings<double, 2><<<grid, block, 0, s>>>(r, sink);
      if (pattern == OLBGFX_PROBE_RING_INPLACE) probe_rings<double, 3><<<grid, block, 0, s>>>(r, sink);
    }
  };

  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (st == OLBGFX_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
    st = OLBGFX_ERR_HIP;
  }
  float ms = 0.f;
  if (st == OLBGFX_OK) {
    launch();  // untimed
    if (hipEventRecord(e0, s) != hipSuccess) {
      st = OLBGFX_ERR_HIP;
    }
    for (uint32_t k = 0; k < reps && st == OLBGFX_OK; ++k) {
      launch();
    }
    if (st == OLBGFX_OK && (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
                            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) {
      st = OLBGFX_ERR_HIP;
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
  } else if (pattern == OLBGFX_PROBE_RING_INPLACE) {
    moved = 2.0 * Q * (double)ncore * (double)elem;
  } else {
    moved = (double)Q * (double)ncore * (double)elem;
  }
  *gbps = moved * reps / (ms * 1e-3) / 1e9;
  return OLBGFX_OK;
}

}  // extern "C"
