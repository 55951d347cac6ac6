// olbgfx.hip -- gfx950 (MI355X) D3Q19 lattice Boltzmann platform behind the
// C ABI declared in include/olbgfx.h.
//
// Storage: 19 structure-of-arrays rings, one per population, in one device
// allocation (ring stride rounded to 256 B).  Population i of logical cell n
// lives at ring slot (n + s_i) mod ncells, the periodic-shift (PS) pattern of
// OpenLB (core/platform/cpu/sisd/column.h:129-204): collision is in place and
// streaming only changes s_i -= d_i, so one step moves exactly 2*19 values per
// cell.  Unlike the reference, whose host rotates 19 pointers and copies them
// to the device every step (core/blockLattice.hh:316-338), the offsets here
// are kernel arguments, or live in device memory so a captured graph can be
// replayed.
//
// The step is one kernel: every core cell loads its 19 populations (coalesced
// along z), applies its dynamics (BGK, TRT, Guo-forced BGK/TRT, BounceBack,
// BounceBackVelocity, selected by a one-byte id per cell or uniformly), stores
// them back, and -- on periodic boundary planes -- also writes the populations
// that cross the boundary into the width-1 ghost layer.  That last part
// replaces the reference's PostCollide overlap copy
// (core/platform/gpu/hip/communicator.hh:529-585): only the populations a core
// cell will pull are written.  Optionally the same pass reduces the lattice
// statistics in FP64 with wavefront shuffles.
//
// The roofline of the step is HBM bandwidth: 2*19*sizeof(storage) bytes per
// cell update against ~400 flops, far below the ridge point, so no MFMA.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/olbgfx.h"

namespace {

constexpr int Q = 19;
// largest workgroup of the step kernel; more waves per block would cap the
// registers per lane below what the D3Q19 collision needs without spilling
constexpr uint32_t kMaxBlock = 256;
constexpr int kMaskedBoundary = -2;  // step-kernel MODE: masked, LocalVelocity kinds compiled in
// step-kernel MODE kWithFaces + K: the dominant launch of a split step whose
// dominant kind is K, with the LocalVelocity and FixedVelocity cells (the
// faces of a velocity boundary and its outer edges and corners, before their
// post-processors) updated inline; every other cell is left to the residual
// launch
constexpr int kWithFaces = 32;

// the boundary kinds that read the cell's discrete normal and run in the
// masked kernel with boundary kinds compiled in
__host__ __device__ constexpr bool is_velocity_boundary(int kind)
{
  return kind == OLBGFX_LOCAL_VELOCITY || kind == OLBGFX_FIXED_VELOCITY || kind == OLBGFX_INNER_EDGE_VELOCITY ||
         kind == OLBGFX_INNER_CORNER_VELOCITY || kind == OLBGFX_LOCAL_PRESSURE ||
         kind == OLBGFX_INTERPOLATED_VELOCITY || kind == OLBGFX_INTERPOLATED_PRESSURE;
}
// the flat boundary kinds whose cells run PlaneFdBoundaryProcessor3D
__host__ __device__ constexpr bool is_plane_fd(int kind)
{
  return kind == OLBGFX_INTERPOLATED_VELOCITY || kind == OLBGFX_INTERPOLATED_PRESSURE;
}
// LDS of one gfx950 compute unit
constexpr uint32_t kLdsPerCu = 160 * 1024;
// persistent blocks per ensemble launch at most (bounds the ticket atomics)
constexpr uint32_t kMaxEnsembleBlocks = 8192;

// D3Q19 velocity set, descriptor/definition/common.h:166-199.
// Index 0 rests, 1..9 are the negative half, i+9 is the opposite of i.
__host__ __device__ constexpr int CX(int i)
{
  constexpr int v[Q] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
  return v[i];
}
__host__ __device__ constexpr int CY(int i)
{
  constexpr int v[Q] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
  return v[i];
}
__host__ __device__ constexpr int CZ(int i)
{
  constexpr int v[Q] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
  return v[i];
}
__host__ __device__ constexpr int TDEN(int i)
{
  return i == 0 ? 3 : ((i % 9 >= 1 && i % 9 <= 3) ? 18 : 36);
}

// ---------------------------------------------------------------------------
// storage types

struct bf16_t {
  uint16_t bits;
};

template <typename S>
struct Storage;

template <>
struct Storage<float> {
  using A = float;
  static __device__ __forceinline__ float load(const float* p) { return *p; }
  static __device__ __forceinline__ float pack(float v) { return v; }
  static __device__ __forceinline__ float unpack(float v) { return v; }
};

template <>
struct Storage<double> {
  using A = double;
  static __device__ __forceinline__ double load(const double* p) { return *p; }
  static __device__ __forceinline__ double pack(double v) { return v; }
  static __device__ __forceinline__ double unpack(double v) { return v; }
};

// FP16 storage in a named wrapper (so kernel names demangle readably)
struct f16_t {
  _Float16 v;
};

template <>
struct Storage<f16_t> {
  using A = float;
  // v_cvt_f16_f32 rounds to nearest even and keeps FP16 subnormals.  The
  // compiler folds these conversions into mixed-precision FMAs (v_fma_mix*)
  // where the instruction graph allows: exact on loads, and on a store it
  // rounds a fused FMA once to FP16 where the unfused form rounds to FP32
  // first.  Kept, because the uniform FP16 kernel is partly issue-bound and
  // runs 3-4 % slower with the conversions forced apart; so a kernel that
  // fuses and one that does not can differ in the last bit of FP16 storage.
  static __device__ __forceinline__ float unpack(f16_t h) { return (float)h.v; }
  static __device__ __forceinline__ float load(const f16_t* p) { return (float)p->v; }
  static __device__ __forceinline__ f16_t pack(float v) { return f16_t{(_Float16)v}; }
};

template <>
struct Storage<bf16_t> {
  using A = float;
  static __device__ __forceinline__ float unpack(bf16_t b)
  {
    return __uint_as_float((uint32_t)b.bits << 16);
  }
  static __device__ __forceinline__ float load(const bf16_t* p) { return unpack(*p); }
  // round to nearest even: gfx950's v_cvt_pk_bf16_f32, one instruction in
  // place of the integer rounding sequence
  static __device__ __forceinline__ bf16_t pack(float v)
  {
    return bf16_t{__builtin_bit_cast(uint16_t, (__bf16)v)};
  }
};

// ---------------------------------------------------------------------------
// kernel arguments

template <typename A>
struct DynEntry {
  int32_t kind;
  A omega;
  A omega2;      // TRT odd relaxation, formed on the host like collision.h:408
  A force[3];
  A wall_u[3];
  A wall_rho;
  A smago_pre;   // Smagorinsky prefactor C^2 invCs2^2 2 sqrt(2) (collisionLES.h:71-73)
};

template <typename A>
struct StepArgs {
  void* pop;                 // ring 0; ring i starts at pop + i*stride elements
  uint64_t stride;
  uint32_t count;            // ring length = padded cells
  uint32_t shift[Q];         // used when shift_dev is null
  const uint32_t* shift_dev; // device copy of the offsets (graph replay)
  const uint8_t* mask;       // dynamics id per padded cell
  int32_t uniform_id;        // >= 0: every core cell has this id, mask unused
  int32_t skip_others;       // 1: cells whose id is not uniform_id are left to the residual launch
  int32_t nx, ny, nz, pad;
  uint32_t py, pz;           // padded extent in y and z
  uint32_t nyz;              // ny*nz
  uint64_t magic_nyz;        // ceil(2^64 / nyz), 0 if nyz == 1
  uint64_t magic_nz;         // ceil(2^64 / nz),  0 if nz == 1
  uint32_t ncore;            // nx*ny*nz
  uint32_t nblocks;          // blocks that carry cells
  uint32_t blocks_per_xcd;   // ceil(nblocks / 8), rounded up to whole runs
  int32_t xcd_run_log2;      // <0: one contiguous part per XCD; k: runs of 2^k blocks
  int32_t periodic;          // bit a set: axis a periodic
  const A* cell_u;           // optional VELOCITY field (3 per cell)
  const A* cell_rho;         // optional RHO field
  const A* cell_force;       // optional FORCE field (3 per cell)
  const uint8_t* normal;     // optional discrete normal code per cell, 9(nx+1)+3(ny+1)+(nz+1)
  double* stats;             // 4 doubles per block when statistics are on
  DynEntry<A> table[OLBGFX_MAX_DYNAMICS];
};

// n / d for 32-bit n, with m = ceil(2^64 / d) (exact for every 32-bit n)
__device__ __forceinline__ uint32_t fast_div(uint32_t n, uint64_t m, uint32_t d)
{
  if (m == 0) {
    return n / d;
  }
  const uint64_t lo = (uint64_t)(uint32_t)m * n;
  const uint64_t hi = (m >> 32) * n;
  return (uint32_t)((hi + (lo >> 32)) >> 32);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v += __shfl_xor(v, o, 64);
  }
  return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v = fmax(v, __shfl_xor(v, o, 64));
  }
  return v;
}

// block-wide FP64 reduction of four values into out[0..3]
__device__ void block_reduce4(double c, double r, double e, double m, double* out, bool write)
{
  __shared__ double part[16][4];
  c = wave_sum(c);
  r = wave_sum(r);
  e = wave_sum(e);
  m = wave_max(m);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) {
    part[wave][0] = c;
    part[wave][1] = r;
    part[wave][2] = e;
    part[wave][3] = m;
  }
  __syncthreads();
  if (wave == 0) {
    const int nw = blockDim.x >> 6;
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    if (lane < nw) {
      a0 = part[lane][0];
      a1 = part[lane][1];
      a2 = part[lane][2];
      a3 = part[lane][3];
    }
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    a3 = wave_max(a3);
    if (lane == 0 && write) {
      out[0] = a0;
      out[1] = a1;
      out[2] = a2;
      out[3] = a3;
    }
  }
}

// ---------------------------------------------------------------------------
// per-cell dynamics on 19 registers

template <typename A>
__device__ __forceinline__ void moments(const A (&f)[Q], A& drho, A& jx, A& jy, A& jz)
{
  // pairwise sums keep the dependency chains short
  A s = (f[0] + f[1]) + (f[2] + f[3]);
  s += (f[4] + f[5]) + (f[6] + f[7]);
  s += (f[8] + f[9]) + (f[10] + f[11]);
  s += (f[12] + f[13]) + (f[14] + f[15]);
  s += (f[16] + f[17]) + f[18];
  drho = s;
  jx = ((f[10] - f[1]) + (f[13] - f[4])) + ((f[14] - f[5]) + (f[15] - f[6])) + (f[16] - f[7]);
  jy = ((f[11] - f[2]) + (f[13] - f[4])) + ((f[5] - f[14]) + (f[17] - f[8])) + (f[18] - f[9]);
  jz = ((f[12] - f[3]) + (f[15] - f[6])) + ((f[7] - f[16]) + (f[17] - f[8])) + (f[9] - f[18]);
}

template <int I, typename A>
__device__ __forceinline__ A cdot(A x, A y, A z)
{
  A r = A(0);
  bool first = true;
  if constexpr (CX(I) != 0) {
    r = CX(I) > 0 ? x : -x;
    first = false;
  }
  if constexpr (CY(I) != 0) {
    r = first ? (CY(I) > 0 ? y : -y) : (CY(I) > 0 ? r + y : r - y);
    first = false;
  }
  if constexpr (CZ(I) != 0) {
    r = first ? (CZ(I) > 0 ? z : -z) : (CZ(I) > 0 ? r + z : r - z);
  }
  return r;
}

template <int I, typename A>
__device__ __forceinline__ constexpr A weight()
{
  return A(1) / A(TDEN(I));
}

// BGK (collision.h:57-80) or TRT (collision.h:381-419) relaxation towards the
// shifted second-order equilibrium (lbm.h:51-64), written in even/odd pairs:
// feq_{i,opp} = sym +- asym with
//   sym  = t_i (drho + rho (4.5 (c.u)^2 - 1.5 u^2)),  asym = 3 t_i rho (c.u).
template <bool TRT, int P, typename A>
__device__ __forceinline__ void relax_pair(A (&f)[Q], A drho, A rho, A ux, A uy, A uz,
                                           A k15, A om, A om2)
{
  const A t = weight<P, A>();
  const A cu = cdot<P>(ux, uy, uz);
  const A sym = t * (drho + rho * (A(4.5) * cu * cu - k15));
  const A asym = t * A(3) * rho * cu;
  const A fi = f[P];
  const A fo = f[P - 9];
  if constexpr (!TRT) {
    f[P] = fi + om * ((sym + asym) - fi);
    f[P - 9] = fo + om * ((sym - asym) - fo);
  } else {
    const A dp = om * (A(0.5) * (fi + fo) - sym);
    const A dm = om2 * (A(0.5) * (fi - fo) - asym);
    f[P] = (fi - dp) - dm;
    f[P - 9] = (fo - dp) + dm;
  }
}

template <bool TRT, typename A>
__device__ __forceinline__ void relax(A (&f)[Q], A drho, A rho, A ux, A uy, A uz, A om, A om2)
{
  const A k15 = A(1.5) * (ux * ux + uy * uy + uz * uz);
  const A feq0 = weight<0, A>() * (drho - rho * k15);
  f[0] = f[0] + om * (feq0 - f[0]);
  relax_pair<TRT, 10>(f, drho, rho, ux, uy, uz, k15, om, om2);
  relax_pair<TRT, 11>(f, drho, rho, ux, uy, uz, k15, om, om2);
  relax_pair<TRT, 12>(f, drho, rho, ux, uy, uz, k15, om, om2);
  relax_pair<TRT, 13>(f, drho, rho, ux, uy, uz, k15, om, om2);
  relax_pair<TRT, 14>(f, drho, rho, ux, uy, uz, k15, om, om2);
  relax_pair<TRT, 15>(f, drho, rho, ux, uy, uz, k15, om, om2);
  relax_pair<TRT, 16>(f, drho, rho, ux, uy, uz, k15, om, om2);
  relax_pair<TRT, 17>(f, drho, rho, ux, uy, uz, k15, om, om2);
  relax_pair<TRT, 18>(f, drho, rho, ux, uy, uz, k15, om, om2);
}

// Guo force term after the collision (lbm.h:567-588):
// f_i += rho (1 - omega/2) t_i [3 (c_i - u).F + 9 (c_i.u)(c_i.F)]
template <int P, typename A>
__device__ __forceinline__ void force_pair(A (&f)[Q], A k, A ux, A uy, A uz, A fx, A fy, A fz, A uf)
{
  const A t = weight<P, A>();
  const A cu = cdot<P>(ux, uy, uz);
  const A cf = cdot<P>(fx, fy, fz);
  const A sym = k * t * (A(9) * cu * cf - A(3) * uf);
  const A asym = k * t * A(3) * cf;
  f[P] += sym + asym;
  f[P - 9] += sym - asym;
}

template <typename A>
__device__ __forceinline__ void add_force(A (&f)[Q], A rho, A om, A ux, A uy, A uz, A fx, A fy, A fz)
{
  const A k = rho * (A(1) - om * A(0.5));
  const A uf = ux * fx + uy * fy + uz * fz;
  f[0] += k * weight<0, A>() * (A(-3) * uf);
  force_pair<10>(f, k, ux, uy, uz, fx, fy, fz, uf);
  force_pair<11>(f, k, ux, uy, uz, fx, fy, fz, uf);
  force_pair<12>(f, k, ux, uy, uz, fx, fy, fz, uf);
  force_pair<13>(f, k, ux, uy, uz, fx, fy, fz, uf);
  force_pair<14>(f, k, ux, uy, uz, fx, fy, fz, uf);
  force_pair<15>(f, k, ux, uy, uz, fx, fy, fz, uf);
  force_pair<16>(f, k, ux, uy, uz, fx, fy, fz, uf);
  force_pair<17>(f, k, ux, uy, uz, fx, fy, fz, uf);
  force_pair<18>(f, k, ux, uy, uz, fx, fy, fz, uf);
}

// Revert (collision.h:465-484) and the Nguyen-Ladd moving-wall correction
// f_i += 6 rho_w t_i (c_i . u_w) (collision.h:494-520)
template <int P, typename A>
__device__ __forceinline__ void revert_pair(A (&f)[Q], bool moving, A rw, A wx, A wy, A wz)
{
  const A a = f[P];
  f[P] = f[P - 9];
  f[P - 9] = a;
  if (moving) {
    const A corr = A(6) * rw * weight<P, A>() * cdot<P>(wx, wy, wz);
    f[P] += corr;
    f[P - 9] -= corr;
  }
}

template <typename A>
__device__ __forceinline__ void revert(A (&f)[Q], bool moving, A rw, A wx, A wy, A wz)
{
  revert_pair<10>(f, moving, rw, wx, wy, wz);
  revert_pair<11>(f, moving, rw, wx, wy, wz);
  revert_pair<12>(f, moving, rw, wx, wy, wz);
  revert_pair<13>(f, moving, rw, wx, wy, wz);
  revert_pair<14>(f, moving, rw, wx, wy, wz);
  revert_pair<15>(f, moving, rw, wx, wy, wz);
  revert_pair<16>(f, moving, rw, wx, wy, wz);
  revert_pair<17>(f, moving, rw, wx, wy, wz);
  revert_pair<18>(f, moving, rw, wx, wy, wz);
}

// apply dynamics KIND to one cell; returns true if it yields a statistic
// (CellStatistic with rho != -1, dynamics/interface.h:44-52)
template <int KIND, typename A>
__device__ __forceinline__ bool apply_kind(A (&f)[Q], const DynEntry<A>& e, uint32_t n,
                                           const StepArgs<A>& a, A& st_rho, A& st_usqr)
{
  if constexpr (KIND >= OLBGFX_BGK && KIND <= OLBGFX_FORCED_TRT) {
    constexpr bool forced = KIND == OLBGFX_FORCED_BGK || KIND == OLBGFX_FORCED_TRT;
    constexpr bool trt = KIND == OLBGFX_TRT || KIND == OLBGFX_FORCED_TRT;
    A drho, jx, jy, jz;
    moments(f, drho, jx, jy, jz);
    const A rho = drho + A(1);
    A ux = jx / rho;
    A uy = jy / rho;
    A uz = jz / rho;
    A fx = A(0), fy = A(0), fz = A(0);
    if constexpr (forced) {
      if (a.cell_force) {
        fx = a.cell_force[3 * (size_t)n + 0];
        fy = a.cell_force[3 * (size_t)n + 1];
        fz = a.cell_force[3 * (size_t)n + 2];
      } else {
        fx = e.force[0];
        fy = e.force[1];
        fz = e.force[2];
      }
      // momenta::ForcedMomentum::computeU shifts u by F/2 (elements.h:1600-1608)
      ux += fx * A(0.5);
      uy += fy * A(0.5);
      uz += fz * A(0.5);
    }
    relax<trt>(f, drho, rho, ux, uy, uz, e.omega, e.omega2);
    if constexpr (forced) {
      add_force(f, rho, e.omega, ux, uy, uz, fx, fy, fz);
    }
    st_rho = rho;
    st_usqr = ux * ux + uy * uy + uz * uz;
    return true;
  } else if constexpr (KIND == OLBGFX_SMAGORINSKY_BGK) {
    // SmagorinskyEffectiveOmega<BGK> (collisionLES.h:59-89): stress of
    // lbm.h:334-352 with rho - 1 = drho, norm of lbm.h:555-564
    A drho, jx, jy, jz;
    moments(f, drho, jx, jy, jz);
    const A rho = drho + A(1);
    const A ux = jx / rho, uy = jy / rho, uz = jz / rho;
    A sxx = A(0), syy = A(0), szz = A(0), sxy = A(0), sxz = A(0), syz = A(0);
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      if constexpr (true) {
        const int cx = CX(i), cy = CY(i), cz = CZ(i);
        if (cx != 0) sxx += f[i];
        if (cy != 0) syy += f[i];
        if (cz != 0) szz += f[i];
        if (cx * cy > 0) sxy += f[i];
        if (cx * cy < 0) sxy -= f[i];
        if (cx * cz > 0) sxz += f[i];
        if (cx * cz < 0) sxz -= f[i];
        if (cy * cz > 0) syz += f[i];
        if (cy * cz < 0) syz -= f[i];
      }
    }
    const A third = A(1) / A(3);
    const A pxx = sxx - rho * ux * ux - third * drho;
    const A pyy = syy - rho * uy * uy - third * drho;
    const A pzz = szz - rho * uz * uz - third * drho;
    const A pxy = sxy - rho * ux * uy;
    const A pxz = sxz - rho * ux * uz;
    const A pyz = syz - rho * uy * uz;
    const A norm2 = pxx * pxx + pyy * pyy + pzz * pzz + A(2) * (pxy * pxy + pxz * pxz + pyz * pyz);
    const A tau_mol = A(1) / e.omega;
    const A tau_turb = A(0.5) * (sqrt(tau_mol * tau_mol + e.smago_pre / rho * sqrt(norm2)) - tau_mol);
    const A om = A(1) / (tau_mol + tau_turb);
    relax<false>(f, drho, rho, ux, uy, uz, om, om);
    st_rho = rho;
    st_usqr = ux * ux + uy * uy + uz * uz;
    return true;
  } else if constexpr (KIND == OLBGFX_BOUNCE_BACK) {
    revert(f, false, A(0), A(0), A(0), A(0));
    return false;
  } else if constexpr (KIND == OLBGFX_BOUNCE_BACK_VELOCITY) {
    const A rw = a.cell_rho ? a.cell_rho[n] : e.wall_rho;
    A wx = e.wall_u[0], wy = e.wall_u[1], wz = e.wall_u[2];
    if (a.cell_u) {
      wx = a.cell_u[3 * (size_t)n + 0];
      wy = a.cell_u[3 * (size_t)n + 1];
      wz = a.cell_u[3 * (size_t)n + 2];
    }
    revert(f, true, rw, wx, wy, wz);
    st_rho = rw;
    st_usqr = wx * wx + wy * wy + wz * wz;
    return true;
  } else {
    return false;
  }
}

template <typename A>
__device__ __forceinline__ bool apply_any(A (&f)[Q], const DynEntry<A>& e, uint32_t n,
                                          const StepArgs<A>& a, A& st_rho, A& st_usqr)
{
  switch (e.kind) {
  case OLBGFX_BGK:
    return apply_kind<OLBGFX_BGK>(f, e, n, a, st_rho, st_usqr);
  case OLBGFX_TRT:
    return apply_kind<OLBGFX_TRT>(f, e, n, a, st_rho, st_usqr);
  case OLBGFX_FORCED_BGK:
    return apply_kind<OLBGFX_FORCED_BGK>(f, e, n, a, st_rho, st_usqr);
  case OLBGFX_FORCED_TRT:
    return apply_kind<OLBGFX_FORCED_TRT>(f, e, n, a, st_rho, st_usqr);
  case OLBGFX_BOUNCE_BACK:
    return apply_kind<OLBGFX_BOUNCE_BACK>(f, e, n, a, st_rho, st_usqr);
  case OLBGFX_BOUNCE_BACK_VELOCITY:
    return apply_kind<OLBGFX_BOUNCE_BACK_VELOCITY>(f, e, n, a, st_rho, st_usqr);
  case OLBGFX_SMAGORINSKY_BGK:
    return apply_kind<OLBGFX_SMAGORINSKY_BGK>(f, e, n, a, st_rho, st_usqr);
  default:
    return false;
  }
}

// ---------------------------------------------------------------------------
// LocalVelocity boundaries: generic loops over the populations (boundary
// cells are few, so clarity wins over pairing here)

__host__ __device__ constexpr int CA(int i, int axis)
{
  return axis == 0 ? CX(i) : (axis == 1 ? CY(i) : CZ(i));
}

__device__ __forceinline__ int normal_component(uint32_t code, int axis)
{
  const int v = axis == 0 ? (int)(code / 9) : (axis == 1 ? (int)(code / 3 % 3) : (int)(code % 3));
  return v - 1;
}

template <typename A>
__device__ __forceinline__ A wall_component(const StepArgs<A>& a, const DynEntry<A>& e, uint32_t n, int axis)
{
  return a.cell_u ? a.cell_u[3 * (size_t)n + axis] : e.wall_u[axis];
}

// shifted second-order equilibrium of population i (lbm.h:51-64)
template <typename A>
__device__ __forceinline__ A feq_i(int i, A rho, const A (&u)[3], A usqr)
{
  const A cu = A(CX(i)) * u[0] + A(CY(i)) * u[1] + A(CZ(i)) * u[2];
  const A t = A(1) / A(TDEN(i));
  return rho * t * (A(1) + A(3) * cu + A(4.5) * cu * cu - A(1.5) * usqr) - t;
}

// velocityBMRho (momenta/elements.h:55-71): density implied by the known
// populations of a fixed-velocity face with the given direction/orientation
template <typename A>
__device__ __forceinline__ A velocity_bm_rho(const A (&f)[Q], int dir, int orient, A un)
{
  A on_wall = A(0), normal = A(0);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int c = CA(i, dir);
    if (c == 0) {
      on_wall += f[i];
    } else if (c == orient) {
      normal += f[i];
    }
  }
  return (A(2) * normal + on_wall + A(1)) / (A(1) + A(orient) * un);
}

// fromPiToFneq (lbm.h:248-270); pi = (xx, xy, xz, yy, yz, zz)
template <typename A>
__device__ __forceinline__ A pi_to_fneq(int i, const A (&pi)[6])
{
  const A cx = A(CX(i)), cy = A(CY(i)), cz = A(CZ(i));
  const A third = A(1) / A(3);
  A r = (cx * cx - third) * pi[0] + A(2) * cx * cy * pi[1] + A(2) * cx * cz * pi[2] +
        (cy * cy - third) * pi[3] + A(2) * cy * cz * pi[4] + (cz * cz - third) * pi[5];
  return r * (A(4.5) / A(TDEN(i)));
}

// CombinedRLBdynamics<BGK, RegularizedVelocityBoundaryTuple<dir, orient>>::collide
// (dynamics.h:506-517, boundary/helper.h:31-81): regularise around the
// boundary momenta, then BGK with the same momenta
template <typename A>
__device__ __forceinline__ void local_velocity(A (&f)[Q], int dir, int orient, const A (&u)[3], A om,
                                               A& rho_out)
{
  const A usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  A rho = velocity_bm_rho(f, dir, orient, u[dir]);
  A pi[6] = {A(0), A(0), A(0), A(0), A(0), A(0)};
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const int c = CA(i, dir);
    if (c == 0 || c == orient) {
      const A w = (c == 0 ? A(1) : A(2)) * (f[i] - feq_i(i, rho, u, usqr));
      const A cx = A(CX(i)), cy = A(CY(i)), cz = A(CZ(i));
      pi[0] += cx * cx * w;
      pi[1] += cx * cy * w;
      pi[2] += cx * cz * w;
      pi[3] += cy * cy * w;
      pi[4] += cy * cz * w;
      pi[5] += cz * cz * w;
    }
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    f[i] = feq_i(i, rho, u, usqr) + pi_to_fneq(i, pi);
  }
  rho = velocity_bm_rho(f, dir, orient, u[dir]);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    f[i] = f[i] + om * (feq_i(i, rho, u, usqr) - f[i]);
  }
  rho_out = rho;
}

// lbm<D>::computeStress (dynamics/lbm.h:334-352) of shifted populations
template <typename A>
__device__ __forceinline__ void bulk_stress(const A (&g)[Q], A rho, const A (&u)[3], A (&pi)[6])
{
  A s[6] = {A(0), A(0), A(0), A(0), A(0), A(0)};
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const A cx = A(CX(i)), cy = A(CY(i)), cz = A(CZ(i));
    s[0] += cx * cx * g[i];
    s[1] += cx * cy * g[i];
    s[2] += cx * cz * g[i];
    s[3] += cy * cy * g[i];
    s[4] += cy * cz * g[i];
    s[5] += cz * cz * g[i];
  }
  const A iso = A(1) / A(3) * (rho - A(1));
  pi[0] = s[0] - rho * u[0] * u[0] - iso;
  pi[1] = s[1] - rho * u[0] * u[1];
  pi[2] = s[2] - rho * u[0] * u[2];
  pi[3] = s[3] - rho * u[1] * u[1] - iso;
  pi[4] = s[4] - rho * u[1] * u[2];
  pi[5] = s[5] - rho * u[2] * u[2] - iso;
}

// InnerEdgeDensity3D / InnerCornerDensity3D (momenta/elements.h:432-482):
// the mean of velocityBMRho over the non-zero axes of the outward normal; an
// edge's axes are plane+1 and plane+2 with plane its zero axis
// (boundary/setBoundary3D.h:156-197)
template <typename A>
__device__ __forceinline__ A inner_density(const A (&f)[Q], const int (&N)[3], const A (&u)[3])
{
  if (N[0] != 0 && N[1] != 0 && N[2] != 0) {
    return (velocity_bm_rho(f, 0, N[0], u[0]) + velocity_bm_rho(f, 1, N[1], u[1]) +
            velocity_bm_rho(f, 2, N[2], u[2])) / A(3);
  }
  const int plane = N[0] == 0 ? 0 : (N[1] == 0 ? 1 : 2);
  const int a1 = (plane + 1) % 3, a2 = (plane + 2) % 3;
  return (velocity_bm_rho(f, a1, N[a1], u[a1]) + velocity_bm_rho(f, a2, N[a2], u[a2])) / A(2);
}

// CombinedRLBdynamics<BGK, InnerEdge / InnerCornerVelocityTuple3D>::collide:
// the stress of InnerEdgeStress3D replaces the population entering from both
// walls by bounce-back of its non-equilibrium part (elements.h:1871-1892);
// an inner corner's entering population -N does not exist in D3Q19
// (elements.h:1843-1869), so its stress is the bulk one
template <typename A>
__device__ __forceinline__ void inner_velocity(A (&f)[Q], const int (&N)[3], const A (&u)[3], A om, A& rho_out)
{
  const A usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  A rho = inner_density(f, N, u);
  A g[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) g[i] = f[i];
  if (N[0] == 0 || N[1] == 0 || N[2] == 0) {
    const int plane = N[0] == 0 ? 0 : (N[1] == 0 ? 1 : 2);
    const int a1 = (plane + 1) % 3, a2 = (plane + 2) % 3;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      if (CA(i, a1) == -N[a1] && CA(i, a2) == -N[a2]) {
        const int o = i <= 9 ? i + 9 : i - 9;
        g[i] = f[o] - feq_i(o, rho, u, usqr) + feq_i(i, rho, u, usqr);
      }
    }
  }
  A pi[6];
  bulk_stress(g, rho, u, pi);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    f[i] = feq_i(i, rho, u, usqr) + pi_to_fneq(i, pi);
  }
  rho = inner_density(f, N, u);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    f[i] = f[i] + om * (feq_i(i, rho, u, usqr) - f[i]);
  }
  rho_out = rho;
}

// FixedPressureMomentum::computeU (momenta/elements.h:668-691): the wall's
// tangential velocity, the normal component from the known populations
template <typename A>
__device__ __forceinline__ void pressure_u(const A (&f)[Q], int dir, int orient, A rho, const A (&uw)[3],
                                           A (&u)[3])
{
  A on_wall = A(0), normal = A(0);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int c = CA(i, dir);
    if (c == 0) {
      on_wall += f[i];
    } else if (c == orient) {
      normal += f[i];
    }
  }
  u[0] = uw[0];
  u[1] = uw[1];
  u[2] = uw[2];
  u[dir] = A(orient) * ((A(2) * normal + on_wall + A(1)) / rho - A(1));
}

// CombinedRLBdynamics<BGK, RegularizedPressureBoundaryTuple<dir, orient>>::collide
template <typename A>
__device__ __forceinline__ void local_pressure(A (&f)[Q], int dir, int orient, A rho, const A (&uw)[3], A om,
                                               A (&u)[3])
{
  pressure_u(f, dir, orient, rho, uw, u);
  A usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  A pi[6] = {A(0), A(0), A(0), A(0), A(0), A(0)};
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const int c = CA(i, dir);
    if (c == 0 || c == orient) {
      const A w = (c == 0 ? A(1) : A(2)) * (f[i] - feq_i(i, rho, u, usqr));
      const A cx = A(CX(i)), cy = A(CY(i)), cz = A(CZ(i));
      pi[0] += cx * cx * w;
      pi[1] += cx * cy * w;
      pi[2] += cx * cz * w;
      pi[3] += cy * cy * w;
      pi[4] += cy * cz * w;
      pi[5] += cz * cz * w;
    }
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    f[i] = feq_i(i, rho, u, usqr) + pi_to_fneq(i, pi);
  }
  pressure_u(f, dir, orient, rho, uw, u);
  usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    f[i] = f[i] + om * (feq_i(i, rho, u, usqr) - f[i]);
  }
}

// the flat normal of a cell: direction and orientation
__device__ __forceinline__ void flat_normal(uint32_t code, int& dir, int& orient)
{
  dir = 0;
  orient = normal_component(code, 0);
  if (normal_component(code, 1) != 0) {
    dir = 1;
    orient = normal_component(code, 1);
  } else if (normal_component(code, 2) != 0) {
    dir = 2;
    orient = normal_component(code, 2);
  }
}

// the boundary kinds that read the discrete normal (the masked kernel with
// boundary kinds); returns true: all yield a statistic
// a flat LocalVelocity face (CombinedRLBdynamics with
// RegularizedVelocityBoundaryTuple): the regularised BGK of local_velocity
template <typename A>
__device__ __forceinline__ void face_local_velocity(A (&f)[Q], const DynEntry<A>& e, uint32_t n,
                                                    const StepArgs<A>& a, const A (&u)[3], A& rho)
{
  int dir, orient;
  flat_normal(a.normal[n], dir, orient);
  local_velocity(f, dir, orient, u, e.omega, rho);
}

// BulkDensity + FixedVelocityMomentumGeneric, BGK (collision.h:57-80)
template <typename A>
__device__ __forceinline__ void fixed_velocity_bgk(A (&f)[Q], const DynEntry<A>& e, const A (&u)[3], A& rho)
{
  A s = A(0);
#pragma unroll
  for (int i = 0; i < Q; ++i) s += f[i];
  rho = s + A(1);
  const A usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    f[i] = f[i] + e.omega * (feq_i(i, rho, u, usqr) - f[i]);
  }
}

template <typename A>
__device__ __forceinline__ bool apply_velocity_boundary(A (&f)[Q], const DynEntry<A>& e, uint32_t n,
                                                        const StepArgs<A>& a, A& st_rho, A& st_usqr)
{
  const A u[3] = {wall_component(a, e, n, 0), wall_component(a, e, n, 1), wall_component(a, e, n, 2)};
  A rho;
  if (e.kind == OLBGFX_LOCAL_VELOCITY) {
    face_local_velocity(f, e, n, a, u, rho);
  } else if (e.kind == OLBGFX_LOCAL_PRESSURE) {
    int dir, orient;
    flat_normal(a.normal[n], dir, orient);
    rho = a.cell_rho ? a.cell_rho[n] : e.wall_rho;
    A up[3];
    local_pressure(f, dir, orient, rho, u, e.omega, up);
    st_rho = rho;
    st_usqr = up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
    return true;
  } else if (e.kind == OLBGFX_INTERPOLATED_VELOCITY || e.kind == OLBGFX_INTERPOLATED_PRESSURE) {
    // BGK (collision.h:57-80) through BasicDirichletVelocityBoundaryTuple
    // (VelocityBoundaryDensity, FixedVelocityMomentumGeneric) or
    // BasicDirichletPressureBoundaryTuple (FixedDensity,
    // FixedPressureMomentum), momenta/aliases.h:173-198
    int dir, orient;
    flat_normal(a.normal[n], dir, orient);
    A uc[3] = {u[0], u[1], u[2]};
    if (e.kind == OLBGFX_INTERPOLATED_VELOCITY) {
      rho = velocity_bm_rho(f, dir, orient, u[dir]);
    } else {
      rho = a.cell_rho ? a.cell_rho[n] : e.wall_rho;
      pressure_u(f, dir, orient, rho, u, uc);
    }
    const A usqr = uc[0] * uc[0] + uc[1] * uc[1] + uc[2] * uc[2];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      f[i] = f[i] + e.omega * (feq_i(i, rho, uc, usqr) - f[i]);
    }
    st_rho = rho;
    st_usqr = usqr;
    return true;
  } else if (e.kind == OLBGFX_INNER_EDGE_VELOCITY || e.kind == OLBGFX_INNER_CORNER_VELOCITY) {
    const uint32_t code = a.normal[n];
    const int N[3] = {normal_component(code, 0), normal_component(code, 1), normal_component(code, 2)};
    inner_velocity(f, N, u, e.omega, rho);
  } else {
    fixed_velocity_bgk(f, e, u, rho);
  }
  st_rho = rho;
  st_usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  return true;
}

// the two kinds the dominant launch of a split step updates inline
// (kWithFaces), with the arithmetic of apply_velocity_boundary
template <typename A>
__device__ __forceinline__ bool apply_face_velocity(A (&f)[Q], const DynEntry<A>& e, uint32_t n,
                                                    const StepArgs<A>& a, A& st_rho, A& st_usqr)
{
  const A u[3] = {wall_component(a, e, n, 0), wall_component(a, e, n, 1), wall_component(a, e, n, 2)};
  A rho;
  if (e.kind == OLBGFX_LOCAL_VELOCITY) {
    face_local_velocity(f, e, n, a, u, rho);
  } else {
    fixed_velocity_bgk(f, e, u, rho);
  }
  st_rho = rho;
  st_usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  return true;
}

// ---------------------------------------------------------------------------
// ring addressing: ring i is a uniform base pointer and the slot a 32-bit
// byte offset, so loads and stores take the scalar-base + vector-offset form

__device__ __forceinline__ uint32_t ring_slot(uint32_t n, uint32_t shift, uint32_t count)
{
  const uint32_t s = n + shift;
  return s >= count ? s - count : s;
}

// The rings are device memory (hipMalloc).  Where the compiler cannot see
// that by itself (the ensemble kernel reads its members' arguments from
// memory) it issues flat loads and stores with a 64-bit address per lane;
// with 2-byte storage the base is therefore declared a global-memory pointer,
// which gives global loads and stores with a scalar base (the 64-cavity
// ensemble in FP16: 937 -> 995 steps/s by graph replay, 937 -> 1030
// persistent).  With 4-byte storage the same ran 1.5 % slower, so it keeps
// the flat form there; the step kernels get the global form either way.
typedef __attribute__((address_space(1))) char global_char;

template <typename S>
__device__ __forceinline__ S* ring_at(void* pop, uint64_t stride, int i, uint32_t slot)
{
  if constexpr (sizeof(S) == 2) {
    global_char* base = (global_char*)pop + (uint64_t)i * stride * sizeof(S);
    return (S*)(base + slot * (uint32_t)sizeof(S));
  } else {
    char* base = (char*)pop + (uint64_t)i * stride * sizeof(S);
    return (S*)(base + slot * (uint32_t)sizeof(S));
  }
}

// ---------------------------------------------------------------------------
// ghost writes of the periodic overlap

// population P of a boundary cell n goes to every ghost image g = n + sum of
// offsets along the axes where it leaves the core through a periodic face
template <int P, typename S>
__device__ __forceinline__ void ghost_pop(void* pop, uint64_t stride, uint32_t count,
                                          uint32_t shift, int64_t n, const bool (&lo)[3],
                                          const bool (&hi)[3], const int64_t (&off)[3], S v)
{
  constexpr int c[3] = {CX(P), CY(P), CZ(P)};
  bool el[3];
  int64_t o[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    el[ax] = (c[ax] > 0 && hi[ax]) || (c[ax] < 0 && lo[ax]);
    o[ax] = c[ax] > 0 ? -off[ax] : off[ax];
  }
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    if (c[ax] != 0 && el[ax]) {
      *ring_at<S>(pop, stride, P, ring_slot((uint32_t)(n + o[ax]), shift, count)) = v;
    }
  }
  // edges (D3Q19 has at most two non-zero components)
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
#pragma unroll
    for (int bx = ax + 1; bx < 3; ++bx) {
      if (c[ax] != 0 && c[bx] != 0 && el[ax] && el[bx]) {
        *ring_at<S>(pop, stride, P, ring_slot((uint32_t)(n + o[ax] + o[bx]), shift, count)) = v;
      }
    }
  }
}

template <typename S, typename A, int... P>
__device__ __forceinline__ void ghost_all(void* pop, uint64_t stride, uint32_t count,
                                          const uint32_t (&sh)[Q], int64_t n, const bool (&lo)[3],
                                          const bool (&hi)[3], const int64_t (&off)[3],
                                          const A (&f)[Q], std::integer_sequence<int, P...>)
{
  (ghost_pop<P + 1, S>(pop, stride, count, sh[P + 1], n, lo, hi, off,
                       Storage<S>::pack(f[P + 1])),
   ...);
}

// ---------------------------------------------------------------------------
// one cell of the step: load, collide, store, periodic ghost writes

template <typename S, int MODE>
__device__ __forceinline__ void step_cell(const StepArgs<typename Storage<S>::A>& a,
                                          const uint32_t (&sh)[Q], uint32_t lin,
                                          typename Storage<S>::A& st_rho,
                                          typename Storage<S>::A& st_usqr, bool& counted)
{
  using A = typename Storage<S>::A;
  using St = Storage<S>;
  const uint32_t x = fast_div(lin, a.magic_nyz, a.nyz);
  const uint32_t r = lin - x * a.nyz;
  const uint32_t y = fast_div(r, a.magic_nz, (uint32_t)a.nz);
  const uint32_t z = r - y * (uint32_t)a.nz;
  const uint32_t n = ((x + a.pad) * a.py + (y + a.pad)) * a.pz + (z + a.pad);

  // In the masked path the populations are loaded before the dynamics id is
  // known, so their loads overlap the mask and table loads instead of
  // waiting behind them; a NoDynamics core cell then stores nothing but,
  // on a periodic plane, still hands its crossing populations to the ghost
  // layer, as the reference's overlap copy does for every cell.
  int kind;
  int id;
  A f[Q];
  if constexpr (MODE >= kWithFaces) {
    id = (int)a.mask[n];
    kind = a.table[id].kind;
    if (id != a.uniform_id && kind != OLBGFX_LOCAL_VELOCITY && kind != OLBGFX_FIXED_VELOCITY) {
      return;
    }
  } else if constexpr (MODE >= 0) {
    id = a.uniform_id;
    kind = MODE;
    // the dominant launch of a split step: other cells belong to the
    // residual launch, which also does their ghost writes
    if (a.skip_others && (int)a.mask[n] != id) {
      return;
    }
  } else {
    id = (int)a.mask[n];
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    f[i] = St::load(ring_at<S>(a.pop, a.stride, i, ring_slot(n, sh[i], a.count)));
  }
  if constexpr (MODE < 0) {
    kind = a.table[id].kind;
  }
  if (kind != OLBGFX_NO_DYNAMICS) {
    if constexpr (MODE >= kWithFaces) {
      if (id == a.uniform_id) {
        counted = apply_kind<MODE - kWithFaces>(f, a.table[id], n, a, st_rho, st_usqr);
      } else {
        counted = apply_face_velocity(f, a.table[id], n, a, st_rho, st_usqr);
      }
    } else if constexpr (MODE >= 0) {
      counted = apply_kind<MODE>(f, a.table[id], n, a, st_rho, st_usqr);
    } else if constexpr (MODE == kMaskedBoundary) {
      const DynEntry<A> e = a.table[id];
      if (is_velocity_boundary(kind)) {
        counted = apply_velocity_boundary(f, e, n, a, st_rho, st_usqr);
      } else {
        counted = apply_any(f, e, n, a, st_rho, st_usqr);
      }
    } else {
      const DynEntry<A> e = a.table[id];
      counted = apply_any(f, e, n, a, st_rho, st_usqr);
    }

    // Masked kernels (and the dominant launch with the faces inline) store
    // through slots computed again from an opaque copy of n, so the 19 load
    // offsets need not stay live across the collision:
    // 65 VGPRs against 103, up to 8 waves per SIMD instead of 4, +17 % on the
    // FP16 porous channel.  The uniform kernels keep their offsets (no gain
    // measured there).
    uint32_t m = n;
    if constexpr (MODE < 0 || MODE >= kWithFaces) {
      asm("" : "+v"(m));
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      *ring_at<S>(a.pop, a.stride, i, ring_slot(m, sh[i], a.count)) = St::pack(f[i]);
    }
  }

  if (a.periodic) {
    const bool lo[3] = {(a.periodic & 1) && x == 0, (a.periodic & 2) && y == 0,
                        (a.periodic & 4) && z == 0};
    const bool hi[3] = {(a.periodic & 1) && x == (uint32_t)a.nx - 1,
                        (a.periodic & 2) && y == (uint32_t)a.ny - 1,
                        (a.periodic & 4) && z == (uint32_t)a.nz - 1};
    if (lo[0] | lo[1] | lo[2] | hi[0] | hi[1] | hi[2]) {
      const int64_t off[3] = {(int64_t)a.nx * a.py * a.pz, (int64_t)a.ny * a.pz, (int64_t)a.nz};
      ghost_all<S>(a.pop, a.stride, a.count, sh, (int64_t)n, lo, hi, off, f,
                   std::make_integer_sequence<int, Q - 1>{});
    }
  }
}

// ---------------------------------------------------------------------------
// PostStream processors of LocalVelocity edges and corners
// (boundary/boundaryPostProcessors3D.hh:164-294), run after the step on the
// streamed populations: a.shift holds the offsets after the rotation

template <typename S>
__device__ __forceinline__ void load_cell(const StepArgs<typename Storage<S>::A>& a, uint32_t n,
                                          typename Storage<S>::A (&f)[Q])
{
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    f[i] = Storage<S>::load(ring_at<S>(a.pop, a.stride, i, ring_slot(n, a.shift[i], a.count)));
  }
}

// computeRho of cell m through its momenta
template <typename S>
__device__ typename Storage<S>::A momenta_rho(const StepArgs<typename Storage<S>::A>& a, uint32_t m)
{
  using A = typename Storage<S>::A;
  A f[Q];
  load_cell<S>(a, m, f);
  const DynEntry<A>& e = a.table[a.mask[m]];
  switch (e.kind) {
  case OLBGFX_NO_DYNAMICS:
    return A(1);  // OneDensity
  case OLBGFX_BOUNCE_BACK:
  case OLBGFX_BOUNCE_BACK_VELOCITY:
  case OLBGFX_LOCAL_PRESSURE:
  case OLBGFX_INTERPOLATED_PRESSURE:
    return a.cell_rho ? a.cell_rho[m] : e.wall_rho;  // FixedDensity
  case OLBGFX_LOCAL_VELOCITY:
  case OLBGFX_INTERPOLATED_VELOCITY: {
    const uint32_t code = a.normal[m];
    for (int ax = 0; ax < 3; ++ax) {
      const int o = normal_component(code, ax);
      if (o != 0) {
        return velocity_bm_rho(f, ax, o, wall_component(a, e, m, ax));
      }
    }
    break;
  }
  case OLBGFX_INNER_EDGE_VELOCITY:
  case OLBGFX_INNER_CORNER_VELOCITY: {
    const uint32_t code = a.normal[m];
    const int N[3] = {normal_component(code, 0), normal_component(code, 1), normal_component(code, 2)};
    const A u[3] = {wall_component(a, e, m, 0), wall_component(a, e, m, 1), wall_component(a, e, m, 2)};
    return inner_density(f, N, u);
  }
  default:
    break;
  }
  A s = A(0);
#pragma unroll
  for (int i = 0; i < Q; ++i) s += f[i];
  return s + A(1);
}

// computeU of cell m through its momenta
template <typename S>
__device__ void momenta_u(const StepArgs<typename Storage<S>::A>& a, uint32_t m, typename Storage<S>::A (&u)[3])
{
  using A = typename Storage<S>::A;
  const DynEntry<A>& e = a.table[a.mask[m]];
  switch (e.kind) {
  case OLBGFX_NO_DYNAMICS:
  case OLBGFX_BOUNCE_BACK:
    u[0] = u[1] = u[2] = A(0);  // ZeroMomentum
    return;
  case OLBGFX_LOCAL_VELOCITY:
  case OLBGFX_FIXED_VELOCITY:
  case OLBGFX_INNER_EDGE_VELOCITY:
  case OLBGFX_INNER_CORNER_VELOCITY:
  case OLBGFX_BOUNCE_BACK_VELOCITY:
  case OLBGFX_INTERPOLATED_VELOCITY:
    for (int ax = 0; ax < 3; ++ax) u[ax] = wall_component(a, e, m, ax);
    return;
  case OLBGFX_LOCAL_PRESSURE:
  case OLBGFX_INTERPOLATED_PRESSURE: {
    A f[Q];
    load_cell<S>(a, m, f);
    int dir, orient;
    flat_normal(a.normal[m], dir, orient);
    const A uw[3] = {wall_component(a, e, m, 0), wall_component(a, e, m, 1), wall_component(a, e, m, 2)};
    pressure_u(f, dir, orient, a.cell_rho ? a.cell_rho[m] : e.wall_rho, uw, u);
    return;
  }
  default:
    break;
  }
  A f[Q];
  load_cell<S>(a, m, f);
  A drho, jx, jy, jz;
  moments(f, drho, jx, jy, jz);
  const A rho = drho + A(1);
  u[0] = jx / rho;
  u[1] = jy / rho;
  u[2] = jz / rho;
  if (e.kind == OLBGFX_FORCED_BGK || e.kind == OLBGFX_FORCED_TRT) {
    for (int ax = 0; ax < 3; ++ax) {
      u[ax] += (a.cell_force ? a.cell_force[3 * (size_t)m + ax] : e.force[ax]) * A(0.5);
    }
  }
}

// DirectedGradients3D along the boundary normal (finiteDifference3D.h:132-155):
// -orient * (-3 u0 + 4 u1 - u2) / 2 with u1, u2 one and two cells inwards
template <typename S>
__device__ void one_sided(const StepArgs<typename Storage<S>::A>& a, uint32_t n, int64_t step, int orient,
                          typename Storage<S>::A (&d)[3])
{
  using A = typename Storage<S>::A;
  A u0[3], u1[3], u2[3];
  momenta_u<S>(a, n, u0);
  momenta_u<S>(a, (uint32_t)((int64_t)n - orient * step), u1);
  momenta_u<S>(a, (uint32_t)((int64_t)n - 2 * orient * step), u2);
  for (int b = 0; b < 3; ++b) {
    d[b] = A(-orient) * ((-A(3) * u0[b] + A(4) * u1[b] - A(1) * u2[b]) / A(2));
  }
}

// The cells whose momenta the edge or corner processor of cell n reads,
// slot by slot.  Edge, plane p (the zero component of the normal), d1 and d2
// the other axes with normal components n1 and n2: 0: n - n1 e_d1,
// 1: n - n2 e_d2, 2: n - 2 n1 e_d1, 3: n - 2 n2 e_d2, 4: n + e_p, 5: n - e_p,
// 6: n.  Corner, normal N: 0-2: n - N_k e_k, 3-5: n - 2 N_k e_k, 6: n.
constexpr int kPostSlots = 7;

__device__ __forceinline__ uint32_t post_neighbor(uint32_t n, const int (&N)[3], const int64_t (&step)[3], int slot)
{
  if (slot == 6) {
    return n;
  }
  if (N[0] == 0 || N[1] == 0 || N[2] == 0) {
    const int p = N[0] == 0 ? 0 : (N[1] == 0 ? 1 : 2);
    const int d1 = (p + 1) % 3, d2 = (p + 2) % 3;
    switch (slot) {
    case 0: return (uint32_t)((int64_t)n - N[d1] * step[d1]);
    case 1: return (uint32_t)((int64_t)n - N[d2] * step[d2]);
    case 2: return (uint32_t)((int64_t)n - 2 * N[d1] * step[d1]);
    case 3: return (uint32_t)((int64_t)n - 2 * N[d2] * step[d2]);
    case 4: return (uint32_t)((int64_t)n + step[p]);
    default: return (uint32_t)((int64_t)n - step[p]);
    }
  }
  const int k = slot % 3;
  const int64_t m = slot < 3 ? 1 : 2;
  return (uint32_t)((int64_t)n - m * N[k] * step[k]);
}

// One launch applies the processors of one list (edges, or corners: a
// corner reads edge cells the edge processors rewrite, so the corners run in
// a launch of their own after the edges).  A block takes 32 cells; lane
// `slot` of a cell evaluates the momenta of one of its 7 neighbour cells
// into LDS, so the ~13 neighbour evaluations of a cell run side by side
// instead of one after another; lane 0 then combines them and writes the
// cell's 19 populations, with the arithmetic of
// OuterVelocityEdgeProcessor3D / OuterVelocityCornerProcessor3D
// (boundary/boundaryPostProcessors3D.hh:164-294).
constexpr uint32_t kPostCellsPerBlock = 32;

template <typename S>
__global__ void __launch_bounds__(256) k_velocity_post(const StepArgs<typename Storage<S>::A> a,
                                                       const uint32_t* cells, uint32_t ncells)
{
  using A = typename Storage<S>::A;
  __shared__ A mom[kPostCellsPerBlock][8][4];
  const uint32_t local = threadIdx.x >> 3;
  const int slot = (int)(threadIdx.x & 7u);
  const uint32_t t = blockIdx.x * kPostCellsPerBlock + local;
  const bool active = t < ncells;
  const uint32_t n = active ? cells[t] : 0u;
  const uint32_t code = active ? a.normal[n] : 13u;
  const int N[3] = {normal_component(code, 0), normal_component(code, 1), normal_component(code, 2)};
  const int64_t step[3] = {(int64_t)a.py * a.pz, (int64_t)a.pz, 1};
  if (active && slot < kPostSlots) {
    const uint32_t m = post_neighbor(n, N, step, slot);
    A u[3];
    momenta_u<S>(a, m, u);
    mom[local][slot][0] = momenta_rho<S>(a, m);
    mom[local][slot][1] = u[0];
    mom[local][slot][2] = u[1];
    mom[local][slot][3] = u[2];
  }
  __syncthreads();
  if (!active || slot != 0) {
    return;
  }
  const DynEntry<A>& e = a.table[a.mask[n]];
  // DirectedGradients3D along a normal (finiteDifference3D.h:132-155):
  // -orient * (-3 u0 + 4 u1 - u2) / 2 with u1, u2 one and two cells inwards
  auto one_sided = [&](int s1, int s2, int orient, A (&d)[3]) {
    for (int b = 0; b < 3; ++b) {
      d[b] = A(-orient) * ((-A(3) * mom[local][6][1 + b] + A(4) * mom[local][s1][1 + b] -
                            A(1) * mom[local][s2][1 + b]) / A(2));
    }
  };
  A rho;
  A du[3][3];
  if (N[0] == 0 || N[1] == 0 || N[2] == 0) {
    // OuterVelocityEdgeProcessor3D<plane, normal1, normal2>
    const int p = N[0] == 0 ? 0 : (N[1] == 0 ? 1 : 2);
    const int d1 = (p + 1) % 3, d2 = (p + 2) % 3;
    const A rho10 = mom[local][0][0], rho01 = mom[local][1][0];
    const A rho20 = mom[local][2][0], rho02 = mom[local][3][0];
    rho = A(2) / A(3) * (rho01 + rho10) - A(1) / A(6) * (rho02 + rho20);
    for (int b = 0; b < 3; ++b) du[p][b] = (mom[local][4][1 + b] - mom[local][5][1 + b]) / A(2);
    one_sided(0, 2, N[d1], du[d1]);
    one_sided(1, 3, N[d2], du[d2]);
  } else {
    // OuterVelocityCornerProcessor3D<xNormal, yNormal, zNormal>
    const A rho100 = mom[local][0][0], rho010 = mom[local][1][0], rho001 = mom[local][2][0];
    const A rho200 = mom[local][3][0], rho020 = mom[local][4][0], rho002 = mom[local][5][0];
    rho = A(4) / A(9) * (rho001 + rho010 + rho100) - A(1) / A(9) * (rho002 + rho020 + rho200);
    for (int ax = 0; ax < 3; ++ax) one_sided(ax, 3 + ax, N[ax], du[ax]);
  }
  // regularised populations (boundaryPostProcessors3D.hh:196-215)
  const A s_to_pi = -rho / A(3) / e.omega;
  const A pi[6] = {A(2) * du[0][0] * s_to_pi, (du[0][1] + du[1][0]) * s_to_pi, (du[0][2] + du[2][0]) * s_to_pi,
                   A(2) * du[1][1] * s_to_pi, (du[1][2] + du[2][1]) * s_to_pi, A(2) * du[2][2] * s_to_pi};
  const A u[3] = {wall_component(a, e, n, 0), wall_component(a, e, n, 1), wall_component(a, e, n, 2)};
  const A usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    *ring_at<S>(a.pop, a.stride, i, ring_slot(n, a.shift[i], a.count)) =
        Storage<S>::pack(feq_i(i, rho, u, usqr) + pi_to_fneq(i, pi));
  }
}

// PlaneFdBoundaryProcessor3D<direction, orientation>
// (boundary/boundaryPostProcessors3D.hh:39-92) on the flat cells of the
// interpolated boundaries: density and velocity through the cell's momenta,
// the velocity gradient one-sided along the normal and central along the
// wall (utilities/finiteDifference3D.h:75-230), then every population set
// to the regularised value feq(rho, u) + fneq(Pi(grad u)).  The new
// populations go to `out` (19 per listed cell) and k_plane_fd_store writes
// them back, so no processor reads a value another one has rewritten.
template <typename S>
__global__ void __launch_bounds__(256) k_plane_fd(const StepArgs<typename Storage<S>::A> a, const uint32_t* cells,
                                                  uint32_t ncells, typename Storage<S>::A* out)
{
  using A = typename Storage<S>::A;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells) {
    return;
  }
  const uint32_t n = cells[t];
  int dir, orient;
  flat_normal(a.normal[n], dir, orient);
  const int64_t step[3] = {(int64_t)a.py * a.pz, (int64_t)a.pz, 1};
  const DynEntry<A>& e = a.table[a.mask[n]];
  const A rho = momenta_rho<S>(a, n);
  A u[3];
  momenta_u<S>(a, n, u);
  A du[3][3];
  for (int d = 0; d < 3; ++d) {
    if (d == dir) {
      one_sided<S>(a, n, step[d], orient, du[d]);
    } else {
      A up[3], um[3];
      momenta_u<S>(a, (uint32_t)((int64_t)n + step[d]), up);
      momenta_u<S>(a, (uint32_t)((int64_t)n - step[d]), um);
      for (int b = 0; b < 3; ++b) du[d][b] = (up[b] - um[b]) / A(2);
    }
  }
  const A s_to_pi = -rho / A(3) / e.omega;
  const A pi[6] = {A(2) * du[0][0] * s_to_pi, (du[0][1] + du[1][0]) * s_to_pi, (du[0][2] + du[2][0]) * s_to_pi,
                   A(2) * du[1][1] * s_to_pi, (du[1][2] + du[2][1]) * s_to_pi, A(2) * du[2][2] * s_to_pi};
  const A usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    out[(size_t)i * ncells + t] = feq_i(i, rho, u, usqr) + pi_to_fneq(i, pi);
  }
}

template <typename S>
__global__ void __launch_bounds__(256) k_plane_fd_store(const StepArgs<typename Storage<S>::A> a,
                                                        const uint32_t* cells, uint32_t ncells,
                                                        const typename Storage<S>::A* out)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells) {
    return;
  }
  const uint32_t n = cells[t];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    *ring_at<S>(a.pop, a.stride, i, ring_slot(n, a.shift[i], a.count)) = Storage<S>::pack(out[(size_t)i * ncells + t]);
  }
}

// ---------------------------------------------------------------------------
// Cell::computeRhoU of every padded cell through the momenta of its dynamics
// (dynamics/momenta/elements.h, through momenta_rho and momenta_u):
// NoDynamics OneDensity and ZeroMomentum (dynamics.h:52-62); BounceBack
// FixedDensity (the cell's RHO, 1 unless defined) and ZeroMomentum
// (dynamics.h:557-568); BounceBackVelocity FixedDensity and
// FixedVelocityMomentum; the LocalVelocity kinds their boundary densities and
// the wall velocity; LocalPressure its fixed density and pressure velocity;
// forced dynamics add F/2; the rest are bulk.
template <typename S, typename H>
__global__ void __launch_bounds__(256) k_cell_rho_u(const StepArgs<typename Storage<S>::A> a, H* rho_out,
                                                    H* u_out)
{
  using A = typename Storage<S>::A;
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.count) {
    return;
  }
  A rho, u[3];
  rho = momenta_rho<S>(a, n);
  momenta_u<S>(a, n, u);
  rho_out[n] = (H)rho;
  for (int ax = 0; ax < 3; ++ax) u_out[3 * (size_t)n + ax] = (H)u[ax];
}

// ---------------------------------------------------------------------------
// PostStream overlap exchange for the post-processors (boundary/setBoundary.h:
// 110-118, addPoints2CommBC): before they run, the periodic padding cells
// they read receive their core image's streamed populations.  pairs holds
// (padding cell, image) ids; padding cells are never images, so the copies
// do not race.
template <typename S>
__global__ void __launch_bounds__(256) k_refresh(const StepArgs<typename Storage<S>::A> a, const uint32_t* pairs,
                                                 uint32_t npairs)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs) {
    return;
  }
  const uint32_t g = pairs[2 * t], m = pairs[2 * t + 1];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    *ring_at<S>(a.pop, a.stride, i, ring_slot(g, a.shift[i], a.count)) =
        *ring_at<S>(a.pop, a.stride, i, ring_slot(m, a.shift[i], a.count));
  }
}

// ---------------------------------------------------------------------------
// BouzidiPostProcessor (boundary/setBouzidiBoundary.h:48-80), PostStream
// priority -1: one thread per listed cell x_b, run on the streamed
// populations.  For i = 1..18 in order, q = BOUZIDI_DISTANCE[i]:
//   q > 0:  x_b[opp i] = (q <= 1/2) (2q x_s[i] + (1 - 2q) x_b[i])
//                      + (q >  1/2) (x_s[i] / 2q + (2q - 1)/2q x_f[opp i])
//           with x_s = x_b + c_i (solid side) and x_f = x_b - c_i;
//   q == 0: x_b[opp i] = x_b[i] (full-way bounce-back).
// The cell's own populations stay in registers, so a later direction reads
// what an earlier one wrote, as in the reference's loop.  The neighbours'
// values read here (x_s[i], x_f[opp i]) streamed from x_b itself, so the
// ring slots of padding cells hold them without an overlap exchange.  The
// host rejects lists in which one cell's update reads a population another
// cell's update writes, so the cells can be updated concurrently.
template <typename S>
__global__ void __launch_bounds__(256) k_bouzidi(const StepArgs<typename Storage<S>::A> a, const uint32_t* cells,
                                                 const typename Storage<S>::A* q, const typename Storage<S>::A* vel,
                                                 uint32_t ncells)
{
  using A = typename Storage<S>::A;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells) {
    return;
  }
  const uint32_t n = cells[t];
  const int64_t step[3] = {(int64_t)a.py * a.pz, (int64_t)a.pz, 1};
  A f[Q];
  load_cell<S>(a, n, f);
  uint32_t written = 0;
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    constexpr int half = 9;
    const int o = i <= half ? i + half : i - half;
    const A qi = q[(size_t)(i - 1) * ncells + t];
    // BouzidiVelocityPostProcessor (setBouzidiBoundary.h:129-164): the wall
    // velocity term BOUZIDI_VELOCITY_i t_i / cs^2
    const A vt = vel ? vel[(size_t)(i - 1) * ncells + t] * (A(1) / A(TDEN(i))) * A(3) : A(0);
    if (qi > A(0)) {
      const int64_t d = CX(i) * step[0] + CY(i) * step[1] + CZ(i) * step[2];
      const A xs = Storage<S>::load(ring_at<S>(a.pop, a.stride, i, ring_slot((uint32_t)((int64_t)n + d),
                                                                              a.shift[i], a.count)));
      if (qi <= A(0.5)) {
        f[o] = A(2) * qi * xs + (A(1) - A(2) * qi) * f[i] - A(2) * vt;
      } else {
        const A xf = Storage<S>::load(ring_at<S>(a.pop, a.stride, o, ring_slot((uint32_t)((int64_t)n - d),
                                                                                a.shift[o], a.count)));
        f[o] = A(0.5) / qi * xs + A(0.5) * (A(2) * qi - A(1)) / qi * xf - A(1) / qi * vt;
      }
      written |= 1u << o;
    } else if (qi == A(0)) {
      f[o] = f[i] - A(2) * vt;
      written |= 1u << o;
    }
  }
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    if (written & (1u << i)) {
      *ring_at<S>(a.pop, a.stride, i, ring_slot(n, a.shift[i], a.count)) = Storage<S>::pack(f[i]);
    }
  }
}

// ---------------------------------------------------------------------------
// the step kernel of one lattice
//
// MODE = -1: every cell reads its dynamics id from the mask;
// MODE = -2: the same with the LocalVelocity kinds compiled in;
// MODE = k: every core cell has the dynamics of table slot uniform_id, whose
//           kind is k, so the mask is not read and only that path is compiled.

// logical block of hardware workgroup b (see the step kernel)
__device__ __forceinline__ uint32_t logical_block(uint32_t b, uint32_t blocks_per_xcd, int32_t run_log2)
{
  if (run_log2 < 0) {
    return (b & 7u) * blocks_per_xcd + (b >> 3);
  }
  const uint32_t k = (uint32_t)run_log2, g = b >> 3;
  return ((g >> k) << (k + 3)) + ((b & 7u) << k) + (g & ((1u << k) - 1u));
}

template <typename S, int MODE, bool STATS>
__global__ void __launch_bounds__(kMaxBlock) step_kernel(const StepArgs<typename Storage<S>::A> a)
{
  using A = typename Storage<S>::A;

  // XCD-aware order: the hardware deals consecutive workgroups round-robin to
  // the 8 XCDs.  By default each XCD sweeps one contiguous part of the block,
  // so the cache lines shared by neighbouring segments stay in one L2; with
  // xcd_run_log2 = k the XCDs take turns on runs of 2^k consecutive blocks,
  // which keeps the cells in flight on the whole chip close together.  Both
  // maps are bijections of [0, 8 * blocks_per_xcd).
  const uint32_t lb = logical_block(blockIdx.x, a.blocks_per_xcd, a.xcd_run_log2);
  const uint32_t lin = lb * blockDim.x + threadIdx.x;
  const bool in_grid = lb < a.nblocks;
  const bool active = in_grid && lin < a.ncore;

  uint32_t sh[Q];
  if (a.shift_dev) {
#pragma unroll
    for (int i = 0; i < Q; ++i) sh[i] = a.shift_dev[i];
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) sh[i] = a.shift[i];
  }

  A st_rho = A(0), st_usqr = A(0);
  bool counted = false;
  if (active) {
    step_cell<S, MODE>(a, sh, lin, st_rho, st_usqr, counted);
  }

  if constexpr (STATS) {
    const double c = counted ? 1.0 : 0.0;
    const double r = counted ? (double)st_rho : 0.0;
    const double u2 = counted ? (double)st_usqr : 0.0;
    block_reduce4(c, r, u2, u2, a.stats + 4 * (size_t)lb, in_grid);
  }
}

// The residual launch of a split step: the core cells (linear core indices,
// ascending) whose dynamics is not the dominant one, with the masked code
// (MODE -1, or -2 when boundary kinds are present).  Together with the
// dominant launch it updates every core cell once, as the reference's CPU
// applyDominant does with its mask (cpu/sisd/operator.h:143-228).
template <typename S, int MODE>
__global__ void __launch_bounds__(256) residual_kernel(const StepArgs<typename Storage<S>::A> a,
                                                       const uint32_t* cells, uint32_t n)
{
  using A = typename Storage<S>::A;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) {
    return;
  }
  uint32_t sh[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) sh[i] = a.shift[i];
  A st_rho = A(0), st_usqr = A(0);
  bool counted = false;
  step_cell<S, MODE>(a, sh, cells[k], st_rho, st_usqr, counted);
}

// ---------------------------------------------------------------------------
// the ensemble kernel: one persistent launch advances one group of member
// lattices by one step.  The work is a list of tiles (member, first core
// cell) of blockDim.x cells each; logical block lb owns a contiguous range of
// tiles_per_block tiles of the group, with the XCD-aware order of the
// single-lattice kernel.  Member descriptors and ring offsets live in device
// memory, so the launch has fixed arguments and can be replayed from a
// captured graph.  The last block to finish (ticket counter) advances the
// ring offsets of the group's members, s_i -= d_i (mod ncells), for the next
// launch; the counter is left at zero for it.

struct Tile {
  uint32_t member;
  uint32_t first;
};

struct GroupArgs {
  uint32_t ticket;           // 1: the last block advances the offsets (else a separate kernel)
  uint32_t tile_begin;       // first tile of the group
  uint32_t ntiles;           // tiles of the group
  uint32_t member_begin;     // members [member_begin, member_end) form the group
  uint32_t member_end;
  uint32_t tiles_per_block;
  uint32_t nblocks;          // logical blocks that carry tiles
  uint32_t blocks_per_xcd;   // rounded up to whole runs
  int32_t xcd_run_log2;      // block order, as in the step kernel
  uint32_t* counter;         // ticket counter of the group (0 between launches)
};

template <typename S>
__global__ void __launch_bounds__(kMaxBlock)
    ensemble_kernel(const StepArgs<typename Storage<S>::A>* members, const Tile* tiles,
                    const GroupArgs g, uint32_t* const* shifts, const uint32_t* dmod,
                    const uint32_t* counts)
{
  using A = typename Storage<S>::A;
  const uint32_t lb = logical_block(blockIdx.x, g.blocks_per_xcd, g.xcd_run_log2);
  if (lb < g.nblocks) {
    const uint32_t end = g.tile_begin + g.ntiles;
    const uint32_t t0 = g.tile_begin + lb * g.tiles_per_block;
    const uint32_t t1 = min(t0 + g.tiles_per_block, end);
    for (uint32_t t = t0; t < t1; ++t) {
      const Tile tile = tiles[t];
      const StepArgs<A>& a = members[tile.member];
      uint32_t sh[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) sh[i] = a.shift_dev[i];
      const uint32_t lin = tile.first + threadIdx.x;
      A st_rho, st_usqr;
      bool counted = false;
      if (lin < a.ncore) {
        step_cell<S, -1>(a, sh, lin, st_rho, st_usqr, counted);
      }
    }
  }
  if (!g.ticket) {
    return;
  }
  // every block has read the offsets it needs before it takes its ticket, so
  // the last one may advance them; the kernel boundary publishes the update
  __shared__ bool last;
  __syncthreads();
  if (threadIdx.x == 0) {
    last = atomicAdd(g.counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (last) {
    for (uint32_t k = g.member_begin * Q + threadIdx.x; k < g.member_end * Q; k += blockDim.x) {
      const uint32_t l = k / Q, i = k % Q;
      const uint64_t s = (uint64_t)shifts[l][i] + dmod[k];
      shifts[l][i] = (uint32_t)(s >= counts[l] ? s - counts[l] : s);
    }
    if (threadIdx.x == 0) {
      atomicExch(g.counter, 0u);
    }
  }
}

// advance the ring offsets of members [begin, end): s_i -= d_i (mod ncells)
__global__ void advance_shifts(uint32_t* const* shifts, const uint32_t* dmod, const uint32_t* counts,
                               uint32_t begin, uint32_t end)
{
  const uint32_t k = begin * Q + blockIdx.x * blockDim.x + threadIdx.x;
  if (k < end * Q) {
    const uint32_t l = k / Q, i = k % Q;
    const uint64_t s = (uint64_t)shifts[l][i] + dmod[k];
    shifts[l][i] = (uint32_t)(s >= counts[l] ? s - counts[l] : s);
  }
}

// sums the per-block statistics in a fixed order (deterministic)
__global__ void __launch_bounds__(256) stats_finalize(const double* partial, uint32_t nblocks,
                                                      double* out)
{
  double c = 0, r = 0, e = 0, m = 0;
  for (uint32_t k = threadIdx.x; k < nblocks; k += blockDim.x) {
    c += partial[4 * (size_t)k + 0];
    r += partial[4 * (size_t)k + 1];
    e += partial[4 * (size_t)k + 2];
    m = fmax(m, partial[4 * (size_t)k + 3]);
  }
  block_reduce4(c, r, e, m, out, true);
}

// ---------------------------------------------------------------------------
// set-up and evaluation kernels (one thread per padded cell)

template <typename S, typename H>
__global__ void k_scatter_pop(S* ring, uint32_t count, uint32_t shift, const H* src)
{
  using A = typename Storage<S>::A;
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < count) {
    uint32_t s = n + shift;
    s = s >= count ? s - count : s;
    ring[s] = Storage<S>::pack((A)src[n]);
  }
}

template <typename S, typename H>
__global__ void k_gather_pop(const S* ring, uint32_t count, uint32_t shift, H* dst)
{
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < count) {
    uint32_t s = n + shift;
    s = s >= count ? s - count : s;
    dst[n] = (H)Storage<S>::unpack(ring[s]);
  }
}

// iniEquilibrium: f_bar_i = feq_i(rho, u) in the reference's own form
// rho t_i (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2) - t_i (lbm.h:51-64)
template <typename S, typename H>
__global__ void k_ini_equilibrium(S* pop, uint64_t stride, uint32_t count, const uint32_t* shift,
                                  const uint8_t* sel, const H* rho_in, const H* u_in)
{
  using A = typename Storage<S>::A;
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= count || (sel && !sel[n])) {
    return;
  }
  const A rho = (A)rho_in[n];
  const A u[3] = {(A)u_in[3 * (size_t)n], (A)u_in[3 * (size_t)n + 1], (A)u_in[3 * (size_t)n + 2]};
  const A usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  constexpr int cx[Q] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
  constexpr int cy[Q] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
  constexpr int cz[Q] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const A t = A(1) / A(TDEN(i));
    const A cu = A(cx[i]) * u[0] + A(cy[i]) * u[1] + A(cz[i]) * u[2];
    const A feq = rho * t * (A(1) + A(3) * cu + A(4.5) * cu * cu - A(1.5) * usqr) - t;
    uint32_t s = n + shift[i];
    s = s >= count ? s - count : s;
    pop[(uint64_t)i * stride + s] = Storage<S>::pack(feq);
  }
}

// FP64 moments of every colliding core cell, summed from the stored values
template <typename S>
__global__ void __launch_bounds__(256) k_moments(const S* pop, uint64_t stride, uint32_t count,
                                                 const uint32_t* shift, const uint8_t* mask,
                                                 const int32_t* kinds, int32_t nx, int32_t ny,
                                                 int32_t nz, int32_t pad, uint32_t py, uint32_t pz,
                                                 double* partial)
{
  const uint64_t ncore = (uint64_t)nx * ny * nz;
  const uint64_t lin = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double c = 0, r = 0, jx = 0, jy = 0, jz = 0, u2 = 0, m = 0;
  if (lin < ncore) {
    const uint32_t z = (uint32_t)(lin % nz);
    const uint32_t y = (uint32_t)((lin / nz) % ny);
    const uint32_t x = (uint32_t)(lin / ((uint64_t)nz * ny));
    const uint32_t n = ((x + pad) * py + (y + pad)) * pz + (z + pad);
    if (kinds[mask[n]] != OLBGFX_NO_DYNAMICS) {
      constexpr int cx[Q] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
      constexpr int cy[Q] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
      constexpr int cz[Q] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        uint32_t s = n + shift[i];
        s = s >= count ? s - count : s;
        const double v = (double)Storage<S>::unpack(pop[(uint64_t)i * stride + s]);
        r += v;
        jx += cx[i] * v;
        jy += cy[i] * v;
        jz += cz[i] * v;
      }
      c = 1;
      const double rho = r + 1.0;
      u2 = (jx * jx + jy * jy + jz * jz) / (rho * rho);
      m = u2;
    }
  }
  __shared__ double part[4][8];
  double vals[7] = {c, r, jx, jy, jz, u2, m};
#pragma unroll
  for (int k = 0; k < 6; ++k) vals[k] = wave_sum(vals[k]);
  vals[6] = wave_max(vals[6]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    for (int k = 0; k < 7; ++k) part[wave][k] = vals[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double o[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      for (int k = 0; k < 6; ++k) o[k] += part[w][k];
      o[6] = fmax(o[6], part[w][6]);
    }
    for (int k = 0; k < 7; ++k) partial[8 * (size_t)blockIdx.x + k] = o[k];
    partial[8 * (size_t)blockIdx.x + 7] = 0;
  }
}

// ---------------------------------------------------------------------------
// fixed scratch of olbgfx_reduce_moments_noalloc: one 8-double partial per
// 256-cell block of k_moments, for lattices of up to 65536 blocks (2^24 core
// cells); 4 MiB on every device the library runs on, held until the process
// ends.  Host code never passes `scratch` itself to a kernel (in host code
// the name is a stand-in, not a device address): it passes the address
// hipGetSymbolAddress returns for the current device.
constexpr uint64_t kNoAllocBlocks = 65536;

// ---------------------------------------------------------------------------
// host side

thread_local std::string g_error;

olbgfx_status fail(olbgfx_status s, const std::string& msg)
{
  g_error = msg;
  return s;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                              \
      return fail(OLBGFX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    }                                                                                    \
  } while (0)

uint64_t magic_for(uint32_t d)
{
  if (d <= 1) {
    return 0;
  }
  // ceil(2^64 / d) = floor((2^64 - 1) / d) + 1 for d not a power of two;
  // for powers of two the same expression is also correct.
  return UINT64_MAX / d + 1;
}

}  // namespace

// the array itself (declared here, at file scope with internal linkage)
static __device__ __align__(256) double scratch[8 * kNoAllocBlocks];

namespace {

// per-device address of `scratch`, looked up on the first no-alloc call on
// that device with the device current; the host copy of the partials
constexpr int kNoAllocMaxDevices = 64;
double* g_noalloc_dev[kNoAllocMaxDevices] = {};
double g_noalloc_host[8 * kNoAllocBlocks];
std::mutex g_noalloc_mutex;

// caller holds g_noalloc_mutex
olbgfx_status noalloc_scratch(double** ptr)
{
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    return fail(OLBGFX_ERR_INVALID, "device ordinal beyond the no-alloc scratch table");
  }
  if (!g_noalloc_dev[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_noalloc_dev[dev] = (double*)p;
  }
  *ptr = g_noalloc_dev[dev];
  return OLBGFX_OK;
}

}  // namespace

struct olbgfx_lattice {
  olbgfx_lattice_desc desc{};
  int elem = 4;               // bytes per stored population
  uint32_t count = 0;         // padded cells
  uint64_t stride = 0;        // ring stride in elements
  uint32_t py = 0, pz = 0;
  int64_t d[Q] = {};
  int64_t dmod[Q] = {};       // (-d_i) mod count
  uint32_t shift[Q] = {};     // host mirror of the offsets
  void* pop = nullptr;
  uint8_t* mask = nullptr;
  uint32_t* shift_dev = nullptr;
  int32_t* kinds_dev = nullptr;
  void* cell_u = nullptr;
  void* cell_rho = nullptr;
  void* cell_force = nullptr;
  double* stats_partial = nullptr;
  double* stats_out = nullptr;
  uint32_t stats_capacity = 0;
  bool stats_on = false;
  bool stats_valid = false;
  int32_t uniform_id = 0;     // all cells start with id 0 = NoDynamics
  uint32_t block_threads = 256;
  uint32_t lds_reserve = 0;    // dynamic LDS per block that limits occupancy
  int32_t xcd_run_log2 = -1;   // block order of the step kernel (olbgfx_set_block_order)
  // LocalVelocity boundaries: discrete normals, and the cells whose
  // post-processors run after each step (edges first, then corners)
  std::vector<int8_t> host_normals;  // 3 per cell, empty = none
  uint8_t* normal_dev = nullptr;     // code 9(nx+1)+3(ny+1)+(nz+1) per cell
  uint32_t* post_dev = nullptr;
  uint32_t post_edges = 0, post_corners = 0;
  uint32_t* fd_dev = nullptr;        // flat interpolated cells (PlaneFdBoundaryProcessor3D)
  void* fd_out = nullptr;            // their new populations, 19 per cell in the arithmetic type
  uint32_t fd_n = 0;
  uint32_t* refresh_dev = nullptr;   // (periodic padding cell, image) pairs the edge/corner processors read
  uint32_t refresh_n = 0;
  bool boundary_kinds = false;       // some core cell has a LocalVelocity kind
  // Bouzidi boundaries: the cells that run the post-processor and their
  // distances, q[(i - 1) * bz_n + k] for direction i of cell k
  uint32_t* bz_cells = nullptr;
  void* bz_q = nullptr;
  uint32_t bz_n = 0;
  // BouzidiVelocityPostProcessor cells: distances as above and the
  // BOUZIDI_VELOCITY coefficients in the same layout; applied after the
  // plain list (same priority, the reference's operator-type order)
  uint32_t* bzv_cells = nullptr;
  void* bzv_q = nullptr;
  void* bzv_v = nullptr;
  uint32_t bzv_n = 0;
  uint64_t boundary_version = ~0ull;
  olbgfx_dynamics_desc table[OLBGFX_MAX_DYNAMICS];
  std::vector<uint8_t> host_mask;
  uint64_t version = 0;        // bumped by every change of dynamics, mask or fields
  bool shift_dev_valid = true; // device copy of the offsets matches the host mirror
  // at most inflight_limit steps queued ahead of the host (0: no limit): an
  // event per step in a ring, the host waits for the one of step t - limit
  uint32_t inflight_limit = 0;
  std::vector<hipEvent_t> inflight_events;
  uint64_t inflight_count = 0;
  // split step: the dominant dynamics runs its own kernel over all core
  // cells (skipping the others), the residual cells a masked kernel over a list
  int32_t split_id = -1;             // dominant table slot, -1: no split
  bool split_faces = false;          // the dominant launch updates the velocity faces inline
  uint32_t* residual_dev = nullptr;  // linear core indices of the other cells
  uint32_t residual_n = 0;
  uint64_t split_version = ~0ull;
  bool split_enabled = true;         // olbgfx_set_split
};

namespace {

bool is_f64(const olbgfx_lattice* L) { return L->desc.storage == OLBGFX_STORE_F64; }

template <typename A>
void fill_table(const olbgfx_lattice* L, DynEntry<A>* t)
{
  for (int k = 0; k < OLBGFX_MAX_DYNAMICS; ++k) {
    const olbgfx_dynamics_desc& s = L->table[k];
    t[k].kind = s.kind;
    t[k].omega = (A)s.omega;
    // collision.h:408: omega2 = 1 / (magic / (1/omega - 0.5) + 0.5), formed in
    // double from the value-type omega and magic
    const double om = (double)(A)s.omega;
    const double mg = (double)(A)s.magic;
    const double om2 = (om != 0.0) ? 1.0 / (mg / ((double)((A)1 / (A)om) - 0.5) + 0.5) : 0.0;
    t[k].omega2 = (A)om2;
    for (int ax = 0; ax < 3; ++ax) {
      t[k].force[ax] = (A)s.force[ax];
      t[k].wall_u[ax] = (A)s.wall_velocity[ax];
    }
    t[k].wall_rho = (A)s.wall_rho;
    // collisionLES.h:71-73: C^2 * invCs2 * invCs2 * 2 in the value type, times
    // sqrt(2) in double
    const A c = (A)s.smagorinsky;
    t[k].smago_pre = (A)((double)(c * c * (A)3 * (A)3 * (A)2) * std::sqrt(2.0));
  }
}

template <typename A>
olbgfx_status fill_args(olbgfx_lattice* L, StepArgs<A>& a, bool device_shifts)
{
  const olbgfx_lattice_desc& D = L->desc;
  a.pop = L->pop;
  a.stride = L->stride;
  a.count = L->count;
  for (int i = 0; i < Q; ++i) a.shift[i] = L->shift[i];
  a.shift_dev = device_shifts ? L->shift_dev : nullptr;
  a.mask = L->mask;
  a.uniform_id = L->uniform_id;
  a.skip_others = 0;
  a.nx = D.nx;
  a.ny = D.ny;
  a.nz = D.nz;
  a.pad = D.padding;
  a.py = L->py;
  a.pz = L->pz;
  a.nyz = (uint32_t)D.ny * (uint32_t)D.nz;
  a.magic_nyz = magic_for(a.nyz);
  a.magic_nz = magic_for((uint32_t)D.nz);
  a.ncore = (uint32_t)D.nx * a.nyz;
  a.nblocks = (a.ncore + L->block_threads - 1) / L->block_threads;
  a.blocks_per_xcd = (a.nblocks + 7) / 8;
  a.xcd_run_log2 = L->xcd_run_log2;
  if (a.xcd_run_log2 >= 0) {
    const uint32_t run = 1u << a.xcd_run_log2;
    a.blocks_per_xcd = (a.blocks_per_xcd + run - 1) / run * run;
  }
  a.periodic = (D.periodic[0] ? 1 : 0) | (D.periodic[1] ? 2 : 0) | (D.periodic[2] ? 4 : 0);
  a.cell_u = (const A*)L->cell_u;
  a.cell_rho = (const A*)L->cell_rho;
  a.cell_force = (const A*)L->cell_force;
  a.normal = L->normal_dev;
  a.stats = L->stats_partial;
  fill_table(L, a.table);
  return OLBGFX_OK;
}

template <typename S, bool STATS>
void launch_mode(int mode, dim3 grid, dim3 block, size_t lds, hipStream_t stream,
                 const StepArgs<typename Storage<S>::A>& a)
{
  switch (mode) {
  case OLBGFX_NO_DYNAMICS:
    break;
  case OLBGFX_BGK:
    step_kernel<S, OLBGFX_BGK, STATS><<<grid, block, lds, stream>>>(a);
    break;
  case OLBGFX_TRT:
    step_kernel<S, OLBGFX_TRT, STATS><<<grid, block, lds, stream>>>(a);
    break;
  case OLBGFX_FORCED_BGK:
    step_kernel<S, OLBGFX_FORCED_BGK, STATS><<<grid, block, lds, stream>>>(a);
    break;
  case OLBGFX_FORCED_TRT:
    step_kernel<S, OLBGFX_FORCED_TRT, STATS><<<grid, block, lds, stream>>>(a);
    break;
  case OLBGFX_SMAGORINSKY_BGK:
    step_kernel<S, OLBGFX_SMAGORINSKY_BGK, STATS><<<grid, block, lds, stream>>>(a);
    break;
  case kMaskedBoundary:
    step_kernel<S, kMaskedBoundary, STATS><<<grid, block, lds, stream>>>(a);
    break;
  default:
    if constexpr (!STATS) {
      // split steps never carry statistics
      switch (mode) {
      case kWithFaces + OLBGFX_BGK:
        step_kernel<S, kWithFaces + OLBGFX_BGK, false><<<grid, block, lds, stream>>>(a);
        return;
      case kWithFaces + OLBGFX_TRT:
        step_kernel<S, kWithFaces + OLBGFX_TRT, false><<<grid, block, lds, stream>>>(a);
        return;
      case kWithFaces + OLBGFX_FORCED_BGK:
        step_kernel<S, kWithFaces + OLBGFX_FORCED_BGK, false><<<grid, block, lds, stream>>>(a);
        return;
      case kWithFaces + OLBGFX_FORCED_TRT:
        step_kernel<S, kWithFaces + OLBGFX_FORCED_TRT, false><<<grid, block, lds, stream>>>(a);
        return;
      case kWithFaces + OLBGFX_SMAGORINSKY_BGK:
        step_kernel<S, kWithFaces + OLBGFX_SMAGORINSKY_BGK, false><<<grid, block, lds, stream>>>(a);
        return;
      default:
        break;
      }
    }
    step_kernel<S, -1, STATS><<<grid, block, lds, stream>>>(a);
    break;
  }
}

olbgfx_status prepare_boundary(olbgfx_lattice* L);

// PostStream processors on the populations at the lattice's current offsets:
// the overlap exchange (exchange = true, as SuperLattice::collideAndStream
// does before them), Bouzidi (priority -1), LocalVelocity edges (priority 0),
// then corners (priority 1), then the flat faces of interpolated boundaries
// (priority 0; their order against the edges changes results only by
// rounding: an edge reads a face's density, which the regularised
// populations the face gets reproduce)
template <typename S>
olbgfx_status launch_post(olbgfx_lattice* L, StepArgs<typename Storage<S>::A>& a, hipStream_t stream,
                          bool exchange)
{
  using A = typename Storage<S>::A;
  for (int i = 0; i < Q; ++i) a.shift[i] = L->shift[i];
  if (exchange && L->refresh_n > 0) {
    k_refresh<S><<<(L->refresh_n + 255) / 256, 256, 0, stream>>>(a, L->refresh_dev, L->refresh_n);
    HIP_TRY(hipGetLastError());
  }
  if (L->bz_n > 0) {
    k_bouzidi<S><<<(L->bz_n + 255) / 256, 256, 0, stream>>>(a, L->bz_cells, (const A*)L->bz_q, nullptr,
                                                           L->bz_n);
    HIP_TRY(hipGetLastError());
  }
  if (L->bzv_n > 0) {
    k_bouzidi<S><<<(L->bzv_n + 255) / 256, 256, 0, stream>>>(a, L->bzv_cells, (const A*)L->bzv_q,
                                                            (const A*)L->bzv_v, L->bzv_n);
    HIP_TRY(hipGetLastError());
  }
  if (L->post_edges) {
    k_velocity_post<S><<<(L->post_edges + kPostCellsPerBlock - 1) / kPostCellsPerBlock, 256, 0, stream>>>(
        a, L->post_dev, L->post_edges);
    HIP_TRY(hipGetLastError());
  }
  if (L->post_corners) {
    k_velocity_post<S><<<(L->post_corners + kPostCellsPerBlock - 1) / kPostCellsPerBlock, 256, 0, stream>>>(
        a, L->post_dev + L->post_edges, L->post_corners);
    HIP_TRY(hipGetLastError());
  }
  if (L->fd_n) {
    k_plane_fd<S><<<(L->fd_n + 255) / 256, 256, 0, stream>>>(a, L->fd_dev, L->fd_n, (A*)L->fd_out);
    HIP_TRY(hipGetLastError());
    k_plane_fd_store<S><<<(L->fd_n + 255) / 256, 256, 0, stream>>>(a, L->fd_dev, L->fd_n, (const A*)L->fd_out);
    HIP_TRY(hipGetLastError());
  }
  return OLBGFX_OK;
}

// the dominant dynamics and the residual list of a split step, recomputed
// when dynamics change.  Split when the lattice holds velocity-boundary kinds
// (whose masked kernel carries all their code and runs at 4 waves per SIMD),
// the most frequent core dynamics has a kernel of its own (BGK, TRT, their
// forced forms, Smagorinsky BGK) and the other core cells are at most a
// quarter.  Without boundary kinds one masked launch is faster than any split
// measured: the bounce-back cavity at 100^3 fell from 33.2 to 24.5 GLUPS with
// a residual list (its z-face wall cells, one per z row, each touch 19 cache
// lines of their own), and a launch specialised to the dominant kind with
// the bounce-back cells inline ran the porous channel 3-7 % and the
// bounce-back cavity 1-6 % below one masked launch.
//
// With 4- and 8-byte storage the dominant launch also updates the
// LocalVelocity faces and the FixedVelocity outer edges and corners inline,
// so that their cells share cache lines with their rows, and the residual
// list holds the other kinds only; with 2-byte storage they stay in the list
// and the dominant launch is the uniform kernel skipping them.  Measured on
// the LocalVelocity cavity at 100^3: inline, 56.1 against 59.9 us per step in
// FP32; listed, 48.4 against 51.9 us in FP16, whose dominant kernel is
// partly issue-bound and slows with the boundary arithmetic inline.
olbgfx_status prepare_split(olbgfx_lattice* L)
{
  if (L->split_version == L->version) {
    return OLBGFX_OK;
  }
  L->split_version = L->version;
  L->split_id = -1;
  if (L->residual_dev) {
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipFree(L->residual_dev));
    L->residual_dev = nullptr;
  }
  L->residual_n = 0;
  if (!L->split_enabled || L->uniform_id >= 0 || !L->boundary_kinds) {
    return OLBGFX_OK;
  }
  const olbgfx_lattice_desc& D = L->desc;
  uint64_t count[OLBGFX_MAX_DYNAMICS] = {};
  for (int x = 0; x < D.nx; ++x) {
    for (int y = 0; y < D.ny; ++y) {
      const size_t base = ((size_t)(x + D.padding) * L->py + (y + D.padding)) * L->pz + D.padding;
      for (int z = 0; z < D.nz; ++z) {
        ++count[L->host_mask[base + z]];
      }
    }
  }
  int dom = 0;
  for (int k = 1; k < OLBGFX_MAX_DYNAMICS; ++k) {
    dom = count[k] > count[dom] ? k : dom;
  }
  const int kind = L->table[dom].kind;
  const bool own_kernel = kind == OLBGFX_BGK || kind == OLBGFX_TRT || kind == OLBGFX_FORCED_BGK ||
                          kind == OLBGFX_FORCED_TRT || kind == OLBGFX_SMAGORINSKY_BGK;
  const uint64_t ncore = (uint64_t)D.nx * D.ny * D.nz;
  if (!own_kernel || 4 * (ncore - count[dom]) > ncore) {
    return OLBGFX_OK;
  }
  const bool faces = L->elem >= 4;
  bool inline_id[OLBGFX_MAX_DYNAMICS];
  uint64_t listed = 0;
  for (int id = 0; id < OLBGFX_MAX_DYNAMICS; ++id) {
    const int k = L->table[id].kind;
    inline_id[id] = id == dom || (faces && (k == OLBGFX_LOCAL_VELOCITY || k == OLBGFX_FIXED_VELOCITY));
    listed += inline_id[id] ? 0 : count[id];
  }
  std::vector<uint32_t> rest;
  rest.reserve(listed);
  uint32_t lin = 0;
  for (int x = 0; x < D.nx; ++x) {
    for (int y = 0; y < D.ny; ++y) {
      const size_t base = ((size_t)(x + D.padding) * L->py + (y + D.padding)) * L->pz + D.padding;
      for (int z = 0; z < D.nz; ++z, ++lin) {
        if (!inline_id[L->host_mask[base + z]]) {
          rest.push_back(lin);
        }
      }
    }
  }
  if (!rest.empty()) {
    HIP_TRY(hipMalloc(&L->residual_dev, sizeof(uint32_t) * rest.size()));
    HIP_TRY(hipMemcpy(L->residual_dev, rest.data(), sizeof(uint32_t) * rest.size(), hipMemcpyHostToDevice));
  }
  L->residual_n = (uint32_t)rest.size();
  L->split_id = dom;
  L->split_faces = faces;
  return OLBGFX_OK;
}

template <typename S>
olbgfx_status launch_steps(olbgfx_lattice* L, uint32_t steps, hipStream_t stream)
{
  using A = typename Storage<S>::A;
  if (prepare_boundary(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  StepArgs<A> a;
  fill_args(L, a, false);
  if (a.ncore == 0) {
    return OLBGFX_OK;
  }
  if (L->stats_on && a.nblocks > L->stats_capacity) {
    if (L->stats_partial) {
      HIP_TRY(hipFree(L->stats_partial));
      L->stats_partial = nullptr;
    }
    HIP_TRY(hipMalloc(&L->stats_partial, 4 * sizeof(double) * (size_t)a.nblocks));
    L->stats_capacity = a.nblocks;
    a.stats = L->stats_partial;
  }
  const dim3 grid(8 * a.blocks_per_xcd);
  const dim3 block(L->block_threads);
  int mode = L->uniform_id >= 0 ? L->table[L->uniform_id].kind : -1;
  if (L->boundary_kinds) {
    mode = kMaskedBoundary;
  }
  const bool split = !L->stats_on && prepare_split(L) == OLBGFX_OK && L->split_id >= 0;
  StepArgs<A> rest = a;
  if (split) {
    // dominant kind over every core cell, the others from the list
    a.uniform_id = L->split_id;
    a.skip_others = 1;
    rest.uniform_id = -1;
  }
  const int dominant =
      split ? (L->split_faces ? kWithFaces : 0) + L->table[L->split_id].kind : mode;
  for (uint32_t k = 0; k < steps; ++k) {
    for (int i = 0; i < Q; ++i) a.shift[i] = L->shift[i];
    if (L->stats_on) {
      launch_mode<S, true>(mode, grid, block, L->lds_reserve, stream, a);
      stats_finalize<<<1, 256, 0, stream>>>(L->stats_partial, a.nblocks, L->stats_out);
    } else {
      launch_mode<S, false>(dominant, grid, block, L->lds_reserve, stream, a);
      if (split && L->residual_n > 0) {
        for (int i = 0; i < Q; ++i) rest.shift[i] = L->shift[i];
        const dim3 rgrid((L->residual_n + 255) / 256);
        if (L->boundary_kinds) {
          residual_kernel<S, kMaskedBoundary><<<rgrid, 256, 0, stream>>>(rest, L->residual_dev, L->residual_n);
        } else {
          residual_kernel<S, -1><<<rgrid, 256, 0, stream>>>(rest, L->residual_dev, L->residual_n);
        }
      }
    }
    HIP_TRY(hipGetLastError());
    // stream(): s_i -= d_i (mod count), core/blockLattice.hh:316-338
    for (int i = 0; i < Q; ++i) {
      const uint64_t s = (uint64_t)L->shift[i] + (uint64_t)L->dmod[i];
      L->shift[i] = (uint32_t)(s >= L->count ? s - L->count : s);
    }
    if (launch_post<S>(L, a, stream, true) != OLBGFX_OK) {
      return OLBGFX_ERR_HIP;
    }
    if (L->inflight_limit) {
      hipEvent_t& e = L->inflight_events[L->inflight_count % L->inflight_limit];
      if (L->inflight_count >= L->inflight_limit) {
        HIP_TRY(hipEventSynchronize(e));
      }
      HIP_TRY(hipEventRecord(e, stream));
      ++L->inflight_count;
    }
  }
  if (steps > 0) {
    L->stats_valid = L->stats_on;
    L->shift_dev_valid = false;
  }
  return OLBGFX_OK;
}

olbgfx_status sync_shifts(olbgfx_lattice* L)
{
  HIP_TRY(hipMemcpy(L->shift_dev, L->shift, sizeof(L->shift), hipMemcpyHostToDevice));
  L->shift_dev_valid = true;
  return OLBGFX_OK;
}

template <typename S, typename H>
olbgfx_status transfer_pop(olbgfx_lattice* L, void* host, bool to_device)
{
  H* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, sizeof(H) * (size_t)L->count));
  const dim3 block(256);
  const dim3 grid((L->count + 255) / 256);
  olbgfx_status st = OLBGFX_OK;
  for (int i = 0; i < Q && st == OLBGFX_OK; ++i) {
    S* ring = (S*)L->pop + (uint64_t)i * L->stride;
    H* h = (H*)host + (size_t)i * L->count;
    if (to_device) {
      if (hipMemcpy(tmp, h, sizeof(H) * (size_t)L->count, hipMemcpyHostToDevice) != hipSuccess) {
        st = fail(OLBGFX_ERR_HIP, "population upload failed");
        break;
      }
      k_scatter_pop<S, H><<<grid, block>>>(ring, L->count, L->shift[i], tmp);
    } else {
      k_gather_pop<S, H><<<grid, block>>>(ring, L->count, L->shift[i], tmp);
      if (hipMemcpy(h, tmp, sizeof(H) * (size_t)L->count, hipMemcpyDeviceToHost) != hipSuccess) {
        st = fail(OLBGFX_ERR_HIP, "population download failed");
        break;
      }
    }
  }
  if (hipDeviceSynchronize() != hipSuccess && st == OLBGFX_OK) {
    st = fail(OLBGFX_ERR_HIP, "population transfer kernel failed");
  }
  hipFree(tmp);
  return st;
}

template <typename S>
olbgfx_status transfer_pop_any(olbgfx_lattice* L, int32_t vt, void* host, bool to_device)
{
  if (vt == OLBGFX_VALUE_F32) {
    return transfer_pop<S, float>(L, host, to_device);
  }
  if (vt == OLBGFX_VALUE_F64) {
    return transfer_pop<S, double>(L, host, to_device);
  }
  return fail(OLBGFX_ERR_INVALID, "unknown value type");
}

template <typename F>
olbgfx_status dispatch_storage(olbgfx_lattice* L, F&& f)
{
  switch (L->desc.storage) {
  case OLBGFX_STORE_F32:
    return f((float*)nullptr);
  case OLBGFX_STORE_F16:
    return f((f16_t*)nullptr);
  case OLBGFX_STORE_BF16:
    return f((bf16_t*)nullptr);
  case OLBGFX_STORE_F64:
    return f((double*)nullptr);
  default:
    return fail(OLBGFX_ERR_INVALID, "unknown storage type");
  }
}

size_t value_size(int32_t vt) { return vt == OLBGFX_VALUE_F64 ? 8 : 4; }

// copies a host array of n values (F32 or F64) into device memory of the
// lattice's arithmetic type
olbgfx_status upload_values(olbgfx_lattice* L, void** dst, int32_t vt, const void* src, size_t n)
{
  if (vt != OLBGFX_VALUE_F32 && vt != OLBGFX_VALUE_F64) {
    return fail(OLBGFX_ERR_INVALID, "unknown value type");
  }
  const bool f64 = is_f64(L);
  const size_t esz = f64 ? 8 : 4;
  std::vector<unsigned char> conv(n * esz);
  for (size_t k = 0; k < n; ++k) {
    const double v = vt == OLBGFX_VALUE_F64 ? ((const double*)src)[k] : (double)((const float*)src)[k];
    if (f64) {
      ((double*)conv.data())[k] = v;
    } else {
      ((float*)conv.data())[k] = vt == OLBGFX_VALUE_F32 ? ((const float*)src)[k] : (float)v;
    }
  }
  if (!*dst) {
    HIP_TRY(hipMalloc(dst, n * esz));
  }
  HIP_TRY(hipMemcpy(*dst, conv.data(), n * esz, hipMemcpyHostToDevice));
  return OLBGFX_OK;
}

olbgfx_status check(const olbgfx_lattice* L)
{
  if (!L) {
    return fail(OLBGFX_ERR_INVALID, "null lattice handle");
  }
  return OLBGFX_OK;
}

void recompute_uniform(olbgfx_lattice* L)
{
  const olbgfx_lattice_desc& D = L->desc;
  int id = -2;
  for (int x = 0; x < D.nx && id != -1; ++x) {
    for (int y = 0; y < D.ny && id != -1; ++y) {
      const size_t base = ((size_t)(x + D.padding) * L->py + (y + D.padding)) * L->pz + D.padding;
      for (int z = 0; z < D.nz; ++z) {
        const int v = L->host_mask[base + z];
        if (id == -2) {
          id = v;
        } else if (v != id) {
          id = -1;
          break;
        }
      }
    }
  }
  L->uniform_id = id < 0 ? -1 : id;
}

// find the LocalVelocity cells and the post-processed edges and corners;
// checks that every LOCAL_VELOCITY cell has a flat normal and that the
// processors' neighbours (up to two cells inwards) lie in the padded block
olbgfx_status prepare_boundary(olbgfx_lattice* L)
{
  if (L->boundary_version == L->version) {
    return OLBGFX_OK;
  }
  bool any = false;
  for (int k = 0; k < OLBGFX_MAX_DYNAMICS; ++k) {
    any = any || is_velocity_boundary(L->table[k].kind);
  }
  // the momenta of these kinds read the normal of every cell that holds
  // them, padding included (computeRhoU, the post-processors' neighbours)
  if (any && L->host_normals.empty()) {
    for (size_t n = 0; n < L->count; ++n) {
      if (is_velocity_boundary(L->table[L->host_mask[n]].kind)) {
        return fail(OLBGFX_ERR_INVALID,
                    "a cell holds a boundary kind that reads the discrete normal, but no normals were set "
                    "(olbgfx_set_cell_normals)");
      }
    }
  }
  std::vector<uint32_t> edges, corners, fd;
  bool used = false;
  const olbgfx_lattice_desc& D = L->desc;
  const int P[3] = {D.nx + 2 * D.padding, D.ny + 2 * D.padding, D.nz + 2 * D.padding};
  for (int x = 0; any && x < D.nx; ++x) {
    for (int y = 0; y < D.ny; ++y) {
      for (int z = 0; z < D.nz; ++z) {
        const int c[3] = {x + D.padding, y + D.padding, z + D.padding};
        const uint32_t n = ((uint32_t)c[0] * L->py + (uint32_t)c[1]) * L->pz + (uint32_t)c[2];
        const int kind = L->table[L->host_mask[n]].kind;
        if (!is_velocity_boundary(kind)) {
          continue;
        }
        used = true;
        int nz = 0;
        int N[3] = {0, 0, 0};
        if (!L->host_normals.empty()) {
          for (int a = 0; a < 3; ++a) {
            N[a] = L->host_normals[3 * (size_t)n + a];
            nz += N[a] != 0;
          }
        }
        if ((kind == OLBGFX_LOCAL_VELOCITY || kind == OLBGFX_LOCAL_PRESSURE || is_plane_fd(kind)) && nz != 1) {
          return fail(OLBGFX_ERR_INVALID,
                      "a LOCAL_VELOCITY, LOCAL_PRESSURE or INTERPOLATED_* cell needs a flat discrete normal");
        }
        if (is_plane_fd(kind)) {
          // reach: two cells inwards along the normal, one along the wall
          for (int a = 0; a < 3; ++a) {
            const int lo = N[a] > 0 ? 2 : 1, hi = N[a] < 0 ? 2 : 1;
            if (c[a] - lo < 0 || c[a] + hi >= P[a]) {
              return fail(OLBGFX_ERR_INVALID, "an interpolated boundary processor reaches outside the block");
            }
          }
          fd.push_back(n);
        }
        if (kind == OLBGFX_INNER_EDGE_VELOCITY && nz != 2) {
          return fail(OLBGFX_ERR_INVALID, "an INNER_EDGE_VELOCITY cell needs a normal with two non-zero entries");
        }
        if (kind == OLBGFX_INNER_CORNER_VELOCITY && nz != 3) {
          return fail(OLBGFX_ERR_INVALID, "an INNER_CORNER_VELOCITY cell needs a normal with three non-zero entries");
        }
        if (kind == OLBGFX_FIXED_VELOCITY && nz >= 2) {
          for (int a = 0; a < 3; ++a) {
            const int reach = N[a] == 0 ? (nz == 2 ? 1 : 0) : 2;
            if (c[a] - reach < 0 || c[a] + reach >= P[a]) {
              return fail(OLBGFX_ERR_INVALID, "an edge or corner processor reaches outside the block");
            }
          }
          (nz == 2 ? edges : corners).push_back(n);
        }
      }
    }
  }
  // The reference runs the processors of one priority one after another in
  // ascending cell id; here they run concurrently.  They agree unless one
  // processor reads the density of another cell of the same list (an edge
  // two cells inward of an edge, a corner of a corner), which the staircase
  // walls seen so far never produce: such a configuration is refused.
  {
    std::vector<uint8_t> in(L->count, 0);
    for (uint32_t n : edges) in[n] = 1;
    for (uint32_t n : corners) in[n] = 2;
    const int64_t step[3] = {(int64_t)L->py * L->pz, (int64_t)L->pz, 1};
    for (int list = 1; list <= 2; ++list) {
      for (uint32_t n : (list == 1 ? edges : corners)) {
        int N[3];
        for (int ax = 0; ax < 3; ++ax) N[ax] = L->host_normals[3 * (size_t)n + ax];
        for (int ax = 0; ax < 3; ++ax) {
          if (N[ax] == 0) continue;
          for (int k = 1; k <= 2; ++k) {
            const int64_t m = (int64_t)n - (int64_t)k * N[ax] * step[ax];
            if (in[(size_t)m] == list) {
              return fail(OLBGFX_ERR_UNSUPPORTED,
                          "a LocalVelocity edge or corner processor reads the density of another one of its kind");
            }
          }
        }
      }
    }
  }
  // periodic padding cells within reach (2) of an edge or corner processor
  std::vector<uint32_t> pairs;
  if (D.periodic[0] || D.periodic[1] || D.periodic[2]) {
    std::vector<uint8_t> seen(L->count, 0);
    const int ext[3] = {D.nx, D.ny, D.nz};
    auto visit = [&](uint32_t n) {
      const int c[3] = {(int)(n / (L->py * L->pz)) - D.padding, (int)(n / L->pz % L->py) - D.padding,
                        (int)(n % L->pz) - D.padding};
      for (int dx = -2; dx <= 2; ++dx)
        for (int dy = -2; dy <= 2; ++dy)
          for (int dz = -2; dz <= 2; ++dz) {
            int g[3] = {c[0] + dx, c[1] + dy, c[2] + dz}, m[3];
            bool pad = false, ok = true;
            for (int ax = 0; ax < 3; ++ax) {
              m[ax] = g[ax];
              if (g[ax] < 0 || g[ax] >= ext[ax]) {
                pad = true;
                if (!D.periodic[ax] || g[ax] < -D.padding || g[ax] >= ext[ax] + D.padding) {
                  ok = false;
                } else {
                  m[ax] = g[ax] < 0 ? g[ax] + ext[ax] : g[ax] - ext[ax];
                }
              }
            }
            if (!pad || !ok) continue;
            const uint32_t gn = ((uint32_t)(g[0] + D.padding) * L->py + (uint32_t)(g[1] + D.padding)) * L->pz +
                                (uint32_t)(g[2] + D.padding);
            if (seen[gn]) continue;
            seen[gn] = 1;
            const uint32_t mn = ((uint32_t)(m[0] + D.padding) * L->py + (uint32_t)(m[1] + D.padding)) * L->pz +
                                (uint32_t)(m[2] + D.padding);
            pairs.push_back(gn);
            pairs.push_back(mn);
          }
    };
    for (uint32_t n : edges) visit(n);
    for (uint32_t n : corners) visit(n);
    for (uint32_t n : fd) visit(n);
  }
  if (L->refresh_dev) {
    HIP_TRY(hipFree(L->refresh_dev));
    L->refresh_dev = nullptr;
  }
  L->refresh_n = (uint32_t)(pairs.size() / 2);
  if (!pairs.empty()) {
    HIP_TRY(hipMalloc(&L->refresh_dev, sizeof(uint32_t) * pairs.size()));
    HIP_TRY(hipMemcpy(L->refresh_dev, pairs.data(), sizeof(uint32_t) * pairs.size(), hipMemcpyHostToDevice));
  }
  if (L->post_dev) {
    HIP_TRY(hipFree(L->post_dev));
    L->post_dev = nullptr;
  }
  L->post_edges = (uint32_t)edges.size();
  L->post_corners = (uint32_t)corners.size();
  if (!edges.empty() || !corners.empty()) {
    edges.insert(edges.end(), corners.begin(), corners.end());
    HIP_TRY(hipMalloc(&L->post_dev, sizeof(uint32_t) * edges.size()));
    HIP_TRY(hipMemcpy(L->post_dev, edges.data(), sizeof(uint32_t) * edges.size(), hipMemcpyHostToDevice));
  }
  if (L->fd_dev) {
    HIP_TRY(hipFree(L->fd_dev));
    L->fd_dev = nullptr;
  }
  if (L->fd_out) {
    HIP_TRY(hipFree(L->fd_out));
    L->fd_out = nullptr;
  }
  L->fd_n = (uint32_t)fd.size();
  if (!fd.empty()) {
    const size_t value = L->desc.storage == OLBGFX_STORE_F64 ? sizeof(double) : sizeof(float);
    HIP_TRY(hipMalloc(&L->fd_dev, sizeof(uint32_t) * fd.size()));
    HIP_TRY(hipMemcpy(L->fd_dev, fd.data(), sizeof(uint32_t) * fd.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&L->fd_out, value * Q * fd.size()));
  }
  L->boundary_kinds = used;
  L->boundary_version = L->version;
  return OLBGFX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char* olbgfx_last_error(void) { return g_error.c_str(); }

olbgfx_status olbgfx_lattice_create(const olbgfx_lattice_desc* desc, olbgfx_lattice** out)
{
  if (!desc || !out) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  *out = nullptr;
  const olbgfx_lattice_desc& D = *desc;
  if (D.nx <= 0 || D.ny <= 0 || D.nz <= 0) {
    return fail(OLBGFX_ERR_INVALID, "lattice extent must be positive");
  }
  if (D.padding < 1) {
    return fail(OLBGFX_ERR_INVALID, "padding must be at least 1 for D3Q19");
  }
  if (D.storage < OLBGFX_STORE_F32 || D.storage > OLBGFX_STORE_F64) {
    return fail(OLBGFX_ERR_INVALID, "unknown storage type");
  }
  const uint64_t px = (uint64_t)D.nx + 2 * (uint64_t)D.padding;
  const uint64_t py = (uint64_t)D.ny + 2 * (uint64_t)D.padding;
  const uint64_t pz = (uint64_t)D.nz + 2 * (uint64_t)D.padding;
  const uint64_t count = px * py * pz;
  // the reference limits CellID to uint32 (core/blockStructure.h:53-55); the
  // kernels add two cell indices in 32 bits, so stay below 2^31
  if (count >= (1ull << 31)) {
    return fail(OLBGFX_ERR_INVALID, "cell count must stay below 2^31");
  }
  // ring slots are addressed with 32-bit byte offsets
  const uint64_t elem = D.storage == OLBGFX_STORE_F32 ? 4 : (D.storage == OLBGFX_STORE_F64 ? 8 : 2);
  if (count * elem >= (1ull << 32)) {
    return fail(OLBGFX_ERR_INVALID, "one population ring must stay below 4 GiB");
  }
  olbgfx_lattice* L = new (std::nothrow) olbgfx_lattice();
  if (!L) {
    return fail(OLBGFX_ERR_NOMEM, "host allocation failed");
  }
  L->desc = D;
  L->elem = D.storage == OLBGFX_STORE_F32 ? 4 : (D.storage == OLBGFX_STORE_F64 ? 8 : 2);
  L->count = (uint32_t)count;
  L->py = (uint32_t)py;
  L->pz = (uint32_t)pz;
  const uint64_t align = 256 / L->elem;
  L->stride = (count + align - 1) / align * align;
  const int64_t proj[3] = {(int64_t)(py * pz), (int64_t)pz, 1};
  for (int i = 0; i < Q; ++i) {
    L->d[i] = CX(i) * proj[0] + CY(i) * proj[1] + CZ(i) * proj[2];
    int64_t m = (-L->d[i]) % (int64_t)count;
    if (m < 0) m += (int64_t)count;
    L->dmod[i] = m;
    L->shift[i] = 0;
  }
  for (int k = 0; k < OLBGFX_MAX_DYNAMICS; ++k) {
    L->table[k] = olbgfx_dynamics_desc{};
    L->table[k].kind = OLBGFX_NO_DYNAMICS;
    L->table[k].wall_rho = 1.0;
  }
  L->host_mask.assign(count, 0);
  L->uniform_id = 0;

  auto cleanup = [&](olbgfx_status s) {
    olbgfx_lattice_destroy(L);
    return s;
  };
  const size_t bytes = (size_t)Q * L->stride * L->elem;
  if (hipMalloc(&L->pop, bytes) != hipSuccess) {
    return cleanup(fail(OLBGFX_ERR_NOMEM, "population allocation failed"));
  }
  if (hipMemset(L->pop, 0, bytes) != hipSuccess ||
      hipMalloc(&L->mask, count) != hipSuccess || hipMemset(L->mask, 0, count) != hipSuccess ||
      hipMalloc(&L->shift_dev, sizeof(L->shift)) != hipSuccess ||
      hipMemset(L->shift_dev, 0, sizeof(L->shift)) != hipSuccess ||
      hipMalloc(&L->kinds_dev, sizeof(int32_t) * 256) != hipSuccess ||
      hipMemset(L->kinds_dev, 0, sizeof(int32_t) * 256) != hipSuccess ||
      hipMalloc(&L->stats_out, 4 * sizeof(double)) != hipSuccess ||
      hipMemset(L->stats_out, 0, 4 * sizeof(double)) != hipSuccess) {
    return cleanup(fail(OLBGFX_ERR_NOMEM, "device allocation failed"));
  }
  *out = L;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_lattice_destroy(olbgfx_lattice* L)
{
  if (!L) {
    return OLBGFX_OK;
  }
  hipDeviceSynchronize();
  hipFree(L->pop);
  hipFree(L->mask);
  hipFree(L->shift_dev);
  hipFree(L->kinds_dev);
  hipFree(L->cell_u);
  hipFree(L->cell_rho);
  hipFree(L->cell_force);
  hipFree(L->normal_dev);
  hipFree(L->post_dev);
  hipFree(L->fd_dev);
  hipFree(L->fd_out);
  hipFree(L->refresh_dev);
  hipFree(L->bz_cells);
  hipFree(L->bz_q);
  hipFree(L->bzv_cells);
  hipFree(L->bzv_q);
  hipFree(L->bzv_v);
  hipFree(L->stats_partial);
  hipFree(L->stats_out);
  hipFree(L->residual_dev);
  for (hipEvent_t e : L->inflight_events) {
    hipEventDestroy(e);
  }
  delete L;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_set_inflight_limit(olbgfx_lattice* L, uint32_t steps)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  if (steps > 64) {
    return fail(OLBGFX_ERR_INVALID, "at most 64 steps in flight");
  }
  HIP_TRY(hipDeviceSynchronize());
  for (hipEvent_t e : L->inflight_events) {
    HIP_TRY(hipEventDestroy(e));
  }
  L->inflight_events.assign(steps, nullptr);
  for (uint32_t k = 0; k < steps; ++k) {
    HIP_TRY(hipEventCreateWithFlags(&L->inflight_events[k], hipEventDisableTiming));
  }
  L->inflight_limit = steps;
  L->inflight_count = 0;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_set_split(olbgfx_lattice* L, int32_t on)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  L->split_enabled = on != 0;
  L->split_version = ~0ull;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_get_split(olbgfx_lattice* L, int32_t* kind, uint32_t* residual)
{
  if (check(L) != OLBGFX_OK || !kind || !residual) {
    return OLBGFX_ERR_INVALID;
  }
  *kind = -1;
  *residual = 0;
  if (prepare_boundary(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  if (prepare_split(L) != OLBGFX_OK) {
    return OLBGFX_ERR_HIP;
  }
  if (!L->stats_on && L->split_id >= 0) {
    *kind = L->table[L->split_id].kind;
    *residual = L->residual_n;
  }
  return OLBGFX_OK;
}

olbgfx_status olbgfx_device_count(int32_t* count)
{
  if (!count) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    n = 0;
  }
  *count = n;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_lattice_ncells(const olbgfx_lattice* L, uint64_t* n)
{
  if (check(L) != OLBGFX_OK || !n) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  *n = L->count;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_neighbor_distances(const olbgfx_lattice* L, int64_t out[19])
{
  if (check(L) != OLBGFX_OK || !out) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  for (int i = 0; i < Q; ++i) out[i] = L->d[i];
  return OLBGFX_OK;
}

olbgfx_status olbgfx_ring_shifts(olbgfx_lattice* L, int64_t out[19])
{
  if (check(L) != OLBGFX_OK || !out) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  HIP_TRY(hipDeviceSynchronize());
  for (int i = 0; i < Q; ++i) out[i] = L->shift[i];
  return OLBGFX_OK;
}

olbgfx_status olbgfx_set_dynamics(olbgfx_lattice* L, uint32_t id, const olbgfx_dynamics_desc* dyn)
{
  if (L) {
    ++L->version;
  }
  if (check(L) != OLBGFX_OK || !dyn) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  if (id >= OLBGFX_MAX_DYNAMICS) {
    return fail(OLBGFX_ERR_INVALID, "dynamics id out of range");
  }
  if (dyn->kind < OLBGFX_NO_DYNAMICS || dyn->kind > OLBGFX_INTERPOLATED_PRESSURE) {
    return fail(OLBGFX_ERR_UNSUPPORTED, "dynamics kind not provided by the gfx950 platform");
  }
  L->table[id] = *dyn;
  int32_t kinds[256];
  for (int k = 0; k < 256; ++k) kinds[k] = k < OLBGFX_MAX_DYNAMICS ? L->table[k].kind : 0;
  HIP_TRY(hipMemcpy(L->kinds_dev, kinds, sizeof(kinds), hipMemcpyHostToDevice));
  return OLBGFX_OK;
}

olbgfx_status olbgfx_set_dynamics_mask(olbgfx_lattice* L, const uint8_t* ids)
{
  if (L) {
    ++L->version;
  }
  if (check(L) != OLBGFX_OK || !ids) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  for (uint32_t n = 0; n < L->count; ++n) {
    if (ids[n] >= OLBGFX_MAX_DYNAMICS) {
      return fail(OLBGFX_ERR_INVALID, "dynamics id out of range in mask");
    }
  }
  std::memcpy(L->host_mask.data(), ids, L->count);
  HIP_TRY(hipMemcpy(L->mask, ids, L->count, hipMemcpyHostToDevice));
  recompute_uniform(L);
  return OLBGFX_OK;
}

olbgfx_status olbgfx_set_cell_normals(olbgfx_lattice* L, const int8_t* normals)
{
  if (L) {
    ++L->version;
  }
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  if (!normals) {
    L->host_normals.clear();
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipFree(L->normal_dev));
    L->normal_dev = nullptr;
    return OLBGFX_OK;
  }
  std::vector<uint8_t> codes(L->count);
  for (uint32_t n = 0; n < L->count; ++n) {
    int code = 0;
    for (int a = 0; a < 3; ++a) {
      const int v = normals[3 * (size_t)n + a];
      if (v < -1 || v > 1) {
        return fail(OLBGFX_ERR_INVALID, "discrete normal entries must be -1, 0 or 1");
      }
      code = 3 * code + (v + 1);
    }
    codes[n] = (uint8_t)code;
  }
  L->host_normals.assign(normals, normals + 3 * (size_t)L->count);
  if (!L->normal_dev) {
    HIP_TRY(hipMalloc(&L->normal_dev, L->count));
  }
  HIP_TRY(hipMemcpy(L->normal_dev, codes.data(), L->count, hipMemcpyHostToDevice));
  return OLBGFX_OK;
}

}  // extern "C"

namespace {

// one Bouzidi list: validated, then uploaded direction-major; vel (may be
// null) holds the BOUZIDI_VELOCITY coefficients of the velocity variant
olbgfx_status set_bouzidi_list(olbgfx_lattice* L, int64_t n, const int64_t* cells, int32_t vt, const void* q,
                               const void* vel, uint32_t*& dev_cells, void*& dev_q, void*& dev_v, uint32_t& dev_n)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  if (n < 0 || (n > 0 && (!cells || !q))) {
    return fail(OLBGFX_ERR_INVALID, "null argument or negative count");
  }
  if (vt != OLBGFX_VALUE_F32 && vt != OLBGFX_VALUE_F64) {
    return fail(OLBGFX_ERR_INVALID, "unknown value type");
  }
  if (n > (int64_t)L->count) {
    return fail(OLBGFX_ERR_INVALID, "more Bouzidi cells than cells");
  }
  const olbgfx_lattice_desc& D = L->desc;
  const int64_t P[3] = {D.nx + 2 * D.padding, D.ny + 2 * D.padding, D.nz + 2 * D.padding};
  std::vector<double> qd(Q * (size_t)n);
  for (size_t k = 0; k < qd.size(); ++k) {
    qd[k] = vt == OLBGFX_VALUE_F64 ? ((const double*)q)[k] : (double)((const float*)q)[k];
    if (!std::isfinite(qd[k])) {
      return fail(OLBGFX_ERR_INVALID, "Bouzidi distances must be finite");
    }
  }
  std::vector<double> vd(vel ? Q * (size_t)n : 0);
  for (size_t k = 0; k < vd.size(); ++k) {
    vd[k] = vt == OLBGFX_VALUE_F64 ? ((const double*)vel)[k] : (double)((const float*)vel)[k];
    if (!std::isfinite(vd[k])) {
      return fail(OLBGFX_ERR_INVALID, "Bouzidi velocity coefficients must be finite");
    }
  }
  for (int64_t k = 0; k < n; ++k) {
    const int64_t c = cells[k];
    if (c < 0 || c >= (int64_t)L->count || (k > 0 && c <= cells[k - 1])) {
      return fail(OLBGFX_ERR_INVALID, "Bouzidi cells must be valid ids in ascending order");
    }
    const int64_t r[3] = {c / ((int64_t)L->py * L->pz), (c / L->pz) % L->py, c % L->pz};
    for (int ax = 0; ax < 3; ++ax) {
      if (r[ax] < D.padding || r[ax] >= P[ax] - D.padding) {
        return fail(OLBGFX_ERR_INVALID, "Bouzidi cells must be core cells");
      }
    }
  }
  // a cell's update reads x_s[i] (x_s = x_b + c_i) and, for q > 1/2,
  // x_f[opp i] (x_f = x_b - c_i); neither may be written by another update
  // of the same list, whose cells run concurrently.  The plain and the
  // velocity list are not checked against each other: they run as two
  // launches, the plain list first, which is the reference's own order (all
  // cells of BouzidiPostProcessor, then all of BouzidiVelocityPostProcessor,
  // in the one priority of core/blockPostProcessorMap.h), so a cell in both
  // lists, or one list reading what the other wrote, sees what it sees there
  auto entry = [&](int64_t c) -> int64_t {
    const int64_t* it = std::lower_bound(cells, cells + n, c);
    return (it != cells + n && *it == c) ? it - cells : -1;
  };
  for (int64_t k = 0; k < n; ++k) {
    for (int i = 1; i < Q; ++i) {
      const double qi = qd[Q * k + i];
      if (qi < 0) {
        continue;
      }
      const int o = i <= 9 ? i + 9 : i - 9;
      const int64_t ys = entry(cells[k] + L->d[i]);
      if (ys >= 0 && qd[Q * ys + o] >= 0) {
        return fail(OLBGFX_ERR_INVALID, "a Bouzidi update reads a population another one writes");
      }
      if (qi > 0.5) {
        const int64_t yf = entry(cells[k] - L->d[i]);
        if (yf >= 0 && qd[Q * yf + i] >= 0) {
          return fail(OLBGFX_ERR_INVALID, "a Bouzidi update reads a population another one writes");
        }
      }
    }
  }
  ++L->version;
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipFree(dev_cells));
  HIP_TRY(hipFree(dev_q));
  HIP_TRY(hipFree(dev_v));
  dev_cells = nullptr;
  dev_q = nullptr;
  dev_v = nullptr;
  dev_n = 0;
  if (n == 0) {
    return OLBGFX_OK;
  }
  std::vector<uint32_t> ids(cells, cells + n);
  // direction-major on the device so the threads of a wavefront read q coalesced
  const size_t as = is_f64(L) ? 8 : 4;
  auto direction_major = [&](const std::vector<double>& src) {
    std::vector<char> out(as * 18 * (size_t)n);
    for (int64_t k = 0; k < n; ++k) {
      for (int i = 1; i < Q; ++i) {
        const size_t at = (size_t)(i - 1) * n + k;
        if (as == 8) {
          ((double*)out.data())[at] = src[Q * k + i];
        } else {
          ((float*)out.data())[at] = (float)src[Q * k + i];
        }
      }
    }
    return out;
  };
  const std::vector<char> qa = direction_major(qd);
  HIP_TRY(hipMalloc(&dev_cells, sizeof(uint32_t) * ids.size()));
  HIP_TRY(hipMalloc(&dev_q, qa.size()));
  HIP_TRY(hipMemcpy(dev_cells, ids.data(), sizeof(uint32_t) * ids.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dev_q, qa.data(), qa.size(), hipMemcpyHostToDevice));
  if (vel) {
    const std::vector<char> va = direction_major(vd);
    HIP_TRY(hipMalloc(&dev_v, va.size()));
    HIP_TRY(hipMemcpy(dev_v, va.data(), va.size(), hipMemcpyHostToDevice));
  }
  dev_n = (uint32_t)n;
  return OLBGFX_OK;
}

}  // namespace

extern "C" {

olbgfx_status olbgfx_set_bouzidi(olbgfx_lattice* L, int64_t n, const int64_t* cells, int32_t vt, const void* q)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  void* unused = nullptr;
  return set_bouzidi_list(L, n, cells, vt, q, nullptr, L->bz_cells, L->bz_q, unused, L->bz_n);
}

olbgfx_status olbgfx_set_bouzidi_velocity(olbgfx_lattice* L, int64_t n, const int64_t* cells, int32_t vt,
                                          const void* q, const void* velocity)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  if (n > 0 && !velocity) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  return set_bouzidi_list(L, n, cells, vt, q, velocity, L->bzv_cells, L->bzv_q, L->bzv_v, L->bzv_n);
}



olbgfx_status olbgfx_set_cell_field(olbgfx_lattice* L, int32_t field, int32_t vt, const void* values)
{
  if (L) {
    ++L->version;
  }
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  void** dst = nullptr;
  size_t comps = 0;
  switch (field) {
  case OLBGFX_FIELD_VELOCITY:
    dst = &L->cell_u;
    comps = 3;
    break;
  case OLBGFX_FIELD_RHO:
    dst = &L->cell_rho;
    comps = 1;
    break;
  case OLBGFX_FIELD_FORCE:
    dst = &L->cell_force;
    comps = 3;
    break;
  default:
    return fail(OLBGFX_ERR_INVALID, "unknown field");
  }
  if (!values) {
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipFree(*dst));
    *dst = nullptr;
    return OLBGFX_OK;
  }
  return upload_values(L, dst, vt, values, comps * (size_t)L->count);
}

olbgfx_status olbgfx_ini_equilibrium(olbgfx_lattice* L, const uint8_t* sel, int32_t vt,
                                     const void* rho, const void* u)
{
  if (check(L) != OLBGFX_OK || !rho || !u) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  if (vt != OLBGFX_VALUE_F32 && vt != OLBGFX_VALUE_F64) {
    return fail(OLBGFX_ERR_INVALID, "unknown value type");
  }
  const size_t n = L->count;
  const size_t vs = value_size(vt);
  void* d_rho = nullptr;
  void* d_u = nullptr;
  uint8_t* d_sel = nullptr;
  olbgfx_status st = OLBGFX_OK;
  if (hipMalloc(&d_rho, n * vs) != hipSuccess || hipMalloc(&d_u, 3 * n * vs) != hipSuccess ||
      (sel && hipMalloc(&d_sel, n) != hipSuccess)) {
    st = fail(OLBGFX_ERR_NOMEM, "temporary allocation failed");
  }
  if (st == OLBGFX_OK &&
      (hipMemcpy(d_rho, rho, n * vs, hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(d_u, u, 3 * n * vs, hipMemcpyHostToDevice) != hipSuccess ||
       (sel && hipMemcpy(d_sel, sel, n, hipMemcpyHostToDevice) != hipSuccess) ||
       sync_shifts(L) != OLBGFX_OK)) {
    st = fail(OLBGFX_ERR_HIP, "upload failed");
  }
  if (st == OLBGFX_OK) {
    const dim3 block(256), grid((L->count + 255) / 256);
    st = dispatch_storage(L, [&](auto* tag) -> olbgfx_status {
      using S = std::remove_pointer_t<decltype(tag)>;
      if (vt == OLBGFX_VALUE_F32) {
        k_ini_equilibrium<S, float><<<grid, block>>>((S*)L->pop, L->stride, L->count, L->shift_dev,
                                                     d_sel, (const float*)d_rho, (const float*)d_u);
      } else {
        k_ini_equilibrium<S, double><<<grid, block>>>((S*)L->pop, L->stride, L->count,
                                                      L->shift_dev, d_sel, (const double*)d_rho,
                                                      (const double*)d_u);
      }
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipDeviceSynchronize());
      return OLBGFX_OK;
    });
  }
  hipFree(d_rho);
  hipFree(d_u);
  hipFree(d_sel);
  return st;
}

olbgfx_status olbgfx_get_populations(olbgfx_lattice* L, int32_t vt, void* host)
{
  if (check(L) != OLBGFX_OK || !host) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  HIP_TRY(hipDeviceSynchronize());
  return dispatch_storage(L, [&](auto* tag) {
    using S = std::remove_pointer_t<decltype(tag)>;
    return transfer_pop_any<S>(L, vt, host, false);
  });
}

olbgfx_status olbgfx_set_populations(olbgfx_lattice* L, int32_t vt, const void* host)
{
  if (check(L) != OLBGFX_OK || !host) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  HIP_TRY(hipDeviceSynchronize());
  return dispatch_storage(L, [&](auto* tag) {
    using S = std::remove_pointer_t<decltype(tag)>;
    return transfer_pop_any<S>(L, vt, const_cast<void*>(host), true);
  });
}

olbgfx_status olbgfx_collide_and_stream(olbgfx_lattice* L, uint32_t steps, void* stream)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  hipStream_t s = (hipStream_t)stream;
  return dispatch_storage(L, [&](auto* tag) {
    using S = std::remove_pointer_t<decltype(tag)>;
    return launch_steps<S>(L, steps, s);
  });
}

olbgfx_status olbgfx_post_process(olbgfx_lattice* L, void* stream)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  if (prepare_boundary(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  hipStream_t s = (hipStream_t)stream;
  return dispatch_storage(L, [&](auto* tag) -> olbgfx_status {
    using S = std::remove_pointer_t<decltype(tag)>;
    StepArgs<typename Storage<S>::A> a;
    fill_args(L, a, false);
    // BlockLattice::postProcess runs without the super-level exchange
    return launch_post<S>(L, a, s, false);
  });
}

olbgfx_status olbgfx_set_statistics(olbgfx_lattice* L, int32_t on)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  L->stats_on = on != 0;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_get_statistics(olbgfx_lattice* L, double out[4])
{
  if (check(L) != OLBGFX_OK || !out) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  if (!L->stats_valid) {
    for (int k = 0; k < 4; ++k) out[k] = 0;
    return OLBGFX_OK;
  }
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(out, L->stats_out, 4 * sizeof(double), hipMemcpyDeviceToHost));
  return OLBGFX_OK;
}

olbgfx_status olbgfx_reduce_moments(olbgfx_lattice* L, double out[8], void* stream)
{
  if (check(L) != OLBGFX_OK || !out) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  const olbgfx_lattice_desc& D = L->desc;
  const uint64_t ncore = (uint64_t)D.nx * D.ny * D.nz;
  const uint32_t nb = (uint32_t)((ncore + 255) / 256);
  double* partial = nullptr;
  HIP_TRY(hipMalloc(&partial, 8 * sizeof(double) * (size_t)nb));
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipStreamSynchronize(s));
  if (sync_shifts(L) != OLBGFX_OK) {
    hipFree(partial);
    return OLBGFX_ERR_HIP;
  }
  olbgfx_status st = dispatch_storage(L, [&](auto* tag) -> olbgfx_status {
    using S = std::remove_pointer_t<decltype(tag)>;
    k_moments<S><<<nb, 256, 0, s>>>((const S*)L->pop, L->stride, L->count, L->shift_dev, L->mask,
                                    L->kinds_dev, D.nx, D.ny, D.nz, D.padding, L->py, L->pz,
                                    partial);
    HIP_TRY(hipGetLastError());
    return OLBGFX_OK;
  });
  if (st == OLBGFX_OK) {
    std::vector<double> h(8 * (size_t)nb);
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpy(h.data(), partial, h.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
      st = fail(OLBGFX_ERR_HIP, "moment reduction failed");
    } else {
      double o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (uint32_t b = 0; b < nb; ++b) {
        for (int k = 0; k < 6; ++k) o[k] += h[8 * (size_t)b + k];
        o[6] = std::max(o[6], h[8 * (size_t)b + 6]);
      }
      for (int k = 0; k < 8; ++k) out[k] = o[k];
    }
  }
  hipFree(partial);
  return st;
}

olbgfx_status olbgfx_reduce_moments_noalloc(olbgfx_lattice* L, double out[8], void* stream)
{
  if (check(L) != OLBGFX_OK || !out) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  const olbgfx_lattice_desc& D = L->desc;
  const uint64_t ncore = (uint64_t)D.nx * D.ny * D.nz;
  const uint64_t nb = (ncore + 255) / 256;
  if (nb > kNoAllocBlocks) {
    return fail(OLBGFX_ERR_INVALID,
                "olbgfx_reduce_moments_noalloc: the lattice has more than 16777216 core cells "
                "(65536 blocks of 256), the size of its fixed device scratch; use olbgfx_reduce_moments");
  }
  // one call at a time per process: the device scratch and the host copy of
  // the partials are shared by every lattice
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  double* partial = nullptr;
  if (noalloc_scratch(&partial) != OLBGFX_OK) {
    return OLBGFX_ERR_HIP;
  }
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipStreamSynchronize(s));
  if (sync_shifts(L) != OLBGFX_OK) {
    return OLBGFX_ERR_HIP;
  }
  // k_moments writes all 8 slots of every block it runs, so nothing left in
  // the scratch by an earlier call is read
  olbgfx_status st = dispatch_storage(L, [&](auto* tag) -> olbgfx_status {
    using S = std::remove_pointer_t<decltype(tag)>;
    k_moments<S><<<(uint32_t)nb, 256, 0, s>>>((const S*)L->pop, L->stride, L->count, L->shift_dev, L->mask,
                                              L->kinds_dev, D.nx, D.ny, D.nz, D.padding, L->py, L->pz,
                                              partial);
    HIP_TRY(hipGetLastError());
    return OLBGFX_OK;
  });
  if (st != OLBGFX_OK) {
    return st;
  }
  if (hipStreamSynchronize(s) != hipSuccess ||
      hipMemcpy(g_noalloc_host, partial, 8 * sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost) != hipSuccess) {
    return fail(OLBGFX_ERR_HIP, "moment reduction failed");
  }
  // the same order of sums as olbgfx_reduce_moments, so both agree bit for bit
  double o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint64_t b = 0; b < nb; ++b) {
    for (int k = 0; k < 6; ++k) o[k] += g_noalloc_host[8 * b + k];
    o[6] = std::max(o[6], g_noalloc_host[8 * b + 6]);
  }
  for (int k = 0; k < 8; ++k) out[k] = o[k];
  return OLBGFX_OK;
}

olbgfx_status olbgfx_compute_rho_u(olbgfx_lattice* L, int32_t vt, void* rho, void* u)
{
  if (check(L) != OLBGFX_OK || !rho || !u) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  if (vt != OLBGFX_VALUE_F32 && vt != OLBGFX_VALUE_F64) {
    return fail(OLBGFX_ERR_INVALID, "unknown value type");
  }
  HIP_TRY(hipDeviceSynchronize());
  const size_t n = L->count;
  const size_t vs = value_size(vt);
  void* d_rho = nullptr;
  void* d_u = nullptr;
  HIP_TRY(hipMalloc(&d_rho, n * vs));
  if (hipMalloc(&d_u, 3 * n * vs) != hipSuccess) {
    hipFree(d_rho);
    return fail(OLBGFX_ERR_NOMEM, "temporary allocation failed");
  }
  olbgfx_status st = prepare_boundary(L);
  if (st == OLBGFX_OK) {
    const dim3 block(256), grid((L->count + 255) / 256);
    st = dispatch_storage(L, [&](auto* tag) -> olbgfx_status {
      using S = std::remove_pointer_t<decltype(tag)>;
      using A = typename Storage<S>::A;
      StepArgs<A> a;
      fill_args(L, a, false);
      if (vt == OLBGFX_VALUE_F32) {
        k_cell_rho_u<S, float><<<grid, block>>>(a, (float*)d_rho, (float*)d_u);
      } else {
        k_cell_rho_u<S, double><<<grid, block>>>(a, (double*)d_rho, (double*)d_u);
      }
      HIP_TRY(hipGetLastError());
      return OLBGFX_OK;
    });
  }
  if (st == OLBGFX_OK &&
      (hipMemcpy(rho, d_rho, n * vs, hipMemcpyDeviceToHost) != hipSuccess ||
       hipMemcpy(u, d_u, 3 * n * vs, hipMemcpyDeviceToHost) != hipSuccess)) {
    st = fail(OLBGFX_ERR_HIP, "download failed");
  }
  hipFree(d_rho);
  hipFree(d_u);
  return st;
}

olbgfx_status olbgfx_set_launch(olbgfx_lattice* L, uint32_t block_threads, uint32_t waves_per_simd)
{
  if (L) {
    ++L->version;
  }
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  if (block_threads == 0) {
    block_threads = 256;
  }
  // the step kernel is compiled with __launch_bounds__(kMaxBlock)
  if (block_threads < 64 || block_threads > kMaxBlock || block_threads % 64 != 0) {
    return fail(OLBGFX_ERR_INVALID, "block size must be a multiple of 64 in [64, 256]");
  }
  if (waves_per_simd > 8) {
    return fail(OLBGFX_ERR_INVALID, "waves per SIMD must be in [0, 8]");
  }
  uint32_t lds = 0;
  if (waves_per_simd > 0) {
    // a CU has 4 SIMDs and 160 KiB of LDS: reserving LDS per block caps the
    // blocks that fit on one CU, and with them the waves per SIMD
    const uint32_t waves_per_block = block_threads / 64;
    const uint32_t blocks_per_cu = std::max(1u, 4 * waves_per_simd / waves_per_block);
    lds = kLdsPerCu / blocks_per_cu / 1024 * 1024;
    if (lds * (blocks_per_cu + 1) <= kLdsPerCu) {
      lds += 1024;
    }
    lds = std::min(lds, kLdsPerCu - 4096);  // room for the kernel's own reduction scratch
  }
  L->block_threads = block_threads;
  L->lds_reserve = lds;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_set_block_order(olbgfx_lattice* L, int32_t xcd_run_log2)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  if (xcd_run_log2 < -1 || xcd_run_log2 > 12) {
    return fail(OLBGFX_ERR_INVALID, "block order must be -1 or a run length 2^k with k in [0, 12]");
  }
  L->xcd_run_log2 = xcd_run_log2;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_synchronize(olbgfx_lattice* L)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  HIP_TRY(hipDeviceSynchronize());
  return OLBGFX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// ensemble of independent lattices

struct olbgfx_ensemble {
  std::vector<olbgfx_lattice*> members;
  int32_t storage = OLBGFX_STORE_F32;
  void* args_dev = nullptr;          // StepArgs<A>[members]
  Tile* tiles_dev = nullptr;
  uint32_t ntiles = 0;
  uint32_t tiles_capacity = 0;
  uint32_t** shift_ptrs_dev = nullptr;
  uint32_t* dmod_dev = nullptr;
  uint32_t* counts_dev = nullptr;
  uint32_t block_threads = 256;
  uint32_t waves_per_simd = 0;
  uint32_t tiles_per_block = 1;
  int32_t xcd_run_log2 = -1;         // block order (olbgfx_ensemble_set_block_order)
  size_t lds = 0;
  std::vector<uint64_t> versions;
  uint32_t tiled_block = 0;
  // schedule: members are packed into groups of at most group_bytes of
  // population storage (0 = one group), and each group advances
  // steps_per_group steps before the next group runs, so a group that fits
  // the 256 MB Infinity Cache is re-read from it instead of from HBM
  uint64_t group_bytes = 0;
  uint32_t steps_per_group = 1;
  uint32_t ticket = 0;               // 1: last block advances the offsets; 0: separate kernel
  uint64_t grouped_bytes = ~0ull;
  std::vector<uint32_t> order;       // member index of each descriptor slot
  std::vector<GroupArgs> groups;
  uint32_t* counters_dev = nullptr;
  uint32_t counters_capacity = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture_stream = nullptr;
  uint32_t graph_steps = 0;
};

namespace {

size_t lds_for(uint32_t block_threads, uint32_t waves_per_simd)
{
  if (waves_per_simd == 0) {
    return 0;
  }
  const uint32_t waves_per_block = block_threads / 64;
  const uint32_t blocks_per_cu = std::max(1u, 4 * waves_per_simd / waves_per_block);
  uint32_t lds = kLdsPerCu / blocks_per_cu / 1024 * 1024;
  if (lds * (blocks_per_cu + 1) <= kLdsPerCu) {
    lds += 1024;
  }
  return std::min(lds, kLdsPerCu - 4096);
}

// host mirror of the offsets after `steps` more steps
void advance_host_shifts(olbgfx_lattice* L, uint64_t steps)
{
  const uint64_t k = steps % L->count;
  for (int i = 0; i < Q; ++i) {
    const uint64_t add = (k * (uint64_t)L->dmod[i]) % L->count;
    const uint64_t s = (uint64_t)L->shift[i] + add;
    L->shift[i] = (uint32_t)(s >= L->count ? s - L->count : s);
  }
}

// (re)build the device descriptors and the tile list of the ensemble
template <typename S>
olbgfx_status ensemble_prepare(olbgfx_ensemble* E, bool force)
{
  using A = typename Storage<S>::A;
  const size_t m = E->members.size();
  for (olbgfx_lattice* L : E->members) {
    if (prepare_boundary(L) != OLBGFX_OK) {
      return OLBGFX_ERR_INVALID;
    }
    if (L->boundary_kinds) {
      return fail(OLBGFX_ERR_UNSUPPORTED, "ensembles do not take members with LocalVelocity boundaries");
    }
    if (L->bz_n > 0 || L->bzv_n > 0) {
      return fail(OLBGFX_ERR_UNSUPPORTED, "ensembles do not take members with Bouzidi boundaries");
    }
  }
  bool changed = force || E->tiled_block != E->block_threads || E->grouped_bytes != E->group_bytes;
  for (size_t k = 0; k < m; ++k) {
    changed = changed || E->versions[k] != E->members[k]->version;
  }
  for (size_t k = 0; k < m; ++k) {
    if (!E->members[k]->shift_dev_valid) {
      if (sync_shifts(E->members[k]) != OLBGFX_OK) {
        return OLBGFX_ERR_HIP;
      }
    }
  }
  if (!changed) {
    return OLBGFX_OK;
  }
  // groups: members in the given order, a new group whenever the next member
  // would push the group's population storage over group_bytes
  std::vector<std::vector<uint32_t>> packs;
  uint64_t bytes = 0;
  for (uint32_t k = 0; k < (uint32_t)m; ++k) {
    const olbgfx_lattice* L = E->members[k];
    const uint64_t b = (uint64_t)Q * L->stride * L->elem;
    if (packs.empty() || (E->group_bytes > 0 && bytes + b > E->group_bytes && !packs.back().empty())) {
      packs.emplace_back();
      bytes = 0;
    }
    packs.back().push_back(k);
    bytes += b;
  }
  std::vector<StepArgs<A>> args;
  std::vector<Tile> tiles;
  std::vector<uint32_t*> shift_ptrs;
  std::vector<uint32_t> dmod, counts;
  E->order.clear();
  E->groups.clear();
  for (const auto& pack : packs) {
    GroupArgs g{};
    g.tile_begin = (uint32_t)tiles.size();
    g.member_begin = (uint32_t)args.size();
    for (uint32_t k : pack) {
      olbgfx_lattice* L = E->members[k];
      const uint32_t slot = (uint32_t)args.size();
      args.emplace_back();
      fill_args(L, args.back(), true);
      args.back().uniform_id = -1;
      args.back().stats = nullptr;
      shift_ptrs.push_back(L->shift_dev);
      counts.push_back(L->count);
      for (int i = 0; i < Q; ++i) dmod.push_back((uint32_t)L->dmod[i]);
      for (uint32_t first = 0; first < args.back().ncore; first += E->block_threads) {
        tiles.push_back(Tile{slot, first});
      }
      E->order.push_back(k);
      E->versions[k] = L->version;
    }
    g.member_end = (uint32_t)args.size();
    g.ntiles = (uint32_t)tiles.size() - g.tile_begin;
    E->groups.push_back(g);
  }
  if (!E->args_dev) {
    HIP_TRY(hipMalloc(&E->args_dev, sizeof(StepArgs<A>) * m));
    HIP_TRY(hipMalloc(&E->shift_ptrs_dev, sizeof(uint32_t*) * m));
    HIP_TRY(hipMalloc(&E->dmod_dev, sizeof(uint32_t) * m * Q));
    HIP_TRY(hipMalloc(&E->counts_dev, sizeof(uint32_t) * m));
  }
  if (tiles.size() > E->tiles_capacity) {
    if (E->tiles_dev) {
      HIP_TRY(hipFree(E->tiles_dev));
    }
    HIP_TRY(hipMalloc(&E->tiles_dev, sizeof(Tile) * tiles.size()));
    E->tiles_capacity = (uint32_t)tiles.size();
  }
  if (E->groups.size() > E->counters_capacity) {
    if (E->counters_dev) {
      HIP_TRY(hipFree(E->counters_dev));
    }
    HIP_TRY(hipMalloc(&E->counters_dev, sizeof(uint32_t) * E->groups.size()));
    E->counters_capacity = (uint32_t)E->groups.size();
  }
  // counters start at zero; the kernels leave them at zero (no launch of
  // this ensemble may be in flight while they are reset)
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemset(E->counters_dev, 0, sizeof(uint32_t) * E->groups.size()));
  for (size_t gi = 0; gi < E->groups.size(); ++gi) {
    E->groups[gi].counter = E->counters_dev + gi;
  }
  HIP_TRY(hipMemcpy(E->args_dev, args.data(), sizeof(StepArgs<A>) * m, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(E->tiles_dev, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(E->shift_ptrs_dev, shift_ptrs.data(), sizeof(uint32_t*) * m, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(E->dmod_dev, dmod.data(), sizeof(uint32_t) * m * Q, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(E->counts_dev, counts.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice));
  E->ntiles = (uint32_t)tiles.size();
  E->tiled_block = E->block_threads;
  E->grouped_bytes = E->group_bytes;
  return OLBGFX_OK;
}

// enqueue `steps` ensemble steps: in rounds of steps_per_group steps, each
// group advances that many steps (one persistent launch per step) before the
// next group runs
template <typename S>
olbgfx_status ensemble_enqueue(olbgfx_ensemble* E, uint32_t steps, hipStream_t stream)
{
  using A = typename Storage<S>::A;
  uint32_t done = 0;
  while (done < steps) {
    const uint32_t k = std::min(E->steps_per_group, steps - done);
    for (GroupArgs g : E->groups) {
      // at most kMaxEnsembleBlocks persistent blocks per launch
      g.tiles_per_block = std::max(E->tiles_per_block,
                                   (g.ntiles + kMaxEnsembleBlocks - 1) / kMaxEnsembleBlocks);
      g.nblocks = (g.ntiles + g.tiles_per_block - 1) / g.tiles_per_block;
      g.blocks_per_xcd = (g.nblocks + 7) / 8;
      g.xcd_run_log2 = E->xcd_run_log2;
      if (g.xcd_run_log2 >= 0) {
        const uint32_t run = 1u << g.xcd_run_log2;
        g.blocks_per_xcd = (g.blocks_per_xcd + run - 1) / run * run;
      }
      g.ticket = E->ticket;
      const dim3 grid(8 * g.blocks_per_xcd), block(E->block_threads);
      const uint32_t nm = (g.member_end - g.member_begin) * Q;
      for (uint32_t j = 0; j < k; ++j) {
        ensemble_kernel<S><<<grid, block, E->lds, stream>>>((const StepArgs<A>*)E->args_dev,
                                                            E->tiles_dev, g, E->shift_ptrs_dev,
                                                            E->dmod_dev, E->counts_dev);
        if (!E->ticket) {
          advance_shifts<<<(nm + 255) / 256, 256, 0, stream>>>(E->shift_ptrs_dev, E->dmod_dev,
                                                               E->counts_dev, g.member_begin,
                                                               g.member_end);
        }
      }
    }
    done += k;
  }
  HIP_TRY(hipGetLastError());
  return OLBGFX_OK;
}

template <typename F>
olbgfx_status dispatch_ensemble(olbgfx_ensemble* E, F&& f)
{
  switch (E->storage) {
  case OLBGFX_STORE_F32:
    return f((float*)nullptr);
  case OLBGFX_STORE_F16:
    return f((f16_t*)nullptr);
  case OLBGFX_STORE_BF16:
    return f((bf16_t*)nullptr);
  case OLBGFX_STORE_F64:
    return f((double*)nullptr);
  default:
    return fail(OLBGFX_ERR_INVALID, "unknown storage type");
  }
}

void drop_graph(olbgfx_ensemble* E)
{
  if (E->exec) {
    (void)hipGraphExecDestroy(E->exec);
    E->exec = nullptr;
  }
  if (E->graph) {
    (void)hipGraphDestroy(E->graph);
    E->graph = nullptr;
  }
  E->graph_steps = 0;
}

}  // namespace

extern "C" {

olbgfx_status olbgfx_ensemble_create(olbgfx_lattice* const* members, uint32_t count,
                                     olbgfx_ensemble** out)
{
  if (!members || !out || count == 0) {
    return fail(OLBGFX_ERR_INVALID, "an ensemble needs at least one member");
  }
  *out = nullptr;
  const int32_t storage = members[0] ? members[0]->desc.storage : -1;
  for (uint32_t k = 0; k < count; ++k) {
    if (!members[k]) {
      return fail(OLBGFX_ERR_INVALID, "null member lattice");
    }
    if (members[k]->desc.storage != storage) {
      return fail(OLBGFX_ERR_INVALID, "all members of an ensemble need the same storage type");
    }
  }
  olbgfx_ensemble* E = new (std::nothrow) olbgfx_ensemble();
  if (!E) {
    return fail(OLBGFX_ERR_NOMEM, "host allocation failed");
  }
  E->members.assign(members, members + count);
  E->versions.assign(count, ~0ull);
  E->storage = storage;
  *out = E;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_ensemble_destroy(olbgfx_ensemble* E)
{
  if (!E) {
    return OLBGFX_OK;
  }
  (void)hipDeviceSynchronize();
  drop_graph(E);
  if (E->capture_stream) {
    (void)hipStreamDestroy(E->capture_stream);
  }
  (void)hipFree(E->args_dev);
  (void)hipFree(E->tiles_dev);
  (void)hipFree(E->shift_ptrs_dev);
  (void)hipFree(E->dmod_dev);
  (void)hipFree(E->counts_dev);
  (void)hipFree(E->counters_dev);
  delete E;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_ensemble_set_launch(olbgfx_ensemble* E, uint32_t block_threads,
                                         uint32_t waves_per_simd, uint32_t tiles_per_block)
{
  if (!E) {
    return fail(OLBGFX_ERR_INVALID, "null ensemble");
  }
  if (block_threads == 0) {
    block_threads = 256;
  }
  if (block_threads < 64 || block_threads > kMaxBlock || block_threads % 64 != 0) {
    return fail(OLBGFX_ERR_INVALID, "block size must be a multiple of 64 in [64, 256]");
  }
  if (waves_per_simd > 8 || tiles_per_block == 0 || tiles_per_block > 4096) {
    return fail(OLBGFX_ERR_INVALID, "waves per SIMD in [0, 8], tiles per block in [1, 4096]");
  }
  E->block_threads = block_threads;
  E->waves_per_simd = waves_per_simd;
  E->tiles_per_block = tiles_per_block;
  E->lds = lds_for(block_threads, waves_per_simd);
  drop_graph(E);
  return OLBGFX_OK;
}

olbgfx_status olbgfx_ensemble_set_block_order(olbgfx_ensemble* E, int32_t xcd_run_log2)
{
  if (!E) {
    return fail(OLBGFX_ERR_INVALID, "null ensemble");
  }
  if (xcd_run_log2 < -1 || xcd_run_log2 > 12) {
    return fail(OLBGFX_ERR_INVALID, "block order must be -1 or a run length 2^k with k in [0, 12]");
  }
  E->xcd_run_log2 = xcd_run_log2;
  drop_graph(E);
  return OLBGFX_OK;
}

olbgfx_status olbgfx_ensemble_set_schedule(olbgfx_ensemble* E, uint64_t group_bytes,
                                           uint32_t steps_per_group, int32_t ticket)
{
  if (!E) {
    return fail(OLBGFX_ERR_INVALID, "null ensemble");
  }
  if (steps_per_group == 0 || steps_per_group > 1024) {
    return fail(OLBGFX_ERR_INVALID, "steps per group must be in [1, 1024]");
  }
  E->group_bytes = group_bytes;
  E->steps_per_group = steps_per_group;
  E->ticket = ticket ? 1 : 0;
  drop_graph(E);
  return OLBGFX_OK;
}

olbgfx_status olbgfx_ensemble_step(olbgfx_ensemble* E, uint32_t steps, void* stream)
{
  if (!E) {
    return fail(OLBGFX_ERR_INVALID, "null ensemble");
  }
  hipStream_t s = (hipStream_t)stream;
  olbgfx_status st = dispatch_ensemble(E, [&](auto* tag) -> olbgfx_status {
    using S = std::remove_pointer_t<decltype(tag)>;
    if (const olbgfx_status p = ensemble_prepare<S>(E, false); p != OLBGFX_OK) {
      return p;
    }
    return ensemble_enqueue<S>(E, steps, s);
  });
  if (st == OLBGFX_OK) {
    for (olbgfx_lattice* L : E->members) {
      advance_host_shifts(L, steps);
      L->stats_valid = false;
    }
  }
  return st;
}

olbgfx_status olbgfx_ensemble_step_separately(olbgfx_ensemble* E, uint32_t steps, void* stream)
{
  if (!E) {
    return fail(OLBGFX_ERR_INVALID, "null ensemble");
  }
  for (uint32_t k = 0; k < steps; ++k) {
    for (olbgfx_lattice* L : E->members) {
      const olbgfx_status st = olbgfx_collide_and_stream(L, 1, stream);
      if (st != OLBGFX_OK) {
        return st;
      }
    }
  }
  return OLBGFX_OK;
}

olbgfx_status olbgfx_ensemble_capture(olbgfx_ensemble* E, uint32_t steps_per_launch)
{
  if (!E || steps_per_launch == 0) {
    return fail(OLBGFX_ERR_INVALID, "null ensemble or zero steps");
  }
  drop_graph(E);
  if (!E->capture_stream) {
    HIP_TRY(hipStreamCreateWithFlags(&E->capture_stream, hipStreamNonBlocking));
  }
  olbgfx_status st = dispatch_ensemble(E, [&](auto* tag) -> olbgfx_status {
    using S = std::remove_pointer_t<decltype(tag)>;
    if (const olbgfx_status p = ensemble_prepare<S>(E, false); p != OLBGFX_OK) {
      return p;
    }
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipStreamBeginCapture(E->capture_stream, hipStreamCaptureModeThreadLocal));
    const olbgfx_status q = ensemble_enqueue<S>(E, steps_per_launch, E->capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(E->capture_stream, &g);
    if (q != OLBGFX_OK) {
      if (g) {
        (void)hipGraphDestroy(g);
      }
      return q;
    }
    if (end != hipSuccess) {
      return fail(OLBGFX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(end));
    }
    E->graph = g;
    HIP_TRY(hipGraphInstantiate(&E->exec, E->graph, nullptr, nullptr, 0));
    return OLBGFX_OK;
  });
  if (st == OLBGFX_OK) {
    E->graph_steps = steps_per_launch;
  } else {
    drop_graph(E);
  }
  return st;
}

olbgfx_status olbgfx_ensemble_launch_graph(olbgfx_ensemble* E, uint32_t launches, void* stream)
{
  if (!E || !E->exec) {
    return fail(OLBGFX_ERR_INVALID, "no captured graph (call olbgfx_ensemble_capture)");
  }
  for (size_t k = 0; k < E->members.size(); ++k) {
    if (E->versions[k] != E->members[k]->version || !E->members[k]->shift_dev_valid) {
      return fail(OLBGFX_ERR_INVALID, "a member changed since the capture; capture again");
    }
  }
  hipStream_t s = (hipStream_t)stream;
  for (uint32_t k = 0; k < launches; ++k) {
    HIP_TRY(hipGraphLaunch(E->exec, s));
  }
  for (olbgfx_lattice* L : E->members) {
    advance_host_shifts(L, (uint64_t)launches * E->graph_steps);
    L->stats_valid = false;
  }
  return OLBGFX_OK;
}

olbgfx_status olbgfx_ensemble_cells(const olbgfx_ensemble* E, uint64_t* cells)
{
  if (!E || !cells) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  uint64_t c = 0;
  for (const olbgfx_lattice* L : E->members) {
    c += (uint64_t)L->desc.nx * L->desc.ny * L->desc.nz;
  }
  *cells = c;
  return OLBGFX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// timing helpers for the autotuner and benchmarks

namespace {

template <typename F>
olbgfx_status time_on_stream(hipStream_t s, F&& enqueue, float* ms)
{
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail(OLBGFX_ERR_HIP, "hipEventCreate failed");
  }
  olbgfx_status st = OLBGFX_OK;
  if (hipEventRecord(a, s) != hipSuccess) {
    st = fail(OLBGFX_ERR_HIP, "hipEventRecord failed");
  }
  if (st == OLBGFX_OK) {
    st = enqueue();
  }
  if (st == OLBGFX_OK &&
      (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
       hipEventElapsedTime(ms, a, b) != hipSuccess)) {
    st = fail(OLBGFX_ERR_HIP, "event timing failed");
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return st;
}

}  // namespace

extern "C" {

olbgfx_status olbgfx_time_steps(olbgfx_lattice* L, uint32_t steps, void* stream, float* ms)
{
  if (check(L) != OLBGFX_OK || !ms) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  hipStream_t s = (hipStream_t)stream;
  return time_on_stream(s, [&] { return olbgfx_collide_and_stream(L, steps, stream); }, ms);
}

olbgfx_status olbgfx_ensemble_time(olbgfx_ensemble* E, int32_t mode, uint32_t steps, void* stream,
                                   float* ms)
{
  if (!E || !ms) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  hipStream_t s = (hipStream_t)stream;
  switch (mode) {
  case 0:
    return time_on_stream(s, [&] { return olbgfx_ensemble_step(E, steps, stream); }, ms);
  case 1:
    if (!E->exec || E->graph_steps == 0 || steps % E->graph_steps != 0) {
      return fail(OLBGFX_ERR_INVALID, "graph mode needs a capture whose steps divide `steps`");
    }
    return time_on_stream(
        s, [&] { return olbgfx_ensemble_launch_graph(E, steps / E->graph_steps, stream); }, ms);
  case 2:
    return time_on_stream(s, [&] { return olbgfx_ensemble_step_separately(E, steps, stream); }, ms);
  case 3:
    // each member's step followed by a synchronisation, the way
    // SuperLattice::collideAndStream waits for the device between stages
    // (core/superLattice.hh:663-725, gpu/hip/operator.hh:597-624)
    return time_on_stream(
        s,
        [&]() -> olbgfx_status {
          for (uint32_t k = 0; k < steps; ++k) {
            for (olbgfx_lattice* L : E->members) {
              const olbgfx_status st = olbgfx_collide_and_stream(L, 1, stream);
              if (st != OLBGFX_OK) {
                return st;
              }
              HIP_TRY(hipStreamSynchronize(s));
            }
          }
          return OLBGFX_OK;
        },
        ms);
  default:
    return fail(OLBGFX_ERR_INVALID, "mode: 0 persistent, 1 graph, 2 separate launches, 3 separate and synchronised");
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// halo planes of a non-periodic face, for a block split over several GPUs
//
// After a step, the populations that left the core through a face sit in the
// ghost plane beyond it (the periodic images along the other axes included);
// the neighbouring block needs exactly these as the incoming populations of
// its own boundary plane, cell for cell.  pack gathers them from the ghost
// plane, unpack writes them into the boundary plane; for a face whose outward
// normal is +e_a the populations are those with c_a = +1 on both sides.

namespace {

struct HaloArgs {
  uint32_t shift[5];      // ring offsets of the five populations
  uint32_t pops[5];       // their indices
  int64_t base;           // CellID of plane cell (0, 0)
  int64_t step_u, step_v; // CellID strides of the two in-plane axes
  uint32_t ext_v;         // cells along v
  uint32_t plane;         // cells per plane
};

template <typename S, bool PACK>
__global__ void __launch_bounds__(256) k_halo(void* pop, uint64_t stride, uint32_t count, const HaloArgs h,
                                              S* buf)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 5ull * h.plane) {
    return;
  }
  const uint32_t k = (uint32_t)(t / h.plane), c = (uint32_t)(t - (uint64_t)k * h.plane);
  const uint32_t u = c / h.ext_v, v = c - u * h.ext_v;
  const int64_t cell = h.base + (int64_t)u * h.step_u + (int64_t)v * h.step_v;
  S* at = ring_at<S>(pop, stride, (int)h.pops[k], ring_slot((uint32_t)cell, h.shift[k], count));
  if constexpr (PACK) {
    buf[t] = *at;
  } else {
    *at = buf[t];
  }
}

// plane of the face (axis, side): the ghost plane for packing, the core
// boundary plane for unpacking
olbgfx_status halo_args(const olbgfx_lattice* L, int32_t axis, int32_t side, bool pack, HaloArgs& h)
{
  const olbgfx_lattice_desc& D = L->desc;
  if (axis < 0 || axis > 2 || (side != 0 && side != 1)) {
    return fail(OLBGFX_ERR_INVALID, "halo face: axis in [0, 2], side 0 (low) or 1 (high)");
  }
  if (D.periodic[axis]) {
    return fail(OLBGFX_ERR_INVALID, "halo faces exist only on non-periodic axes");
  }
  const int32_t ext[3] = {D.nx, D.ny, D.nz};
  const int64_t cs[3] = {(int64_t)L->py * L->pz, (int64_t)L->pz, 1};
  const int ua = axis == 0 ? 1 : 0, va = axis == 2 ? 1 : 2;
  // pack reads the ghost plane beyond the face, unpack the core plane inside it
  const int32_t pos = pack ? (side ? ext[axis] : -1) : (side ? ext[axis] - 1 : 0);
  int32_t at[3] = {0, 0, 0};
  at[axis] = pos;
  h.base = 0;
  for (int a = 0; a < 3; ++a) {
    h.base += (int64_t)(at[a] + D.padding) * cs[a];
  }
  h.step_u = cs[ua];
  h.step_v = cs[va];
  h.ext_v = (uint32_t)ext[va];
  h.plane = (uint32_t)ext[ua] * (uint32_t)ext[va];
  // populations crossing the face: c_a = +1 for the high face when packing
  // and for the low face when unpacking, c_a = -1 otherwise
  const int want = (pack == (side == 1)) ? 1 : -1;
  int k = 0;
  for (int i = 1; i < Q; ++i) {
    const int c[3] = {CX(i), CY(i), CZ(i)};
    if (c[axis] == want) {
      h.pops[k] = (uint32_t)i;
      h.shift[k] = L->shift[i];
      ++k;
    }
  }
  return OLBGFX_OK;
}

template <typename S>
olbgfx_status halo_launch(olbgfx_lattice* L, const HaloArgs& h, bool pack, void* buf, hipStream_t s)
{
  const uint64_t n = 5ull * h.plane;
  if (n == 0) {
    return OLBGFX_OK;
  }
  const dim3 grid((uint32_t)((n + 255) / 256)), block(256);
  if (pack) {
    k_halo<S, true><<<grid, block, 0, s>>>(L->pop, L->stride, L->count, h, (S*)buf);
  } else {
    k_halo<S, false><<<grid, block, 0, s>>>(L->pop, L->stride, L->count, h, (S*)buf);
  }
  HIP_TRY(hipGetLastError());
  return OLBGFX_OK;
}

olbgfx_status halo(olbgfx_lattice* L, int32_t axis, int32_t side, void* buf, void* stream, bool pack)
{
  if (check(L) != OLBGFX_OK || !buf) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  HaloArgs h{};
  if (halo_args(L, axis, side, pack, h) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  hipStream_t s = (hipStream_t)stream;
  switch (L->desc.storage) {
  case OLBGFX_STORE_F32:
    return halo_launch<float>(L, h, pack, buf, s);
  case OLBGFX_STORE_F64:
    return halo_launch<double>(L, h, pack, buf, s);
  default:  // F16 and BF16 move as raw 16-bit words
    return halo_launch<uint16_t>(L, h, pack, buf, s);
  }
}

}  // namespace

extern "C" {

olbgfx_status olbgfx_halo_bytes(const olbgfx_lattice* L, int32_t axis, uint64_t* bytes)
{
  if (!L || !bytes || axis < 0 || axis > 2) {
    return fail(OLBGFX_ERR_INVALID, "null argument or axis outside [0, 2]");
  }
  const int32_t ext[3] = {L->desc.nx, L->desc.ny, L->desc.nz};
  const int ua = axis == 0 ? 1 : 0, va = axis == 2 ? 1 : 2;
  *bytes = 5ull * (uint64_t)ext[ua] * (uint64_t)ext[va] * (uint64_t)L->elem;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_halo_buffer_create(const olbgfx_lattice* L, int32_t axis, void** device_buffer)
{
  uint64_t bytes = 0;
  if (!device_buffer || olbgfx_halo_bytes(L, axis, &bytes) != OLBGFX_OK) {
    return fail(OLBGFX_ERR_INVALID, "null argument or axis outside [0, 2]");
  }
  *device_buffer = nullptr;
  if (hipMalloc(device_buffer, std::max<uint64_t>(bytes, 256)) != hipSuccess) {
    (void)hipGetLastError();
    *device_buffer = nullptr;
    return fail(OLBGFX_ERR_NOMEM, "halo buffer allocation failed");
  }
  return OLBGFX_OK;
}

olbgfx_status olbgfx_halo_buffer_destroy(void* device_buffer)
{
  if (device_buffer) {
    HIP_TRY(hipFree(device_buffer));
  }
  return OLBGFX_OK;
}

olbgfx_status olbgfx_halo_pack(olbgfx_lattice* L, int32_t axis, int32_t side, void* device_buffer,
                               void* stream)
{
  return halo(L, axis, side, device_buffer, stream, true);
}

olbgfx_status olbgfx_halo_unpack(olbgfx_lattice* L, int32_t axis, int32_t side, const void* device_buffer,
                                 void* stream)
{
  return halo(L, axis, side, const_cast<void*>(device_buffer), stream, false);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// cell lists: all 19 populations of listed cells to and from a device buffer
// (the data a communicator moves for an overlap of arbitrary shape)

struct olbgfx_cell_list {
  uint32_t* cells = nullptr;  // device copy of the CellIDs
  uint32_t n = 0;
  uint32_t count = 0;         // padded cells of the lattice the list was made for
};

namespace {

struct CellCopyArgs {
  uint32_t shift[Q];
  uint32_t count;
  uint32_t n;
};

template <typename S, bool GATHER>
__global__ void __launch_bounds__(256) k_cells(void* pop, uint64_t stride, const CellCopyArgs c,
                                               const uint32_t* cells, S* buf)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)Q * c.n) {
    return;
  }
  const uint32_t i = (uint32_t)(t / c.n), k = (uint32_t)(t - (uint64_t)i * c.n);
  const uint32_t cell = cells[k];
  if (cell >= c.count) {
    return;  // lists are checked when made; a foreign list cannot reach outside the rings
  }
  S* at = ring_at<S>(pop, stride, (int)i, ring_slot(cell, c.shift[i], c.count));
  if constexpr (GATHER) {
    buf[t] = *at;
  } else {
    *at = buf[t];
  }
}

template <typename S>
void cells_launch(olbgfx_lattice* L, const olbgfx_cell_list* list, void* buf, bool gather, hipStream_t s)
{
  CellCopyArgs c{};
  for (int i = 0; i < Q; ++i) c.shift[i] = L->shift[i];
  c.count = L->count;
  c.n = list->n;
  const uint64_t total = (uint64_t)Q * list->n;
  const dim3 grid((uint32_t)((total + 255) / 256)), block(256);
  if (gather) {
    k_cells<S, true><<<grid, block, 0, s>>>(L->pop, L->stride, c, list->cells, (S*)buf);
  } else {
    k_cells<S, false><<<grid, block, 0, s>>>(L->pop, L->stride, c, list->cells, (S*)buf);
  }
}

olbgfx_status cells_copy(olbgfx_lattice* L, const olbgfx_cell_list* list, void* buf, void* stream, bool gather)
{
  if (check(L) != OLBGFX_OK || !list || !buf) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  if (list->count != L->count) {
    return fail(OLBGFX_ERR_INVALID, "the cell list was made for a lattice of another size");
  }
  if (list->n == 0) {
    return OLBGFX_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  switch (L->desc.storage) {
  case OLBGFX_STORE_F32:
    cells_launch<float>(L, list, buf, gather, s);
    break;
  case OLBGFX_STORE_F64:
    cells_launch<double>(L, list, buf, gather, s);
    break;
  default:
    cells_launch<uint16_t>(L, list, buf, gather, s);
    break;
  }
  HIP_TRY(hipGetLastError());
  return OLBGFX_OK;
}

}  // namespace

extern "C" {

olbgfx_status olbgfx_cell_list_create(const olbgfx_lattice* L, const uint32_t* cells, uint32_t n,
                                      olbgfx_cell_list** out)
{
  if (!L || !out || (n > 0 && !cells)) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  *out = nullptr;
  for (uint32_t k = 0; k < n; ++k) {
    if (cells[k] >= L->count) {
      return fail(OLBGFX_ERR_INVALID, "cell id outside the padded block");
    }
  }
  olbgfx_cell_list* list = new (std::nothrow) olbgfx_cell_list();
  if (!list) {
    return fail(OLBGFX_ERR_NOMEM, "out of host memory");
  }
  list->n = n;
  list->count = L->count;
  if (n > 0) {
    if (hipMalloc(&list->cells, sizeof(uint32_t) * n) != hipSuccess) {
      (void)hipGetLastError();
      delete list;
      return fail(OLBGFX_ERR_NOMEM, "cell list allocation failed");
    }
    if (hipMemcpy(list->cells, cells, sizeof(uint32_t) * n, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(list->cells);
      delete list;
      return fail(OLBGFX_ERR_HIP, "cell list upload failed");
    }
  }
  *out = list;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_cell_list_destroy(olbgfx_cell_list* list)
{
  if (list) {
    if (list->cells) {
      HIP_TRY(hipFree(list->cells));
    }
    delete list;
  }
  return OLBGFX_OK;
}

olbgfx_status olbgfx_cell_list_bytes(const olbgfx_lattice* L, const olbgfx_cell_list* list, uint64_t* bytes)
{
  if (!L || !list || !bytes) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  *bytes = (uint64_t)Q * list->n * (uint64_t)L->elem;
  return OLBGFX_OK;
}

olbgfx_status olbgfx_gather_cells(olbgfx_lattice* L, const olbgfx_cell_list* list, void* device_buffer,
                                  void* stream)
{
  return cells_copy(L, list, device_buffer, stream, true);
}

olbgfx_status olbgfx_scatter_cells(olbgfx_lattice* L, const olbgfx_cell_list* list, const void* device_buffer,
                                   void* stream)
{
  return cells_copy(L, list, const_cast<void*>(device_buffer), stream, false);
}

olbgfx_status olbgfx_device_buffer_create(uint64_t bytes, void** device_buffer)
{
  if (!device_buffer) {
    return fail(OLBGFX_ERR_INVALID, "null argument");
  }
  *device_buffer = nullptr;
  if (hipMalloc(device_buffer, std::max<uint64_t>(bytes, 256)) != hipSuccess) {
    (void)hipGetLastError();
    *device_buffer = nullptr;
    return fail(OLBGFX_ERR_NOMEM, "device buffer allocation failed");
  }
  return OLBGFX_OK;
}

olbgfx_status olbgfx_device_buffer_destroy(void* device_buffer)
{
  if (device_buffer) {
    HIP_TRY(hipFree(device_buffer));
  }
  return OLBGFX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// autotuner: candidate launch configurations timed on the lattice or ensemble
// itself; the populations, offsets and statistics are saved before the trials
// and put back after them, so tuning does not advance the simulation

namespace {

struct SavedState {
  olbgfx_lattice* L = nullptr;
  void* pop = nullptr;
  uint32_t shift[Q] = {};
  double stats[4] = {};
  bool stats_valid = false;
};

void release(std::vector<SavedState>& saved)
{
  for (SavedState& s : saved) {
    if (s.pop) {
      (void)hipFree(s.pop);
      s.pop = nullptr;
    }
  }
}

olbgfx_status save_state(olbgfx_lattice* L, hipStream_t stream, SavedState& out)
{
  out.L = L;
  const size_t bytes = (size_t)Q * L->stride * L->elem;
  if (hipMalloc(&out.pop, bytes) != hipSuccess) {
    (void)hipGetLastError();
    out.pop = nullptr;
    return fail(OLBGFX_ERR_NOMEM, "not enough device memory to keep the lattice state while tuning");
  }
  HIP_TRY(hipMemcpyAsync(out.pop, L->pop, bytes, hipMemcpyDeviceToDevice, stream));
  HIP_TRY(hipMemcpyAsync(out.stats, L->stats_out, sizeof(out.stats), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  std::memcpy(out.shift, L->shift, sizeof(out.shift));
  out.stats_valid = L->stats_valid;
  return OLBGFX_OK;
}

olbgfx_status restore_state(const SavedState& s, hipStream_t stream)
{
  olbgfx_lattice* L = s.L;
  const size_t bytes = (size_t)Q * L->stride * L->elem;
  HIP_TRY(hipMemcpyAsync(L->pop, s.pop, bytes, hipMemcpyDeviceToDevice, stream));
  HIP_TRY(hipMemcpyAsync(L->stats_out, s.stats, sizeof(s.stats), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  std::memcpy(L->shift, s.shift, sizeof(L->shift));
  L->stats_valid = s.stats_valid;
  return sync_shifts(L);
}

uint64_t fnv1a(uint64_t h, const void* data, size_t n)
{
  const unsigned char* p = (const unsigned char*)data;
  for (size_t i = 0; i < n; ++i) {
    h = (h ^ p[i]) * 1099511628211ull;
  }
  return h;
}

// extent, padding, periodicity, storage and dynamics mix of a lattice
std::string tune_key(const olbgfx_lattice* L)
{
  uint64_t h = 1469598103934665603ull;
  h = fnv1a(h, &L->uniform_id, sizeof(L->uniform_id));
  if (L->uniform_id < 0) {
    h = fnv1a(h, L->host_mask.data(), L->host_mask.size());
  }
  for (int k = 0; k < OLBGFX_MAX_DYNAMICS; ++k) {
    h = fnv1a(h, &L->table[k].kind, sizeof(L->table[k].kind));
  }
  const olbgfx_lattice_desc& D = L->desc;
  char buf[160];
  std::snprintf(buf, sizeof(buf), "lattice:%ux%ux%u:p%u:%d%d%d:s%d:%016llx", D.nx, D.ny, D.nz, D.padding,
                D.periodic[0] ? 1 : 0, D.periodic[1] ? 1 : 0, D.periodic[2] ? 1 : 0, D.storage,
                (unsigned long long)h);
  return buf;
}

std::string tune_key(const olbgfx_ensemble* E)
{
  uint64_t h = 1469598103934665603ull;
  for (const olbgfx_lattice* L : E->members) {
    const std::string k = tune_key(L);
    h = fnv1a(h, k.data(), k.size());
  }
  char buf[96];
  std::snprintf(buf, sizeof(buf), "ensemble:%zu:s%d:%016llx", E->members.size(), E->storage,
                (unsigned long long)h);
  return buf;
}

// cache file: one line per key, "key block waves order tiles group_bytes
// steps_per_group ticket ms"
bool cache_lookup(const char* path, const std::string& key, olbgfx_tune_trial& t)
{
  if (!path || !*path) {
    return false;
  }
  FILE* f = std::fopen(path, "r");
  if (!f) {
    return false;
  }
  char line[512];
  bool hit = false;
  while (std::fgets(line, sizeof(line), f)) {
    char k[256];
    unsigned long long gb = 0;
    olbgfx_tune_trial r{};
    if (std::sscanf(line, "%255s %u %u %d %u %llu %u %d %lf", k, &r.block_threads, &r.waves_per_simd,
                    &r.block_order, &r.tiles_per_block, &gb, &r.steps_per_group, &r.ticket,
                    &r.ms_per_step) == 9 &&
        key == k) {
      r.group_bytes = gb;
      t = r;
      hit = true;
    }
  }
  std::fclose(f);
  return hit;
}

void cache_store(const char* path, const std::string& key, const olbgfx_tune_trial& t)
{
  if (!path || !*path) {
    return;
  }
  std::vector<std::string> keep;
  if (FILE* f = std::fopen(path, "r")) {
    char line[512];
    while (std::fgets(line, sizeof(line), f)) {
      char k[256];
      if (std::sscanf(line, "%255s", k) == 1 && key != k) {
        keep.emplace_back(line);
      }
    }
    std::fclose(f);
  }
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) {
    return;  // an unwritable cache only costs the next run its trials
  }
  for (const std::string& l : keep) {
    std::fputs(l.c_str(), f);
  }
  std::fprintf(f, "%s %u %u %d %u %llu %u %d %.6f\n", key.c_str(), t.block_threads, t.waves_per_simd,
               t.block_order, t.tiles_per_block, (unsigned long long)t.group_bytes, t.steps_per_group,
               t.ticket, t.ms_per_step);
  const bool ok = std::fclose(f) == 0;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
  }
}

template <typename V, size_t N, size_t M>
std::vector<V> choices(uint32_t n, const V (&given)[N], const V (&defaults)[M])
{
  if (n == 0) {
    return std::vector<V>(defaults, defaults + M);
  }
  return std::vector<V>(given, given + std::min<size_t>(n, N));
}

struct Trials {
  std::vector<olbgfx_tune_trial> all;
  size_t best = 0;
  void add(const olbgfx_tune_trial& t)
  {
    all.push_back(t);
    if (t.ms_per_step < all[best].ms_per_step) {
      best = all.size() - 1;
    }
  }
};

std::vector<olbgfx_tune_trial> fastest(std::vector<olbgfx_tune_trial> all, size_t n)
{
  std::stable_sort(all.begin(), all.end(), [](const olbgfx_tune_trial& x, const olbgfx_tune_trial& y) {
    return x.ms_per_step < y.ms_per_step;
  });
  all.resize(std::min(n, all.size()));
  return all;
}

void report(const Trials& tr, bool cached, const olbgfx_tune_trial& best, olbgfx_tune_result* out)
{
  if (!out) {
    return;
  }
  std::memset(out, 0, sizeof(*out));
  out->best = best;
  out->cached = cached ? 1 : 0;
  out->ntrials = (uint32_t)std::min<size_t>(tr.all.size(), OLBGFX_TUNE_TRIALS);
  for (uint32_t k = 0; k < out->ntrials; ++k) {
    out->trials[k] = tr.all[k];
  }
}

}  // namespace

extern "C" {

olbgfx_status olbgfx_autotune_lattice(olbgfx_lattice* L, uint32_t steps, const olbgfx_tune_space* space,
                                      const char* cache_path, void* stream, olbgfx_tune_result* out)
{
  if (check(L) != OLBGFX_OK) {
    return OLBGFX_ERR_INVALID;
  }
  if (steps == 0) {
    return fail(OLBGFX_ERR_INVALID, "the autotuner needs at least one timed step per trial");
  }
  static const uint32_t kBlocks[] = {64, 128, 256};
  static const uint32_t kWaves[] = {0, 2, 3, 4, 5, 6};
  static const int32_t kOrders[] = {-1, 0, 2, 3, 4, 5, 6, 8};
  const olbgfx_tune_space none{};
  const olbgfx_tune_space& sp = space ? *space : none;
  const std::vector<uint32_t> blocks = choices(sp.n_block_threads, sp.block_threads, kBlocks);
  const std::vector<uint32_t> waves = choices(sp.n_waves, sp.waves_per_simd, kWaves);
  const std::vector<int32_t> orders = choices(sp.n_orders, sp.block_order, kOrders);
  const std::string key = tune_key(L);
  Trials tr;
  olbgfx_tune_trial best{};
  if (cache_lookup(cache_path, key, best)) {
    olbgfx_status st = olbgfx_set_launch(L, best.block_threads, best.waves_per_simd);
    if (st == OLBGFX_OK) {
      st = olbgfx_set_block_order(L, best.block_order);
    }
    if (st == OLBGFX_OK) {
      report(tr, true, best, out);
    }
    return st;
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<SavedState> saved(1);
  olbgfx_status st = save_state(L, s, saved[0]);
  auto trial = [&](uint32_t b, uint32_t w, int32_t o) -> olbgfx_status {
    olbgfx_status r = olbgfx_set_launch(L, b, w);
    if (r == OLBGFX_OK) r = olbgfx_set_block_order(L, o);
    if (r == OLBGFX_OK) r = olbgfx_collide_and_stream(L, 1, stream);
    float ms = 0;
    if (r == OLBGFX_OK) r = olbgfx_time_steps(L, steps, stream, &ms);
    if (r == OLBGFX_OK) {
      olbgfx_tune_trial t{};
      t.block_threads = b;
      t.waves_per_simd = w;
      t.block_order = o;
      t.tiles_per_block = 1;
      t.steps_per_group = 1;
      t.ms_per_step = (double)ms / steps;
      tr.add(t);
    }
    return r;
  };
  for (size_t i = 0; i < blocks.size() && st == OLBGFX_OK; ++i) {
    for (size_t j = 0; j < waves.size() && st == OLBGFX_OK; ++j) {
      st = trial(blocks[i], waves[j], -1);
    }
  }
  // block orders for the three fastest shapes
  const std::vector<olbgfx_tune_trial> shapes = fastest(tr.all, 3);
  for (size_t m = 0; m < shapes.size() && st == OLBGFX_OK; ++m) {
    for (size_t k = 0; k < orders.size() && st == OLBGFX_OK; ++k) {
      if (orders[k] != -1) {
        st = trial(shapes[m].block_threads, shapes[m].waves_per_simd, orders[k]);
      }
    }
  }
  if (saved[0].pop) {
    const std::string err = st == OLBGFX_OK ? std::string() : g_error;
    const olbgfx_status rs = restore_state(saved[0], s);
    if (st == OLBGFX_OK) {
      st = rs;
    } else {
      g_error = err;
    }
  }
  release(saved);
  if (st != OLBGFX_OK) {
    return st;
  }
  if (tr.all.empty()) {
    return fail(OLBGFX_ERR_INVALID, "no candidate configurations");
  }
  best = tr.all[tr.best];
  st = olbgfx_set_launch(L, best.block_threads, best.waves_per_simd);
  if (st == OLBGFX_OK) st = olbgfx_set_block_order(L, best.block_order);
  if (st == OLBGFX_OK) {
    cache_store(cache_path, key, best);
    report(tr, false, best, out);
  }
  return st;
}

olbgfx_status olbgfx_autotune_ensemble(olbgfx_ensemble* E, uint32_t steps, const olbgfx_tune_space* space,
                                       const char* cache_path, void* stream, olbgfx_tune_result* out)
{
  if (!E || E->members.empty()) {
    return fail(OLBGFX_ERR_INVALID, "null or empty ensemble");
  }
  static const uint32_t kBlocks[] = {128, 256};
  static const uint32_t kWaves[] = {0, 3, 4, 5};
  static const uint32_t kTiles[] = {1, 2, 4};
  static const uint64_t kGroups[] = {0, 96ull << 20, 160ull << 20};
  static const uint32_t kPerGroup[] = {1, 4, 8};
  static const int32_t kTickets[] = {0, 1};
  const olbgfx_tune_space none{};
  const olbgfx_tune_space& sp = space ? *space : none;
  const std::vector<uint32_t> blocks = choices(sp.n_block_threads, sp.block_threads, kBlocks);
  const std::vector<uint32_t> waves = choices(sp.n_waves, sp.waves_per_simd, kWaves);
  const std::vector<uint32_t> tiles = choices(sp.n_tiles, sp.tiles_per_block, kTiles);
  const std::vector<uint64_t> groups = choices(sp.n_group_bytes, sp.group_bytes, kGroups);
  const std::vector<uint32_t> per_group = choices(sp.n_steps_per_group, sp.steps_per_group, kPerGroup);
  const std::vector<int32_t> tickets = choices(sp.n_tickets, sp.ticket, kTickets);
  static const int32_t kOrders[] = {-1, 0, 2, 3, 4, 5, 6, 8};
  const std::vector<int32_t> orders = choices(sp.n_orders, sp.block_order, kOrders);
  const std::string key = tune_key(E);
  Trials tr;
  olbgfx_tune_trial best{};
  auto apply = [&](const olbgfx_tune_trial& t) -> olbgfx_status {
    olbgfx_status r = olbgfx_ensemble_set_launch(E, t.block_threads, t.waves_per_simd, t.tiles_per_block);
    if (r == OLBGFX_OK) r = olbgfx_ensemble_set_block_order(E, t.block_order);
    if (r == OLBGFX_OK) r = olbgfx_ensemble_set_schedule(E, t.group_bytes, t.steps_per_group, t.ticket);
    return r;
  };
  if (cache_lookup(cache_path, key, best)) {
    const olbgfx_status st = apply(best);
    if (st == OLBGFX_OK) {
      report(tr, true, best, out);
    }
    return st;
  }
  hipStream_t s = (hipStream_t)stream;
  const uint32_t span = std::max(8u, steps);
  std::vector<SavedState> saved(E->members.size());
  olbgfx_status st = OLBGFX_OK;
  for (size_t k = 0; k < E->members.size() && st == OLBGFX_OK; ++k) {
    st = save_state(E->members[k], s, saved[k]);
  }
  auto trial = [&](const olbgfx_tune_trial& cand) -> olbgfx_status {
    olbgfx_status r = apply(cand);
    if (r == OLBGFX_OK) r = olbgfx_ensemble_step(E, cand.steps_per_group, stream);
    float ms = 0;
    if (r == OLBGFX_OK) r = olbgfx_ensemble_time(E, 0, span, stream, &ms);
    if (r == OLBGFX_OK) {
      olbgfx_tune_trial t = cand;
      t.ms_per_step = (double)ms / span;
      tr.add(t);
    }
    return r;
  };
  for (size_t i = 0; i < blocks.size() && st == OLBGFX_OK; ++i) {
    for (size_t j = 0; j < waves.size() && st == OLBGFX_OK; ++j) {
      for (size_t k = 0; k < tiles.size() && st == OLBGFX_OK; ++k) {
        olbgfx_tune_trial c{};
        c.block_threads = blocks[i];
        c.waves_per_simd = waves[j];
        c.block_order = -1;
        c.tiles_per_block = tiles[k];
        c.group_bytes = 0;
        c.steps_per_group = 1;
        c.ticket = 0;
        st = trial(c);
      }
    }
  }
  const std::vector<olbgfx_tune_trial> shapes = fastest(tr.all, 3);
  for (size_t m = 0; m < shapes.size() && st == OLBGFX_OK; ++m) {
    for (size_t k = 0; k < orders.size() && st == OLBGFX_OK; ++k) {
      if (orders[k] != -1) {
        olbgfx_tune_trial c = shapes[m];
        c.block_order = orders[k];
        st = trial(c);
      }
    }
  }
  if (st == OLBGFX_OK && !tr.all.empty()) {
    const olbgfx_tune_trial first = tr.all[tr.best];
    for (size_t i = 0; i < groups.size() && st == OLBGFX_OK; ++i) {
      for (size_t j = 0; j < per_group.size() && st == OLBGFX_OK; ++j) {
        for (size_t k = 0; k < tickets.size() && st == OLBGFX_OK; ++k) {
          if (groups[i] == 0 && per_group[j] == 1 && tickets[k] == 0) {
            continue;  // the launch phase already timed this schedule
          }
          olbgfx_tune_trial c = first;
          c.group_bytes = groups[i];
          c.steps_per_group = per_group[j];
          c.ticket = tickets[k] ? 1 : 0;
          st = trial(c);
        }
      }
    }
  }
  const std::string err = st == OLBGFX_OK ? std::string() : g_error;
  for (size_t k = 0; k < saved.size(); ++k) {
    if (saved[k].pop) {
      const olbgfx_status rs = restore_state(saved[k], s);
      if (st == OLBGFX_OK) {
        st = rs;
      }
    }
  }
  release(saved);
  if (st != OLBGFX_OK) {
    if (!err.empty()) {
      g_error = err;
    }
    return st;
  }
  if (tr.all.empty()) {
    return fail(OLBGFX_ERR_INVALID, "no candidate configurations");
  }
  best = tr.all[tr.best];
  st = apply(best);
  if (st == OLBGFX_OK) {
    cache_store(cache_path, key, best);
    report(tr, false, best, out);
  }
  return st;
}

}  // extern "C"
