/*
 * olb_oracle.c -- CPU restatement of OpenLB's D3Q19 collide-and-stream step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a test oracle: it is loaded by the
 * tests under tests/, by __graft_entry__.smoke() and by the cpu_baseline leg
 * of bench.py, and nowhere else.  The product library (libolbgfx) never links,
 * loads or calls it.
 *
 * The file is compiled twice (see oracle/Makefile):
 *   -DORC_DOUBLE  -> symbols orcd_*, value type double  (the FP64 oracle)
 *   (default)     -> symbols orcf_*, value type float   (the FP32 oracle; it
 *                    can also emulate FP16/BF16 population storage by rounding
 *                    every stored population, see orc_quantize)
 *
 * What it restates (reference = OpenLB 1.8r1, paths relative to its root):
 *   - D3Q19 lattice constants c, t, opposite:
 *       src/descriptor/definition/common.h:152-200
 *   - cell indexing CellID = ((x+p)*Py + (y+p))*Pz + (z+p) and neighbour
 *     distance d_i = c_i . (Py*Pz, Pz, 1):
 *       src/core/blockStructure.h:49-62, 150-182
 *   - periodic-shift (PS) population storage: logical cell n of population i
 *     lives at ring slot (n + shift_i) mod count, and streaming is
 *     shift_i -= d_i:
 *       src/core/platform/cpu/sisd/column.h:129-204,
 *       src/core/blockLattice.hh:316-338
 *   - one step = collide all core cells in place, copy post-collision
 *     populations into the width-1 periodic overlap, rotate the rings:
 *       src/core/superLattice.hh:663-725,
 *       src/core/platform/cpu/sisd/operator.h:143-228 (core cells only),
 *       src/communication/blockCommunicationNeighborhood.hh:67-107 (overlap 1)
 *   - densities and velocities: src/dynamics/lbm.h:289-330
 *     (rho = sum f_bar + 1 because populations are stored shifted by t_i)
 *   - second-order equilibrium (shifted): src/dynamics/lbm.h:51-64
 *   - BGK: src/dynamics/collision.h:57-80
 *   - TRT: src/dynamics/collision.h:381-419
 *   - Guo / PlainGuo forcing (velocity shifted by F/2, force term after the
 *     collision, uses omega even for TRT): src/dynamics/forcing.h:36-151,
 *     src/dynamics/momenta/elements.h:1593-1633,
 *     src/dynamics/lbm.h:567-588
 *   - BounceBack = Revert: src/dynamics/collision.h:465-484,
 *     src/dynamics/dynamics.h:557-568
 *   - BounceBackVelocity = Revert + Nguyen-Ladd correction with the per-cell
 *     RHO (FixedDensity) and VELOCITY fields:
 *       src/dynamics/collision.h:494-520, src/dynamics/dynamics.h:634-640,
 *       src/dynamics/momenta/elements.h:270-297, 797-830
 *   - LocalVelocity boundaries (boundary/localVelocity3D.h:36-120): flat
 *     faces collide as CombinedRLBdynamics<BGK, RegularizedVelocityBoundaryTuple>
 *     (dynamics/dynamics.h:476-526; density momenta/elements.h:55-71, 358-365;
 *     stress boundary/helper.h:31-81; fromPiToFneq dynamics/lbm.h:248-270),
 *     edges and corners as BGK with FixedVelocityBoundaryTuple
 *     (momenta/aliases.h:215-220) followed after streaming by
 *     OuterVelocityEdgeProcessor3D (priority 0) and
 *     OuterVelocityCornerProcessor3D (priority 1)
 *     (boundary/boundaryPostProcessors3D.hh:164-294; finite differences
 *     utilities/finiteDifference.h:35-45, finiteDifference3D.h:72-191)
 *   - Smagorinsky BGK (SmagorinskyBGKdynamics, dynamics/smagorinskyBGKdynamics.h:37-44):
 *     effective omega of dynamics/collisionLES.h:59-89 from the stress of
 *     dynamics/lbm.h:334-352 and its norm dynamics/lbm.h:555-564
 *   - lattice statistics (cells whose collision returns rho != -1 count):
 *       src/dynamics/interface.h:44-52, src/core/latticeStatistics.h:43-70
 *
 * Arithmetic follows the generic template formulas of the reference in the
 * order the reference writes them; it is built with -ffp-contract=off so every
 * operation is rounded as written.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORC_DOUBLE
typedef double real;
#define ORC(name) orcd_##name
#else
typedef float real;
#define ORC(name) orcf_##name
#endif

#define Q 19

enum {
  KIND_NONE = 0,
  KIND_BGK = 1,
  KIND_TRT = 2,
  KIND_BGK_GUO = 3,
  KIND_TRT_GUO = 4,
  KIND_BOUNCE_BACK = 5,
  KIND_BOUNCE_BACK_VELOCITY = 6,
  KIND_LOCAL_VELOCITY = 7,   /* flat LocalVelocity face: CombinedRLB<BGK, RegularizedVelocityBoundaryTuple> */
  KIND_FIXED_VELOCITY = 8,   /* LocalVelocity edge/corner: BGK<FixedVelocityBoundaryTuple> + post-processor */
  KIND_SMAGORINSKY_BGK = 9,  /* SmagorinskyBGKdynamics */
  KIND_INNER_EDGE = 10,      /* LocalVelocity inner edge: CombinedRLB<BGK, InnerEdgeVelocityTuple3D> */
  KIND_INNER_CORNER = 11,    /* LocalVelocity inner corner: CombinedRLB<BGK, InnerCornerVelocityTuple3D> */
  KIND_LOCAL_PRESSURE = 12,  /* LocalPressure face: CombinedRLB<BGK, RegularizedPressureBoundaryTuple> */
  KIND_INTERP_VELOCITY = 13, /* InterpolatedVelocity face: BGK<BasicDirichletVelocityBoundaryTuple> + PlaneFd */
  KIND_INTERP_PRESSURE = 14, /* InterpolatedPressure face: BGK<BasicDirichletPressureBoundaryTuple> + PlaneFd */
};

enum { STORE_NATIVE = 0, STORE_F16 = 1, STORE_BF16 = 2 };

static const int C[Q][3] = {
  { 0, 0, 0},
  {-1, 0, 0}, { 0,-1, 0}, { 0, 0,-1},
  {-1,-1, 0}, {-1, 1, 0}, {-1, 0,-1},
  {-1, 0, 1}, { 0,-1,-1}, { 0,-1, 1},
  { 1, 0, 0}, { 0, 1, 0}, { 0, 0, 1},
  { 1, 1, 0}, { 1,-1, 0}, { 1, 0, 1},
  { 1, 0,-1}, { 0, 1, 1}, { 0, 1,-1}
};
static const int TNUM[Q] = {1, 1,1,1, 1,1,1, 1,1,1, 1,1,1, 1,1,1, 1,1,1};
static const int TDEN[Q] = {3, 18,18,18, 36,36,36, 36,36,36,
                            18,18,18, 36,36,36, 36,36,36};

static int opp(int i) { return i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9); }

typedef struct {
  int kind;
  real omega;
  real magic;
  real force[3];
  real wall_u[3];
  real wall_rho;
  real smagorinsky;
} dyn_t;

typedef struct {
  int nx, ny, nz, pad;
  int periodic[3];
  int64_t size[3];      /* padded extent */
  int64_t proj[3];      /* (Py*Pz, Pz, 1) */
  uint64_t count;       /* ring length = number of padded cells */
  int64_t d[Q];         /* neighbour distance of population i */
  uint64_t shift[Q];    /* ring offset in [0, count) */
  real* ring[Q];
  uint8_t* mask;        /* dynamics id per cell (0 = NoDynamics) */
  dyn_t table[256];
  real* cell_velocity;  /* optional per-cell VELOCITY (count*3) */
  real* cell_rho;       /* optional per-cell RHO (count) */
  real* cell_force;     /* optional per-cell FORCE (count*3) */
  int8_t* normals;      /* optional per-cell discrete normal (count*3) */
  int64_t bz_n;         /* cells running BouzidiPostProcessor */
  int64_t* bz_cells;    /* their ids, in the order they are applied */
  real* bz_q;           /* their BOUZIDI_DISTANCE, 19 per cell */
  int64_t bzv_n;        /* cells running BouzidiVelocityPostProcessor */
  int64_t* bzv_cells;
  real* bzv_q;
  real* bzv_v;          /* their BOUZIDI_VELOCITY, 19 per cell */
  int store;
  int stats_on;
  double stats[4];      /* nCells, sum rho, sum uSqr, max uSqr (last step) */
  uint64_t time;
} lattice_t;

static real weight(int i) { return (real)TNUM[i] / (real)TDEN[i]; }

/* ---- storage rounding (FP16 / BF16 emulation, round to nearest even) ---- */

static uint16_t f32_to_f16_bits(float f)
{
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t absx = x & 0x7fffffffu;
  if (absx >= 0x7f800000u) {
    return (uint16_t)(sign | (absx > 0x7f800000u ? 0x7e00u : 0x7c00u));
  }
  if (absx >= 0x477ff000u) {           /* >= 65520 rounds to infinity */
    return (uint16_t)(sign | 0x7c00u);
  }
  if (absx < 0x38800000u) {            /* below 2^-14: half subnormal */
    float v;
    memcpy(&v, &absx, 4);
    float scaled = v * 16777216.0f;    /* exact: times 2^24 */
    uint32_t m = (uint32_t)nearbyintf(scaled);   /* ties to even */
    return (uint16_t)(sign | m);
  }
  uint32_t mant = absx & 0x7fffffu;
  uint32_t exp = ((absx >> 23) & 0xffu) - 127u + 15u;
  uint32_t h = (exp << 10) | (mant >> 13);
  uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) {
    h += 1u;
  }
  return (uint16_t)(sign | h);
}

static float f16_bits_to_f32(uint16_t h)
{
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t exp = ((uint32_t)h >> 10) & 0x1fu;
  uint32_t mant = (uint32_t)h & 0x3ffu;
  uint32_t x;
  if (exp == 0) {
    float v = (float)mant * 5.9604644775390625e-08f;   /* mant * 2^-24 */
    memcpy(&x, &v, 4);
    x |= sign;
  } else if (exp == 31) {
    x = sign | 0x7f800000u | (mant << 13);
  } else {
    x = sign | ((exp - 15u + 127u) << 23) | (mant << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

static uint16_t f32_to_bf16_bits(float f)
{
  uint32_t x;
  memcpy(&x, &f, 4);
  if ((x & 0x7fffffffu) > 0x7f800000u) {
    return (uint16_t)((x >> 16) | 0x40u);
  }
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}

static float bf16_bits_to_f32(uint16_t b)
{
  uint32_t x = (uint32_t)b << 16;
  float f;
  memcpy(&f, &x, 4);
  return f;
}

static real orc_quantize(const lattice_t* L, real v)
{
#ifdef ORC_DOUBLE
  (void)L;
  return v;
#else
  if (L->store == STORE_F16) {
    return f16_bits_to_f32(f32_to_f16_bits(v));
  }
  if (L->store == STORE_BF16) {
    return bf16_bits_to_f32(f32_to_bf16_bits(v));
  }
  return v;
#endif
}

/* exported conversion helpers so tests can pin the device conversions */
uint16_t ORC(f32_to_f16)(float f) { return f32_to_f16_bits(f); }
float ORC(f16_to_f32)(uint16_t h) { return f16_bits_to_f32(h); }
uint16_t ORC(f32_to_bf16)(float f) { return f32_to_bf16_bits(f); }
float ORC(bf16_to_f32)(uint16_t b) { return bf16_bits_to_f32(b); }

/* ---- ring access (cpu/sisd/column.h:162-170) ---- */

static inline uint64_t slot(const lattice_t* L, int i, uint64_t n)
{
  uint64_t s = n + L->shift[i];
  return s >= L->count ? s - L->count : s;
}

/* ---- lattice lifetime ---- */

void* ORC(create)(int nx, int ny, int nz, int pad, int px, int py, int pz)
{
  if (nx <= 0 || ny <= 0 || nz <= 0 || pad < 1) {
    return NULL;
  }
  lattice_t* L = (lattice_t*)calloc(1, sizeof(lattice_t));
  if (!L) return NULL;
  L->nx = nx; L->ny = ny; L->nz = nz; L->pad = pad;
  L->periodic[0] = px; L->periodic[1] = py; L->periodic[2] = pz;
  L->size[0] = nx + 2 * pad; L->size[1] = ny + 2 * pad; L->size[2] = nz + 2 * pad;
  L->proj[0] = L->size[1] * L->size[2];
  L->proj[1] = L->size[2];
  L->proj[2] = 1;
  L->count = (uint64_t)(L->size[0] * L->size[1] * L->size[2]);
  for (int i = 0; i < Q; ++i) {
    L->d[i] = C[i][0] * L->proj[0] + C[i][1] * L->proj[1] + C[i][2] * L->proj[2];
    L->shift[i] = 0;
    L->ring[i] = (real*)calloc(L->count, sizeof(real));
    if (!L->ring[i]) return NULL;
  }
  L->mask = (uint8_t*)calloc(L->count, 1);
  for (int k = 0; k < 256; ++k) {
    L->table[k].kind = KIND_NONE;
    L->table[k].wall_rho = 1;
  }
  return L;
}

void ORC(destroy)(void* h)
{
  lattice_t* L = (lattice_t*)h;
  if (!L) return;
  for (int i = 0; i < Q; ++i) free(L->ring[i]);
  free(L->mask);
  free(L->cell_velocity);
  free(L->cell_rho);
  free(L->cell_force);
  free(L->normals);
  free(L->bz_cells);
  free(L->bz_q);
  free(L->bzv_cells);
  free(L->bzv_q);
  free(L->bzv_v);
  free(L);
}

uint64_t ORC(ncells)(void* h) { return ((lattice_t*)h)->count; }

uint32_t ORC(cell_id)(void* h, int x, int y, int z)
{
  lattice_t* L = (lattice_t*)h;
  return (uint32_t)((x + L->pad) * L->proj[0] + (y + L->pad) * L->proj[1] + (z + L->pad));
}

void ORC(neighbor_distances)(void* h, int64_t* out)
{
  lattice_t* L = (lattice_t*)h;
  for (int i = 0; i < Q; ++i) out[i] = L->d[i];
}

void ORC(shifts)(void* h, int64_t* out)
{
  lattice_t* L = (lattice_t*)h;
  for (int i = 0; i < Q; ++i) out[i] = (int64_t)L->shift[i];
}

void ORC(set_storage)(void* h, int mode) { ((lattice_t*)h)->store = mode; }

void ORC(set_statistics)(void* h, int on) { ((lattice_t*)h)->stats_on = on; }

void ORC(get_statistics)(void* h, double* out)
{
  lattice_t* L = (lattice_t*)h;
  for (int k = 0; k < 4; ++k) out[k] = L->stats[k];
}

void ORC(set_dynamics_mask)(void* h, const uint8_t* ids)
{
  lattice_t* L = (lattice_t*)h;
  memcpy(L->mask, ids, L->count);
}

void ORC(set_dynamics)(void* h, int id, int kind, double omega, double magic,
                       const double* force, const double* wall_u, double wall_rho, double smagorinsky)
{
  lattice_t* L = (lattice_t*)h;
  dyn_t* D = &L->table[id & 255];
  D->kind = kind;
  D->omega = (real)omega;
  D->magic = (real)magic;
  for (int a = 0; a < 3; ++a) {
    D->force[a] = force ? (real)force[a] : (real)0;
    D->wall_u[a] = wall_u ? (real)wall_u[a] : (real)0;
  }
  D->wall_rho = (real)wall_rho;
  D->smagorinsky = (real)smagorinsky;
}

static void set_field(real** dst, uint64_t n, const real* src)
{
  if (!src) {
    free(*dst);
    *dst = NULL;
    return;
  }
  if (!*dst) *dst = (real*)malloc(n * sizeof(real));
  memcpy(*dst, src, n * sizeof(real));
}

void ORC(set_cell_velocity)(void* h, const real* u)
{
  lattice_t* L = (lattice_t*)h;
  set_field(&L->cell_velocity, L->count * 3, u);
}

void ORC(set_cell_rho)(void* h, const real* rho)
{
  lattice_t* L = (lattice_t*)h;
  set_field(&L->cell_rho, L->count, rho);
}

void ORC(set_cell_force)(void* h, const real* f)
{
  lattice_t* L = (lattice_t*)h;
  set_field(&L->cell_force, L->count * 3, f);
}

/* populations in logical order: out[i*count + n] = f_bar_i(n) */
void ORC(get_populations)(void* h, real* out)
{
  lattice_t* L = (lattice_t*)h;
  for (int i = 0; i < Q; ++i) {
    for (uint64_t n = 0; n < L->count; ++n) {
      out[(uint64_t)i * L->count + n] = L->ring[i][slot(L, i, n)];
    }
  }
}

void ORC(set_populations)(void* h, const real* in)
{
  lattice_t* L = (lattice_t*)h;
  for (int i = 0; i < Q; ++i) {
    for (uint64_t n = 0; n < L->count; ++n) {
      L->ring[i][slot(L, i, n)] = orc_quantize(L, in[(uint64_t)i * L->count + n]);
    }
  }
}

/* ---- physics (reference formulas) ---- */

/* lbm.h:51-64 equilibrium<D>::secondOrder, shifted by -t_i */
static real eq_second_order(int i, real rho, const real u[3], real usqr)
{
  real cu = 0;
  for (int a = 0; a < 3; ++a) cu += (real)C[i][a] * u[a];
  const real inv = 3;
  const real t = weight(i);
  return rho * t * ((real)1 + inv * cu + inv * inv * (real)0.5 * cu * cu
                    - inv * (real)0.5 * usqr) - t;
}

/* lbm.h:289-330 computeRho / computeJ / computeRhoU */
static void rho_u(const real f[Q], real* rho_out, real u[3])
{
  real rho = 0;
  for (int i = 0; i < Q; ++i) rho += f[i];
  rho += (real)1;
  real j[3] = {0, 0, 0};
  for (int i = 0; i < Q; ++i) {
    for (int a = 0; a < 3; ++a) j[a] += f[i] * (real)C[i][a];
  }
  for (int a = 0; a < 3; ++a) u[a] = j[a] / rho;
  *rho_out = rho;
}

static real norm_sqr(const real u[3]) { return u[0] * u[0] + u[1] * u[1] + u[2] * u[2]; }

/* collision.h:57-80 BGK with the equilibrium of (rho, u) */
static void bgk(real f[Q], real rho, const real u[3], real omega)
{
  const real usqr = norm_sqr(u);
  for (int i = 0; i < Q; ++i) {
    const real feq = eq_second_order(i, rho, u, usqr);
    f[i] *= (real)1 - omega;
    f[i] += omega * feq;
  }
}

/* collision.h:381-419 TRT; omega2 is formed in double as the reference's
 * literal 0.5 promotes it (collision.h:408) */
static void trt(real f[Q], real rho, const real u[3], real omega, real magic)
{
  const real usqr = norm_sqr(u);
  real feq[Q], fp[Q], fm[Q], ep[Q], em[Q];
  for (int i = 0; i < Q; ++i) feq[i] = eq_second_order(i, rho, u, usqr);
  for (int i = 0; i < Q; ++i) {
    fp[i] = (real)(0.5 * (f[i] + f[opp(i)]));
    fm[i] = (real)(0.5 * (f[i] - f[opp(i)]));
  }
  for (int i = 0; i < Q; ++i) {
    ep[i] = (real)(0.5 * (feq[i] + feq[opp(i)]));
    em[i] = (real)(0.5 * (feq[i] - feq[opp(i)]));
  }
  const real omega2 = (real)((real)1 / (magic / ((real)1 / omega - 0.5) + (real)0.5));
  for (int i = 0; i < Q; ++i) {
    f[i] -= omega * (fp[i] - ep[i]) + omega2 * (fm[i] - em[i]);
  }
}

/* lbm.h:567-588 addExternalForce */
static void add_force(real f[Q], real rho, const real u[3], real omega, const real F[3])
{
  const real inv = 3;
  const real inv4 = inv * inv;
  for (int i = 0; i < Q; ++i) {
    real cu = 0;
    for (int a = 0; a < 3; ++a) cu += (real)C[i][a] * u[a];
    cu *= inv4;
    real term = 0;
    for (int a = 0; a < 3; ++a) {
      term += (((real)C[i][a] - u[a]) * inv + cu * (real)C[i][a]) * F[a];
    }
    term *= (rho * ((real)1 - omega * (real)0.5) * weight(i));
    f[i] += term;
  }
}

/* collision.h:465-484 Revert */
static void revert(real f[Q])
{
  for (int i = 1; i <= Q / 2; ++i) {
    real tmp = f[i];
    f[i] = f[opp(i)];
    f[opp(i)] = tmp;
  }
}

/* ---- LocalVelocity boundaries ---- */

static const real* wall_velocity(const lattice_t* L, const dyn_t* D, uint64_t n)
{
  return L->cell_velocity ? &L->cell_velocity[3 * n] : D->wall_u;
}

/* direction and orientation of a flat discrete normal (one non-zero entry) */
static int flat_normal(const lattice_t* L, uint64_t n, int* dir, int* orient)
{
  if (!L->normals) return 0;
  int k = 0;
  for (int a = 0; a < 3; ++a) {
    if (L->normals[3 * n + a] != 0) {
      *dir = a;
      *orient = L->normals[3 * n + a];
      ++k;
    }
  }
  return k == 1;
}

/* momenta/elements.h:55-71 velocityBMRho */
static real velocity_bm_rho(const real f[Q], int dir, int orient, const real u[3])
{
  real on_wall = 0, normal = 0;
  for (int i = 0; i < Q; ++i) if (C[i][dir] == 0) on_wall += f[i];
  for (int i = 0; i < Q; ++i) if (C[i][dir] == orient) normal += f[i];
  return ((real)2 * normal + on_wall + (real)1) / ((real)1 + (real)orient * u[dir]);
}

/* boundary/helper.h:31-81 BoundaryHelpers::computeStress; pi in the order
 * xx, xy, xz, yy, yz, zz */
static void boundary_stress(const real f[Q], int dir, int orient, real rho, const real u[3], real pi[6])
{
  const real usqr = norm_sqr(u);
  real fneq[Q];
  for (int i = 0; i < Q; ++i) {
    fneq[i] = (i > 0 && (C[i][dir] == 0 || C[i][dir] == orient))
                ? f[i] - eq_second_order(i, rho, u, usqr) : (real)0;
  }
  int k = 0;
  for (int a = 0; a < 3; ++a) {
    for (int b = a; b < 3; ++b) {
      real v = 0;
      for (int i = 0; i < Q; ++i) if (C[i][dir] == 0) v += (real)(C[i][a] * C[i][b]) * fneq[i];
      for (int i = 0; i < Q; ++i) if (C[i][dir] == orient) v += (real)2 * (real)(C[i][a] * C[i][b]) * fneq[i];
      pi[k++] = v;
    }
  }
}

/* lbm.h:248-270 fromPiToFneq */
static real pi_to_fneq(int i, const real pi[6])
{
  real fneq = 0;
  int k = 0;
  for (int a = 0; a < 3; ++a) {
    for (int b = a; b < 3; ++b) {
      real add = (real)(C[i][a] * C[i][b]);
      if (a == b) {
        add -= (real)1 / (real)3;
      } else {
        add *= (real)2;
      }
      add *= pi[k++];
      fneq += add;
    }
  }
  fneq *= weight(i) * (real)3 * (real)3 / (real)2;
  return fneq;
}

/* one cell; returns 1 if the cell contributes to the statistics */
/* lbm<D>::computeStress (dynamics/lbm.h:334-352): sum c_a c_b f - rho u_a u_b
 * - (rho - 1)/3 delta_ab, on shifted populations */
static void bulk_stress(const real f[Q], real rho, const real u[3], real pi[6])
{
  int k = 0;
  for (int a = 0; a < 3; ++a) {
    for (int b = a; b < 3; ++b) {
      real v = 0;
      for (int i = 0; i < Q; ++i) v += (real)(C[i][a] * C[i][b]) * f[i];
      v -= rho * u[a] * u[b];
      if (a == b) v -= (real)1 / (real)3 * (rho - (real)1);
      pi[k++] = v;
    }
  }
}

/* InnerEdgeDensity3D / InnerCornerDensity3D (momenta/elements.h:432-482):
 * the mean of velocityBMRho over the non-zero axes of the outward normal */
static real inner_density(const real f[Q], const int N[3], const real u[3])
{
  real s = 0;
  int k = 0;
  if (N[0] != 0 && N[1] != 0 && N[2] != 0) {
    const real rx = velocity_bm_rho(f, 0, N[0], u);
    const real ry = velocity_bm_rho(f, 1, N[1], u);
    const real rz = velocity_bm_rho(f, 2, N[2], u);
    return (rx + ry + rz) / (real)3;
  }
  /* edge: plane = the zero axis, normal1 = N[plane+1], normal2 = N[plane+2]
   * (setBoundary3D.h:156-197) */
  int plane = N[0] == 0 ? 0 : (N[1] == 0 ? 1 : 2);
  const int a1 = (plane + 1) % 3, a2 = (plane + 2) % 3;
  s = velocity_bm_rho(f, a1, N[a1], u) + velocity_bm_rho(f, a2, N[a2], u);
  (void)k;
  return s / (real)2;
}

/* InnerEdgeStress3D / InnerCornerStress3D (momenta/elements.h:1844-1898):
 * the populations entering from both walls are replaced by bounce-back of
 * their non-equilibrium part, then the bulk stress */
static void inner_stress(const real f[Q], const int N[3], real rho, const real u[3], real pi[6])
{
  real g[Q];
  for (int i = 0; i < Q; ++i) g[i] = f[i];
  const real usqr = norm_sqr(u);
  if (!(N[0] != 0 && N[1] != 0 && N[2] != 0)) {
    const int plane = N[0] == 0 ? 0 : (N[1] == 0 ? 1 : 2);
    const int a1 = (plane + 1) % 3, a2 = (plane + 2) % 3;
    for (int i = 0; i < Q; ++i) {
      if (C[i][a1] == -N[a1] && C[i][a2] == -N[a2]) {
        const int o = opp(i);
        g[i] = g[o] - eq_second_order(o, rho, u, usqr) + eq_second_order(i, rho, u, usqr);
      }
    }
  }
  /* an inner corner's entering population c = -N does not exist in D3Q19 */
  bulk_stress(g, rho, u, pi);
}

/* FixedPressureMomentum::computeU (momenta/elements.h:653-691): the wall's
 * tangential velocity, the normal component from the known populations */
static void pressure_u(const real f[Q], int dir, int orient, real rho, const real uw[3], real u[3])
{
  real on_wall = 0, normal = 0;
  for (int i = 0; i < Q; ++i) if (C[i][dir] == 0) on_wall += f[i];
  for (int i = 0; i < Q; ++i) if (C[i][dir] == orient) normal += f[i];
  for (int a = 0; a < 3; ++a) u[a] = uw[a];
  u[dir] = (real)orient * (((real)2 * normal + on_wall + (real)1) / rho - (real)1);
}

static real wall_density(const lattice_t* L, const dyn_t* D, uint64_t n)
{
  return L->cell_rho ? L->cell_rho[n] : D->wall_rho;
}

static void cell_normal(const lattice_t* L, uint64_t n, int N[3])
{
  for (int a = 0; a < 3; ++a) N[a] = L->normals ? L->normals[3 * n + a] : 0;
}

static int collide_cell(const lattice_t* L, uint64_t n, real f[Q], real* st_rho, real* st_usqr)
{
  const dyn_t* D = &L->table[L->mask[n]];
  real rho, u[3];
  switch (D->kind) {
  case KIND_BGK:
    rho_u(f, &rho, u);
    bgk(f, rho, u, D->omega);
    *st_rho = rho; *st_usqr = norm_sqr(u);
    return 1;
  case KIND_TRT:
    rho_u(f, &rho, u);
    trt(f, rho, u, D->omega, D->magic);
    *st_rho = rho; *st_usqr = norm_sqr(u);
    return 1;
  case KIND_BGK_GUO:
  case KIND_TRT_GUO: {
    const real* F = L->cell_force ? &L->cell_force[3 * n] : D->force;
    rho_u(f, &rho, u);
    for (int a = 0; a < 3; ++a) u[a] += F[a] * (real)0.5;   /* elements.h:1603-1607 */
    if (D->kind == KIND_BGK_GUO) {
      bgk(f, rho, u, D->omega);
    } else {
      trt(f, rho, u, D->omega, D->magic);
    }
    add_force(f, rho, u, D->omega, F);
    *st_rho = rho; *st_usqr = norm_sqr(u);
    return 1;
  }
  case KIND_BOUNCE_BACK:
    revert(f);
    return 0;
  case KIND_BOUNCE_BACK_VELOCITY: {
    revert(f);
    const real r = L->cell_rho ? L->cell_rho[n] : D->wall_rho;
    const real* uw = L->cell_velocity ? &L->cell_velocity[3 * n] : D->wall_u;
    for (int i = 1; i < Q; ++i) {
      for (int a = 0; a < 3; ++a) {
        f[i] += 2 * r * uw[a] * (real)C[i][a] * weight(i) * (real)3;
      }
    }
    *st_rho = r; *st_usqr = norm_sqr(uw);
    return 1;
  }
  case KIND_SMAGORINSKY_BGK: {
    /* SmagorinskyEffectiveOmega<BGK> (collisionLES.h:59-89) with the stress
     * of lbm.h:334-352 and the norm of lbm.h:555-564 */
    rho_u(f, &rho, u);
    real pi[6];
    bulk_stress(f, rho, u, pi);
    real norm_sqr_pi = pi[0] * pi[0] + (real)(2. * pi[1] * pi[1]) + pi[2] * pi[2];
    norm_sqr_pi += pi[2] * pi[2] + pi[3] * pi[3] + (real)(2. * pi[4] * pi[4]) + pi[5] * pi[5];
    const real c = D->smagorinsky;
    const real pre = (real)((double)(c * c * (real)3 * (real)3 * (real)2) * sqrt(2.0));
    const real tau_mol = (real)1 / D->omega;
    const real tau_turb = (real)0.5 * ((real)sqrt(tau_mol * tau_mol + pre / rho * (real)sqrt(norm_sqr_pi)) - tau_mol);
    const real omega_eff = (real)1 / (tau_mol + tau_turb);
    bgk(f, rho, u, omega_eff);
    *st_rho = rho; *st_usqr = norm_sqr(u);
    return 1;
  }
  case KIND_LOCAL_VELOCITY: {
    /* CombinedRLBdynamics::collide (dynamics.h:506-517): regularise around
     * (rho, u, pi) of the boundary momenta, then BGK with the same momenta */
    int dir = 0, orient = 0;
    if (!flat_normal(L, n, &dir, &orient)) return 0;
    const real* uw = wall_velocity(L, D, n);
    real uu[3] = {uw[0], uw[1], uw[2]};
    real pi[6];
    rho = velocity_bm_rho(f, dir, orient, uu);
    boundary_stress(f, dir, orient, rho, uu, pi);
    const real usqr = norm_sqr(uu);
    for (int i = 0; i < Q; ++i) f[i] = eq_second_order(i, rho, uu, usqr) + pi_to_fneq(i, pi);
    rho = velocity_bm_rho(f, dir, orient, uu);
    bgk(f, rho, uu, D->omega);
    *st_rho = rho; *st_usqr = usqr;
    return 1;
  }
  case KIND_INNER_EDGE:
  case KIND_INNER_CORNER: {
    /* CombinedRLBdynamics<BGK, InnerEdge/InnerCornerVelocityTuple3D>
     * (boundary/localVelocity3D.h:86-99) */
    int N[3];
    cell_normal(L, n, N);
    const real* uw = wall_velocity(L, D, n);
    real uu[3] = {uw[0], uw[1], uw[2]};
    real pi[6];
    rho = inner_density(f, N, uu);
    inner_stress(f, N, rho, uu, pi);
    const real usqr = norm_sqr(uu);
    for (int i = 0; i < Q; ++i) f[i] = eq_second_order(i, rho, uu, usqr) + pi_to_fneq(i, pi);
    rho = inner_density(f, N, uu);
    bgk(f, rho, uu, D->omega);
    *st_rho = rho; *st_usqr = usqr;
    return 1;
  }
  case KIND_LOCAL_PRESSURE: {
    /* CombinedRLBdynamics<BGK, RegularizedPressureBoundaryTuple<dir,orient>>
     * (boundary/localPressure3D.h:47-58, momenta/aliases.h:152-159) */
    int dir = 0, orient = 0;
    if (!flat_normal(L, n, &dir, &orient)) return 0;
    const real* uw = wall_velocity(L, D, n);
    real pi[6], uu[3];
    rho = wall_density(L, D, n);
    pressure_u(f, dir, orient, rho, uw, uu);
    boundary_stress(f, dir, orient, rho, uu, pi);
    real usqr = norm_sqr(uu);
    for (int i = 0; i < Q; ++i) f[i] = eq_second_order(i, rho, uu, usqr) + pi_to_fneq(i, pi);
    pressure_u(f, dir, orient, rho, uw, uu);
    bgk(f, rho, uu, D->omega);
    *st_rho = rho; *st_usqr = norm_sqr(uu);
    return 1;
  }
  case KIND_INTERP_VELOCITY:
  case KIND_INTERP_PRESSURE: {
    /* BGK (collision.h:57-80) through BasicDirichletVelocityBoundaryTuple or
     * BasicDirichletPressureBoundaryTuple (momenta/aliases.h:173-198) */
    int dir = 0, orient = 0;
    if (!flat_normal(L, n, &dir, &orient)) return 0;
    const real* uw = wall_velocity(L, D, n);
    real uu[3] = {uw[0], uw[1], uw[2]};
    if (D->kind == KIND_INTERP_VELOCITY) {
      rho = velocity_bm_rho(f, dir, orient, uu);
    } else {
      rho = wall_density(L, D, n);
      pressure_u(f, dir, orient, rho, uw, uu);
    }
    bgk(f, rho, uu, D->omega);
    *st_rho = rho; *st_usqr = norm_sqr(uu);
    return 1;
  }
  case KIND_FIXED_VELOCITY: {
    /* BulkDensity + FixedVelocityMomentumGeneric, BGK */
    const real* uw = wall_velocity(L, D, n);
    real uu[3] = {uw[0], uw[1], uw[2]};
    rho_u(f, &rho, u);
    bgk(f, rho, uu, D->omega);
    *st_rho = rho; *st_usqr = norm_sqr(uu);
    return 1;
  }
  default:
    return 0;
  }
}

static void collide(lattice_t* L)
{
  double n_tot = 0, rho_tot = 0, e_tot = 0, max_u = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+:n_tot,rho_tot,e_tot) reduction(max:max_u)
#endif
  for (int x = 0; x < L->nx; ++x) {
    for (int y = 0; y < L->ny; ++y) {
      uint64_t n = (uint64_t)((x + L->pad) * L->proj[0] + (y + L->pad) * L->proj[1] + L->pad);
      for (int z = 0; z < L->nz; ++z, ++n) {
        if (L->table[L->mask[n]].kind == KIND_NONE) continue;
        real f[Q];
        uint64_t s[Q];
        for (int i = 0; i < Q; ++i) {
          s[i] = slot(L, i, n);
          f[i] = L->ring[i][s[i]];
        }
        real r = 0, e = 0;
        if (collide_cell(L, n, f, &r, &e) && L->stats_on) {
          n_tot += 1;
          rho_tot += (double)r;
          e_tot += (double)e;
          if ((double)e > max_u) max_u = (double)e;
        }
        for (int i = 0; i < Q; ++i) L->ring[i][s[i]] = orc_quantize(L, f[i]);
      }
    }
  }
  L->stats[0] = n_tot;
  L->stats[1] = rho_tot;
  L->stats[2] = e_tot;
  L->stats[3] = max_u;
}

/* PostCollide overlap copy: every padding cell within distance 1 of the core
 * whose periodic image lies in the core receives all its populations. */
static void periodic_copy(lattice_t* L)
{
  const int n3[3] = {L->nx, L->ny, L->nz};
  for (int x = -1; x <= L->nx; ++x) {
    for (int y = -1; y <= L->ny; ++y) {
      for (int z = -1; z <= L->nz; ++z) {
        int g[3] = {x, y, z};
        int img[3];
        int outside = 0, ok = 1;
        for (int a = 0; a < 3; ++a) {
          img[a] = g[a];
          if (g[a] < 0 || g[a] >= n3[a]) {
            outside = 1;
            if (!L->periodic[a]) { ok = 0; break; }
            img[a] = g[a] < 0 ? g[a] + n3[a] : g[a] - n3[a];
          }
        }
        if (!outside || !ok) continue;
        uint64_t gn = (uint64_t)((g[0] + L->pad) * L->proj[0] + (g[1] + L->pad) * L->proj[1] + (g[2] + L->pad));
        uint64_t in = (uint64_t)((img[0] + L->pad) * L->proj[0] + (img[1] + L->pad) * L->proj[1] + (img[2] + L->pad));
        for (int i = 0; i < Q; ++i) {
          L->ring[i][slot(L, i, gn)] = L->ring[i][slot(L, i, in)];
        }
      }
    }
  }
}

/* PostStream overlap exchange of the post-processor neighbourhood
 * (boundary/setBoundary3D.h:57-110 addPoints2CommBC requests the populations
 * of the padding cells near the boundary): the padding cells within reach
 * (2) of an edge, corner or plane finite-difference processor -- the ones
 * those processors read --
 * whose periodic image lies in the core get the image's streamed
 * populations, before any post-processor runs. */
static int near_processor(const lattice_t* L, const int g[3])
{
  if (!L->normals) return 0;
  for (int dx = -2; dx <= 2; ++dx) {
    for (int dy = -2; dy <= 2; ++dy) {
      for (int dz = -2; dz <= 2; ++dz) {
        const int c[3] = {g[0] + dx, g[1] + dy, g[2] + dz};
        if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= L->nx || c[1] >= L->ny || c[2] >= L->nz) continue;
        const uint64_t n = (uint64_t)((c[0] + L->pad) * L->proj[0] + (c[1] + L->pad) * L->proj[1] + (c[2] + L->pad));
        const int kind = L->table[L->mask[n]].kind;
        if (kind == KIND_INTERP_VELOCITY || kind == KIND_INTERP_PRESSURE) return 1;
        if (kind != KIND_FIXED_VELOCITY) continue;
        const int nz = (L->normals[3 * n] != 0) + (L->normals[3 * n + 1] != 0) + (L->normals[3 * n + 2] != 0);
        if (nz >= 2) return 1;
      }
    }
  }
  return 0;
}

static void periodic_refresh(lattice_t* L)
{
  const int n3[3] = {L->nx, L->ny, L->nz};
  const int p = L->pad;
  for (int x = -p; x < L->nx + p; ++x) {
    for (int y = -p; y < L->ny + p; ++y) {
      for (int z = -p; z < L->nz + p; ++z) {
        int g[3] = {x, y, z};
        int img[3];
        int outside = 0, ok = 1;
        for (int a = 0; a < 3; ++a) {
          img[a] = g[a];
          if (g[a] < 0 || g[a] >= n3[a]) {
            outside = 1;
            if (!L->periodic[a]) { ok = 0; break; }
            img[a] = g[a] < 0 ? g[a] + n3[a] : g[a] - n3[a];
          }
        }
        if (!outside || !ok || !near_processor(L, g)) continue;
        uint64_t gn = (uint64_t)((g[0] + p) * L->proj[0] + (g[1] + p) * L->proj[1] + (g[2] + p));
        uint64_t in = (uint64_t)((img[0] + p) * L->proj[0] + (img[1] + p) * L->proj[1] + (img[2] + p));
        for (int i = 0; i < Q; ++i) {
          L->ring[i][slot(L, i, gn)] = L->ring[i][slot(L, i, in)];
        }
      }
    }
  }
}

/* cpu/sisd/column.h:193-204 rotate(d_i): shift -= d_i (mod count) */
static void stream(lattice_t* L)
{
  for (int i = 0; i < Q; ++i) {
    int64_t c = (int64_t)L->count;
    int64_t s = ((int64_t)L->shift[i] - L->d[i]) % c;
    if (s < 0) s += c;
    L->shift[i] = (uint64_t)s;
  }
}

/* ---- PostStream processors of LocalVelocity edges and corners ---- */

static void load_cell(const lattice_t* L, uint64_t n, real f[Q])
{
  for (int i = 0; i < Q; ++i) f[i] = L->ring[i][slot(L, i, n)];
}

/* computeRho / computeU of a cell through its momenta */
static real momenta_rho(const lattice_t* L, uint64_t n)
{
  const dyn_t* D = &L->table[L->mask[n]];
  real f[Q], rho, u[3];
  load_cell(L, n, f);
  switch (D->kind) {
  case KIND_NONE:
    return (real)1;                      /* OneDensity */
  case KIND_BOUNCE_BACK:
  case KIND_BOUNCE_BACK_VELOCITY:
  case KIND_LOCAL_PRESSURE:
  case KIND_INTERP_PRESSURE:
    return wall_density(L, D, n);        /* FixedDensity */
  case KIND_LOCAL_VELOCITY:
  case KIND_INTERP_VELOCITY: {
    int dir = 0, orient = 0;
    flat_normal(L, n, &dir, &orient);
    const real* uw = wall_velocity(L, D, n);
    real uu[3] = {uw[0], uw[1], uw[2]};
    return velocity_bm_rho(f, dir, orient, uu);
  }
  case KIND_INNER_EDGE:
  case KIND_INNER_CORNER: {
    int N[3];
    cell_normal(L, n, N);
    const real* uw = wall_velocity(L, D, n);
    real uu[3] = {uw[0], uw[1], uw[2]};
    return inner_density(f, N, uu);
  }
  default:
    break;
  }
  rho_u(f, &rho, u);
  return rho;
}

static void momenta_u(const lattice_t* L, uint64_t n, real u[3])
{
  const dyn_t* D = &L->table[L->mask[n]];
  switch (D->kind) {
  case KIND_NONE:
  case KIND_BOUNCE_BACK:
    for (int a = 0; a < 3; ++a) u[a] = 0;   /* ZeroMomentum */
    return;
  case KIND_LOCAL_VELOCITY:
  case KIND_FIXED_VELOCITY:
  case KIND_INNER_EDGE:
  case KIND_INNER_CORNER:
  case KIND_BOUNCE_BACK_VELOCITY:
  case KIND_INTERP_VELOCITY: {
    const real* uw = wall_velocity(L, D, n);
    for (int a = 0; a < 3; ++a) u[a] = uw[a];
    return;
  }
  case KIND_LOCAL_PRESSURE:
  case KIND_INTERP_PRESSURE: {
    real f[Q];
    int dir = 0, orient = 0;
    load_cell(L, n, f);
    flat_normal(L, n, &dir, &orient);
    pressure_u(f, dir, orient, wall_density(L, D, n), wall_velocity(L, D, n), u);
    return;
  }
  default:
    break;
  }
  real f[Q], rho;
  load_cell(L, n, f);
  rho_u(f, &rho, u);
  if (D->kind == KIND_BGK_GUO || D->kind == KIND_TRT_GUO) {
    const real* F = L->cell_force ? &L->cell_force[3 * n] : D->force;
    for (int a = 0; a < 3; ++a) u[a] += F[a] * (real)0.5;
  }
}

static uint64_t step_cell(const lattice_t* L, uint64_t n, int axis, int k)
{
  return (uint64_t)((int64_t)n + (int64_t)k * L->proj[axis]);
}

/* finiteDifference.h:42-45 */
static real boundary_gradient(real u0, real u1, real u2)
{
  return (-(real)3 * u0 + (real)4 * u1 - (real)1 * u2) / (real)2;
}

/* DirectedGradients3D<..., orthogonal = true>: one-sided along axis with the
 * outward orientation (finiteDifference3D.h:132-155) */
static void one_sided(const lattice_t* L, uint64_t n, int axis, int orient, real d[3])
{
  real u0[3], u1[3], u2[3];
  momenta_u(L, n, u0);
  momenta_u(L, step_cell(L, n, axis, -orient), u1);
  momenta_u(L, step_cell(L, n, axis, -2 * orient), u2);
  for (int a = 0; a < 3; ++a) d[a] = (real)(-orient) * boundary_gradient(u0[a], u1[a], u2[a]);
}

/* regularised populations from rho, the cell's velocity and the velocity
 * gradients dA_uB[A][B] = d_A u_B (boundaryPostProcessors3D.hh:196-215) */
static void regularise(lattice_t* L, uint64_t n, real rho, real du[3][3])
{
  const dyn_t* D = &L->table[L->mask[n]];
  const real s_to_pi = -rho / (real)3 / D->omega;
  real pi[6];
  pi[0] = (real)2 * du[0][0] * s_to_pi;
  pi[3] = (real)2 * du[1][1] * s_to_pi;
  pi[5] = (real)2 * du[2][2] * s_to_pi;
  pi[1] = (du[0][1] + du[1][0]) * s_to_pi;
  pi[2] = (du[0][2] + du[2][0]) * s_to_pi;
  pi[4] = (du[1][2] + du[2][1]) * s_to_pi;
  real u[3];
  momenta_u(L, n, u);
  const real usqr = norm_sqr(u);
  for (int i = 0; i < Q; ++i) {
    L->ring[i][slot(L, i, n)] = orc_quantize(L, eq_second_order(i, rho, u, usqr) + pi_to_fneq(i, pi));
  }
}

/* OuterVelocityEdgeProcessor3D<plane, normal1, normal2>::apply */
static void post_edge(lattice_t* L, uint64_t n, const int N[3])
{
  const int p = N[0] == 0 ? 0 : (N[1] == 0 ? 1 : 2);
  const int d1 = (p + 1) % 3, d2 = (p + 2) % 3;
  const int n1 = N[d1], n2 = N[d2];
  const real rho10 = momenta_rho(L, step_cell(L, n, d1, -n1));
  const real rho01 = momenta_rho(L, step_cell(L, n, d2, -n2));
  const real rho20 = momenta_rho(L, step_cell(L, n, d1, -2 * n1));
  const real rho02 = momenta_rho(L, step_cell(L, n, d2, -2 * n2));
  const real rho = (real)2 / (real)3 * (rho01 + rho10) - (real)1 / (real)6 * (rho02 + rho20);
  real du[3][3], up[3], um[3];
  momenta_u(L, step_cell(L, n, p, 1), up);
  momenta_u(L, step_cell(L, n, p, -1), um);
  for (int b = 0; b < 3; ++b) du[p][b] = (up[b] - um[b]) / (real)2;
  one_sided(L, n, d1, n1, du[d1]);
  one_sided(L, n, d2, n2, du[d2]);
  regularise(L, n, rho, du);
}

/* OuterVelocityCornerProcessor3D<xNormal, yNormal, zNormal>::apply */
static void post_corner(lattice_t* L, uint64_t n, const int N[3])
{
  const real rho100 = momenta_rho(L, step_cell(L, n, 0, -N[0]));
  const real rho010 = momenta_rho(L, step_cell(L, n, 1, -N[1]));
  const real rho001 = momenta_rho(L, step_cell(L, n, 2, -N[2]));
  const real rho200 = momenta_rho(L, step_cell(L, n, 0, -2 * N[0]));
  const real rho020 = momenta_rho(L, step_cell(L, n, 1, -2 * N[1]));
  const real rho002 = momenta_rho(L, step_cell(L, n, 2, -2 * N[2]));
  const real rho = (real)4 / (real)9 * (rho001 + rho010 + rho100)
                 - (real)1 / (real)9 * (rho002 + rho020 + rho200);
  real du[3][3];
  for (int a = 0; a < 3; ++a) one_sided(L, n, a, N[a], du[a]);
  regularise(L, n, rho, du);
}

/* PlaneFdBoundaryProcessor3D<direction, orientation>::apply
 * (boundary/boundaryPostProcessors3D.hh:39-92): density and velocity
 * through the momenta, the gradient one-sided along the normal and central
 * along the wall (finiteDifference3D.h:75-230) */
static void post_plane_fd(lattice_t* L, uint64_t n)
{
  int dir = 0, orient = 0;
  flat_normal(L, n, &dir, &orient);
  const real rho = momenta_rho(L, n);
  real du[3][3], up[3], um[3];
  for (int a = 0; a < 3; ++a) {
    if (a == dir) {
      one_sided(L, n, a, orient, du[a]);
    } else {
      momenta_u(L, step_cell(L, n, a, 1), up);
      momenta_u(L, step_cell(L, n, a, -1), um);
      for (int b = 0; b < 3; ++b) du[a][b] = (up[b] - um[b]) / (real)2;
    }
  }
  regularise(L, n, rho, du);
}

/* BouzidiPostProcessor::apply (boundary/setBouzidiBoundary.h:48-80) on every
 * listed cell in the listed order (the reference's ascending cell id,
 * core/platform/cpu/sisd/operator.h:368-386), in place.  Neighbours are read
 * from their own ring slots: the values read (x_s[i], x_f[opp i] and x_b[i])
 * all streamed from a core cell, so a padding cell's slot already holds what
 * the reference's PostStream overlap exchange would put there. */
static void bouzidi_list(lattice_t* L, int64_t count, const int64_t* cells, const real* qs, const real* vs)
{
  for (int64_t e = 0; e < count; ++e) {
    const uint64_t n = (uint64_t)cells[e];
    const real* q = &qs[Q * e];
    for (int i = 1; i < Q; ++i) {
      const int o = opp(i);
      real* xb_o = &L->ring[o][slot(L, o, n)];
      const real xb_i = L->ring[i][slot(L, i, n)];
      /* BouzidiVelocityPostProcessor's wall term (setBouzidiBoundary.h:149) */
      const real vt = vs ? vs[Q * e + i] * weight(i) * (real)3 : (real)0;
      if (q[i] > (real)0) {
        const int64_t dn = L->d[i];
        const real xs = L->ring[i][slot(L, i, (uint64_t)((int64_t)n + dn))];
        const real xf = L->ring[o][slot(L, o, (uint64_t)((int64_t)n - dn))];
        const real lo = (real)(q[i] <= (real)0.5), hi = (real)(q[i] > (real)0.5);
        *xb_o = orc_quantize(L, lo * ((real)2 * q[i] * xs + ((real)1 - (real)2 * q[i]) * xb_i - (real)2 * vt)
                              + hi * ((real)0.5 / q[i] * xs + (real)0.5 * ((real)2 * q[i] - (real)1) / q[i] * xf
                                      - (real)1 / q[i] * vt));
      } else if (q[i] == (real)0) {
        *xb_o = orc_quantize(L, xb_i - (real)2 * vt);
      }
    }
  }
}

/* BouzidiPostProcessor, then BouzidiVelocityPostProcessor (priority -1,
 * in the reference's operator-type order) */
static void post_bouzidi(lattice_t* L)
{
  bouzidi_list(L, L->bz_n, L->bz_cells, L->bz_q, NULL);
  bouzidi_list(L, L->bzv_n, L->bzv_cells, L->bzv_q, L->bzv_v);
}

/* PostStream stage: Bouzidi (priority -1), LocalVelocity edges (priority 0),
 * corners (priority 1), then the flat faces of the interpolated boundaries
 * (priority 0), each in ascending cell id and in place */
static void post_stream(lattice_t* L)
{
  post_bouzidi(L);
  if (!L->normals) return;
  for (int pass = 2; pass <= 4; ++pass) {
    for (int x = 0; x < L->nx; ++x) {
      for (int y = 0; y < L->ny; ++y) {
        for (int z = 0; z < L->nz; ++z) {
          const uint64_t n = (uint64_t)((x + L->pad) * L->proj[0] + (y + L->pad) * L->proj[1] + (z + L->pad));
          const int kind = L->table[L->mask[n]].kind;
          if (pass == 4) {
            if (kind == KIND_INTERP_VELOCITY || kind == KIND_INTERP_PRESSURE) post_plane_fd(L, n);
            continue;
          }
          if (kind != KIND_FIXED_VELOCITY) continue;
          const int N[3] = {L->normals[3 * n], L->normals[3 * n + 1], L->normals[3 * n + 2]};
          const int nz = (N[0] != 0) + (N[1] != 0) + (N[2] != 0);
          if (nz != pass) continue;
          if (pass == 2) post_edge(L, n, N); else post_corner(L, n, N);
        }
      }
    }
  }
}

void ORC(set_cell_normals)(void* h, const int8_t* normals)
{
  lattice_t* L = (lattice_t*)h;
  free(L->normals);
  L->normals = NULL;
  if (normals) {
    L->normals = (int8_t*)malloc(3 * L->count);
    memcpy(L->normals, normals, 3 * L->count);
  }
}

/* the cells that run BouzidiPostProcessor and their BOUZIDI_DISTANCE
 * (19 per cell); n = 0 removes them */
void ORC(set_bouzidi)(void* h, int64_t n, const int64_t* cells, const real* q)
{
  lattice_t* L = (lattice_t*)h;
  free(L->bz_cells);
  free(L->bz_q);
  L->bz_cells = NULL;
  L->bz_q = NULL;
  L->bz_n = 0;
  if (n <= 0) return;
  L->bz_cells = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  L->bz_q = (real*)malloc(sizeof(real) * Q * (size_t)n);
  memcpy(L->bz_cells, cells, sizeof(int64_t) * (size_t)n);
  memcpy(L->bz_q, q, sizeof(real) * Q * (size_t)n);
  L->bz_n = n;
}

/* the cells running BouzidiVelocityPostProcessor with BOUZIDI_DISTANCE and
 * BOUZIDI_VELOCITY (19 each per cell); n = 0 removes them */
void ORC(set_bouzidi_velocity)(void* h, int64_t n, const int64_t* cells, const real* q, const real* v)
{
  lattice_t* L = (lattice_t*)h;
  free(L->bzv_cells);
  free(L->bzv_q);
  free(L->bzv_v);
  L->bzv_cells = NULL;
  L->bzv_q = NULL;
  L->bzv_v = NULL;
  L->bzv_n = 0;
  if (n <= 0) return;
  L->bzv_cells = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  L->bzv_q = (real*)malloc(sizeof(real) * Q * (size_t)n);
  L->bzv_v = (real*)malloc(sizeof(real) * Q * (size_t)n);
  memcpy(L->bzv_cells, cells, sizeof(int64_t) * (size_t)n);
  memcpy(L->bzv_q, q, sizeof(real) * Q * (size_t)n);
  memcpy(L->bzv_v, v, sizeof(real) * Q * (size_t)n);
  L->bzv_n = n;
}

/* BlockLattice::postProcess() (core/blockLattice.hh:56-67): the PostStream
 * processors on the current populations, as initialize() runs them once */
void ORC(post_process)(void* h)
{
  post_stream((lattice_t*)h);
}

void ORC(collide_and_stream)(void* h, int steps)
{
  lattice_t* L = (lattice_t*)h;
  for (int k = 0; k < steps; ++k) {
    collide(L);
    periodic_copy(L);
    stream(L);
    if (L->bz_n > 0 || L->bzv_n > 0 || L->normals) {
      periodic_refresh(L);
    }
    post_stream(L);
    L->time += 1;
  }
}

/* Dynamics::iniEquilibrium for the selected cells (interface.h:111-121):
 * f_bar_i = feq_i(rho, u); sel may be NULL (all cells). */
void ORC(ini_equilibrium)(void* h, const uint8_t* sel, const real* rho, const real* u)
{
  lattice_t* L = (lattice_t*)h;
  for (uint64_t n = 0; n < L->count; ++n) {
    if (sel && !sel[n]) continue;
    const real* un = &u[3 * n];
    const real usqr = norm_sqr(un);
    for (int i = 0; i < Q; ++i) {
      L->ring[i][slot(L, i, n)] = orc_quantize(L, eq_second_order(i, rho[n], un, usqr));
    }
  }
}

/* density and velocity of every cell from its populations (BulkMomentum,
 * forced variants add F/2 when a force is given) */
void ORC(compute_rho_u)(void* h, real* rho_out, real* u_out)
{
  lattice_t* L = (lattice_t*)h;
  for (uint64_t n = 0; n < L->count; ++n) {
    real f[Q], rho, u[3];
    for (int i = 0; i < Q; ++i) f[i] = L->ring[i][slot(L, i, n)];
    rho_u(f, &rho, u);
    rho_out[n] = rho;
    for (int a = 0; a < 3; ++a) u_out[3 * n + a] = u[a];
  }
}

void ORC(set_threads)(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int ORC(get_threads)(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
