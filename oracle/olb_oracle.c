/*
 * olb_oracle.c -- CPU restatement of OpenLB's D3Q19 collide-and-stream step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
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
                       const double* force, const double* wall_u, double wall_rho)
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

/* one cell; returns 1 if the cell contributes to the statistics */
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

void ORC(collide_and_stream)(void* h, int steps)
{
  lattice_t* L = (lattice_t*)h;
  for (int k = 0; k < steps; ++k) {
    collide(L);
    periodic_copy(L);
    stream(L);
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
