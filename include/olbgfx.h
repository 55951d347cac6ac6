/*
 * olbgfx.h -- C ABI of the gfx950 (MI355X) lattice Boltzmann platform.
 *
 * The library owns all device memory; callers own host buffers.  Every call
 * is stream-ordered on the stream given (NULL = the default stream) unless it
 * says it synchronises.  A handle must not be used from two threads at once.
 * No C++ types and no HIP types cross this boundary: streams are passed as
 * `void*` (a hipStream_t).
 *
 * Each entry point names the OpenLB 1.8r1 interface it stands in for
 * (paths relative to the OpenLB source root).
 */
#ifndef OLBGFX_H
#define OLBGFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct olbgfx_lattice olbgfx_lattice;
typedef struct olbgfx_ensemble olbgfx_ensemble;
typedef struct olbgfx_cell_list olbgfx_cell_list;
typedef int32_t olbgfx_status;

#define OLBGFX_OK              0
#define OLBGFX_ERR_INVALID     1  /* bad argument (std::invalid_argument)   */
#define OLBGFX_ERR_UNSUPPORTED 2  /* dynamics not provided (std::domain_error) */
#define OLBGFX_ERR_HIP         3  /* HIP runtime error (device::check)       */
#define OLBGFX_ERR_NOMEM       4  /* device allocation failed               */

/* population storage precision; arithmetic is FP32 for F32/F16/BF16 and
 * FP64 for F64.  F16/BF16 hold the shifted populations f_i - t_i. */
#define OLBGFX_STORE_F32  0
#define OLBGFX_STORE_F16  1
#define OLBGFX_STORE_BF16 2
#define OLBGFX_STORE_F64  3

/* host value types for field and population transfers */
#define OLBGFX_VALUE_F32 0
#define OLBGFX_VALUE_F64 1

/* dynamics kinds (the closed set this platform provides) */
#define OLBGFX_NO_DYNAMICS          0  /* dynamics.h:52-63 NoDynamics            */
#define OLBGFX_BGK                  1  /* dynamics.h:90-99 BGKdynamics           */
#define OLBGFX_TRT                  2  /* dynamics.h:531-537 TRTdynamics         */
#define OLBGFX_FORCED_BGK           3  /* dynamics.h:223-230 ForcedBGKdynamics   */
#define OLBGFX_FORCED_TRT           4  /* dynamics.h:543-549 ForcedTRTdynamics   */
#define OLBGFX_BOUNCE_BACK          5  /* dynamics.h:558-568 BounceBack          */
#define OLBGFX_BOUNCE_BACK_VELOCITY 6  /* dynamics.h:635-640 BounceBackVelocity  */
/* LocalVelocity boundaries (boundary/localVelocity3D.h:36-120); both need the
 * cell's discrete normal (olbgfx_set_cell_normals) and take the wall
 * velocity from wall_velocity or the VELOCITY field:
 *   LOCAL_VELOCITY: a flat face, CombinedRLBdynamics<T,D,BGKdynamics,
 *     RegularizedVelocityBoundaryTuple<direction,orientation>>
 *     (dynamics.h:476-526) with the direction and orientation of its normal;
 *   FIXED_VELOCITY: BGKdynamics with FixedVelocityBoundaryTuple
 *     (momenta/aliases.h:215-220); on a cell whose normal has two non-zero
 *     entries (external edge) or three (external corner) the step also runs
 *     OuterVelocityEdgeProcessor3D / OuterVelocityCornerProcessor3D after
 *     streaming (boundary/boundaryPostProcessors3D.hh:164-294), edges before
 *     corners as their priorities 0 and 1 order them. */
#define OLBGFX_LOCAL_VELOCITY       7
#define OLBGFX_FIXED_VELOCITY       8
/* SmagorinskyBGKdynamics (dynamics/smagorinskyBGKdynamics.h:37-44): BGK with
 * the effective relaxation of collision::SmagorinskyEffectiveOmega
 * (dynamics/collisionLES.h:59-89), constant `smagorinsky` */
#define OLBGFX_SMAGORINSKY_BGK      9
/* LocalVelocity on other boundary shapes (boundary/localVelocity3D.h:86-99)
 * and LocalPressure faces (boundary/localPressure3D.h:47-58), all
 * CombinedRLBdynamics around BGK that need the cell's discrete normal:
 *   INNER_EDGE_VELOCITY: InnerEdgeVelocityTuple3D (normal with two non-zero
 *     entries): density the mean of velocityBMRho over both walls, stress
 *     with the populations entering from both walls bounced back
 *     (momenta/elements.h:459-482, 1871-1892);
 *   INNER_CORNER_VELOCITY: InnerCornerVelocityTuple3D (three non-zero
 *     entries; elements.h:432-457, 1843-1869);
 *   LOCAL_PRESSURE: RegularizedPressureBoundaryTuple (flat normal): density
 *     fixed at wall_rho or the RHO field, velocity the wall's tangential
 *     VELOCITY with the normal component from the known populations
 *     (momenta/aliases.h:152-159, elements.h:653-691). */
#define OLBGFX_INNER_EDGE_VELOCITY   10
#define OLBGFX_INNER_CORNER_VELOCITY 11
#define OLBGFX_LOCAL_PRESSURE        12
/* InterpolatedVelocity and InterpolatedPressure faces
 * (boundary/interpolatedVelocity3D.h:45-100, interpolatedPressure3D.h:45-85),
 * flat normal only: BGK through BasicDirichletVelocityBoundaryTuple (density
 * from the known populations, the wall VELOCITY) or
 * BasicDirichletPressureBoundaryTuple (fixed density, normal velocity from
 * the known populations) (momenta/aliases.h:173-198), and after streaming
 * PlaneFdBoundaryProcessor3D (boundary/boundaryPostProcessors3D.hh:39-92):
 * every population set to feq + fneq of the finite-difference strain.  The
 * edges and corners of InterpolatedVelocity are the LocalVelocity kinds
 * (FIXED_VELOCITY, INNER_EDGE_VELOCITY, INNER_CORNER_VELOCITY). */
#define OLBGFX_INTERPOLATED_VELOCITY 13
#define OLBGFX_INTERPOLATED_PRESSURE 14

/* per-cell fields */
#define OLBGFX_FIELD_VELOCITY 0  /* descriptors::VELOCITY, 3 values per cell */
#define OLBGFX_FIELD_RHO      1  /* momenta::FixedDensity::RHO, 1 value      */
#define OLBGFX_FIELD_FORCE    2  /* descriptors::FORCE, 3 values per cell    */

#define OLBGFX_MAX_DYNAMICS 16

typedef struct {
  int32_t nx, ny, nz;    /* core extent (BlockStructureD::_core)            */
  int32_t padding;       /* overlap width, >= 1 (BlockStructureD::_padding) */
  int32_t periodic[3];   /* CuboidDecomposition::setPeriodicity             */
  int32_t storage;       /* OLBGFX_STORE_*                                   */
} olbgfx_lattice_desc;

typedef struct {
  int32_t kind;              /* OLBGFX_* dynamics kind                       */
  double omega;              /* descriptors::OMEGA                            */
  double magic;              /* collision::TRT::MAGIC                         */
  double force[3];           /* uniform FORCE unless a FORCE field is set     */
  double wall_velocity[3];   /* uniform VELOCITY unless a VELOCITY field is set */
  double wall_rho;           /* uniform RHO unless a RHO field is set (1)     */
  double smagorinsky;        /* collision::LES::SMAGORINSKY                    */
} olbgfx_dynamics_desc;

/* ---- lattice lifetime --------------------------------------------------
 * Replaces ConcreteBlockLattice<T,D,GPU_*> construction, which allocates one
 * cyclic column per population (core/blockLattice.hh:392-415,
 * core/platform/gpu/hip/column.hh:217-271).  All populations start at 0
 * (rest equilibrium with rho = 1) and every cell has NoDynamics. */
olbgfx_status olbgfx_lattice_create(const olbgfx_lattice_desc* desc, olbgfx_lattice** out);
olbgfx_status olbgfx_lattice_destroy(olbgfx_lattice* lattice);

/* split steps on (default) or off.  A lattice with velocity-boundary cells
 * (LocalVelocity, InterpolatedVelocity and their pressure, edge and corner
 * forms, FixedVelocity) whose most frequent core dynamics is BGK, TRT,
 * ForcedBGK, ForcedTRT or Smagorinsky BGK and whose other core cells are at
 * most a quarter steps in two launches: that dynamics' own kernel over every
 * core cell (with 4- and 8-byte storage also updating the LocalVelocity and
 * FixedVelocity cells), skipping the others, and the masked kernel over the
 * list of the others (the GPU form of the reference's CPU applyDominant,
 * core/platform/cpu/sisd/operator.h:143-228).  Off: one masked launch.  Both
 * update every core cell once with the same arithmetic and agree bit for bit
 * in FP32, BF16 and FP64 storage (in FP16 to the last bit: a store may fuse
 * its narrowing with an FMA in one kernel and not in the other); steps with
 * statistics on always take one launch.  No reference counterpart for the
 * switch itself. */
olbgfx_status olbgfx_set_split(olbgfx_lattice* lattice, int32_t on);

/* the launch the next step takes: *kind the dominant dynamics kind of a
 * split step and *residual the number of cells of its list, or -1 and 0 for
 * one launch.  No reference counterpart. */
olbgfx_status olbgfx_get_split(olbgfx_lattice* lattice, int32_t* kind, uint32_t* residual);

/* number of HIP devices visible (0 when none); checkPlatform<GPU_HIP>
 * (core/platform/gpu/hip/communicator.hh:45-81) checks the same */
olbgfx_status olbgfx_device_count(int32_t* count);

/* bound the steps queued ahead of the host: olbgfx_collide_and_stream then
 * waits, before returning, until step t - steps has finished (0 = no bound,
 * the default).  A host timer around a run of steps, as the reference's
 * examples keep (cavity3d.cpp:200-221), then misses at most that many steps.
 * No reference counterpart (the reference's GPU platform synchronises per
 * stage, core/superLattice.hh:800-815). */
olbgfx_status olbgfx_set_inflight_limit(olbgfx_lattice* lattice, uint32_t steps);

/* number of cells of the padded block (BlockStructureD::getNcells,
 * core/blockStructure.h:137-148) */
olbgfx_status olbgfx_lattice_ncells(const olbgfx_lattice* lattice, uint64_t* ncells);

/* neighbour distances d_i = c_i . projection
 * (BlockStructureD::getNeighborDistance, core/blockStructure.h:180-182) */
olbgfx_status olbgfx_neighbor_distances(const olbgfx_lattice* lattice, int64_t out[19]);

/* current ring offsets s_i in [0, ncells): population i of cell n is stored
 * at slot (n + s_i) mod ncells (CyclicColumn::operator[],
 * core/platform/cpu/sisd/column.h:162-170).  Synchronises. */
olbgfx_status olbgfx_ring_shifts(olbgfx_lattice* lattice, int64_t out[19]);

/* ---- dynamics ------------------------------------------------------------
 * Replaces DynamicsPromise::realize<PLATFORM> + ConcreteBlockCollisionO::set
 * (core/blockDynamicsMap.h:282-350, core/platform/gpu/hip/operator.h:85-97):
 * table slot `id` (< OLBGFX_MAX_DYNAMICS) holds one dynamics with its
 * parameters (setParameter<OMEGA>, TRT::MAGIC, FORCE); an unknown kind is
 * rejected with OLBGFX_ERR_UNSUPPORTED. */
olbgfx_status olbgfx_set_dynamics(olbgfx_lattice* lattice, uint32_t id,
                                  const olbgfx_dynamics_desc* dynamics);

/* per-cell dynamics ids, one byte per padded cell in CellID order; cells of
 * the padding never collide (core/platform/cpu/sisd/operator.h:175-187). */
olbgfx_status olbgfx_set_dynamics_mask(olbgfx_lattice* lattice, const uint8_t* ids);

/* per-cell field (BlockLattice::defineField, core/blockLattice.h:373-390);
 * values holds ncells*components entries in CellID order, NULL removes the
 * field so the uniform dynamics value applies again. */
olbgfx_status olbgfx_set_cell_field(olbgfx_lattice* lattice, int32_t field,
                                    int32_t value_type, const void* values);

/* discrete normal of every cell (3 x int8 per padded cell, CellID order,
 * entries in {-1, 0, 1}, pointing out of the fluid), as boundary::set
 * computes it (geometry/discreteNormals.hh:176, boundary/setBoundary.h:76-100);
 * NULL removes them.  Read by LOCAL_VELOCITY and FIXED_VELOCITY cells; a step
 * fails with OLBGFX_ERR_INVALID if a LOCAL_VELOCITY cell has no flat normal. */
olbgfx_status olbgfx_set_cell_normals(olbgfx_lattice* lattice, const int8_t* normals);

/* setBouzidiBoundary<T, DESCRIPTOR, BouzidiPostProcessor>
 * (boundary/setBouzidiBoundary.h:193-330): the cells that run
 * BouzidiPostProcessor (setBouzidiBoundary.h:48-80) after each stream, before
 * the LocalVelocity processors (its priority -1), and their BOUZIDI_DISTANCE
 * field (boundary/bouzidiFields.h:32-36).  cells: n core cell ids in
 * ascending order; q: 19 distances per cell (q[19*k + i] for direction i of
 * cells[k]; < 0 none, 0 full-way bounce-back, > 0 interpolated) in value_type.
 * The reference applies the cells one after another; here they are updated
 * concurrently, so a list in which one cell's update reads a population that
 * another's writes is rejected (OLBGFX_ERR_INVALID) -- the setter's q = 0.5
 * pairs between two solid cells resolve in ascending-id order before the
 * call.  n = 0 removes them.  Lattices with Bouzidi cells cannot join an
 * ensemble (OLBGFX_ERR_UNSUPPORTED). */
olbgfx_status olbgfx_set_bouzidi(olbgfx_lattice* lattice, int64_t n, const int64_t* cells,
                                 int32_t value_type, const void* q);

/* the cells that run BouzidiVelocityPostProcessor (setBouzidiBoundary.h:
 * 129-164, the inflow of the Poiseuille KAT's non-forced Bouzidi rows) with
 * their BOUZIDI_DISTANCE and BOUZIDI_VELOCITY fields (19 values each per
 * cell, the coefficient c_i . u_wall that setBouzidiVelocity writes,
 * setBouzidiBoundary.h:382-420): the Bouzidi update of olbgfx_set_bouzidi
 * minus the wall-velocity term (2 or 1/q) t_i coefficient / cs^2.  Applied
 * after the olbgfx_set_bouzidi list each step (same priority, later in the
 * reference's operator-type order); replaces an earlier list, n = 0
 * removes it.  Same rules for the cells as olbgfx_set_bouzidi. */
olbgfx_status olbgfx_set_bouzidi_velocity(olbgfx_lattice* lattice, int64_t n, const int64_t* cells,
                                          int32_t value_type, const void* q, const void* velocity);

/* ---- populations -----------------------------------------------------------
 * iniEquilibrium(indicator, rho, u) (core/blockLattice.hh:202-236,
 * dynamics/interface.h:111-121): f_bar_i = feq_i(rho, u) on selected cells.
 * sel: one byte per cell (NULL = all), rho: ncells, u: 3*ncells. */
olbgfx_status olbgfx_ini_equilibrium(olbgfx_lattice* lattice, const uint8_t* sel,
                                     int32_t value_type, const void* rho, const void* u);

/* setProcessingContext(Evaluation / Simulation) of the population columns
 * (core/platform/gpu/hip/column.hh:289-300): copy the populations in logical
 * order, host[i*ncells + n] = f_bar_i(n).  Synchronise. */
olbgfx_status olbgfx_get_populations(olbgfx_lattice* lattice, int32_t value_type, void* host);
olbgfx_status olbgfx_set_populations(olbgfx_lattice* lattice, int32_t value_type, const void* host);

/* ---- the step ---------------------------------------------------------------
 * SuperLattice::collideAndStream (core/superLattice.hh:663-725) for one block:
 * collide every core cell, fill the periodic overlap, rotate the rings.
 * Enqueued on `stream`; does not synchronise. */
olbgfx_status olbgfx_collide_and_stream(olbgfx_lattice* lattice, uint32_t steps, void* stream);

/* BlockLattice::postProcess() of the PostStream stage on the current
 * populations, without a step: the Bouzidi, LocalVelocity edge and corner
 * processors.  BlockLattice::initialize (core/blockLattice.hh:56-67) runs it
 * once before the first step, so a case that starts from its own initial
 * values calls it once after setting them.  Asynchronous on `stream`. */
olbgfx_status olbgfx_post_process(olbgfx_lattice* lattice, void* stream);

/* ---- statistics and reductions -------------------------------------------
 * BlockLattice::setStatisticsEnabled (core/blockLattice.h:239-244).  With statistics on,
 * each step accumulates, in FP64, the LatticeStatistics aggregate
 * (core/latticeStatistics.h:50-70): out = {nCells, sum rho, sum uSqr,
 * max uSqr} of the last step.  get_statistics synchronises. */
olbgfx_status olbgfx_set_statistics(olbgfx_lattice* lattice, int32_t on);
olbgfx_status olbgfx_get_statistics(olbgfx_lattice* lattice, double out[4]);

/* FP64 conservation check over all colliding core cells, computed from the
 * stored populations: out = {nCells, sum (rho-1), sum j_x, sum j_y, sum j_z,
 * sum |u|^2, max |u|^2, 0}.  Replaces the thrust reductions of
 * core/platform/gpu/hip/statistics.hh:79-115.  Synchronises. */
olbgfx_status olbgfx_reduce_moments(olbgfx_lattice* lattice, double out[8], void* stream);

/* The same reduction for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  It runs the kernel of
 * olbgfx_reduce_moments into a fixed device array of the library (4 MiB per
 * device, held until the process ends) instead of a temporary allocation,
 * and returns the same values bit for bit.  Lattices of more than 2^24 core
 * cells (65536 blocks of 256) are refused with OLBGFX_ERR_INVALID.  It calls
 * no hipMalloc or hipFree.  Calls within one process are serialised on a
 * lock, since every call on a device shares that device's array, which
 * keeps the last call's partial sums.  Synchronises. */
olbgfx_status olbgfx_reduce_moments_noalloc(olbgfx_lattice* lattice, double out[8], void* stream);

/* density and velocity of every padded cell through the momenta of its
 * dynamics (Cell::computeRhoU, dynamics/momenta/elements.h): NoDynamics
 * 1 and 0, BounceBack its RHO (wall_rho) and 0, BounceBackVelocity its RHO
 * and wall velocity, LocalVelocity the boundary density and wall velocity,
 * forced dynamics u + F/2, the rest bulk moments.  FP32 or FP64 host arrays:
 * rho[ncells], u[3*ncells].  Synchronises. */
olbgfx_status olbgfx_compute_rho_u(olbgfx_lattice* lattice, int32_t value_type, void* rho, void* u);

/* ---- launch configuration ---------------------------------------------------
 * threads per block of the step kernel (64..256, multiple of 64; 0 = 256) and
 * an occupancy cap in waves per SIMD (1..8; 0 = as many as registers allow),
 * enforced by reserving LDS per block.  Used by the autotuner; no reference
 * counterpart (the reference fixes 64-thread blocks,
 * core/platform/gpu/hip/operator.hh:426-433). */
olbgfx_status olbgfx_set_launch(olbgfx_lattice* lattice, uint32_t block_threads,
                                uint32_t waves_per_simd);

/* order in which the step kernel's blocks visit the cells: -1 (default) gives
 * each of the 8 XCDs one contiguous part of the block; k in [0, 12] lets the
 * XCDs take turns on runs of 2^k consecutive blocks.  Results do not depend
 * on it.  Used by the autotuner; no reference counterpart. */
olbgfx_status olbgfx_set_block_order(olbgfx_lattice* lattice, int32_t xcd_run_log2);

/* ---- ensembles of independent lattices ------------------------------------
 * A batch of lattices (same storage type, any extents and dynamics) advanced
 * together: one persistent kernel per step walks a tile list that spans every
 * member, and the ring offsets advance on the device so a captured hipGraph
 * can be replayed.  Members stay owned by the caller and must outlive the
 * ensemble.  No reference counterpart: OpenLB advances each
 * SuperLattice::collideAndStream (core/superLattice.hh:663-725) on its own. */
olbgfx_status olbgfx_ensemble_create(olbgfx_lattice* const* members, uint32_t count,
                                     olbgfx_ensemble** out);
olbgfx_status olbgfx_ensemble_destroy(olbgfx_ensemble* ensemble);

/* threads per block (64..256), occupancy cap in waves per SIMD (0 = none) and
 * tiles of block_threads cells per persistent block (>= 1); drops a captured
 * graph.  Used by the autotuner. */
olbgfx_status olbgfx_ensemble_set_launch(olbgfx_ensemble* ensemble, uint32_t block_threads,
                                         uint32_t waves_per_simd, uint32_t tiles_per_block);

/* schedule of ensemble steps: members are packed, in order, into groups of at
 * most group_bytes of population storage (0 = all members in one group) and
 * each group advances steps_per_group steps (1..1024) before the next group
 * runs; members are independent, so every member still ends a call
 * `steps` steps further.  A group that fits the 256 MB Infinity Cache is then
 * re-read from it.  ticket = 1 lets the last block of each launch advance the
 * ring offsets (one launch per group step); 0 uses a separate small kernel.
 * Drops a captured graph.  Used by the autotuner. */
olbgfx_status olbgfx_ensemble_set_schedule(olbgfx_ensemble* ensemble, uint64_t group_bytes,
                                           uint32_t steps_per_group, int32_t ticket);

/* block order of the persistent kernel, as olbgfx_set_block_order; drops a
 * captured graph.  Used by the autotuner. */
olbgfx_status olbgfx_ensemble_set_block_order(olbgfx_ensemble* ensemble, int32_t xcd_run_log2);

/* advance every member by `steps` steps, one persistent launch per step */
olbgfx_status olbgfx_ensemble_step(olbgfx_ensemble* ensemble, uint32_t steps, void* stream);

/* the comparison path: each member launched on its own, member after member
 * (olbgfx_collide_and_stream per member and step) */
olbgfx_status olbgfx_ensemble_step_separately(olbgfx_ensemble* ensemble, uint32_t steps,
                                              void* stream);

/* capture `steps_per_launch` ensemble steps into a hipGraph (synchronises);
 * launch_graph replays it `launches` times on `stream`.  A change to any
 * member's dynamics, mask or fields requires a new capture. */
olbgfx_status olbgfx_ensemble_capture(olbgfx_ensemble* ensemble, uint32_t steps_per_launch);
olbgfx_status olbgfx_ensemble_launch_graph(olbgfx_ensemble* ensemble, uint32_t launches,
                                           void* stream);

/* number of core cells over all members (cell updates per ensemble step) */
olbgfx_status olbgfx_ensemble_cells(const olbgfx_ensemble* ensemble, uint64_t* cells);

/* ---- timing (HIP events on `stream`; these synchronise) -------------------
 * elapsed milliseconds of `steps` steps; for ensembles mode 0 = persistent
 * kernel per step, 1 = replay of the captured graph, 2 = separate launches,
 * 3 = separate launches with a synchronisation after each member's step (the
 * way SuperLattice::collideAndStream waits between its stages,
 * core/superLattice.hh:663-725). */
olbgfx_status olbgfx_time_steps(olbgfx_lattice* lattice, uint32_t steps, void* stream, float* ms);
olbgfx_status olbgfx_ensemble_time(olbgfx_ensemble* ensemble, int32_t mode, uint32_t steps,
                                   void* stream, float* ms);

/* ---- halo planes of a block split over several GPUs -------------------------
 * For a non-periodic face (axis 0..2, side 0 = low, 1 = high) of a block whose
 * neighbour block lives elsewhere.  After a step, the five populations that
 * left the core through the face sit in the ghost plane beyond it;
 * olbgfx_halo_pack gathers them into a device buffer (population-major, the
 * plane's cells in CellID order, raw storage words).  The neighbour passes
 * that buffer to olbgfx_halo_unpack for its facing side, which writes them as
 * the incoming populations of its boundary plane, so the next step sees what
 * one undivided block would.  Both sides must have the same cross-section.
 * Stands in for the overlap copy between blocks of
 * ConcreteBlockCommunicator (communication/blockCommunicationNeighborhood.hh
 * :67-107, core/platform/gpu/hip/communicator.hh:529-585), reduced to the
 * populations that cross the face.  Enqueued on `stream`. */
olbgfx_status olbgfx_halo_bytes(const olbgfx_lattice* lattice, int32_t axis, uint64_t* bytes);
/* a device buffer of olbgfx_halo_bytes for faces normal to `axis`, owned by
 * the library until olbgfx_halo_buffer_destroy (a caller may pass its own
 * device memory to pack/unpack instead, e.g. an RCCL receive buffer) */
olbgfx_status olbgfx_halo_buffer_create(const olbgfx_lattice* lattice, int32_t axis, void** device_buffer);
olbgfx_status olbgfx_halo_buffer_destroy(void* device_buffer);
olbgfx_status olbgfx_halo_pack(olbgfx_lattice* lattice, int32_t axis, int32_t side,
                               void* device_buffer, void* stream);
olbgfx_status olbgfx_halo_unpack(olbgfx_lattice* lattice, int32_t axis, int32_t side,
                                 const void* device_buffer, void* stream);

/* ---- cell lists ----------------------------------------------------------------
 * ConcreteCommunicatable<CyclicColumn>::size / serialize / deserialize for
 * all 19 populations (core/platform/gpu/hip/column.h:205-254; the GPU
 * communicator's gather and scatter of overlap cells,
 * core/platform/gpu/hip/communicator.hh:529-585): a list of CellIDs (copied
 * to the device once; ids beyond the padded block are rejected), and the
 * populations of those cells at the current offsets to or from a device
 * buffer, population-major (buffer[i * n + k] = f_bar_i(cells[k])), raw
 * storage words.  Gather and scatter are enqueued on `stream`. */
olbgfx_status olbgfx_cell_list_create(const olbgfx_lattice* lattice, const uint32_t* cells, uint32_t n,
                                      olbgfx_cell_list** out);
olbgfx_status olbgfx_cell_list_destroy(olbgfx_cell_list* list);
olbgfx_status olbgfx_cell_list_bytes(const olbgfx_lattice* lattice, const olbgfx_cell_list* list,
                                     uint64_t* bytes);
olbgfx_status olbgfx_gather_cells(olbgfx_lattice* lattice, const olbgfx_cell_list* list,
                                  void* device_buffer, void* stream);
olbgfx_status olbgfx_scatter_cells(olbgfx_lattice* lattice, const olbgfx_cell_list* list,
                                   const void* device_buffer, void* stream);

/* device memory owned by the library, for halo planes, gathered cells or
 * RCCL transfers (no reference counterpart: OpenLB allocates these buffers
 * inside its communicator) */
olbgfx_status olbgfx_device_buffer_create(uint64_t bytes, void** device_buffer);
olbgfx_status olbgfx_device_buffer_destroy(void* device_buffer);

/* ---- autotuner ---------------------------------------------------------------
 * Times candidate launch configurations on the lattice (ensemble) itself
 * with HIP events on `stream`, applies the fastest, and leaves the simulation
 * state as it found it: populations, offsets and statistics are saved in
 * device memory before the trials and put back after them (this needs a
 * second copy of the populations; ERR_NOMEM if it does not fit).
 *
 * Lattices: first every (block_threads x waves_per_simd) pair with the
 * default block order, then every block order with the three fastest pairs.
 * Ensembles: first every (block_threads x waves_per_simd x tiles_per_block)
 * with one group, then every block order with the three fastest, then every
 * (group_bytes x steps_per_group x ticket) with the best launch.  A list with n == 0 takes the built-in candidates; a null
 * `space` takes them all.  Each trial runs one untimed step (steps_per_group
 * for schedules) and then `steps` timed steps (ensembles: at least 8).
 *
 * `cache_path` (may be null) names a text file of earlier choices keyed by
 * extent, padding, periodicity, storage and the dynamics mix (ensembles: by
 * all members); a hit is applied without trials and reported with cached = 1.
 * No reference counterpart: OpenLB fixes its launch shapes
 * (core/platform/gpu/hip/operator.hh:426-433). */
#define OLBGFX_TUNE_CHOICES 8
#define OLBGFX_TUNE_TRIALS 128

typedef struct olbgfx_tune_space {
  uint32_t n_block_threads;
  uint32_t block_threads[OLBGFX_TUNE_CHOICES];
  uint32_t n_waves;
  uint32_t waves_per_simd[OLBGFX_TUNE_CHOICES];
  uint32_t n_orders;
  int32_t block_order[OLBGFX_TUNE_CHOICES];
  uint32_t n_tiles;
  uint32_t tiles_per_block[OLBGFX_TUNE_CHOICES];
  uint32_t n_group_bytes;
  uint64_t group_bytes[OLBGFX_TUNE_CHOICES];
  uint32_t n_steps_per_group;
  uint32_t steps_per_group[OLBGFX_TUNE_CHOICES];
  uint32_t n_tickets;
  int32_t ticket[2];
} olbgfx_tune_space;

typedef struct olbgfx_tune_trial {
  uint32_t block_threads;
  uint32_t waves_per_simd;
  int32_t block_order;        /* lattices */
  uint32_t tiles_per_block;   /* ensembles */
  uint64_t group_bytes;       /* ensembles */
  uint32_t steps_per_group;   /* ensembles */
  int32_t ticket;             /* ensembles */
  double ms_per_step;
} olbgfx_tune_trial;

typedef struct olbgfx_tune_result {
  olbgfx_tune_trial best;
  int32_t cached;
  uint32_t ntrials;
  olbgfx_tune_trial trials[OLBGFX_TUNE_TRIALS];
} olbgfx_tune_result;

olbgfx_status olbgfx_autotune_lattice(olbgfx_lattice* lattice, uint32_t steps,
                                      const olbgfx_tune_space* space, const char* cache_path,
                                      void* stream, olbgfx_tune_result* out);
olbgfx_status olbgfx_autotune_ensemble(olbgfx_ensemble* ensemble, uint32_t steps,
                                       const olbgfx_tune_space* space, const char* cache_path,
                                       void* stream, olbgfx_tune_result* out);

/* block until all work of this lattice is done */
olbgfx_status olbgfx_synchronize(olbgfx_lattice* lattice);

/* message of the last failed call on this thread ("" if none) */
const char* olbgfx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* OLBGFX_H */
