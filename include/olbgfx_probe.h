/*
 * olbgfx_probe.h -- bandwidth probes of the access patterns of the step
 * kernel, for the measurement (the ceilings its HBM fraction is read
 * against).  No reference counterpart: OpenLB reports MLUPs only
 * (utilities/timer.hh:72-76).
 *
 * Each probe allocates its own device memory, runs `reps` timed repetitions
 * after one untimed one for each launch shape of a small sweep (the copy over
 * 256..16384 looping workgroups and one workgroup per 256 elements, with
 * plain and with nontemporal accesses; the ring patterns over the occupancy
 * caps the step kernel's autotuner tries), frees the memory and returns the best
 * bandwidth in GB/s (10^9 bytes per second) of the bytes the pattern moves,
 * timed with HIP events on `stream` (synchronises).
 */
#ifndef OLBGFX_PROBE_H
#define OLBGFX_PROBE_H

#include <stdint.h>

#include "olbgfx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* plain device copy of 19 rings' worth of storage (19 x padded cells x
 * element size), 16-byte loads and stores; moves 2 x that */
#define OLBGFX_PROBE_COPY          0
/* every core cell loads its 19 populations from the PS rings (the slots of
 * the step kernel) and keeps one sum per block; moves 19 elements per cell */
#define OLBGFX_PROBE_RING_READ     1
/* every core cell stores 19 values into its ring slots; 19 elements per cell */
#define OLBGFX_PROBE_RING_WRITE    2
/* every core cell loads its 19 populations and stores them back scaled, in
 * place: the step kernel's traffic without its arithmetic; 2 x 19 elements */
#define OLBGFX_PROBE_RING_INPLACE  3
/* every core cell loads its 19 populations from the rings of one buffer and
 * stores them, scaled, into the rings of a second buffer at slot n (offset
 * zero): the traffic of a two-buffer step, whose write offsets are free;
 * 2 x 19 elements per cell */
#define OLBGFX_PROBE_RING_TWO_BUFFER 4

/* desc: the lattice shape (extent, padding, storage F32/F16/BF16/F64; the
 * periodic flags are ignored).  shift_steps: the ring offsets are those after
 * that many steps, s_i = -t d_i mod count (0 = every ring aligned with the
 * cell index).  block_order: as olbgfx_set_block_order.  The indexing of
 * core cells, the block shape (256 threads) and the ring layout are the step
 * kernel's. */
olbgfx_status olbgfx_probe_bandwidth(const olbgfx_lattice_desc* desc, int32_t pattern, uint32_t shift_steps,
                                     int32_t block_order, uint32_t reps, void* stream, double* gbps);

#ifdef __cplusplus
}
#endif

#endif /* OLBGFX_PROBE_H */
