"""openlb_amd -- the gfx950 (MI355X) lattice Boltzmann platform, Python host.

The compute path is ``libolbgfx.so`` (hand-written HIP kernels behind the C ABI
of ``include/olbgfx.h``).  This module loads it with ctypes and mirrors the
block-lattice calls of OpenLB's platform interface: dynamics per cell,
``setParameter``-style dynamics parameters, ``iniEquilibrium``,
``collideAndStream``, statistics and the evaluation view of the populations.

There is no CPU fallback: if the library is missing or cannot be loaded, the
import of :class:`Lattice` fails with an error that says so.

A process that also uses PyTorch on the GPU must import torch first: torch's
wheel carries its own HIP runtime under the same soname, and the library then
binds to it, so both share one runtime (streams and events interoperate).
Loaded the other way round, two HIP runtimes end up in the process and
torch's finds no GPU.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libolbgfx.so")

# C ABI constants (include/olbgfx.h)
STORE = {"f32": 0, "f16": 1, "bf16": 2, "f64": 3}
VALUE_F32, VALUE_F64 = 0, 1
NO_DYNAMICS, BGK, TRT, FORCED_BGK, FORCED_TRT, BOUNCE_BACK, BOUNCE_BACK_VELOCITY = range(7)
LOCAL_VELOCITY, FIXED_VELOCITY, SMAGORINSKY_BGK = 7, 8, 9
INNER_EDGE_VELOCITY, INNER_CORNER_VELOCITY, LOCAL_PRESSURE = 10, 11, 12
INTERPOLATED_VELOCITY, INTERPOLATED_PRESSURE = 13, 14
FIELD_VELOCITY, FIELD_RHO, FIELD_FORCE = 0, 1, 2
MAX_DYNAMICS = 16
Q = 19

KIND_NAMES = {
    "NoDynamics": NO_DYNAMICS,
    "BGKdynamics": BGK,
    "TRTdynamics": TRT,
    "ForcedBGKdynamics": FORCED_BGK,
    "ForcedTRTdynamics": FORCED_TRT,
    "BounceBack": BOUNCE_BACK,
    "BounceBackVelocity": BOUNCE_BACK_VELOCITY,
}


class OlbgfxError(RuntimeError):
    """A C ABI call returned an error status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"olbgfx status {status}: {message}")
        self.status = status


class _LatticeDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("padding", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("storage", ctypes.c_int32),
    ]


class _DynamicsDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("omega", ctypes.c_double),
        ("magic", ctypes.c_double),
        ("force", ctypes.c_double * 3),
        ("wall_velocity", ctypes.c_double * 3),
        ("wall_rho", ctypes.c_double),
        ("smagorinsky", ctypes.c_double),
    ]


_lib = None


def library() -> ctypes.CDLL:
    """Load libolbgfx.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(the gfx950 kernels have no CPU fallback)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32
    sigs = {
        "olbgfx_lattice_create": [ctypes.POINTER(_LatticeDesc), ctypes.POINTER(vp)],
        "olbgfx_lattice_destroy": [vp],
        "olbgfx_device_count": [ctypes.POINTER(i32)],
        "olbgfx_set_split": [vp, i32],
        "olbgfx_get_split": [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint32)],
        "olbgfx_probe_bandwidth": [ctypes.POINTER(_LatticeDesc), i32, u32, i32, u32, vp,
                                   ctypes.POINTER(ctypes.c_double)],
        "olbgfx_set_inflight_limit": [vp, u32],
        "olbgfx_lattice_ncells": [vp, ctypes.POINTER(ctypes.c_uint64)],
        "olbgfx_neighbor_distances": [vp, ctypes.POINTER(ctypes.c_int64)],
        "olbgfx_ring_shifts": [vp, ctypes.POINTER(ctypes.c_int64)],
        "olbgfx_set_dynamics": [vp, u32, ctypes.POINTER(_DynamicsDesc)],
        "olbgfx_set_dynamics_mask": [vp, vp],
        "olbgfx_set_cell_field": [vp, i32, i32, vp],
        "olbgfx_set_cell_normals": [vp, vp],
        "olbgfx_set_bouzidi": [vp, ctypes.c_int64, vp, i32, vp],
        "olbgfx_set_bouzidi_velocity": [vp, ctypes.c_int64, vp, i32, vp, vp],
        "olbgfx_post_process": [vp, vp],
        "olbgfx_ini_equilibrium": [vp, vp, i32, vp, vp],
        "olbgfx_get_populations": [vp, i32, vp],
        "olbgfx_set_populations": [vp, i32, vp],
        "olbgfx_collide_and_stream": [vp, u32, vp],
        "olbgfx_set_statistics": [vp, i32],
        "olbgfx_get_statistics": [vp, ctypes.POINTER(ctypes.c_double)],
        "olbgfx_reduce_moments": [vp, ctypes.POINTER(ctypes.c_double), vp],
        "olbgfx_reduce_moments_noalloc": [vp, ctypes.POINTER(ctypes.c_double), vp],
        "olbgfx_compute_rho_u": [vp, i32, vp, vp],
        "olbgfx_set_launch": [vp, u32, u32],
        "olbgfx_set_block_order": [vp, i32],
        "olbgfx_synchronize": [vp],
        "olbgfx_ensemble_create": [ctypes.POINTER(vp), u32, ctypes.POINTER(vp)],
        "olbgfx_ensemble_destroy": [vp],
        "olbgfx_ensemble_set_launch": [vp, u32, u32, u32],
        "olbgfx_ensemble_step": [vp, u32, vp],
        "olbgfx_ensemble_set_schedule": [vp, ctypes.c_uint64, u32, i32],
        "olbgfx_ensemble_set_block_order": [vp, i32],
        "olbgfx_ensemble_step_separately": [vp, u32, vp],
        "olbgfx_ensemble_capture": [vp, u32],
        "olbgfx_ensemble_launch_graph": [vp, u32, vp],
        "olbgfx_ensemble_cells": [vp, ctypes.POINTER(ctypes.c_uint64)],
        "olbgfx_time_steps": [vp, u32, vp, ctypes.POINTER(ctypes.c_float)],
        "olbgfx_ensemble_time": [vp, i32, u32, vp, ctypes.POINTER(ctypes.c_float)],
        "olbgfx_autotune_lattice": [vp, u32, vp, ctypes.c_char_p, vp, vp],
        "olbgfx_halo_bytes": [vp, i32, ctypes.POINTER(ctypes.c_uint64)],
        "olbgfx_cell_list_create": [vp, vp, u32, ctypes.POINTER(vp)],
        "olbgfx_cell_list_destroy": [vp],
        "olbgfx_cell_list_bytes": [vp, vp, ctypes.POINTER(ctypes.c_uint64)],
        "olbgfx_gather_cells": [vp, vp, vp, vp],
        "olbgfx_scatter_cells": [vp, vp, vp, vp],
        "olbgfx_device_buffer_create": [ctypes.c_uint64, ctypes.POINTER(vp)],
        "olbgfx_device_buffer_destroy": [vp],
        "olbgfx_halo_buffer_create": [vp, i32, ctypes.POINTER(vp)],
        "olbgfx_halo_buffer_destroy": [vp],
        "olbgfx_halo_pack": [vp, i32, i32, vp, vp],
        "olbgfx_halo_unpack": [vp, i32, i32, vp, vp],
        "olbgfx_autotune_ensemble": [vp, u32, vp, ctypes.c_char_p, vp, vp],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int32
    lib.olbgfx_last_error.argtypes = []
    lib.olbgfx_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        msg = library().olbgfx_last_error().decode(errors="replace")
        raise OlbgfxError(status, msg)


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Dynamics:
    """One dynamics with its parameters (a DynamicsPromise plus setParameter)."""

    kind: int
    omega: float = 0.0
    magic: float = 0.0
    force: tuple = (0.0, 0.0, 0.0)
    wall_velocity: tuple = (0.0, 0.0, 0.0)
    wall_rho: float = 1.0
    smagorinsky: float = 0.0

    def _desc(self) -> _DynamicsDesc:
        d = _DynamicsDesc()
        d.kind = int(self.kind)
        d.omega = float(self.omega)
        d.magic = float(self.magic)
        for a in range(3):
            d.force[a] = float(self.force[a])
            d.wall_velocity[a] = float(self.wall_velocity[a])
        d.wall_rho = float(self.wall_rho)
        d.smagorinsky = float(self.smagorinsky)
        return d


class Lattice:
    """One D3Q19 block lattice resident in HBM (ConcreteBlockLattice on gfx950).

    Arrays indexed by cell follow OpenLB's CellID order over the padded block
    (core/blockStructure.h:150-160): CellID = ((x+p)*Py + (y+p))*Pz + (z+p),
    i.e. a C-order ravel of a (Px, Py, Pz) array.
    """

    def __init__(self, extent, padding: int = 3, periodic=(False, False, False), storage: str = "f32"):
        lib = library()
        nx, ny, nz = (int(v) for v in extent)
        self.extent = (nx, ny, nz)
        self.padding = int(padding)
        self.periodic = tuple(bool(p) for p in periodic)
        self.storage = storage
        self.shape = (nx + 2 * self.padding, ny + 2 * self.padding, nz + 2 * self.padding)
        desc = _LatticeDesc()
        desc.nx, desc.ny, desc.nz = nx, ny, nz
        desc.padding = self.padding
        for a in range(3):
            desc.periodic[a] = int(self.periodic[a])
        desc.storage = STORE[storage]
        h = ctypes.c_void_p()
        _check(lib.olbgfx_lattice_create(ctypes.byref(desc), ctypes.byref(h)))
        self._h = h
        n = ctypes.c_uint64()
        _check(lib.olbgfx_lattice_ncells(self._h, ctypes.byref(n)))
        self.ncells = int(n.value)
        self._value = VALUE_F64 if storage == "f64" else VALUE_F32
        self._dtype = np.float64 if storage == "f64" else np.float32

    # -- lifetime -------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            library().olbgfx_lattice_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- structure ----------------------------------------------------------------
    def cell_id(self, x: int, y: int, z: int) -> int:
        p = self.padding
        return ((x + p) * self.shape[1] + (y + p)) * self.shape[2] + (z + p)

    def neighbor_distances(self) -> np.ndarray:
        out = np.zeros(Q, dtype=np.int64)
        _check(library().olbgfx_neighbor_distances(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def ring_shifts(self) -> np.ndarray:
        out = np.zeros(Q, dtype=np.int64)
        _check(library().olbgfx_ring_shifts(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    # -- dynamics -------------------------------------------------------------------
    def set_dynamics(self, slot: int, dynamics: Dynamics) -> None:
        d = dynamics._desc()
        _check(library().olbgfx_set_dynamics(self._h, int(slot), ctypes.byref(d)))

    def set_dynamics_mask(self, ids: np.ndarray) -> None:
        ids = np.ascontiguousarray(ids, dtype=np.uint8).reshape(-1)
        if ids.size != self.ncells:
            raise ValueError(f"mask has {ids.size} entries, lattice has {self.ncells} cells")
        _check(library().olbgfx_set_dynamics_mask(self._h, _ptr(ids)))

    def set_cell_field(self, field: int, values: np.ndarray | None) -> None:
        if values is None:
            _check(library().olbgfx_set_cell_field(self._h, field, self._value, None))
            return
        comps = 1 if field == FIELD_RHO else 3
        v = np.ascontiguousarray(values, dtype=self._dtype).reshape(-1)
        if v.size != comps * self.ncells:
            raise ValueError("field array has the wrong size")
        _check(library().olbgfx_set_cell_field(self._h, field, self._value, _ptr(v)))

    # -- populations -------------------------------------------------------------
    def set_cell_normals(self, normals: np.ndarray | None) -> None:
        """discrete normal per cell (3 x int8 in {-1, 0, 1}, CellID order), as
        boundary::set computes it; None removes them"""
        if normals is None:
            _check(library().olbgfx_set_cell_normals(self._h, None))
            return
        v = np.ascontiguousarray(normals, dtype=np.int8).reshape(-1)
        if v.size != 3 * self.ncells:
            raise ValueError("normals must hold 3 entries per cell")
        _check(library().olbgfx_set_cell_normals(self._h, _ptr(v)))

    def post_process(self, stream: int | None = None) -> None:
        """BlockLattice::postProcess(): the PostStream processors (Bouzidi,
        LocalVelocity edges and corners) on the current populations; the
        reference runs it once in initialize()"""
        _check(library().olbgfx_post_process(self._h, stream))

    def set_bouzidi(self, cells, q) -> None:
        """setBouzidiBoundary: the core cells (ascending ids) that run
        BouzidiPostProcessor after each stream and their BOUZIDI_DISTANCE
        (n x 19, < 0 where unset); openlb_amd.bouzidi computes both from a
        geometry and an indicator.  An empty list removes them."""
        cells = np.ascontiguousarray(cells, dtype=np.int64).reshape(-1)
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        if q.size != 19 * cells.size:
            raise ValueError("q must hold 19 distances per cell")
        _check(library().olbgfx_set_bouzidi(self._h, int(cells.size), _ptr(cells), VALUE_F64, _ptr(q)))

    def set_bouzidi_velocity(self, cells, q, velocity) -> None:
        """setBouzidiBoundary<BouzidiVelocityPostProcessor> with
        setBouzidiVelocity: the core cells (ascending ids), their
        BOUZIDI_DISTANCE and BOUZIDI_VELOCITY coefficients (n x 19 each),
        applied after the set_bouzidi list.  An empty list removes them."""
        cells = np.ascontiguousarray(cells, dtype=np.int64).reshape(-1)
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(velocity, dtype=np.float64).reshape(-1)
        if q.size != 19 * cells.size or v.size != 19 * cells.size:
            raise ValueError("q and velocity must hold 19 values per cell")
        _check(library().olbgfx_set_bouzidi_velocity(self._h, int(cells.size), _ptr(cells), VALUE_F64, _ptr(q),
                                                     _ptr(v)))

    def ini_equilibrium(self, rho: np.ndarray, u: np.ndarray, sel: np.ndarray | None = None) -> None:
        rho = np.ascontiguousarray(rho, dtype=self._dtype).reshape(-1)
        u = np.ascontiguousarray(u, dtype=self._dtype).reshape(-1)
        if rho.size != self.ncells or u.size != 3 * self.ncells:
            raise ValueError("rho/u arrays have the wrong size")
        s = None if sel is None else np.ascontiguousarray(sel, dtype=np.uint8).reshape(-1)
        if s is not None and s.size != self.ncells:
            raise ValueError("the selection mask needs one entry per padded cell")
        _check(library().olbgfx_ini_equilibrium(self._h, _ptr(s), self._value, _ptr(rho), _ptr(u)))

    def get_populations(self) -> np.ndarray:
        """Populations in logical order, shape (19, ncells), shifted by -t_i."""
        out = np.empty((Q, self.ncells), dtype=self._dtype)
        _check(library().olbgfx_get_populations(self._h, self._value, _ptr(out)))
        return out

    def set_populations(self, f: np.ndarray) -> None:
        f = np.ascontiguousarray(f, dtype=self._dtype)
        if f.shape != (Q, self.ncells):
            raise ValueError("populations must have shape (19, ncells)")
        _check(library().olbgfx_set_populations(self._h, self._value, _ptr(f)))

    # -- the step ------------------------------------------------------------------
    def collide_and_stream(self, steps: int = 1, stream: int | None = None) -> None:
        _check(library().olbgfx_collide_and_stream(self._h, int(steps), stream))

    collideAndStream = collide_and_stream

    def synchronize(self) -> None:
        _check(library().olbgfx_synchronize(self._h))

    def time_steps(self, steps: int, stream: int | None = None) -> float:
        """milliseconds of `steps` steps, HIP events on `stream` (synchronises)"""
        ms = ctypes.c_float()
        _check(library().olbgfx_time_steps(self._h, int(steps), stream, ctypes.byref(ms)))
        return float(ms.value)

    def set_launch(self, block_threads: int, waves_per_simd: int = 0) -> None:
        """threads per block (64..256) and an occupancy cap in waves per SIMD
        (0 = no cap)"""
        _check(library().olbgfx_set_launch(self._h, int(block_threads), int(waves_per_simd)))

    def set_split(self, on: bool = True) -> None:
        """split steps (the dominant dynamics' own kernel plus a masked
        launch over the other cells) on or off; results do not change (FP16
        storage: up to the last bit)"""
        _check(library().olbgfx_set_split(self._h, 1 if on else 0))

    def get_split(self) -> tuple:
        """(dominant dynamics kind, residual cells) of the split the next
        step takes, or (-1, 0) for one launch"""
        kind, rest = ctypes.c_int32(), ctypes.c_uint32()
        _check(library().olbgfx_get_split(self._h, ctypes.byref(kind), ctypes.byref(rest)))
        return kind.value, rest.value

    def set_block_order(self, xcd_run_log2: int = -1) -> None:
        """-1: each XCD sweeps one contiguous part; k: XCDs take turns on runs
        of 2^k blocks (results do not change)"""
        _check(library().olbgfx_set_block_order(self._h, int(xcd_run_log2)))

    # -- halo planes of a block split over several GPUs ------------------------------

    def halo_bytes(self, axis: int) -> int:
        n = ctypes.c_uint64()
        _check(library().olbgfx_halo_bytes(self._h, int(axis), ctypes.byref(n)))
        return n.value

    def halo_buffer(self, axis: int) -> "HaloBuffer":
        """a library-owned device buffer for one packed face normal to axis"""
        p = ctypes.c_void_p()
        _check(library().olbgfx_halo_buffer_create(self._h, int(axis), ctypes.byref(p)))
        return HaloBuffer(p.value, self.halo_bytes(axis))

    def cell_list(self, cells) -> "CellList":
        """CellIDs whose 19 populations gather_cells / scatter_cells move"""
        ids = np.ascontiguousarray(cells, dtype=np.uint32).reshape(-1)
        h = ctypes.c_void_p()
        _check(library().olbgfx_cell_list_create(self._h, _ptr(ids), int(ids.size), ctypes.byref(h)))
        n = ctypes.c_uint64()
        _check(library().olbgfx_cell_list_bytes(self._h, h, ctypes.byref(n)))
        return CellList(h.value, int(ids.size), n.value)

    def gather_cells(self, cells: "CellList", device_ptr: int, stream=None) -> None:
        _check(library().olbgfx_gather_cells(self._h, ctypes.c_void_p(cells.handle), ctypes.c_void_p(device_ptr),
                                             stream))

    def scatter_cells(self, cells: "CellList", device_ptr: int, stream=None) -> None:
        _check(library().olbgfx_scatter_cells(self._h, ctypes.c_void_p(cells.handle), ctypes.c_void_p(device_ptr),
                                              stream))

    def halo_pack(self, axis: int, side: int, device_ptr: int, stream=None) -> None:
        """gather the populations that left through face (axis, side) into the
        device buffer at device_ptr"""
        _check(library().olbgfx_halo_pack(self._h, int(axis), int(side), ctypes.c_void_p(device_ptr), stream))

    def halo_unpack(self, axis: int, side: int, device_ptr: int, stream=None) -> None:
        """write a neighbour's packed plane as the incoming populations of
        face (axis, side)"""
        _check(library().olbgfx_halo_unpack(self._h, int(axis), int(side), ctypes.c_void_p(device_ptr), stream))

    # -- statistics -----------------------------------------------------------------

    def set_statistics(self, on: bool) -> None:
        _check(library().olbgfx_set_statistics(self._h, int(bool(on))))

    def get_statistics(self) -> dict:
        out = (ctypes.c_double * 4)()
        _check(library().olbgfx_get_statistics(self._h, out))
        n = out[0]
        return {
            "nCells": n,
            "sumRho": out[1],
            "sumUSqr": out[2],
            "maxUSqr": out[3],
            "averageRho": out[1] / n if n else 0.0,
            # energy is 0.5 u^2 (core/latticeStatistics.hh:89)
            "averageEnergy": 0.5 * out[2] / n if n else 0.0,
            "maxU": float(np.sqrt(out[3])),
        }

    def reduce_moments(self, stream: int | None = None) -> dict:
        out = (ctypes.c_double * 8)()
        _check(library().olbgfx_reduce_moments(self._h, out, stream))
        return {
            "nCells": out[0],
            "mass": out[1],
            "momentum": (out[2], out[3], out[4]),
            "sumUSqr": out[5],
            "maxUSqr": out[6],
        }

    def compute_rho_u(self):
        rho = np.empty(self.ncells, dtype=self._dtype)
        u = np.empty(3 * self.ncells, dtype=self._dtype)
        _check(library().olbgfx_compute_rho_u(self._h, self._value, _ptr(rho), _ptr(u)))
        return rho, u.reshape(self.ncells, 3)


class CellList:
    """a device copy of CellIDs (olbgfx_cell_list_create)"""

    def __init__(self, handle: int, n: int, nbytes: int):
        self.handle = handle
        self.n = n
        self.nbytes = nbytes

    def close(self) -> None:
        if self.handle:
            library().olbgfx_cell_list_destroy(ctypes.c_void_p(self.handle))
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PROBES = {"copy": 0, "ring_read": 1, "ring_write": 2, "ring_inplace": 3, "ring_two_buffer": 4}


def probe_bandwidth(extent, padding: int, storage: str, pattern: str, shift_steps: int = 1,
                    block_order: int = -1, reps: int = 10, stream: int | None = None) -> float:
    """GB/s of one access pattern of the step kernel (include/olbgfx_probe.h)
    on a lattice of that shape and storage: a plain 16-byte copy of the rings,
    the 19 rings read, written, or read-modified-written in place, or read
    from one buffer and written to a second (ring_two_buffer)"""
    d = _LatticeDesc()
    d.nx, d.ny, d.nz = (int(v) for v in extent)
    d.padding = int(padding)
    d.storage = STORE[storage]
    out = ctypes.c_double()
    _check(library().olbgfx_probe_bandwidth(ctypes.byref(d), PROBES[pattern], int(shift_steps),
                                            int(block_order), int(reps), stream, ctypes.byref(out)))
    return out.value


def device_buffer(nbytes: int) -> "HaloBuffer":
    """library-owned device memory (olbgfx_device_buffer_create)"""
    p = ctypes.c_void_p()
    _check(library().olbgfx_device_buffer_create(int(nbytes), ctypes.byref(p)))
    return HaloBuffer(p.value, int(nbytes))


class HaloBuffer:
    """device memory for one packed halo plane or gathered cells"""

    def __init__(self, ptr: int, nbytes: int):
        self.ptr = ptr
        self.nbytes = nbytes

    def close(self) -> None:
        if self.ptr:
            library().olbgfx_halo_buffer_destroy(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """A batch of independent lattices advanced by one persistent kernel per
    step (optionally replayed from a captured hipGraph)."""

    def __init__(self, members):
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        h = ctypes.c_void_p()
        _check(library().olbgfx_ensemble_create(arr, len(self.members), ctypes.byref(h)))
        self._h = h
        n = ctypes.c_uint64()
        _check(library().olbgfx_ensemble_cells(self._h, ctypes.byref(n)))
        self.cells = int(n.value)

    def close(self) -> None:
        if getattr(self, "_h", None):
            library().olbgfx_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, block_threads: int = 256, waves_per_simd: int = 0, tiles_per_block: int = 1) -> None:
        _check(library().olbgfx_ensemble_set_launch(self._h, int(block_threads), int(waves_per_simd),
                                                     int(tiles_per_block)))

    def set_block_order(self, xcd_run_log2: int = -1) -> None:
        """block order of the persistent kernel (see Lattice.set_block_order)"""
        _check(library().olbgfx_ensemble_set_block_order(self._h, int(xcd_run_log2)))

    def set_schedule(self, group_bytes: int = 0, steps_per_group: int = 1, ticket: bool = False) -> None:
        """pack members into groups of at most group_bytes of populations (0 =
        one group); each group advances steps_per_group steps at a time; with
        ticket the last block of a launch advances the ring offsets"""
        _check(library().olbgfx_ensemble_set_schedule(self._h, int(group_bytes), int(steps_per_group),
                                                       int(bool(ticket))))

    def step(self, steps: int = 1, stream: int | None = None) -> None:
        _check(library().olbgfx_ensemble_step(self._h, int(steps), stream))

    def step_separately(self, steps: int = 1, stream: int | None = None) -> None:
        _check(library().olbgfx_ensemble_step_separately(self._h, int(steps), stream))

    def capture(self, steps_per_launch: int) -> None:
        _check(library().olbgfx_ensemble_capture(self._h, int(steps_per_launch)))

    def launch_graph(self, launches: int = 1, stream: int | None = None) -> None:
        _check(library().olbgfx_ensemble_launch_graph(self._h, int(launches), stream))

    def time(self, mode: str, steps: int, stream: int | None = None) -> float:
        """milliseconds of `steps` ensemble steps; mode 'persistent', 'graph',
        'separate' or 'separate_synchronized' (a synchronisation after each
        member's step) (HIP events on `stream`, synchronises)"""
        code = {"persistent": 0, "graph": 1, "separate": 2, "separate_synchronized": 3}[mode]
        ms = ctypes.c_float()
        _check(library().olbgfx_ensemble_time(self._h, code, int(steps), stream, ctypes.byref(ms)))
        return float(ms.value)

