"""ctypes access to the CPU restatement in olb_oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the oracle.  The product path
(openlb_amd, libolbgfx.so) never imports it.

``OracleLattice`` has the same methods as ``openlb_amd.Lattice`` so a test
sets up both sides with the same code and compares them.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
Q = 19
STORE_MODE = {"f32": 0, "f16": 1, "bf16": 2, "f64": 0}

_lib = None


def library() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        for pfx in ("orcf_", "orcd_"):
            g = lambda n: getattr(lib, pfx + n)  # noqa: E731
            g("create").argtypes = [i32] * 7
            g("create").restype = vp
            g("destroy").argtypes = [vp]
            g("ncells").argtypes = [vp]
            g("ncells").restype = ctypes.c_uint64
            g("cell_id").argtypes = [vp, i32, i32, i32]
            g("cell_id").restype = ctypes.c_uint32
            g("neighbor_distances").argtypes = [vp, vp]
            g("shifts").argtypes = [vp, vp]
            g("set_storage").argtypes = [vp, i32]
            g("set_statistics").argtypes = [vp, i32]
            g("get_statistics").argtypes = [vp, vp]
            g("set_dynamics_mask").argtypes = [vp, vp]
            g("set_dynamics").argtypes = [vp, i32, i32, ctypes.c_double, ctypes.c_double, vp, vp,
                                          ctypes.c_double, ctypes.c_double]
            g("set_cell_velocity").argtypes = [vp, vp]
            g("set_cell_rho").argtypes = [vp, vp]
            g("set_cell_force").argtypes = [vp, vp]
            g("set_cell_normals").argtypes = [vp, vp]
            g("set_bouzidi").argtypes = [vp, ctypes.c_int64, vp, vp]
            g("set_bouzidi_velocity").argtypes = [vp, ctypes.c_int64, vp, vp, vp]
            g("post_process").argtypes = [vp]
            g("get_populations").argtypes = [vp, vp]
            g("set_populations").argtypes = [vp, vp]
            g("collide_and_stream").argtypes = [vp, i32]
            g("ini_equilibrium").argtypes = [vp, vp, vp, vp]
            g("compute_rho_u").argtypes = [vp, vp, vp]
            g("set_threads").argtypes = [i32]
            g("get_threads").restype = i32
            g("f32_to_f16").argtypes = [ctypes.c_float]
            g("f32_to_f16").restype = ctypes.c_uint16
            g("f16_to_f32").argtypes = [ctypes.c_uint16]
            g("f16_to_f32").restype = ctypes.c_float
            g("f32_to_bf16").argtypes = [ctypes.c_float]
            g("f32_to_bf16").restype = ctypes.c_uint16
            g("bf16_to_f32").argtypes = [ctypes.c_uint16]
            g("bf16_to_f32").restype = ctypes.c_float
        _lib = lib
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n: int) -> None:
    library().orcf_set_threads(int(n))
    library().orcd_set_threads(int(n))


def get_threads() -> int:
    return int(library().orcf_get_threads())


class OracleLattice:
    """CPU restatement of one D3Q19 block lattice (float or double)."""

    def __init__(self, extent, padding: int = 3, periodic=(False, False, False), storage: str = "f32"):
        lib = library()
        self._double = storage == "f64"
        self._pfx = "orcd_" if self._double else "orcf_"
        self._dtype = np.float64 if self._double else np.float32
        nx, ny, nz = (int(v) for v in extent)
        self.extent = (nx, ny, nz)
        self.padding = int(padding)
        self.periodic = tuple(bool(p) for p in periodic)
        self.storage = storage
        self.shape = (nx + 2 * padding, ny + 2 * padding, nz + 2 * padding)
        self._h = self._f("create")(nx, ny, nz, padding, *(int(p) for p in self.periodic))
        if not self._h:
            raise ValueError("invalid oracle lattice")
        self._f("set_storage")(self._h, STORE_MODE[storage])
        self.ncells = int(self._f("ncells")(self._h))

    def _f(self, name):
        return getattr(library(), self._pfx + name)

    def close(self):
        if self._h:
            self._f("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cell_id(self, x, y, z) -> int:
        return int(self._f("cell_id")(self._h, x, y, z))

    def neighbor_distances(self) -> np.ndarray:
        out = np.zeros(Q, dtype=np.int64)
        self._f("neighbor_distances")(self._h, _ptr(out))
        return out

    def ring_shifts(self) -> np.ndarray:
        out = np.zeros(Q, dtype=np.int64)
        self._f("shifts")(self._h, _ptr(out))
        return out

    def set_dynamics(self, slot: int, dynamics) -> None:
        force = np.asarray(dynamics.force, dtype=np.float64)
        wall = np.asarray(dynamics.wall_velocity, dtype=np.float64)
        self._f("set_dynamics")(self._h, int(slot), int(dynamics.kind), float(dynamics.omega),
                                float(dynamics.magic), _ptr(force), _ptr(wall),
                                float(dynamics.wall_rho), float(getattr(dynamics, "smagorinsky", 0.0)))

    def set_dynamics_mask(self, ids) -> None:
        ids = np.ascontiguousarray(ids, dtype=np.uint8).reshape(-1)
        assert ids.size == self.ncells
        self._f("set_dynamics_mask")(self._h, _ptr(ids))

    def set_cell_field(self, field: int, values) -> None:
        name = {0: "set_cell_velocity", 1: "set_cell_rho", 2: "set_cell_force"}[field]
        if values is None:
            self._f(name)(self._h, None)
            return
        v = np.ascontiguousarray(values, dtype=self._dtype).reshape(-1)
        self._f(name)(self._h, _ptr(v))

    def set_cell_normals(self, normals) -> None:
        """discrete normal per cell (3 x int8, CellID order), None removes them"""
        if normals is None:
            self._f("set_cell_normals")(self._h, None)
            return
        v = np.ascontiguousarray(normals, dtype=np.int8).reshape(-1)
        assert v.size == 3 * self.ncells
        self._f("set_cell_normals")(self._h, _ptr(v))

    def post_process(self, stream=None) -> None:
        """the PostStream processors on the current populations"""
        self._f("post_process")(self._h)

    def set_bouzidi(self, cells, q) -> None:
        """the cells running BouzidiPostProcessor (applied in the given order)
        and their BOUZIDI_DISTANCE (n x 19); an empty list removes them"""
        cells = np.ascontiguousarray(cells, dtype=np.int64).reshape(-1)
        q = np.ascontiguousarray(q, dtype=self._dtype).reshape(-1)
        assert q.size == Q * cells.size
        self._f("set_bouzidi")(self._h, int(cells.size), _ptr(cells), _ptr(q))

    def set_bouzidi_velocity(self, cells, q, velocity) -> None:
        """the cells running BouzidiVelocityPostProcessor, their
        BOUZIDI_DISTANCE and BOUZIDI_VELOCITY (n x 19 each); applied after the
        set_bouzidi list; an empty list removes them"""
        cells = np.ascontiguousarray(cells, dtype=np.int64).reshape(-1)
        q = np.ascontiguousarray(q, dtype=self._dtype).reshape(-1)
        v = np.ascontiguousarray(velocity, dtype=self._dtype).reshape(-1)
        assert q.size == Q * cells.size and v.size == Q * cells.size
        self._f("set_bouzidi_velocity")(self._h, int(cells.size), _ptr(cells), _ptr(q), _ptr(v))

    def ini_equilibrium(self, rho, u, sel=None) -> None:
        rho = np.ascontiguousarray(rho, dtype=self._dtype).reshape(-1)
        u = np.ascontiguousarray(u, dtype=self._dtype).reshape(-1)
        s = None if sel is None else np.ascontiguousarray(sel, dtype=np.uint8).reshape(-1)
        self._f("ini_equilibrium")(self._h, _ptr(s), _ptr(rho), _ptr(u))

    def get_populations(self) -> np.ndarray:
        out = np.empty((Q, self.ncells), dtype=self._dtype)
        self._f("get_populations")(self._h, _ptr(out))
        return out

    def set_populations(self, f) -> None:
        f = np.ascontiguousarray(f, dtype=self._dtype)
        self._f("set_populations")(self._h, _ptr(f))

    def collide_and_stream(self, steps: int = 1, stream=None) -> None:
        self._f("collide_and_stream")(self._h, int(steps))

    collideAndStream = collide_and_stream

    def set_statistics(self, on: bool) -> None:
        self._f("set_statistics")(self._h, int(bool(on)))

    def get_statistics(self) -> dict:
        out = np.zeros(4, dtype=np.float64)
        self._f("get_statistics")(self._h, _ptr(out))
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

    def compute_rho_u(self):
        rho = np.empty(self.ncells, dtype=self._dtype)
        u = np.empty(3 * self.ncells, dtype=self._dtype)
        self._f("compute_rho_u")(self._h, _ptr(rho), _ptr(u))
        return rho, u.reshape(self.ncells, 3)


def f32_to_f16_bits(values: np.ndarray) -> np.ndarray:
    f = library().orcf_f32_to_f16
    return np.array([f(float(v)) for v in np.asarray(values, dtype=np.float32).ravel()], dtype=np.uint16)


def f32_to_bf16_bits(values: np.ndarray) -> np.ndarray:
    f = library().orcf_f32_to_bf16
    return np.array([f(float(v)) for v in np.asarray(values, dtype=np.float32).ravel()], dtype=np.uint16)
