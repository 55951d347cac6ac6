"""Ragged and degenerate lattice shapes with a random mix of dynamics, set
up identically on any lattice with the Lattice interface (the GPU library or
the oracle): extents of one cell along an axis, core counts that are no
multiple of the 256-cell block, every padding width 1-3, periodic and
closed axes mixed.

A closed axis gets a bounce-back layer on its outermost core cells, as
every case of the reference closes a non-periodic face: a cell next to a
closed face pulls populations from the padding, which no implementation
defines (with periodic shifts they are whatever an earlier step left in
those slots); bounce-back sends them straight back out, so the cells inside
the layers (`compared_mask`) never see them."""
from __future__ import annotations

import numpy as np

from openlb_amd import BGK, BOUNCE_BACK, FORCED_BGK, NO_DYNAMICS, TRT, Dynamics

# (extent, padding, periodic)
SHAPES = [
    ((1, 1, 1), 1, (True, True, True)),
    ((1, 1, 1), 3, (True, True, True)),
    ((2, 3, 5), 2, (True, False, True)),
    ((7, 1, 13), 1, (False, True, False)),
    ((3, 17, 1), 3, (True, True, True)),
    ((5, 7, 257), 1, (True, True, True)),
    ((33, 4, 9), 2, (False, False, False)),
    ((64, 1, 1), 1, (True, True, True)),
]


def shape_id(shape) -> str:
    (x, y, z), p, per = shape
    return f"{x}x{y}x{z}_p{p}_" + "".join("p" if q else "c" for q in per)


def dynamics(forced: bool = True) -> dict:
    """table slot -> Dynamics; slot 0 stays NoDynamics"""
    slots = {1: Dynamics(BGK, omega=1.0 / 0.61), 2: Dynamics(TRT, omega=1.0 / 0.8, magic=0.25),
             3: Dynamics(BOUNCE_BACK), 4: Dynamics(NO_DYNAMICS)}
    if forced:
        slots[5] = Dynamics(FORCED_BGK, omega=1.0 / 0.7, force=(2e-4, -1e-4, 5e-5))
    return slots


def setup(lat, shape, seed: int = 7, forced: bool = True) -> np.ndarray:
    """random dynamics mask over the core (padding NoDynamics) and random
    shifted populations (0.002 standard normal, padding included, so every
    unshifted population stays far from zero) on `lat`; returns the mask"""
    extent, p, _ = shape
    P = tuple(e + 2 * p for e in extent)
    rng = np.random.default_rng(seed)
    slots = dynamics(forced)
    for k, d in slots.items():
        lat.set_dynamics(k, d)
    mask = np.zeros(P, dtype=np.uint8)
    mask[p:p + extent[0], p:p + extent[1], p:p + extent[2]] = rng.choice(
        np.array(sorted(slots), dtype=np.uint8), size=extent)
    core = mask[p:p + extent[0], p:p + extent[1], p:p + extent[2]]
    for ax, per in enumerate(shape[2]):
        if not per:
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[ax], hi[ax] = 0, extent[ax] - 1
            core[tuple(lo)] = 3
            core[tuple(hi)] = 3
    lat.set_dynamics_mask(mask.reshape(-1))
    f = 0.002 * rng.standard_normal((19, int(np.prod(P))))
    lat.set_populations(f)
    return mask.reshape(-1)


def core_mask(shape) -> np.ndarray:
    extent, p, _ = shape
    P = tuple(e + 2 * p for e in extent)
    core = np.zeros(P, dtype=bool)
    core[p:p + extent[0], p:p + extent[1], p:p + extent[2]] = True
    return core.reshape(-1)


def compared_mask(shape) -> np.ndarray:
    """the core cells inside the bounce-back layers of the closed axes"""
    extent, p, periodic = shape
    P = tuple(e + 2 * p for e in extent)
    m = np.zeros(P, dtype=bool)
    sl = tuple(slice(p, p + e) if per else slice(p + 1, p + e - 1) for e, per in zip(extent, periodic))
    m[sl] = True
    return m.reshape(-1)
