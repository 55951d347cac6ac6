"""Halo planes of a block split along x: one block whose x faces exchange
packed planes (with itself, or with a second block) must step exactly like
the undivided periodic block, bit for bit, since every cell sees the same
incoming populations and runs the same arithmetic."""
import dataclasses

import numpy as np
import pytest

from openlb_amd import Lattice, OlbgfxError, cases

pytestmark = pytest.mark.gpu


def _x_slice(case, x0, x1):
    """the sub-block x in [x0, x1) of a case, periodic along y and z only"""
    p = case.padding
    sl = slice(x0, x1 + 2 * p)
    vf = case.velocity_field
    if vf is not None:
        vf = vf.reshape(case.shape + (3,))[sl]
    return dataclasses.replace(
        case, extent=(x1 - x0,) + tuple(case.extent[1:]), periodic=(False,) + tuple(case.periodic[1:]),
        materials=np.ascontiguousarray(case.materials[sl]),
        rho=None if case.rho is None else np.ascontiguousarray(case.rho[sl]),
        u=None if case.u is None else np.ascontiguousarray(case.u[sl]), velocity_field=vf)


def _buffer(lat):
    return lat.halo_buffer(0)


def _core(case, lat):
    p = case.padding
    f = lat.get_populations().reshape((19,) + case.shape)
    return f[:, p:-p, p:-p, p:-p]


def _porous(storage):
    dtype = np.float64 if storage == "f64" else np.float32
    return cases.porous_channel((20, 10, 12), radius=2.5, porosity=0.8, seed=11, force=1e-4, dtype=dtype)


@pytest.mark.parametrize("storage", ["f32", "f16", "f64"])
def test_self_exchange_equals_periodic_block(storage):
    """one block, its two x faces wired to each other through pack/unpack"""
    print(f"[halo self] storage={storage}", flush=True)
    c = _porous(storage)
    ref = Lattice(c.extent, c.padding, c.periodic, storage)
    c.apply(ref)
    h = _x_slice(c, 0, c.extent[0])
    lat = Lattice(h.extent, h.padding, h.periodic, storage)
    h.apply(lat)
    lo, hi = _buffer(lat), _buffer(lat)
    for _ in range(25):
        ref.collide_and_stream(1)
        lat.collide_and_stream(1)
        lat.halo_pack(0, 1, hi.ptr)
        lat.halo_pack(0, 0, lo.ptr)
        lat.halo_unpack(0, 0, hi.ptr)
        lat.halo_unpack(0, 1, lo.ptr)
    np.testing.assert_array_equal(_core(h, lat), _core(c, ref))


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_two_blocks_equal_one(storage):
    """the porous box cut at x = 8 into two blocks that exchange halos"""
    print(f"[halo split] storage={storage}", flush=True)
    c = _porous(storage)
    ref = Lattice(c.extent, c.padding, c.periodic, storage)
    c.apply(ref)
    parts = [_x_slice(c, 0, 8), _x_slice(c, 8, c.extent[0])]
    lats = []
    for part in parts:
        lat = Lattice(part.extent, part.padding, part.periodic, storage)
        part.apply(lat)
        lats.append(lat)
    a, b = lats
    bufs = [_buffer(a) for _ in range(4)]
    for _ in range(25):
        ref.collide_and_stream(1)
        a.collide_and_stream(1)
        b.collide_and_stream(1)
        a.halo_pack(0, 1, bufs[0].ptr)   # a -> b across x = 8
        b.halo_pack(0, 0, bufs[1].ptr)   # b -> a across x = 8
        b.halo_pack(0, 1, bufs[2].ptr)   # b -> a across the periodic x face
        a.halo_pack(0, 0, bufs[3].ptr)   # a -> b across the periodic x face
        b.halo_unpack(0, 0, bufs[0].ptr)
        a.halo_unpack(0, 1, bufs[1].ptr)
        a.halo_unpack(0, 0, bufs[2].ptr)
        b.halo_unpack(0, 1, bufs[3].ptr)
    joined = np.concatenate([_core(parts[0], a), _core(parts[1], b)], axis=1)
    np.testing.assert_array_equal(joined, _core(c, ref))


def test_halo_on_periodic_axis_is_rejected():
    c = cases.tgv(8, dtype=np.float32)
    lat = Lattice(c.extent, c.padding, c.periodic, "f32")
    buf = _buffer(lat)
    assert lat.halo_bytes(0) == 5 * 8 * 8 * 4
    with pytest.raises(OlbgfxError):
        lat.halo_pack(0, 1, buf.ptr)
    with pytest.raises(OlbgfxError):
        lat.halo_unpack(2, 0, buf.ptr)
