"""olbgfx_reduce_moments_noalloc, the moment reduction that allocates no
device memory, against olbgfx_reduce_moments on the same lattice: bit for
bit, on two calls in a row (the second on a smaller lattice, so the fixed
scratch still holds the first call's partials beyond its blocks), and the
refusal of a lattice larger than its fixed scratch."""
import ctypes

import numpy as np
import pytest

from openlb_amd import Lattice, cases, library

pytestmark = pytest.mark.gpu


def _noalloc(lat):
    out = (ctypes.c_double * 8)()
    rc = library().olbgfx_reduce_moments_noalloc(lat._h, out, None)
    return rc, np.array(out[:])


def _alloc(lat):
    out = (ctypes.c_double * 8)()
    rc = library().olbgfx_reduce_moments(lat._h, out, None)
    assert rc == 0
    return np.array(out[:])


@pytest.mark.parametrize("storage", ["f32", "f16", "f64"])
def test_noalloc_moments_equal_allocating_ones(storage):
    dtype = np.float64 if storage == "f64" else np.float32
    for n, steps in ((40, 7), (13, 3)):
        print(f"[noalloc] storage={storage} tgv{n} steps={steps}", flush=True)
        case = cases.tgv(n, u0=0.04, tau=0.56, dtype=dtype)
        lat = Lattice(case.extent, case.padding, case.periodic, storage)
        case.apply(lat)
        lat.collide_and_stream(steps)
        rc, got = _noalloc(lat)
        assert rc == 0
        want = _alloc(lat)
        np.testing.assert_array_equal(got, want)
        assert got[0] == n ** 3
        lat.close()


def test_noalloc_moments_on_a_cavity_with_walls():
    case = cases.box_cavity(30, u_lid=0.05, dtype=np.float32)
    lat = Lattice(case.extent, case.padding, case.periodic, "f32")
    case.apply(lat)
    lat.collide_and_stream(11)
    rc, got = _noalloc(lat)
    assert rc == 0
    np.testing.assert_array_equal(got, _alloc(lat))
    lat.close()


def test_noalloc_refuses_lattices_beyond_its_scratch():
    # 2^24 core cells fit exactly (65536 blocks of 256); one more row does not
    lat = Lattice((256, 256, 257), 1, (True, True, True), "f16")
    rc, _ = _noalloc(lat)
    assert rc == 1  # OLBGFX_ERR_INVALID
    assert b"16777216" in library().olbgfx_last_error()
    lat.close()
