"""Cell lists: gather the 19 populations of listed cells from one lattice and
scatter them into another (the communicator's overlap copy); exact for
every storage, other cells untouched, ids beyond the block rejected."""
import numpy as np
import pytest

from openlb_amd import Lattice, OlbgfxError, device_buffer

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16", "f64"])
def test_gather_scatter_moves_exactly_the_listed_cells(storage):
    print(f"[cells] storage={storage}", flush=True)
    ext, pad = (9, 7, 11), 2
    a = Lattice(ext, pad, (True, False, True), storage)
    b = Lattice(ext, pad, (True, False, True), storage)
    rng = np.random.default_rng(5)
    dtype = np.float64 if storage == "f64" else np.float32
    fa = (rng.standard_normal((19, a.ncells)) * 1e-2).astype(dtype)
    fb = (rng.standard_normal((19, b.ncells)) * 1e-2).astype(dtype)
    a.set_populations(fa)
    b.set_populations(fb)
    cells = np.sort(rng.choice(a.ncells, 300, replace=False)).astype(np.uint32)
    la, lb = a.cell_list(cells), b.cell_list(cells)
    itemsize = 8 if storage == "f64" else (4 if storage == "f32" else 2)
    assert la.nbytes == 19 * 300 * itemsize
    buf = device_buffer(la.nbytes)
    a.gather_cells(la, buf.ptr)
    b.scatter_cells(lb, buf.ptr)
    ga, gb = a.get_populations(), b.get_populations()
    np.testing.assert_array_equal(gb[:, cells], ga[:, cells])
    rest = np.setdiff1d(np.arange(b.ncells), cells)
    np.testing.assert_array_equal(gb[:, rest], b_before(fb, storage)[:, rest])
    for x in (la, lb, buf):
        x.close()


def b_before(f, storage):
    """what the device holds after set_populations for this storage (FP16 and
    BF16 round to nearest even, checked against the oracle elsewhere)"""
    if storage in ("f32", "f64"):
        return f
    if storage == "f16":
        return f.astype(np.float16).astype(np.float32)
    from oracle.oracle import f32_to_bf16_bits

    bits = f32_to_bf16_bits(f.reshape(-1)).astype(np.uint32) << 16
    return bits.view(np.float32).reshape(f.shape)


def test_gather_follows_the_ring_offsets():
    """after steps the list still addresses logical cells"""
    ext = (6, 6, 6)
    a = Lattice(ext, 1, (True, True, True), "f32")
    rng = np.random.default_rng(9)
    a.set_populations((rng.standard_normal((19, a.ncells)) * 1e-2).astype(np.float32))
    a.collide_and_stream(3)   # no dynamics: the rings only rotate
    cells = np.arange(0, a.ncells, 7, dtype=np.uint32)
    lst = a.cell_list(cells)
    buf = device_buffer(lst.nbytes)
    a.gather_cells(lst, buf.ptr)
    b = Lattice(ext, 1, (True, True, True), "f32")
    lb = b.cell_list(cells)
    b.scatter_cells(lb, buf.ptr)
    np.testing.assert_array_equal(b.get_populations()[:, cells], a.get_populations()[:, cells])


def test_cell_ids_beyond_the_block_are_rejected():
    a = Lattice((4, 4, 4), 1, (False, False, False), "f32")
    with pytest.raises(OlbgfxError):
        a.cell_list([0, a.ncells])
