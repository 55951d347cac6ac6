"""Ragged and degenerate shapes on the GPU against the oracle (through the C
ABI): one-cell axes, core counts that are no multiple of the 256-cell block,
padding 1-3, periodic and closed axes mixed, a random mix of BGK, TRT,
ForcedBGK, BounceBack and NoDynamics cells, closed axes lined with
bounce-back.  Ring offsets bit for bit; populations of every core cell inside
those layers within 1e-12 (FP64) and 1e-5 (FP32, relative to the unshifted
populations) after 9 steps."""
import numpy as np
import pytest

from openlb_amd import Lattice
from oracle.oracle import OracleLattice
from tests.helpers import rel_err_unshifted
from tests.shape_cases import SHAPES, compared_mask, setup, shape_id

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("shape", SHAPES, ids=[shape_id(s) for s in SHAPES])
def test_shape_matches_oracle(shape, storage):
    extent, p, periodic = shape
    print(f"[shapes] {shape_id(shape)} {storage}", flush=True)
    gpu = Lattice(extent, p, periodic, storage)
    orc = OracleLattice(extent, p, periodic, storage)
    for lat in (gpu, orc):
        setup(lat, shape)
        lat.collide_and_stream(9)
    np.testing.assert_array_equal(gpu.ring_shifts(), orc.ring_shifts())
    cells = compared_mask(shape)
    assert cells.sum() > 0
    assert rel_err_unshifted(gpu.get_populations()[:, cells], orc.get_populations()[:, cells]) <= TOL[storage]
    gpu.close()
