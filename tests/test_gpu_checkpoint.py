"""An OpenLB checkpoint restarts a GPU lattice, and a GPU lattice's
checkpoint reads back through the same format."""
import os

import numpy as np
import pytest

from openlb_amd import BGK, Dynamics, Lattice, checkpoint
from tests.golden_runs import GOLDEN

pytestmark = pytest.mark.gpu


def _tgv4(storage="f32"):
    lat = Lattice((4, 4, 4), 3, (True, True, True), storage)
    lat.set_dynamics(1, Dynamics(BGK, omega=1 / 0.56))
    mask = np.zeros((10, 10, 10), dtype=np.uint8)
    mask[3:-3, 3:-3, 3:-3] = 1
    lat.set_dynamics_mask(mask.reshape(-1))
    return lat


def test_restart_from_reference_checkpoint(tmp_path):
    d = np.load(os.path.join(GOLDEN, "tgv4_checkpoint_f32.npz"))
    a, b = _tgv4(), _tgv4()
    checkpoint.load_lattice(a, os.path.join(GOLDEN, "tgv4_openlb_checkpoint_f32.dat"))
    b.set_populations(d["f1"])
    core = np.zeros((10, 10, 10), dtype=bool)
    core[3:-3, 3:-3, 3:-3] = True
    core = core.reshape(-1)
    np.testing.assert_array_equal(a.get_populations()[:, core], d["f1"][:, core])
    a.collide_and_stream(5)
    b.collide_and_stream(5)
    np.testing.assert_array_equal(a.get_populations()[:, core], b.get_populations()[:, core])
    path = tmp_path / "gpu.dat"
    checkpoint.save_lattice(a, str(path))
    np.testing.assert_array_equal(checkpoint.read(str(path), a.ncells), a.get_populations())
