"""The full-size workloads on the GPU against reductions of the reference's
own FP32 CPU run (tests/golden/make_fixtures.py, FULL): mid-plane velocity,
the populations of 256 seeded fluid cells and FP64 sums over the fluid."""
import zlib

import numpy as np
import pytest

from tests.golden_runs import core_fluid_mask, dynamics_for, load, reduce_state
from tests.helpers import rel_err_unshifted

FULL = ["tgv256_500steps_f32", "porous384x128x128_5000steps_f32", "tgv512_1000steps_f32"]


def _case(meta):
    from openlb_amd import cases

    if meta["case"] == "tgv":
        return cases.tgv(meta["nx"], u0=0.04, tau=0.56, dtype=np.float32)
    return cases.porous_channel((meta["nx"], meta["ny"], meta["nz"]), radius=6.0, porosity=0.7,
                                seed=20261202, force=1e-6, tau=0.8, magic=3.0 / 16.0,
                                dtype=np.float32)


def _have(name):
    import os

    from tests.golden_runs import GOLDEN

    return os.path.exists(os.path.join(GOLDEN, name + ".npz"))


# FP32 tolerances against the reference's FP32 run (different summation
# order and FMA contraction, growing with the step count).  "u" is an
# absolute bound on the mid-plane velocity (u0 = 0.04 for the vortex); None
# bounds it instead by how far the reference's own FP32 velocity lies from an
# FP64 run of the same workload on the GPU (the porous flow's velocities,
# ~1e-4, sit five decades below the populations, where FP32 rounding shows)
TOL = {
    "tgv256_500steps_f32": {"u": 2e-6, "f": 1e-5, "sums": 2e-5},
    "porous384x128x128_5000steps_f32": {"u": None, "f": 1e-5, "sums": 2e-5},
    "tgv512_1000steps_f32": {"u": 2e-6, "f": 1e-5, "sums": 2e-5},
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", [n for n in FULL if _have(n)])
def test_full_size_fp32_matches_reference(name):
    from openlb_amd import Lattice

    meta, d = load(name)
    case = _case(meta)
    p = case.padding
    core = np.ascontiguousarray(case.materials[p:-p, p:-p, p:-p])
    assert zlib.crc32(core.tobytes()) == int(d["core_materials_crc32"])
    lat = Lattice(case.extent, case.padding, case.periodic, "f32")
    case.apply(lat)
    for mat, dyn in dynamics_for(meta, "f32").items():
        assert case.dynamics[mat].kind == dyn.kind
    lat.collide_and_stream(meta["steps"])
    fluid = core_fluid_mask(meta, case.materials)
    f = lat.get_populations()
    up, fs, sums = reduce_state(f, meta, fluid, d["sample_cells"])
    del f
    tol = TOL[name]
    du = float(np.max(np.abs(up - d["u_plane"])))
    ef = rel_err_unshifted(fs, d["f1_sample"])
    print(f"{name}: max |du| plane {du:.3e}, sample rel {ef:.3e}, sums {sums} ref {d['fluid_sums']}")
    if tol["u"] is None:
        lat.close()
        lat64 = Lattice(case.extent, case.padding, case.periodic, "f64")
        case.apply(lat64)
        lat64.collide_and_stream(meta["steps"])
        up64, _, sums64 = reduce_state(lat64.get_populations(), meta, fluid, d["sample_cells"])
        lat64.close()
        ref_noise = float(np.max(np.abs(d["u_plane"] - up64)))
        ours_noise = float(np.max(np.abs(up - up64)))
        print(f"{name}: reference FP32 vs FP64 {ref_noise:.3e}, ours vs FP64 {ours_noise:.3e}; "
              f"sum u^2 FP64 {sums64[4]:.9e}")
        assert du <= 2.0 * ref_noise + 1e-9
        assert ours_noise <= 2.0 * ref_noise + 1e-9
        assert abs(sums[4] - sums64[4]) <= 2.0 * abs(d["fluid_sums"][4] - sums64[4]) + 1e-12
    else:
        assert du <= tol["u"]
        np.testing.assert_allclose(sums[4], d["fluid_sums"][4], rtol=tol["sums"])
    assert ef <= tol["f"]
