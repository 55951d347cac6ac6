"""VTK output from the GPU against the reference's own SuperVTMwriter3D
files: each fixture's lattice is run on the GPU from the reference's t = 0
populations, written with openlb_amd.vtk, and every point value of every
functor compared with the reference's file (FP32; the walls and the lid of
the cavities exercise the momenta of BounceBack, BounceBackVelocity and the
LocalVelocity kinds, the Taylor-Green vortex the periodic overlap and the
converter factors)."""
import numpy as np
import pytest

from openlb_amd import Lattice, vtk
from tests.golden_runs import build, load
from tests.test_vtk import FIXTURES, header

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_vtk_matches_reference(name, tmp_path):
    print(f"[vtk] {name}", flush=True)
    meta, d = load(name)
    lat = build(Lattice, meta, d, "f32")
    lat.collide_and_stream(meta["steps"])
    ref_text = str(d["vti"])
    origin, delta = header(ref_text)
    w = vtk.SuperVTMwriter3D(meta["vtk_name"], out_dir=str(tmp_path), origin=origin, delta=delta,
                             velocity_factor=meta.get("velocity_factor", 1.0),
                             pressure_factor=meta.get("pressure_factor", 1.0))
    for f in vtk.FUNCTORS:
        w.add_functor(f)
    w.create_master_file()
    w.write(lat, meta["steps"])
    stem = vtk.file_name(meta["vtk_name"], meta["steps"], 0)
    path = tmp_path / "data" / (stem + ".vti")
    ours, ref = vtk.read_vti(str(path)), vtk.read_vti_text(ref_text)
    tol = {"velocity": 1e-6, "density": 2e-6, "physVelocity": 2e-6, "physPressure": 1e-5}
    for f in vtk.FUNCTORS:
        err = float(np.max(np.abs(ours[f].astype(np.float64) - ref[f])))
        print(name, f, "max abs err", err)
        assert ours[f].shape == ref[f].shape
        assert err <= tol[f]
    # everything but the encoded values is the reference's text
    strip = lambda t: [ln for ln in t.splitlines() if ln.startswith("<")]  # noqa: E731
    assert strip(path.read_text()) == strip(ref_text)
    assert (tmp_path / "data" / (vtk.file_name(meta["vtk_name"], meta["steps"]) + ".vtm")).read_text() == str(d["vtm"])
    lat.close()
