"""The C++ host mirror (openlb_amd/host/olb_gfx950.h) drives the platform from
a case written like OpenLB's cavity3dBenchmark; its result must match the
oracle on the same case."""
import os
import subprocess

import numpy as np
import pytest

from openlb_amd import cases
from oracle.oracle import OracleLattice
from tests.helpers import rel_err_unshifted

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "openlb_amd", "host", "examples", "cavity3d")


@pytest.mark.parametrize("boundary", ["bounce_back", "local_velocity"])
def test_cpp_cavity_matches_oracle(tmp_path, boundary):
    if not os.path.exists(BINARY):
        pytest.fail("openlb_amd/host/examples/cavity3d is not built (run __graft_entry__.build())")
    res, steps = 16, 90
    dump = tmp_path / "cavity.bin"
    print("[cpp] cavity3d", res, steps, boundary, flush=True)
    out = subprocess.run([BINARY, "--RESOLUTION", str(res), "--TIME_STEPS", str(steps), "--DUMP", str(dump),
                          "--BOUNDARY", boundary], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    fields = out.stdout.strip().split(",")
    assert int(fields[0]) == res and float(fields[-1]) > 0
    make = cases.cavity3d_local_velocity if boundary == "local_velocity" else cases.cavity3d
    c = make(res, dtype=np.float32)
    raw = np.fromfile(dump, dtype=np.uint8)
    mats = raw[:c.ncells]
    np.testing.assert_array_equal(mats, c.materials.reshape(-1))
    f = raw[c.ncells:].view(np.float32).reshape(19, c.ncells)
    orc = OracleLattice(c.extent, c.padding, c.periodic, "f32")
    c.apply(orc)
    orc.collide_and_stream(10 + steps)
    fluid = c.materials.reshape(-1) == 1
    assert rel_err_unshifted(f[:, fluid], orc.get_populations()[:, fluid]) <= 1e-5


def test_cpp_rejects_bad_options():
    out = subprocess.run([BINARY, "--BOGUS", "1"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "unknown option" in out.stderr


def test_cpp_checkpoint_save_and_restart(tmp_path):
    """the mirror's SuperLattice::save writes OpenLB's checkpoint format (the
    Python reader, pinned against the reference, reads the dumped
    populations back), and restarting from it continues the run exactly"""
    from openlb_amd import checkpoint

    res = 8
    c = cases.cavity3d_local_velocity(res, dtype=np.float32)

    def run(steps, *extra):
        dump = tmp_path / f"d{steps}_{len(extra)}.bin"
        out = subprocess.run([BINARY, "--RESOLUTION", str(res), "--TIME_STEPS", str(steps), "--DUMP", str(dump),
                              *extra], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        raw = np.fromfile(dump, dtype=np.uint8)
        return raw[c.ncells:].view(np.float32).reshape(19, c.ncells)

    ck = tmp_path / "ck.dat"
    f5 = run(5, "--SAVE", str(ck))
    np.testing.assert_array_equal(checkpoint.read(str(ck), c.ncells), f5)
    f_restart = run(5, "--LOAD", str(ck))      # 10 warm-up + 5 more after the restart
    f_straight = run(20)                       # 10 warm-up + 20
    core = c.materials.reshape(-1) != 0
    np.testing.assert_array_equal(f_restart[:, core], f_straight[:, core])


@pytest.mark.parametrize("boundary,fixture", [("bounce_back", "vtk_cavity8"), ("local_velocity", "vtk_cavity_lv8")])
def test_cpp_cavity_vtk_matches_reference(tmp_path, boundary, fixture):
    """the C++ mirror's SuperVTMwriter3D after 20 steps of the RESOLUTION 8
    cavity against the reference's own VTK output of the same run: every
    point value of every functor (FP32), the .vtm and the .pvd text"""
    from openlb_amd import vtk
    from tests.golden_runs import load

    meta, d = load(fixture)
    print("[cpp] cavity3d vtk", boundary, flush=True)
    dump = tmp_path / "cavity.bin"
    out = subprocess.run([BINARY, "--RESOLUTION", "8", "--TIME_STEPS", "10", "--BOUNDARY", boundary,
                          "--VTK", str(tmp_path), "--DUMP", str(dump)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    # the run itself against the reference's populations after 20 steps (fluid)
    raw = np.fromfile(dump, dtype=np.uint8)
    ncells = d["materials"].size
    np.testing.assert_array_equal(raw[:ncells], d["materials"])
    f = raw[ncells:].view(np.float32).reshape(19, ncells)
    fluid = d["materials"][d["cells"]] == 1
    assert rel_err_unshifted(f[:, d["cells"]][:, fluid], d["f1"][:, fluid]) <= 1e-5
    name = meta["vtk_name"]
    ours = vtk.read_vti(str(tmp_path / "data" / (vtk.file_name(name, 20, 0) + ".vti")))
    ref = vtk.read_vti_text(str(d["vti"]))
    for fn in vtk.FUNCTORS:
        e = np.abs(ours[fn].astype(np.float64) - ref[fn])
        err = float(e.max())
        print(boundary, fn, "max abs err", err)
        if err > 2e-6:
            k = int(np.argmax(e)) // vtk.NCOMP[fn]
            nx, ny = meta["nx"] + 2, meta["ny"] + 2
            x, y, z = k % nx - 1, (k // nx) % ny - 1, k // (nx * ny) - 1
            print("worst point", (x, y, z), "ours", ours[fn][k * vtk.NCOMP[fn]:(k + 1) * vtk.NCOMP[fn]],
                  "reference", ref[fn][k * vtk.NCOMP[fn]:(k + 1) * vtk.NCOMP[fn]])
        assert err <= 2e-6
    assert (tmp_path / "data" / (vtk.file_name(name, 20) + ".vtm")).read_text() == str(d["vtm"])
    assert (tmp_path / (name + ".pvd")).read_text() == str(d["pvd"])
    head = lambda t: [ln for ln in t.splitlines() if ln.startswith("<")]  # noqa: E731
    assert head((tmp_path / "data" / (vtk.file_name(name, 20, 0) + ".vti")).read_text()) == head(str(d["vti"]))


@pytest.mark.parametrize("boundary", ["local", "interpolated"])
def test_cpp_channel_matches_oracle(tmp_path, boundary):
    """the C++ mirror's channel (BounceBack walls, LocalVelocity inlet and
    LocalPressure outlet with defineRho, or InterpolatedVelocity and
    InterpolatedPressure) against the oracle on the same case after 150
    steps (FP32, 1e-5 relative on the fluid)"""
    binary = os.path.join(ROOT, "openlb_amd", "host", "examples", "channel3d")
    dump = tmp_path / "channel.bin"
    print("[cpp] channel3d", boundary, flush=True)
    out = subprocess.run([binary, "--NX", "24", "--NY", "10", "--NZ", "10", "--TIME_STEPS", "150",
                          "--DUMP", str(dump), "--BOUNDARY", boundary], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    c = cases.channel3d(24, 10, 10, u_in=0.05, tau=0.8, dtype=np.float32, boundary=boundary)
    raw = np.fromfile(dump, dtype=np.uint8)
    np.testing.assert_array_equal(raw[:c.ncells], c.materials.reshape(-1))
    f = raw[c.ncells:].view(np.float32).reshape(19, c.ncells)
    orc = OracleLattice(c.extent, c.padding, c.periodic, "f32")
    c.apply(orc)
    orc.collide_and_stream(150)
    sel = np.isin(c.materials.reshape(-1), (1,))
    err = rel_err_unshifted(f[:, sel], orc.get_populations()[:, sel])
    print("channel fluid vs oracle", err)
    assert err <= 1e-5


def test_cpp_step_cavity_matches_reference(tmp_path):
    """the C++ mirror's cavity with a solid corner cube (--STEP: its
    classifier gives the walls every LocalVelocity shape, BounceBack where no
    normal) after 10 + 90 steps against the reference's FP64 run of the same
    case on every defined population of every core cell (FP32 storage,
    within 2e-5 relative)"""
    import json

    from tests.golden_runs import step_cavity_oracle
    from tests.helpers import T as W

    c, d, cells, _, defined = step_cavity_oracle("f32")
    meta = json.loads(str(d["meta"]))
    dump = tmp_path / "step.bin"
    print("[cpp] cavity3d step", flush=True)
    out = subprocess.run([BINARY, "--RESOLUTION", str(meta["res"]), "--TIME_STEPS", str(meta["steps"] - 10),
                          "--STEP", str(meta["step"]), "--DUMP", str(dump)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    raw = np.fromfile(dump, dtype=np.uint8)
    np.testing.assert_array_equal(raw[:c.ncells], d["materials"])
    f = raw[c.ncells:].view(np.float32).reshape(19, c.ncells)[:, cells].astype(np.float64)
    ref = d["f1"]
    rel = np.where(defined, np.abs(f - ref) / np.abs(ref + W[:, None]), 0.0)
    print("cpp step cavity vs reference", float(rel.max()))
    assert rel.max() <= 2e-5
