"""Host-side logic of the C++ mirror that needs no GPU: its discrete-normal
classification of the cavity walls equals the reference's (the golden
fixture holds the normals OpenLB's boundary::set computed), and its base64
codec reproduces an OpenLB checkpoint."""
import os
import subprocess

import numpy as np

from tests.golden_runs import load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_discrete_normals_match_reference(tmp_path):
    exe = tmp_path / "normals_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "openlb_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "normals_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "cavity", "12"], capture_output=True, text=True, check=True).stdout
    meta, d = load("cavity3d_lv_res12_f32")
    p = meta["padding"]
    shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
    types = d["normal_type"].reshape(shape)
    normals = d["normals"].reshape(shape + (3,))
    rows = np.array([[int(v) for v in line.split()] for line in out.strip().splitlines()])
    assert len(rows) == int(np.count_nonzero(types >= 0))
    for x, y, z, t, nx, ny, nz in rows:
        assert types[x + p, y + p, z + p] == t
        assert tuple(normals[x + p, y + p, z + p]) == (nx, ny, nz)


def test_cpp_classifier_matches_reference_on_random_geometries(tmp_path):
    """the mirror's discrete normals on the classifier fixture's geometries
    (unions of boxes, voxel noise, spheres in a shell) equal the reference's,
    cell for cell, for every cell of materials 2 and 3 it classifies"""
    import json

    exe = tmp_path / "normals_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "openlb_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "normals_check.cpp"), "-o", str(exe)], check=True)
    d = np.load(os.path.join(ROOT, "tests", "golden", "discrete_normals_geometries.npz"))
    meta = json.loads(str(d["meta"]))
    p = meta["padding"]
    for k, kind in enumerate(meta["kinds"]):
        m = d[f"materials_{k}"]
        core = np.ascontiguousarray(m[p:-p, p:-p, p:-p])
        geo = tmp_path / f"geo{k}.bin"
        with open(geo, "wb") as f:
            f.write(np.asarray(core.shape, dtype=np.int32).tobytes())
            f.write(core.astype(np.uint8).tobytes())
        out = subprocess.run([str(exe), "file", str(geo)], capture_output=True, text=True, check=True).stdout
        rows = np.array([[int(v) for v in line.split()] for line in out.strip().splitlines()])
        types, normals = d[f"normal_type_{k}"], d[f"normals_{k}"]
        assert len(rows) == int(np.count_nonzero(types >= 0)), (k, kind)
        got_t = types.copy()
        got_n = normals.copy()
        got_t[...] = -1
        got_n[...] = 0
        for x, y, z, t, nx, ny, nz in rows:
            got_t[x + p, y + p, z + p] = t
            got_n[x + p, y + p, z + p] = (nx, ny, nz)
        np.testing.assert_array_equal(got_t, types, err_msg=f"geometry {k} ({kind})")
        np.testing.assert_array_equal(got_n, normals, err_msg=f"geometry {k} ({kind})")


def test_cpp_base64_reproduces_reference_checkpoint(tmp_path):
    """the mirror's base64 decoder reads OpenLB's checkpoint text and its
    encoder writes the same text back, byte for byte"""
    exe = tmp_path / "base64_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "openlb_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "base64_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), os.path.join(ROOT, "tests", "golden", "tgv4_openlb_checkpoint_f32.dat")],
                         capture_output=True, text=True, check=True).stdout.split()
    assert out[0] == out[1] == str(19 * (16 + 2 * 1024 * 4))
    assert out[2] == "same"


def test_cpp_bouzidi_distances_match_reference(tmp_path):
    """the mirror's setBouzidiBoundary field (IndicatorCylinder3D and the
    bisection ray distance in C++) equals the field the reference's setter
    wrote for its forced Bouzidi Poiseuille pipe, entry for entry"""
    from tests.golden_runs import bouzidi_field

    exe = tmp_path / "bouzidi_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "openlb_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "bouzidi_check.cpp"), "-o", str(exe)], check=True)
    meta, d = load("poiseuille11_forced_bouzidi_f64")
    mats = tmp_path / "materials.bin"
    d["materials"].astype(np.uint8).tofile(mats)
    origin = str(meta["origin"]).split(",")
    out = subprocess.run([str(exe), str(mats), str(meta["nx"]), str(meta["ny"]), str(meta["nz"]),
                          str(meta["padding"]), *origin, repr(meta["delta"])],
                         capture_output=True, text=True, check=True).stdout
    ref = bouzidi_field(meta, d)
    got = np.full_like(ref, -1.0)
    for line in out.strip().splitlines():
        c, i, q = line.split()
        got[int(i), int(c)] = float(q)
    np.testing.assert_array_equal(got >= 0, ref >= 0)
    assert np.max(np.abs(got - ref)) <= 1e-15
