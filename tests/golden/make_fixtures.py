"""Writes the golden fixtures of tests/golden/ by running OpenLB's own CPU
path (oracle/_ref, built by `make -C oracle ref` from the OpenLB sources).

Run in the build container (the OpenLB sources are not on the GPU host):
    python tests/golden/make_fixtures.py            # the small fixtures
    python tests/golden/make_fixtures.py NAME ...   # full-size ones (FULL), or
                                                    # local_velocity, smagorinsky, checkpoint
Each fixture holds the materials of the padded block (CellID order), the
neighbour distances, the populations at t = 0 and after the run for the core
cells (logical order, shifted by -t_i), and the run's parameters.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from openlb_amd import cases  # noqa: E402
from tests.golden.olbdump import read_dump  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
RUN = os.path.join(REF, "run")

POROUS = {"extent": (16, 12, 12), "radius": 2.5, "porosity": 0.7, "seed": 7, "tau": 0.8,
          "magic": 3.0 / 16.0, "fx": 1e-4, "steps": 200}


def run(binary: str, args: list[str], threads: int = 4) -> None:
    os.makedirs(RUN, exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run([os.path.join(REF, binary), *args], cwd=RUN, env=env, check=True,
                   stdout=subprocess.DEVNULL)


def core_ids(meta) -> np.ndarray:
    p = meta["padding"]
    shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
    idx = np.arange(np.prod(shape)).reshape(shape)
    return idx[p:-p, p:-p, p:-p].reshape(-1)


def save(name: str, meta: dict, arrays: dict, extra: dict | None = None, cells=None, texts=None) -> None:
    if cells is None:
        cells = core_ids(meta)
    q = 19
    n = arrays["f0"].size // q
    out = {
        "meta": np.array(json.dumps(dict(meta, **(extra or {})))),
        "materials": arrays["materials"].astype(np.uint8),
        "distances": arrays["distances"],
        "cells": cells.astype(np.int64),
        "f0": arrays["f0"].reshape(q, n)[:, cells],
        "f1": arrays["f1"].reshape(q, n)[:, cells],
    }
    for k in ("energy", "usol"):
        if k in arrays:
            out[k] = arrays[k]
    if "normal_type" in arrays:
        out["normal_type"] = np.asarray(arrays["normal_type"]).astype(np.int8)
        out["normals"] = np.asarray(arrays["normals"]).astype(np.int8)
    for k, v in (texts or {}).items():
        out[k] = np.array(v)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, {k: v.shape for k, v in out.items() if hasattr(v, "shape")})


def local_velocity_fixtures() -> None:
    """cavity3dBenchmark's default boundaries (LocalVelocity on walls and
    lid) at RESOLUTION 12, 200 steps, FP32 and FP64, plus the reference's
    discrete normals of the boundary cells"""
    for tag, binary in (("f32", "olb_ref_f"), ("f64", "olb_ref_d")):
        path = os.path.join(RUN, f"cavity_lv_{tag}.bin")
        run(binary, ["cavity_lv", "12", "200", "0.1", "100", path])
        save(f"cavity3d_lv_res12_{tag}", *read_dump(path))


def smagorinsky_fixtures() -> None:
    """periodic Taylor-Green 12^3 with the bulk dynamics of tgv3d
    (SmagorinskyBGKdynamics, C = 0.1), tau = 0.505 so the eddy viscosity
    dominates, 100 steps, FP32 and FP64"""
    for tag, binary in (("f32", "olb_ref_f"), ("f64", "olb_ref_d")):
        path = os.path.join(RUN, f"tgv_smagorinsky_{tag}.bin")
        run(binary, ["tgv_smagorinsky", "12", "100", "0.04", "0.505", "0.1", path])
        save(f"tgv12_smagorinsky_{tag}", *read_dump(path))


def checkpoint_fixtures() -> None:
    """OpenLB's checkpoint format (SuperLattice::save / load, CPU_SIMD
    columns, FP32) on a periodic TGV 4^3 (10^3 padded cells):
      - the reference saves after 3 steps: tgv4_openlb_checkpoint_f32.dat,
        with its populations in tgv4_checkpoint_f32.npz ("f1");
      - the reference loads a checkpoint that openlb_amd.checkpoint.write
        made from seeded populations: the npz keeps the seed, the SHA-256 of
        that file and what the reference read back ("restored")."""
    import hashlib
    import shutil

    from openlb_amd import checkpoint

    tmp = os.path.join(RUN, "tmp")
    os.makedirs(tmp, exist_ok=True)
    dump = os.path.join(RUN, "ck_ref.bin")
    run("olb_ref_f", ["checkpoint", "4", "3", dump, "ck_ref"])
    shutil.copyfile(os.path.join(tmp, "ck_ref_rank0_size1.dat"), os.path.join(HERE, "tgv4_openlb_checkpoint_f32.dat"))
    meta, arrays = read_dump(dump)
    n = np.asarray(arrays["materials"]).size
    seed = 20261202
    f = (np.random.default_rng(seed).standard_normal((19, n)) * 1e-2).astype(np.float32)
    ours = os.path.join(tmp, "ck_ours_rank0_size1.dat")
    checkpoint.write(ours, f, "cpu_simd", np.float32)
    digest = hashlib.sha256(open(ours, "rb").read()).hexdigest()
    back = os.path.join(RUN, "ck_back.bin")
    run("olb_ref_f", ["restore", "4", back, "ck_ours"])
    _, restored = read_dump(back)
    np.savez_compressed(os.path.join(HERE, "tgv4_checkpoint_f32.npz"),
                        meta=np.array(json.dumps(dict(meta, seed=seed, sha256=digest))),
                        f1=np.asarray(arrays["f1"]).reshape(19, n),
                        restored=np.asarray(restored["f1"]).reshape(19, n))
    print("wrote checkpoint fixtures", meta)


def cavity_reduced(binary: str, tag: str, res: int = 64, steps: int = 1000, mode: str = "cavity",
                   ulid: float = 0.1, re: float = 100.0, name: str | None = None, threads: int = 4) -> None:
    """the cavity3dBenchmark geometry at RESOLUTION 64, Re = 100 (tau = 0.692),
    lid u = 0.1, 1000 steps from rest; mode "cavity": BounceBack walls and a
    BounceBackVelocity lid, mode "cavity_lv": LocalVelocity walls and lid.
    Too large to keep whole, so the fixture holds the
    velocity on the mid-plane z = res/2, the populations of 256 seeded fluid
    cells, and FP64 sums over the fluid."""
    path = os.path.join(RUN, f"{mode}{res}_{tag}.bin")
    run(binary, [mode, str(res), str(steps), repr(float(ulid)), repr(float(re)), path], threads=threads)
    meta, arrays = read_dump(path)
    q = 19
    n = arrays["f1"].size // q
    p = meta["padding"]
    shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
    f = arrays["f1"].reshape(q, n).astype(np.float64)
    from tests.helpers import C
    rho = f.sum(0) + 1.0
    u = (C.T @ f / rho).T.reshape(shape + (3,))
    mats = arrays["materials"].reshape(shape)
    fluid = mats == 1
    zc = p + res // 2
    core = (slice(p, -p), slice(p, -p))
    mid = u[core[0], core[1], zc, :].astype(np.float32)
    fl = np.flatnonzero(fluid.reshape(-1))
    rng = np.random.default_rng(64)
    sample = np.sort(rng.choice(fl, 256, replace=False))
    uf = u.reshape(-1, 3)[fluid.reshape(-1)]
    sums = np.array([(rho.reshape(-1)[fluid.reshape(-1)] - 1).sum(), *uf.sum(0), (uf ** 2).sum()])
    if name is None:
        name = f"cavity3d{'_lv' if mode == 'cavity_lv' else ''}_res{res}_{steps}steps_{tag}"
    np.savez_compressed(os.path.join(HERE, name + ".npz"),
                        meta=np.array(json.dumps(meta)), materials=arrays["materials"].astype(np.uint8),
                        u_midplane=mid, sample_cells=sample.astype(np.int64),
                        f1_sample=f[:, sample].astype(np.float32 if tag == "f32" else np.float64),
                        fluid_sums=sums)
    print("wrote", name, sums)


ENSEMBLE_STEPS = 2000


def ensemble_member_picks(members) -> dict:
    """the smallest, the median and the largest member of the 64-cavity
    ensemble by edge length (first index on ties)"""
    edges = [m.extent[0] for m in members]
    order = sorted(range(len(edges)), key=lambda k: (edges[k], k))
    return {"smallest": order[0], "median": order[len(order) // 2], "largest": order[-1]}


def ensemble_members() -> None:
    """three members of the 64-cavity ensemble (cases.ensemble_cavities, seed
    20261202) through the reference's FP32 path, ENSEMBLE_STEPS (2000) steps from rest: a
    member of edge n is the cavity3dBenchmark geometry at RESOLUTION n - 1
    (n^3 core cells) with its own lid speed U; the reference forms
    nu = U RES / Re, so Re' = Re (n - 1) / n gives the member's nu = U n / Re"""
    members = cases.ensemble_cavities(dtype=np.float32)
    for label, k in ensemble_member_picks(members).items():
        m = members[k]
        n = m.extent[0]
        u = m.info["u_lid"]
        cavity_reduced("olb_ref_f", "f32", res=n - 1, steps=ENSEMBLE_STEPS, ulid=u, re=100.0 * (n - 1) / n,
                       name=f"ensemble_member_{label}_{ENSEMBLE_STEPS}steps_f32", threads=8)


# full-size workloads, kept as reductions (tests/golden_runs.reduce_state)
FULL = {
    # periodic Taylor-Green 256^3, u0 = 0.04, tau = 0.56, 500 steps
    "tgv256_500steps_f32": {"case": "tgv", "n": 256, "steps": 500, "stride": 4},
    # porous channel 384x128x128, spheres r = 6 to porosity 0.7 (seed 20261202),
    # ForcedTRT Lambda = 3/16, tau = 0.8, g = 1e-6, 5000 steps from rest
    "porous384x128x128_5000steps_f32": {"case": "porous", "extent": (384, 128, 128), "radius": 6.0,
                                        "porosity": 0.7, "seed": 20261202, "tau": 0.8,
                                        "magic": 3.0 / 16.0, "fx": 1e-6, "steps": 5000, "stride": 2},
    # periodic Taylor-Green 512^3, 1000 steps
    "tgv512_1000steps_f32": {"case": "tgv", "n": 512, "steps": 1000, "stride": 8},
}


def full_size(name: str) -> None:
    """one FULL workload through the reference's FP32 path, reduced to the
    strided mid-plane velocity, 256 seeded fluid cells and fluid sums; the
    core materials are kept as a CRC-32 and per-material counts"""
    import zlib

    from tests.golden_runs import core_fluid_mask, reduce_state

    P = FULL[name]
    path = os.path.join(RUN, name + ".bin")
    extra = {}
    if P["case"] == "tgv":
        args = ["tgv", str(P["n"]), str(P["steps"]), "0.04", "0.56", path]
    else:
        centres = cases.sphere_pack(P["extent"], P["radius"], P["porosity"], P["seed"])
        sph = os.path.join(RUN, "spheres_full.txt")
        np.savetxt(sph, centres, fmt="%.17g")
        args = ["porous", *map(str, P["extent"]), sph, str(P["radius"]), str(P["tau"]),
                repr(P["magic"]), repr(P["fx"]), str(P["steps"]), path]
        extra = {"nspheres": len(centres)}
    run("olb_ref_f", args, threads=8)
    meta, arrays = read_dump(path)
    meta = dict(meta, stride=P["stride"], **extra)
    p = meta["padding"]
    shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
    mats = np.asarray(arrays["materials"]).astype(np.uint8)
    core = np.ascontiguousarray(mats.reshape(shape)[p:-p, p:-p, p:-p])
    fluid = core_fluid_mask(meta, mats)
    sample = np.sort(np.random.default_rng(64).choice(np.flatnonzero(fluid), 256, replace=False))
    n = mats.size
    up, fs, sums = reduce_state(arrays["f1"].reshape(19, n), meta, fluid, sample)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), meta=np.array(json.dumps(meta)),
                        core_materials_crc32=np.int64(zlib.crc32(core.tobytes())),
                        material_counts=np.bincount(core.reshape(-1), minlength=4).astype(np.int64),
                        u_plane=up.astype(np.float32), sample_cells=sample.astype(np.int64),
                        f1_sample=fs.astype(np.float32), fluid_sums=sums)
    print("wrote", name, sums)
    del arrays
    os.remove(path)


def main() -> None:
    if len(sys.argv) > 1:
        for name in sys.argv[1:]:
            if name == "local_velocity":
                local_velocity_fixtures()
                cavity_reduced("olb_ref_f", "f32", mode="cavity_lv")
            elif name == "smagorinsky":
                smagorinsky_fixtures()
            elif name == "checkpoint":
                checkpoint_fixtures()
            elif name == "bouzidi":
                poiseuille_bouzidi()
            elif name == "poiseuille_bounceback":
                poiseuille_bounceback()
            elif name == "vtk":
                vtk_fixtures()
            elif name == "poiseuille_local":
                poiseuille_local()
            elif name == "poiseuille_interpolated":
                poiseuille_interpolated()
            elif name == "poiseuille_bouzidi_nonforced":
                poiseuille_bouzidi_nonforced()
            elif name == "poiseuille_bounceback_nonforced":
                # the non-forced bounceBack row: BounceBack wall, interpolated
                # inflow and outflow (one thread, as the interpolated rows)
                poiseuille_rows((("poiseuille_bounceback_nonforced", "poiseuille11_nonforced_bounceback_f64"),),
                                threads=1)
            elif name == "normals":
                normals_fixture()
            elif name == "step_cavity":
                step_cavity_fixture()
            elif name == "ensemble_members":
                ensemble_members()
            else:
                full_size(name)
        return
    local_velocity_fixtures()
    smagorinsky_fixtures()
    checkpoint_fixtures()
    cavity_reduced("olb_ref_f", "f32", mode="cavity_lv")
    cavity_reduced("olb_ref_f", "f32")
    for tag, binary in (("f32", "olb_ref_f"), ("f64", "olb_ref_d")):
        path = os.path.join(RUN, f"cavity_{tag}.bin")
        run(binary, ["cavity", "12", "200", "0.1", "100", path])
        save(f"cavity3d_res12_{tag}", *read_dump(path))

        path = os.path.join(RUN, f"tgv_{tag}.bin")
        run(binary, ["tgv", "12", "100", "0.04", "0.56", path])
        save(f"tgv12_{tag}", *read_dump(path))

        P = POROUS
        centres = cases.sphere_pack(P["extent"], P["radius"], P["porosity"], P["seed"])
        sph = os.path.join(RUN, "spheres.txt")
        np.savetxt(sph, centres, fmt="%.17g")
        path = os.path.join(RUN, f"porous_{tag}.bin")
        run(binary, ["porous", *map(str, P["extent"]), sph, str(P["radius"]), str(P["tau"]),
                     repr(P["magic"]), repr(P["fx"]), str(P["steps"]), path])
        save(f"porous16x12x12_{tag}", *read_dump(path), extra={"spheres": centres.tolist(), **P})

    poiseuille_bounceback()
    poiseuille_bouzidi()


def poiseuille_bounceback() -> None:
    """the forced bounceBack row of testPoiseuille3DBGK.cpp at N = 11"""
    path = os.path.join(RUN, "poiseuille.bin")
    run("olb_ref_d", ["poiseuille", "11", path])
    meta, arrays = read_dump(path)
    fluid = np.flatnonzero(arrays["materials"] == 1)
    keep = np.flatnonzero(arrays["materials"] != 0)
    q = 19
    n = arrays["f0"].size // q
    out = {
        "meta": np.array(json.dumps(meta)),
        "materials": arrays["materials"].astype(np.uint8),
        "distances": arrays["distances"],
        "cells": keep.astype(np.int64),
        "f0": arrays["f0"].reshape(q, n)[:, keep],
        "fluid": fluid.astype(np.int64),
        "f1_fluid": arrays["f1"].reshape(q, n)[:, fluid],
        "energy": arrays["energy"],
        "usol_fluid": arrays["usol"].reshape(n, 3)[fluid],
        "ssol_fluid": arrays["ssol"].reshape(n, 9)[fluid],
    }
    np.savez_compressed(os.path.join(HERE, "poiseuille11_forced_bounceback_f64.npz"), **out)
    print("wrote poiseuille", meta)


def vtk_fixtures() -> None:
    """SuperVTMwriter3D output of the reference (velocity, density,
    physVelocity, physPressure; compressed and plain binary) for a periodic
    Taylor-Green vortex 6^3 after 10 steps (converter factors 2 and 4.8) and
    the bounce-back and LocalVelocity cavities at RESOLUTION 8 after 20
    steps, FP32, with the populations they were written from"""
    runs = [("vtk_tgv6", "tgv", ["vtk_tgv", "6", "10"]),
            ("vtk_cavity8", "cavity", ["cavity", "8", "20", "0.1", "100"]),
            ("vtk_cavity_lv8", "cavity_lv", ["cavity_lv", "8", "20", "0.1", "100"])]
    for fixture, name, args in runs:
        path = os.path.join(RUN, fixture + ".bin")
        out_dir = os.path.join(RUN, fixture)
        shutil.rmtree(out_dir, ignore_errors=True)
        run("olb_ref_f", [*args, path, out_dir])
        meta, arrays = read_dump(path)
        steps = int(args[2])
        vtk = os.path.join(out_dir, "vtkData")
        texts = {}
        for key, stem in (("vti", name), ("vti_raw", name + "_raw")):
            with open(os.path.join(vtk, "data", f"{stem}_iT{steps:07d}iC00000.vti")) as fh:
                texts[key] = fh.read()
        with open(os.path.join(vtk, "data", f"{name}_iT{steps:07d}.vtm")) as fh:
            texts["vtm"] = fh.read()
        with open(os.path.join(vtk, f"{name}.pvd")) as fh:
            texts["pvd"] = fh.read()
        save(fixture, meta, arrays, extra={"vtk_name": name}, texts=texts)


def poiseuille_local() -> None:
    """testPoiseuille3DBGK's local rows at N = 11, FP64, run to convergence by
    the reference: forced (LocalVelocity wall, periodic) and non-forced
    (LocalVelocity wall and inflow, LocalPressure outflow)"""
    poiseuille_rows((("poiseuille_local", "poiseuille11_forced_local_f64"),
                     ("poiseuille_local_nonforced", "poiseuille11_nonforced_local_f64")))


def poiseuille_interpolated() -> None:
    """testPoiseuille3DBGK's interpolated rows at N = 11, FP64, run to
    convergence by the reference: forced (InterpolatedVelocity wall,
    periodic) and non-forced (InterpolatedVelocity wall and inflow,
    InterpolatedPressure outflow)"""
    # one thread: the reference applies PlaneFdBoundaryProcessor3D under
    # omp parallel for, and a face cell reads the velocity of its neighbours
    # on the outlet from populations another thread may be rewriting
    poiseuille_rows((("poiseuille_interpolated", "poiseuille11_forced_interpolated_f64"),
                     ("poiseuille_interpolated_nonforced", "poiseuille11_nonforced_interpolated_f64")), threads=1)


def poiseuille_rows(rows, threads: int = 4) -> None:
    for mode, name in rows:
        path = os.path.join(RUN, mode + ".bin")
        run("olb_ref_d", [mode, "11", path], threads=threads)
        meta, arrays = read_dump(path)
        mats = arrays["materials"]
        keep = np.flatnonzero(mats != 0)
        fluid = np.flatnonzero(mats == 1)
        q = 19
        n = arrays["f0"].size // q
        p = meta["padding"]
        shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
        core = np.zeros(shape, dtype=bool)
        core[p:-p, p:-p, p:-p] = True
        wall = np.flatnonzero(core.reshape(-1) & (mats >= 2))
        out = {
            "meta": np.array(json.dumps(meta)),
            "materials": mats.astype(np.uint8),
            "distances": arrays["distances"],
            "cells": keep.astype(np.int64),
            "f0": arrays["f0"].reshape(q, n)[:, keep],
            "fluid": fluid.astype(np.int64),
            "f1_fluid": arrays["f1"].reshape(q, n)[:, fluid],
            "wall": wall.astype(np.int64),
            "f1_wall": arrays["f1"].reshape(q, n)[:, wall],
            "energy": arrays["energy"],
            "usol_fluid": arrays["usol"].reshape(n, 3)[fluid],
            "ssol_fluid": arrays["ssol"].reshape(n, 9)[fluid],
            "psol_fluid": arrays["psol"][fluid],
            "rho0": arrays["rho0"][keep],
            "u0": arrays["u0"].reshape(n, 3)[keep],
            "normal_type": np.asarray(arrays["normal_type"]).astype(np.int8),
            "normals": np.asarray(arrays["normals"]).astype(np.int8),
        }
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print("wrote", name, meta)


def poiseuille_bouzidi() -> None:
    """testPoiseuille3DBGK's forced Bouzidi row at N = 11, FP64, run to
    convergence by the reference (one thread: the reference applies its
    post-processors in ascending cell id only when OpenMP runs one thread)"""
    from openlb_amd import bouzidi

    path = os.path.join(RUN, "poiseuille_bouzidi.bin")
    run("olb_ref_d", ["poiseuille_bouzidi", "11", path], threads=1)
    meta, arrays = read_dump(path)
    p = meta["padding"]
    shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
    mats = arrays["materials"]
    keep = np.flatnonzero(mats != 0)
    fluid = np.flatnonzero(mats == 1)
    q = 19
    n = arrays["f0"].size // q
    field = np.asarray(arrays["bouzidi"]).reshape(q, n)
    # cells that run the post-processor: every non-padding cell with an entry
    core = np.zeros(shape, dtype=bool)
    core[p:-p, p:-p, p:-p] = True
    bz_cells = np.flatnonzero((field >= 0).any(axis=0) & core.reshape(-1))
    entries = np.flatnonzero((field >= 0).any(axis=0))
    out = {
        "meta": np.array(json.dumps(meta)),
        "materials": mats.astype(np.uint8),
        "distances": arrays["distances"],
        "cells": keep.astype(np.int64),
        "f0": arrays["f0"].reshape(q, n)[:, keep],
        "fluid": fluid.astype(np.int64),
        "f1_fluid": arrays["f1"].reshape(q, n)[:, fluid],
        "f1_bz": arrays["f1"].reshape(q, n)[:, bz_cells],
        "energy": arrays["energy"],
        "usol_fluid": arrays["usol"].reshape(n, 3)[fluid],
        "ssol_fluid": arrays["ssol"].reshape(n, 9)[fluid],
        "bz_cells": bz_cells.astype(np.int64),
        "bz_q": np.ascontiguousarray(field[:, bz_cells].T),
        "field_cells": entries.astype(np.int64),
        "field_q": np.ascontiguousarray(field[:, entries].T),
    }
    np.savez_compressed(os.path.join(HERE, "poiseuille11_forced_bouzidi_f64.npz"), **out)
    print("wrote poiseuille bouzidi", meta)



def normal_geometries(seed: int = 20261202):
    """core materials of the classifier geometries: unions of boxes with a
    wall shell (some walls material 3, some stray wall cells left far from
    the fluid), random voxel noise, and spheres in a shell (curved walls);
    together they reach every pattern of the classifier"""
    rng = np.random.default_rng(seed)
    out = []

    def shell(fluid, keep=0.0):
        m = np.where(fluid, 1, 0).astype(np.uint8)
        near = np.zeros_like(fluid)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    near |= np.roll(np.roll(np.roll(fluid, dx, 0), dy, 1), dz, 2)
        m[near & ~fluid] = 2
        far = ~near & ~fluid
        m[far & (rng.random(m.shape) < keep)] = 2
        return m

    for _ in range(4):
        n = rng.integers(12, 17, size=3)
        fluid = np.zeros(n, bool)
        for _ in range(rng.integers(2, 5)):
            lo = [rng.integers(1, s - 3) for s in n]
            hi = [rng.integers(l + 2, s - 1) for l, s in zip(lo, n)]
            fluid[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = True
        m = shell(fluid, keep=0.3)
        m[(m == 2) & (rng.random(m.shape) < 0.2)] = 3
        out.append(("boxes", m))
    for _ in range(2):
        n = rng.integers(10, 14, size=3)
        out.append(("noise", rng.choice(np.array([0, 1, 2, 3], np.uint8), size=n, p=[0.1, 0.5, 0.3, 0.1])))
    for _ in range(2):
        n = rng.integers(14, 18, size=3)
        c = (n - 1) / 2 + rng.uniform(-0.5, 0.5, 3)
        r = rng.uniform(4.0, min(n) / 2 - 1.5)
        x, y, z = np.meshgrid(*[np.arange(s) for s in n], indexing="ij")
        out.append(("sphere", shell((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 < r * r)))
    return out


def normals_fixture() -> None:
    """the reference classifier's types and normals (materials 2 and 3,
    fluid 1, outside 0) on the geometries of normal_geometries, overlap 3"""
    arrays = {}
    kinds = []
    for k, (kind, m) in enumerate(normal_geometries()):
        src = os.path.join(RUN, f"normals_geo{k}.bin")
        with open(src, "wb") as f:
            f.write(np.asarray(m.shape, dtype=np.int32).tobytes())
            f.write(np.ascontiguousarray(m, dtype=np.uint8).tobytes())
        path = os.path.join(RUN, f"normals_{k}.bin")
        run("olb_ref_d", ["normals", src, path], threads=1)
        meta, a = read_dump(path)
        padded = np.pad(m, 3)
        if not np.array_equal(a["materials"].reshape(padded.shape), padded):
            raise RuntimeError("the reference's padded materials differ from the geometry written")
        arrays[f"materials_{k}"] = padded.astype(np.uint8)
        arrays[f"normal_type_{k}"] = np.asarray(a["normal_type"]).astype(np.int8).reshape(padded.shape)
        arrays[f"normals_{k}"] = np.asarray(a["normals"]).astype(np.int8).reshape(padded.shape + (3,))
        kinds.append(kind)
    arrays["meta"] = np.array(json.dumps({"kinds": kinds, "padding": 3, "boundary": [2, 3], "fluid": 1,
                                          "outside": 0, "seed": 20261202}))
    np.savez_compressed(os.path.join(HERE, "discrete_normals_geometries.npz"), **arrays)
    print("wrote discrete_normals_geometries", kinds)



def step_cavity_fixture(res: int = 10, steps: int = 100) -> None:
    """the reference's LocalVelocity cavity with a solid corner cube
    (cases.step_cavity: every LocalVelocity shape, BounceBack where no
    normal) in FP64: t = 0 and t = steps populations of the core cells,
    materials and the reference's own normals"""
    c = cases.step_cavity(res)
    p = c.padding
    core_m = np.ascontiguousarray(c.materials[p:-p, p:-p, p:-p])
    src = os.path.join(RUN, "step_cavity_geo.bin")
    with open(src, "wb") as f:
        f.write(np.asarray(core_m.shape, dtype=np.int32).tobytes())
        f.write(core_m.astype(np.uint8).tobytes())
    path = os.path.join(RUN, "step_cavity.bin")
    run("olb_ref_d", ["lv_geometry", src, str(steps), "0.1", "100", path])
    meta, a = read_dump(path)
    if not np.array_equal(a["materials"], c.materials.reshape(-1)):
        raise RuntimeError("the reference's padded materials differ from the case")
    q = 19
    n = a["f0"].size // q
    core = np.flatnonzero(c.core_mask().reshape(-1))
    out = {
        "meta": np.array(json.dumps({**meta, "step": c.info["step"], "u_lid": 0.1, "re": 100.0})),
        "materials": c.materials.reshape(-1).astype(np.uint8),
        "distances": a["distances"],
        "cells": core.astype(np.int64),
        "f0": a["f0"].reshape(q, n)[:, core],
        "f1": a["f1"].reshape(q, n)[:, core],
        "normal_type": np.asarray(a["normal_type"]).astype(np.int8),
        "normals": np.asarray(a["normals"]).astype(np.int8),
    }
    np.savez_compressed(os.path.join(HERE, f"step_cavity_res{res}_f64.npz"), **out)
    print("wrote step_cavity", meta)



def poiseuille_bouzidi_nonforced(n: int = 11) -> None:
    """testPoiseuille3DBGK's non-forced bouzidi row at N = 11 (FP64, one
    thread, run to convergence): BouzidiPostProcessor on the wall,
    BouzidiVelocityPostProcessor on the inflow disc (setBouzidiVelocity with
    the Poiseuille profile), InterpolatedPressure on the outflow disc.
    Besides the arrays of the other rows it holds the two processor lists:
    a cell runs BouzidiPostProcessor when one of its distances points at a
    wall cell (material 2) and BouzidiVelocityPostProcessor when one points
    at an inflow cell (material 3); each applies all of the cell's
    distances, plain before velocity (the reference's operator-type order)"""
    from openlb_amd import bouzidi as bz

    name = f"poiseuille{n}_nonforced_bouzidi_f64"
    path = os.path.join(RUN, "poiseuille_bouzidi_nonforced.bin")
    run("olb_ref_d", ["poiseuille_bouzidi_nonforced", str(n), path], threads=1)
    meta, arrays = read_dump(path)
    mats = arrays["materials"]
    p = meta["padding"]
    shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
    count = int(np.prod(shape))
    q = np.asarray(arrays["bouzidi"]).reshape(19, count)
    v = np.asarray(arrays["bouzidi_velocity"]).reshape(19, count)
    lists = {}
    for solid, key in ((2, "bz"), (3, "bzv")):
        cells, qq = bz.processor_entries(q, mats.reshape(shape), p, solid)
        lists[key + "_cells"] = cells.astype(np.int64)
        lists[key + "_q"] = qq
        if key == "bzv":
            lists["bzv_v"] = v[:, cells].T.copy()
    keep = np.flatnonzero(mats != 0)
    fluid = np.flatnonzero(mats == 1)
    nq = arrays["f0"].size // 19
    out = {
        "meta": np.array(json.dumps(meta)),
        "materials": mats.astype(np.uint8),
        "distances": arrays["distances"],
        "cells": keep.astype(np.int64),
        "f0": arrays["f0"].reshape(19, nq)[:, keep],
        "fluid": fluid.astype(np.int64),
        "f1_fluid": arrays["f1"].reshape(19, nq)[:, fluid],
        "energy": arrays["energy"],
        "usol_fluid": arrays["usol"].reshape(nq, 3)[fluid],
        "ssol_fluid": arrays["ssol"].reshape(nq, 9)[fluid],
        "psol_fluid": arrays["psol"][fluid],
        "rho0": arrays["rho0"][keep],
        "u0": arrays["u0"].reshape(nq, 3)[keep],
        "normal_type": np.asarray(arrays["normal_type"]).astype(np.int8),
        "normals": np.asarray(arrays["normals"]).astype(np.int8),
        **lists,
    }
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, meta, {k: v.shape for k, v in lists.items()})


def poiseuille_bouzidi_nonforced_light(n: int = 41) -> None:
    """testPoiseuille3DBGK's non-forced bouzidi row at N = 41 (FP64, one
    thread, run to convergence), kept small: the geometry (materials and the
    discrete normals of the outflow's InterpolatedPressure), the energy
    series and the converged errors.  Its start, Bouzidi lists and analytic
    solutions are restated from the geometry (tests/golden_runs.py
    build_nonforced_bouzidi), pinned at N = 11 and 21 against full fixtures"""
    name = f"poiseuille{n}_nonforced_bouzidi_light_f64"
    path = os.path.join(RUN, "poiseuille_bouzidi_nonforced.bin")
    run("olb_ref_d", ["poiseuille_bouzidi_nonforced", str(n), path], threads=1)
    meta, arrays = read_dump(path)
    out = {
        "meta": np.array(json.dumps(meta)),
        "materials": arrays["materials"].astype(np.uint8),
        "normal_type": np.asarray(arrays["normal_type"]).astype(np.int8),
        "normals": np.asarray(arrays["normals"]).astype(np.int8),
        "energy": arrays["energy"],
    }
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, meta)

if __name__ == "__main__":
    main()
