"""Writes the golden fixtures of tests/golden/ by running OpenLB's own CPU
path (oracle/_ref, built by `make -C oracle ref` from the OpenLB sources).

Run in the build container (the OpenLB sources are not on the GPU host):
    python tests/golden/make_fixtures.py
Each fixture holds the materials of the padded block (CellID order), the
neighbour distances, the populations at t = 0 and after the run for the core
cells (logical order, shifted by -t_i), and the run's parameters.
"""
from __future__ import annotations

import json
import os
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


def run(binary: str, args: list[str]) -> None:
    os.makedirs(RUN, exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS="4")
    subprocess.run([os.path.join(REF, binary), *args], cwd=RUN, env=env, check=True,
                   stdout=subprocess.DEVNULL)


def core_ids(meta) -> np.ndarray:
    p = meta["padding"]
    shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
    idx = np.arange(np.prod(shape)).reshape(shape)
    return idx[p:-p, p:-p, p:-p].reshape(-1)


def save(name: str, meta: dict, arrays: dict, extra: dict | None = None, cells=None) -> None:
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
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, {k: v.shape for k, v in out.items() if hasattr(v, "shape")})


def cavity_reduced(binary: str, tag: str, res: int = 64, steps: int = 1000) -> None:
    """config 1 at full size: the cavity3dBenchmark geometry at RESOLUTION 64,
    BounceBack walls, BounceBackVelocity lid u = 0.1, Re = 100 (tau = 0.692),
    1000 steps from rest.  Too large to keep whole, so the fixture holds the
    velocity on the mid-plane z = res/2, the populations of 256 seeded fluid
    cells, and FP64 sums over the fluid."""
    path = os.path.join(RUN, f"cavity{res}_{tag}.bin")
    run(binary, ["cavity", str(res), str(steps), "0.1", "100", path])
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
    np.savez_compressed(os.path.join(HERE, f"cavity3d_res{res}_{steps}steps_{tag}.npz"),
                        meta=np.array(json.dumps(meta)), materials=arrays["materials"].astype(np.uint8),
                        u_midplane=mid, sample_cells=sample.astype(np.int64),
                        f1_sample=f[:, sample].astype(np.float32 if tag == "f32" else np.float64),
                        fluid_sums=sums)
    print("wrote", f"cavity3d_res{res}_{steps}steps_{tag}", sums)


def main() -> None:
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
    }
    np.savez_compressed(os.path.join(HERE, "poiseuille11_forced_bounceback_f64.npz"), **out)
    print("wrote poiseuille", meta)


if __name__ == "__main__":
    main()
