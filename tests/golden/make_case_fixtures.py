"""Writes the fixtures of the OpenLB-case tests: OpenLB's own
cavity3dBenchmark (examples/laminar/cavity3dBenchmark/cavity3d.cpp), built
unchanged for its CPU_SISD and CPU_SIMD platforms by
`python -m openlb_amd.platform.build_cases --cpu` (the *_cpu and *_simd
binaries), run here; the fixture holds the point values of the physVelocity
and physPressure pieces the CPU_SISD run writes after its run (its last VTK
output), the arguments, and how far the CPU_SIMD run lies from it.

    python tests/golden/make_case_fixtures.py [tgv3d]

tgv3d (DNS): the lattice statistics it prints and the dissipation rates it
writes, CPU_SISD, with how far CPU_SIMD lies from them.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from openlb_amd import vtk  # noqa: E402

BIN = os.path.join(ROOT, "openlb_amd", "platform", "bin")

# tgv3d (DNS, BGK) at RESOLUTION 64 with its default MAX_PHYS_T 10 (490
# steps); it reads its DNS reference curves from the working directory
TGV_RUN = ("tgv3d_dns", ["--RESOLUTION", "64"])
TGV_DATA = os.path.join(BIN, "tgv3d_data")

# fixture name: (case binary, arguments)
RUNS = {
    "openlb_cavity3d_bounceback_res16": ("cavity3d_bounceback", ["--RESOLUTION", "16", "--TIME_STEPS", "100"]),
    "openlb_cavity3d_res16": ("cavity3d", ["--RESOLUTION", "16", "--TIME_STEPS", "100"]),
}


def run_case(binary: str, args: list[str], cwd: str, env=None) -> str:
    """run a case binary in cwd; returns its output"""
    proc = subprocess.run([binary, *args], cwd=cwd, env=env, stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True, timeout=600)
    if proc.returncode != 0:
        raise RuntimeError(f"{binary} failed ({proc.returncode}):\n{proc.stdout[-4000:]}")
    return proc.stdout


def final_fields(cwd: str) -> dict:
    """the last physVelocity and physPressure pieces the case wrote"""
    data = os.path.join(cwd, "tmp", "vtkData", "data")
    out = {}
    for name in ("physVelocity", "physPressure"):
        pieces = sorted(f for f in os.listdir(data) if f.startswith(name + "_iT") and f.endswith(".vti"))
        out[name] = vtk.read_vti(os.path.join(data, pieces[-1]))
    return out


def field_errors(fields: dict, ref: dict) -> tuple:
    """max |du| / max |u_ref| and max |dp| / max |p_ref| of two runs"""
    u = fields["physVelocity"]["physVelocity"].astype(np.float64)
    u_ref = ref["physVelocity"]["physVelocity"].astype(np.float64)
    p = fields["physPressure"]["physPressure"].astype(np.float64)
    p_ref = ref["physPressure"]["physPressure"].astype(np.float64)
    return (float(np.abs(u - u_ref).max() / np.abs(u_ref).max()),
            float(np.abs(p - p_ref).max() / np.abs(p_ref).max()))


def tgv_series(cwd: str, text: str) -> dict:
    """what tgv3d reports: the lattice statistics it prints every 12 steps
    (LatticeStatistics::print) and the dissipation rates it writes
    (time, molecular, eddy, effective, DNS; Turbulence_Dissipation_Rate.dat)"""
    import re

    stats = [(int(m.group(1)), float(m.group(2)), float(m.group(3)))
             for m in re.finditer(r"\[LatticeStatistics\] step=(\d+); physT=[^;]+; maxCFL=([^;]+); "
                                  r"avgEnergy=([^;]+);", text)]
    dat = np.loadtxt(os.path.join(cwd, "tmp", "gnuplotData", "data", "Turbulence_Dissipation_Rate.dat"))
    return {"stats": np.array(stats, dtype=np.float64), "dissipation": dat}


def series_error(a: np.ndarray, ref: np.ndarray) -> float:
    """max over the value columns (all but the first, step or time) of
    max |a - ref| / max |ref|, skipping columns that are zero throughout"""
    err = 0.0
    for k in range(1, ref.shape[1]):
        scale = np.abs(ref[:, k]).max()
        if scale > 0:
            err = max(err, float(np.abs(a[:, k] - ref[:, k]).max() / scale))
    return err


def run_tgv(binary: str, env=None) -> dict:
    import shutil

    with tempfile.TemporaryDirectory() as cwd:
        for f in os.listdir(TGV_DATA):
            shutil.copy(os.path.join(TGV_DATA, f), cwd)
        text = run_case(binary, TGV_RUN[1], cwd, env)
        return tgv_series(cwd, text)


def tgv_fixture() -> None:
    env = dict(os.environ, OMP_NUM_THREADS="8")
    sisd = run_tgv(os.path.join(BIN, TGV_RUN[0] + "_cpu"), env)
    simd = run_tgv(os.path.join(BIN, TGV_RUN[0] + "_simd"), env)
    spread = {k: series_error(simd[k], sisd[k]) for k in ("stats", "dissipation")}
    meta = {"case": TGV_RUN[0], "args": TGV_RUN[1], "simd_vs_sisd": spread}
    np.savez_compressed(os.path.join(HERE, "openlb_tgv3d_dns_res64.npz"), meta=np.array(json.dumps(meta)),
                        stats=sisd["stats"], dissipation=sisd["dissipation"])
    print("wrote openlb_tgv3d_dns_res64", meta, sisd["stats"].shape, sisd["dissipation"].shape)


def main() -> None:
    for fixture, (case, args) in RUNS.items():
        runs = {}
        for platform in ("cpu", "simd"):
            with tempfile.TemporaryDirectory() as cwd:
                env = dict(os.environ, OMP_NUM_THREADS="4")
                text = run_case(os.path.join(BIN, f"{case}_{platform}"), args, cwd, env)
                runs[platform] = (text, final_fields(cwd))
        text, fields = runs["cpu"]
        arrays = {}
        for name, piece in fields.items():
            for key, value in piece.items():
                arrays[f"{name}/{key}"] = np.asarray(value)
        # how far OpenLB's own two CPU platforms lie apart on this case: the
        # same FP32 arithmetic in another order (vectorised, generated code)
        spread_u, spread_p = field_errors(runs["simd"][1], fields)
        meta = {"case": case, "args": args, "last_line": text.strip().splitlines()[-1],
                "simd_vs_sisd_u": spread_u, "simd_vs_sisd_p": spread_p}
        np.savez_compressed(os.path.join(HERE, fixture + ".npz"), meta=np.array(json.dumps(meta)), **arrays)
        print("wrote", fixture, {k: np.shape(v) for k, v in arrays.items()})


if __name__ == "__main__":
    if sys.argv[1:] == ["tgv3d"]:
        tgv_fixture()
    else:
        main()
        tgv_fixture()
