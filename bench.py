"""Throughput of the D3Q19 collide-and-stream step on one MI355X.

Workload: the periodic Taylor-Green vortex at 512^3 (D3Q19 BGK, tau = 0.56,
u0 = 0.04), one step = one collideAndStream of the whole lattice.  The timed
region runs with the populations already resident in HBM.  The headline value
is GLUPS with FP32 population storage; the same step with FP16 and BF16
shifted storage is timed beside it.

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X published HBM3E bandwidth
BYTES_PER_POP = {"f32": 4, "f16": 2, "bf16": 2, "f64": 8}


# with --tune-cache, every autotuner call reads and writes that cache, so a
# second run with the same file launches exactly the kernels of the first
# without trial launches (what a kernel trace of the run needs)
USE_TUNE_CACHE = False


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--storages", default="f32,f16,bf16")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-size", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--ens-steps", type=int, default=200)
    ap.add_argument("--ens-storages", default="f32,f16")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--acc-size", type=int, default=512)
    ap.add_argument("--acc-steps", type=int, default=1000)
    ap.add_argument("--no-porous", action="store_true")
    ap.add_argument("--porous-steps", type=int, default=500)
    ap.add_argument("--no-slab", action="store_true")
    ap.add_argument("--no-les", action="store_true")
    ap.add_argument("--no-probes", action="store_true")
    ap.add_argument("--no-openlb-cases", action="store_true")
    ap.add_argument("--tune-cache", default="",
                    help="autotuner cache file; launch choices found there are reused without trials")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a 1 GiB device copy (PMC counter calibration)")
    return ap.parse_args()


def time_storage(torch, size, storage, steps, warmup, block, tune=True):
    from openlb_amd import Lattice, autotune, cases

    dtype = np.float64 if storage == "f64" else np.float32
    case = cases.tgv(size, u0=0.04, tau=0.56, dtype=dtype)
    lat = Lattice(case.extent, case.padding, case.periodic, storage)
    if block:
        lat.set_launch(block)
    case.apply(lat)
    del case
    stream = torch.cuda.current_stream()
    handle = stream.cuda_stream
    tuned = None
    if tune and not block:
        tuned = autotune.tune_lattice(lat, steps=3, stream=handle, use_cache=USE_TUNE_CACHE)
        tuned = {k: tuned[k] for k in ("block", "waves", "order", "ms")}
    lat.collide_and_stream(warmup, handle)
    torch.cuda.synchronize()

    # whole-job timing: K steps between two synchronisations
    t0 = time.perf_counter()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    lat.collide_and_stream(steps, handle)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = start.elapsed_time(stop)

    # per-launch kernel duration: HIP events around single steps on the stream
    # the kernel is launched on
    durs = []
    for _ in range(min(steps, 20)):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        lat.collide_and_stream(1, handle)
        b.record(stream)
        durs.append((a, b))
    torch.cuda.synchronize()
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in durs]))
    moments = lat.reduce_moments(handle)
    lat.close()
    cells = size ** 3
    return {
        "ms_per_step": ms_total / steps,
        "wall_s": wall,
        "glups": cells * steps / (ms_total * 1e-3) / 1e9,
        "kernel_ms": kernel_ms,
        "kernel_glups": cells / (kernel_ms * 1e-3) / 1e9,
        "bytes_per_cell": 2 * 19 * BYTES_PER_POP[storage],
        "launch": tuned,
        "finite": bool(np.isfinite(moments["mass"]) and np.isfinite(moments["sumUSqr"])),
    }


def time_ensemble(torch, storage, steps, warmup, tune=True):
    """the ensemble of 64 lid-driven cavities, N ~ U{32..128}, U ~ U[0.02, 0.1],
    Re = 100; persistent kernel per step (eager and graph-replayed) against
    launching every cavity separately"""
    from openlb_amd import Ensemble, Lattice, autotune, cases

    dtype = np.float64 if storage == "f64" else np.float32
    lats = []
    for c in cases.ensemble_cavities(dtype=dtype):
        lat = Lattice(c.extent, c.padding, c.periodic, storage)
        c.apply(lat)
        lats.append(lat)
    ens = Ensemble(lats)
    handle = torch.cuda.current_stream().cuda_stream
    tuned = None
    if tune:
        tuned = autotune.tune_ensemble(ens, steps=3, stream=handle, use_cache=USE_TUNE_CACHE)
        tuned = {k: tuned[k] for k in ("block", "waves", "tiles", "order", "group_bytes", "steps_per_group", "ticket", "ms")}
    ens.step(warmup, handle)
    steps = max(10, steps // 10 * 10)
    ms_persistent = ens.time("persistent", steps, handle)
    ens.capture(10)
    ens.launch_graph(1, handle)
    ms_graph = ens.time("graph", steps, handle)
    ms_sep = ens.time("separate", steps, handle)
    ms_sync = ens.time("separate_synchronized", steps, handle)
    ens.set_schedule(0, 1, False)
    ens.step(2, handle)
    ms_flat = ens.time("persistent", steps, handle)
    cells = ens.cells
    out = {
        "members": len(lats),
        "cells": cells,
        "steps": steps,
        "launch": tuned,
        "graph_steps_per_s": steps / (ms_graph * 1e-3),
        "persistent_steps_per_s": steps / (ms_persistent * 1e-3),
        "separate_steps_per_s": steps / (ms_sep * 1e-3),
        "graph_glups": cells * steps / (ms_graph * 1e-3) / 1e9,
        "persistent_glups": cells * steps / (ms_persistent * 1e-3) / 1e9,
        "separate_glups": cells * steps / (ms_sep * 1e-3) / 1e9,
        "ungrouped_persistent_steps_per_s": steps / (ms_flat * 1e-3),
        "separate_synchronized_steps_per_s": steps / (ms_sync * 1e-3),
    }
    out["speedup_graph_vs_separate"] = ms_sep / ms_graph
    out["speedup_graph_vs_separate_synchronized"] = ms_sync / ms_graph
    ens.close()
    for lat in lats:
        lat.close()
    return out


def accuracy(torch, size, steps):
    """TGV at size^3 run `steps` steps in FP64, FP32, FP16 and BF16
    storage from the same analytic start; velocity error of each against the
    FP64 run (relative L2 over the core, max abs relative to u0), and the FP64
    kinetic-energy decay against the analytic rate"""
    from openlb_amd import Lattice, autotune, cases

    n = size
    core = None
    u_ref = None
    out = {"size": n, "steps": steps}
    for st in ("f64", "f32", "f16", "bf16"):
        case = cases.tgv(n, u0=0.04, tau=0.56, dtype=np.float64 if st == "f64" else np.float32)
        if core is None:
            core = case.core_mask().reshape(-1)
        lat = Lattice(case.extent, case.padding, case.periodic, st)
        case.apply(lat)
        del case
        autotune.tune_lattice(lat, steps=3, use_cache=USE_TUNE_CACHE)
        e0 = lat.reduce_moments()["sumUSqr"]
        t0 = time.perf_counter()
        lat.collide_and_stream(steps)
        lat.synchronize()
        dt = time.perf_counter() - t0
        m = lat.reduce_moments()
        _, u = lat.compute_rho_u()
        lat.close()
        u = u[core].astype(np.float64)
        if st == "f64":
            u_ref = u
            nu = (0.56 - 0.5) / 3.0
            rate = -np.log(m["sumUSqr"] / e0) / steps
            out["f64"] = {"energy_decay_rate": rate,
                          "analytic_rate": 2 * nu * 3 * (2 * np.pi / n) ** 2,
                          "seconds": dt}
            continue
        diff = u - u_ref
        out[st] = {
            "velocity_l2_vs_f64": float(np.sqrt(np.sum(diff ** 2) / np.sum(u_ref ** 2))),
            "velocity_max_vs_f64_over_u0": float(np.max(np.abs(diff)) / 0.04),
            "mass_drift_per_cell": float(m["mass"] / n ** 3),
            "seconds": dt,
        }
        del u, diff
    return out


def time_porous(torch, steps, storages=("f32", "f16")):
    """porous channel: 384x128x128 periodic sphere pack (r = 6, porosity ~0.7),
    ForcedTRT (Lambda = 3/16) with g_x = 1e-6, BounceBack solids"""
    from openlb_amd import Lattice, autotune, cases

    res = {}
    case = cases.porous_channel((384, 128, 128), radius=6.0, porosity=0.7, seed=20261202, force=1e-6,
                                dtype=np.float32)
    handle = torch.cuda.current_stream().cuda_stream
    cells = int(np.prod(case.extent))
    for st in storages:
        lat = Lattice(case.extent, case.padding, case.periodic, st)
        case.apply(lat)
        tuned = autotune.tune_lattice(lat, steps=3, stream=handle, use_cache=USE_TUNE_CACHE)
        lat.collide_and_stream(10, handle)
        ms = lat.time_steps(steps, handle)
        m = lat.reduce_moments(handle)
        res[st] = {"glups": cells * steps / (ms * 1e-3) / 1e9, "ms_per_step": ms / steps,
                   "mean_ux": m["momentum"][0] / m["nCells"],
                   "launch": {k: tuned[k] for k in ("block", "waves", "order")}}
        lat.close()
    res["porosity"] = case.info["porosity"]
    res["spheres"] = int(len(case.info["spheres"]))
    return res


def time_smagorinsky(torch, size, steps, storages=("f32", "f16")):
    """the bulk dynamics of OpenLB's tgv3d example: SmagorinskyBGKdynamics
    (C = 0.1) on the Taylor-Green vortex at size^3, tau = 0.505"""
    from openlb_amd import SMAGORINSKY_BGK, Dynamics, Lattice, autotune, cases

    case = cases.tgv(size, u0=0.04, tau=0.505, dtype=np.float32)
    case.dynamics[1] = Dynamics(SMAGORINSKY_BGK, omega=1 / 0.505, smagorinsky=0.1)
    handle = torch.cuda.current_stream().cuda_stream
    res = {}
    for st in storages:
        lat = Lattice(case.extent, case.padding, case.periodic, st)
        case.apply(lat)
        tuned = autotune.tune_lattice(lat, steps=3, stream=handle, use_cache=USE_TUNE_CACHE)
        lat.collide_and_stream(5, handle)
        ms = lat.time_steps(steps, handle) / steps
        res[st] = {"glups": size ** 3 / (ms * 1e-3) / 1e9, "ms_per_step": ms,
                   "launch": {k: tuned[k] for k in ("block", "waves", "order")}}
        lat.close()
    return res


def time_slab(torch, steps, storage="f32", shape=(64, 512, 512)):
    """one GPU's share of a 512^3 block cut into eight x-slabs: 64x512x512 BGK,
    periodic along y and z, its two x faces fed through halo planes (wired to
    each other here, where a multi-GPU run would receive the neighbours'
    planes); the step with pack/unpack, and the four halo kernels alone"""
    from openlb_amd import BGK, Dynamics, Lattice, autotune, cases

    p = 3
    mats = np.zeros(tuple(n + 2 * p for n in shape), dtype=np.uint8)
    mats[p:-p, p:-p, p:-p] = 1
    case = cases.Case(name="slab", extent=shape, padding=p, periodic=(False, True, True), materials=mats,
                      dynamics={1: Dynamics(BGK, omega=1 / 0.56)})
    lat = Lattice(case.extent, case.padding, case.periodic, storage)
    case.apply(lat)
    stream = torch.cuda.current_stream()
    handle = stream.cuda_stream
    tuned = autotune.tune_lattice(lat, steps=3, stream=handle, use_cache=USE_TUNE_CACHE)
    nbytes = lat.halo_bytes(0)
    lo, hi = lat.halo_buffer(0), lat.halo_buffer(0)

    def exchange():
        lat.halo_pack(0, 1, hi.ptr, handle)
        lat.halo_pack(0, 0, lo.ptr, handle)
        lat.halo_unpack(0, 0, hi.ptr, handle)
        lat.halo_unpack(0, 1, lo.ptr, handle)

    def timed(fn, k):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(k):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / k

    def step():
        lat.collide_and_stream(1, handle)
        exchange()

    for _ in range(5):
        step()
    ms_step = timed(step, steps)
    ms_halo = timed(exchange, steps)
    lo.close()
    hi.close()
    lat.close()
    cells = int(np.prod(shape))
    return {"shape": list(shape), "storage": storage, "glups": cells / (ms_step * 1e-3) / 1e9,
            "ms_per_step": ms_step, "ms_halo_kernels": ms_halo, "halo_bytes_per_face": nbytes,
            "launch": {k: tuned[k] for k in ("block", "waves", "order")}}


def probe_ceilings(size, storages, orders):
    """bandwidth of the step kernel's access patterns on the same lattice
    (include/olbgfx_probe.h): a plain 16-byte copy of the rings, and the 19
    PS rings read only, written only, and read-modified-written in place
    with the step kernel's indexing, block shape and (tuned) block order,
    after one step of ring shifts; ring_inplace_aligned is the in-place
    pattern with every ring offset zero (each ring's lines aligned with the
    cell index), the most any alignment of the rings could give;
    ring_two_buffer reads the rings of one buffer and writes those of a
    second at offset zero, the traffic of a two-buffer step"""
    from openlb_amd import probe_bandwidth

    out = {}
    for st in storages:
        order = orders.get(st, -1)
        out[st] = {pat: round(probe_bandwidth((size, size, size), 3, st, pat, 1, order, 10), 1)
                   for pat in ("copy", "ring_read", "ring_write", "ring_inplace", "ring_two_buffer")}
        out[st]["ring_inplace_aligned"] = round(
            probe_bandwidth((size, size, size), 3, st, "ring_inplace", 0, order, 10), 1)
        out[st]["block_order"] = order
    return out


def openlb_cases(resolution=100, steps=1000):
    """OpenLB's own cavity3dBenchmark built unchanged on the GPU_GFX950
    platform (openlb_amd/platform/bin): the MLUPs it prints itself, the
    metric of the reference's GPU CI (.ci/benchmark_cuda.sh:15-20), for its
    default LocalVelocity walls in FP32 and FP16 storage and its bounce-back
    variant"""
    binary = os.path.join(ROOT, "openlb_amd", "platform", "bin")
    runs = {"cavity3d_f32": ("cavity3d", "f32"), "cavity3d_f16": ("cavity3d", "f16"),
            "cavity3d_bounceback_f32": ("cavity3d_bounceback", "f32")}
    out = {}
    for key, (case, storage) in runs.items():
        path = os.path.join(binary, case)
        if not os.path.exists(path):
            continue
        env = dict(os.environ, OMP_NUM_THREADS="4", OLBGFX_STORAGE=storage)
        try:
            with tempfile.TemporaryDirectory() as scratch:
                text = subprocess.run([path, "--RESOLUTION", str(resolution), "--TIME_STEPS", str(steps),
                                       "--VTK_ENABLED", "0"], env=env, cwd=scratch, capture_output=True,
                                      text=True, timeout=300, check=True).stdout
            last = text.strip().splitlines()[-1].split(",")
            out[key] = {"mlups": float(last[-1]), "resolution": resolution, "time_steps": steps}
        except Exception as exc:
            sys.stderr.write(f"{case} ({storage}) failed: {exc}\n")
    return out


def cpu_baseline(size, steps):
    """The reference's own CPU path (oracle/_ref, compiled from the OpenLB
    sources) when it is present, else the C restatement in oracle/, timed on
    all host cores on a bounded sample of the same workload."""
    ref = os.path.join(ROOT, "oracle", "_ref", "olb_ref_f")
    cores = os.cpu_count() or 1
    if os.path.exists(ref):
        env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close")
        try:
            with tempfile.TemporaryDirectory() as scratch:  # the reference writes ./tmp/
                out = subprocess.run([ref, "bench_tgv", str(size), str(steps)], env=env, cwd=scratch,
                                     capture_output=True, text=True, timeout=300, check=True).stdout
            mlups = float([ln for ln in out.splitlines() if ln.startswith("MLUPS")][-1].split()[1])
            return {"value": mlups / 1000.0, "unit": "GLUPS", "cores": cores, "kind": "reference",
                    "sample": f"OpenLB CPU_SIMD+OpenMP BGK TGV {size}^3 FP32, {steps} steps after 2 warm-up"}
        except Exception as exc:  # fall through to the restatement
            sys.stderr.write(f"reference baseline failed: {exc}\n")
    from openlb_amd import cases
    from oracle.oracle import OracleLattice, set_threads

    set_threads(cores)
    case = cases.tgv(size, dtype=np.float32)
    lat = OracleLattice(case.extent, case.padding, case.periodic, "f32")
    case.apply(lat)
    lat.collide_and_stream(2)
    t0 = time.perf_counter()
    lat.collide_and_stream(steps)
    dt = time.perf_counter() - t0
    return {"value": size ** 3 * steps / dt / 1e9, "unit": "GLUPS", "cores": cores, "kind": "port",
            "sample": f"C restatement (OpenMP) BGK TGV {size}^3 FP32, {steps} steps after 2 warm-up"}


def load_traffic(storage):
    path = os.path.join(ROOT, "profiles", f"traffic_step_{storage}_tgv512.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("bytes_per_launch")
    return None


def main():
    global USE_TUNE_CACHE
    args = parse_args()
    if args.tune_cache:
        os.environ["OLBGFX_TUNE_CACHE"] = os.path.abspath(args.tune_cache)
        USE_TUNE_CACHE = True
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (a lattice shards into x-slabs "
              "whose halo planes olbgfx_halo_pack/unpack build; the transport between GPUs is not built)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    if args.calib_copy:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes_read": a.numel() * 4, "calib_copy_bytes_written": b.numel() * 4}))
        return
    storages = [s for s in args.storages.split(",") if s]
    res = {}
    for st in storages:
        res[st] = time_storage(torch, args.size, st, args.steps, args.warmup, args.block, tune=not args.no_tune)
        print(f"# {st}: {res[st]['glups']:.2f} GLUPS, kernel {res[st]['kernel_ms']:.3f} ms", file=sys.stderr)
    ens = {}
    if not args.no_ensemble:
        for st in [x for x in args.ens_storages.split(",") if x]:
            ens[st] = time_ensemble(torch, st, args.ens_steps, args.warmup, tune=not args.no_tune)
            print(f"# ensemble {st}: graph {ens[st]['graph_steps_per_s']:.1f} steps/s "
                  f"({ens[st]['graph_glups']:.2f} GLUPS), separate {ens[st]['separate_steps_per_s']:.1f} "
                  f"steps/s, x{ens[st]['speedup_graph_vs_separate']:.2f}; separate and synchronised "
                  f"{ens[st]['separate_synchronized_steps_per_s']:.1f} steps/s, "
                  f"x{ens[st]['speedup_graph_vs_separate_synchronized']:.2f}", file=sys.stderr)
    porous = None if args.no_porous else time_porous(torch, args.porous_steps)
    if porous:
        print(f"# porous: f32 {porous['f32']['glups']:.2f} GLUPS, f16 {porous['f16']['glups']:.2f} GLUPS",
              file=sys.stderr)
    les = None if args.no_les else time_smagorinsky(torch, args.size, args.steps)
    if les:
        print("# smagorinsky: " + ", ".join(f"{k} {v['glups']:.2f} GLUPS" for k, v in les.items()),
              file=sys.stderr)
    slab = None if args.no_slab else time_slab(torch, args.steps)
    if slab:
        print(f"# slab {slab['shape']}: {slab['glups']:.2f} GLUPS with halo planes, "
              f"halo kernels {slab['ms_halo_kernels'] * 1e3:.1f} us of {slab['ms_per_step'] * 1e3:.1f} us",
              file=sys.stderr)
    acc = None if args.no_accuracy else accuracy(torch, args.acc_size, args.acc_steps)
    if acc:
        print("# accuracy vs f64: " + ", ".join(f"{k} L2 {acc[k]['velocity_l2_vs_f64']:.2e}"
                                                for k in ("f32", "f16", "bf16")), file=sys.stderr)
    head = res.get("f32") or next(iter(res.values()))
    head_storage = "f32" if "f32" in res else storages[0]
    probes = None
    if not args.no_probes:
        torch.cuda.empty_cache()
        probes = probe_ceilings(args.size, [head_storage] + (["f16"] if "f16" in res and head_storage != "f16" else []),
                                {k: v["launch"]["order"] for k, v in res.items()})
        print("# probes: " + "; ".join(f"{k} " + ", ".join(f"{p} {v[p]:.0f}" for p in
                                                           ("copy", "ring_read", "ring_write", "ring_inplace",
                                                            "ring_two_buffer"))
                                       for k, v in probes.items()) + " GB/s", file=sys.stderr)
    cases = None if args.no_openlb_cases else openlb_cases()
    if cases:
        print("# OpenLB cavity3dBenchmark on GPU_GFX950, RESOLUTION 100, 1000 steps: "
              + ", ".join(f"{k} {v['mlups']:.0f} MLUPs" for k, v in cases.items()), file=sys.stderr)
    cells = args.size ** 3
    alg_bytes = head["bytes_per_cell"] * cells
    achieved = alg_bytes / (head["kernel_ms"] * 1e-3) / 1e9
    traffic = load_traffic(head_storage)
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4),
        "traffic": traffic,
        "kernel": f"step_kernel<{head_storage}, BGK>",
        "bytes_per_cell": head["bytes_per_cell"],
    }
    if probes:
        p = probes[head_storage]
        roofline.update({
            "measured_copy_GBps": p["copy"],
            "frac_of_measured_copy": round(achieved / p["copy"], 4),
            "pattern_ceiling_GBps": p["ring_inplace"],
            "frac_of_pattern_ceiling": round(achieved / p["ring_inplace"], 4),
        })
    base = None if args.no_cpu_baseline else cpu_baseline(args.cpu_size, args.cpu_steps)
    line = {
        "metric": "D3Q19 GLUPS & %HBM roofline at 512^3 (FP32 vs FP16 store); ensemble steps/s",
        "value": round(head["glups"], 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head_storage == "f32" else head_storage,
        "data": "synthetic (analytic Taylor-Green initial field)",
        "config": {"workload": f"taylor_green_d3q19_bgk_{args.size}^3", "cells": cells, "tau": 0.56,
                   "u0": 0.04, "storage": head_storage, "parallelism": "single GPU"},
        "storage_glups": {k: round(v["glups"], 3) for k, v in res.items()},
        "storage_launch": {k: v["launch"] for k, v in res.items()},
        "storage_hbm_frac": {k: round(v["bytes_per_cell"] * cells / (v["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
                             for k, v in res.items()},
        "fp16_over_fp32": round(res["f16"]["glups"] / res["f32"]["glups"], 3) if "f16" in res and "f32" in res else None,
        "ensemble": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                     for k, v in ens.items()},
        "porous": porous,
        "slab": slab,
        "smagorinsky": les,
        "accuracy": acc,
        "probes_GBps": probes,
        "openlb_cavity3d_benchmark": cases,
        "roofline": roofline,
        "cpu_baseline": base,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
