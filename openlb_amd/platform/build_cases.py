"""Builds OpenLB's own cases, unchanged, on Platform::GPU_GFX950.

The OpenLB sources are not part of this repository.  The recipe

1. copies the OpenLB tree (OLB_ROOT, default /root/reference: src/,
   external/ and the examples it builds) into openlb_amd/platform/_tree,
   which git ignores and no GPU run reads;
2. adds GPU_GFX950 at OpenLB's platform switch points in that copy (PATCHES:
   each inserts a few lines next to an anchor line of the file, the way the
   reference adds a platform: the enum value, the dispatch cases, the
   default platform of the load balancer, the includes of
   openlb_amd/platform/olbgfx950/*.h, the start-up check);
3. compiles each case of CASES with g++ against that copy, the headers of
   openlb_amd/platform and include/, and links it to libolbgfx.so, into
   openlb_amd/platform/bin (git-ignored; the binaries travel with the tree).

The case sources are compiled as they are, only selected by the
preprocessor switches their authors provide (cavity3d.cpp:29
LID_DRIVEN_BOUNCE_BACK, tgv3d.cpp:95-110 the bulk dynamics).

    python -m openlb_amd.platform.build_cases [--cpu] [CASE ...]

--cpu also builds each case for OpenLB's own CPU_SISD (name_cpu) and
CPU_SIMD (name_simd) platforms, the reference runs the integration tests
compare with.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
TREE = os.path.join(HERE, "_tree")
BIN = os.path.join(HERE, "bin")
OLB_ROOT = os.environ.get("OLB_ROOT", "/root/reference")

# name: (source relative to the OpenLB root, extra preprocessor switches)
CASES = {
    "cavity3d": ("examples/laminar/cavity3dBenchmark/cavity3d.cpp", []),
    "cavity3d_bounceback": ("examples/laminar/cavity3dBenchmark/cavity3d.cpp", ["-DLID_DRIVEN_BOUNCE_BACK"]),
    "tgv3d_dns": ("examples/turbulence/tgv3d/tgv3d.cpp", ["-DDNS"]),
    "tgv3d": ("examples/turbulence/tgv3d/tgv3d.cpp", []),
}

# data files a case reads from its working directory, copied beside the
# binaries (bin/<dir>/): tgv3d's DNS reference curves (tgv3d.cpp:203-229)
CASE_DATA = {
    "tgv3d_data": ("examples/turbulence/tgv3d", (".inp",)),
}

# OpenLB's core translation units (the rest of the library is header-only)
CORE_SOURCES = [
    "src/communication/mpiManager.cpp",
    "src/communication/ompManager.cpp",
    "src/core/olbInit.cpp",
    "src/core/expr.cpp",
    "src/io/ostreamManager.cpp",
    "external/tinyxml2/tinyxml2.cpp",
]

_HIP_CASE = "#ifdef PLATFORM_GPU_HIP\n"

# (file, anchor, text, where): the text goes "after" or "before" the anchor
# line, or "dup" duplicates every `#ifdef PLATFORM_GPU_HIP ... #endif` block
# of the file with GPU_HIP renamed (the dispatch switches)
PATCHES = [
    ("src/core/platform/platform.h", "  GPU_HIP   /// GPU using HIP\n",
     "  GPU_HIP,  /// GPU using HIP\n  GPU_GFX950 /// MI355X (gfx950) through libolbgfx\n", "replace"),
    ("src/core/platform/platform.h", None, None, "dup"),
    ("src/core/platform/dispatch.h", None, None, "dup"),
    ("src/core/platform/column.h", '#include "gpu/hip/column.h"\n#endif\n',
     '\n#ifdef PLATFORM_GPU_GFX950\n#include "olbgfx950/column.h"\n#endif\n', "after"),
    ("src/communication/loadBalancer.h", "    #if defined(PLATFORM_GPU_CUDA)\n      return Platform::GPU_CUDA;\n",
     "    #if defined(PLATFORM_GPU_GFX950)\n      return Platform::GPU_GFX950;\n    #elif defined(PLATFORM_GPU_CUDA)\n"
     "      return Platform::GPU_CUDA;\n", "replace"),
    ("src/core/initialize.h", None, None, "dup"),
    ("src/olb.h", '#include "cse/cse.h"\n',
     '\n#ifdef PLATFORM_GPU_GFX950\n#include "olbgfx950/platform.h"\n#endif\n', "after"),
]


def _dup_hip_blocks(text: str) -> str:
    """append to every `#ifdef PLATFORM_GPU_HIP` ... `#endif` block a copy
    for GPU_GFX950"""
    out, i = [], 0
    while True:
        j = text.find(_HIP_CASE, i)
        if j < 0:
            out.append(text[i:])
            return "".join(out)
        k = text.find("#endif\n", j) + len("#endif\n")
        block = text[j:k]
        out.append(text[i:k])
        out.append(block.replace("GPU_HIP", "GPU_GFX950"))
        i = k


def prepare_tree() -> str:
    """the patched copy of the OpenLB tree"""
    src = os.path.join(OLB_ROOT, "src")
    if not os.path.isdir(src):
        raise RuntimeError(f"the OpenLB sources are not at {OLB_ROOT}")
    if os.path.isdir(TREE):
        shutil.rmtree(TREE)
    os.makedirs(TREE)
    shutil.copytree(src, os.path.join(TREE, "src"))
    shutil.copytree(os.path.join(OLB_ROOT, "external", "tinyxml2"), os.path.join(TREE, "external", "tinyxml2"))
    for rel, _ in CASES.values():
        d = os.path.dirname(rel)
        if not os.path.isdir(os.path.join(TREE, d)):
            shutil.copytree(os.path.join(OLB_ROOT, d), os.path.join(TREE, d))
    for rel, anchor, text, where in PATCHES:
        path = os.path.join(TREE, rel)
        body = open(path).read()
        if where == "dup":
            new = _dup_hip_blocks(body)
            if new == body:
                raise RuntimeError(f"no GPU_HIP block in {rel}")
        else:
            if body.count(anchor) != 1:
                raise RuntimeError(f"anchor not found once in {rel}: {anchor!r}")
            if where == "replace":
                new = body.replace(anchor, text)
            elif where == "after":
                new = body.replace(anchor, anchor + text)
            else:
                new = body.replace(anchor, text + anchor)
        with open(path, "w") as f:
            f.write(new)
    return TREE


# the platforms a case is built for: GPU_GFX950, and OpenLB's own CPU_SISD
# and CPU_SIMD (AVX-512) for the reference runs the tests compare with
PLATFORM_FLAGS = {
    "gfx950": ["-march=x86-64-v3", "-DPLATFORM_CPU_SISD", "-DPLATFORM_GPU_GFX950"],
    "cpu": ["-march=x86-64-v3", "-DPLATFORM_CPU_SISD"],
    "simd": ["-march=x86-64-v4", "-DPLATFORM_CPU_SISD", "-DPLATFORM_CPU_SIMD"],
}


def _flags(platform: str = "gfx950") -> list[str]:
    return ["-O3", "-std=c++20", "-fopenmp", "-DPARALLEL_MODE_OMP",
            *PLATFORM_FLAGS[platform], '-DOLB_VERSION="1.8r1"',
            "-DDEFAULT_FLOATING_POINT_TYPE=float", "-Wno-deprecated-declarations",
            "-I", os.path.join(TREE, "src"), "-I", os.path.join(TREE, "external", "tinyxml2"),
            "-I", HERE, "-I", os.path.join(ROOT, "include")]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout[-20000:])
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... {cmd[-1]}")


def build_core(platform: str = "gfx950") -> list[str]:
    objdir = os.path.join(TREE, "obj_" + platform)
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for rel in CORE_SOURCES:
        obj = os.path.join(objdir, os.path.basename(rel).replace(".cpp", ".o"))
        _run(["g++", *_flags(platform), "-c", os.path.join(TREE, rel), "-o", obj])
        objs.append(obj)
    return objs


def build_case(name: str, objs: list[str], platform: str = "gfx950") -> str:
    """the case on GPU_GFX950, or the same source on OpenLB's own CPU_SISD
    ("cpu", name + "_cpu") or CPU_SIMD ("simd", name + "_simd") platform, as
    the reference runs it"""
    rel, switches = CASES[name]
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, name if platform == "gfx950" else f"{name}_{platform}")
    link = ["-L", PKG, "-lolbgfx", "-Wl,-rpath,$ORIGIN/../.."] if platform == "gfx950" else []
    _run(["g++", *_flags(platform), *switches, os.path.join(TREE, rel), *objs, *link, "-lz", "-o", out])
    return out


def copy_case_data() -> None:
    for target, (rel, suffixes) in CASE_DATA.items():
        out = os.path.join(BIN, target)
        os.makedirs(out, exist_ok=True)
        src = os.path.join(OLB_ROOT, rel)
        for f in sorted(os.listdir(src)):
            if f.endswith(suffixes):
                shutil.copy(os.path.join(src, f), os.path.join(out, f))


def main(names: list[str], jobs: int = 4) -> list[str]:
    """patch the tree and build the cases (all by default), `jobs` compilers
    at a time; --cpu also builds them for CPU_SISD and CPU_SIMD"""
    from concurrent.futures import ThreadPoolExecutor

    cpu = "--cpu" in names
    names = [n for n in names if n != "--cpu"] or list(CASES)
    prepare_tree()
    copy_case_data()
    objs = build_core()
    work = [(n, objs, "gfx950") for n in names]
    if cpu:
        for platform in ("cpu", "simd"):
            cpu_objs = build_core(platform)
            work += [(n, cpu_objs, platform) for n in names]
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        built = list(pool.map(lambda w: build_case(*w), work))
    for path in built:
        print("built", path, flush=True)
    return built


if __name__ == "__main__":
    main(sys.argv[1:])
