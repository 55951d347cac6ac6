"""Build recipes for the gfx950 platform library and the CPU oracle.

``build_library`` compiles ``openlb_amd/csrc/olbgfx.hip`` with hipcc for
gfx950 into ``openlb_amd/libolbgfx.so`` (in-tree, so it travels with the
source tree).  ``build_oracle`` compiles the C restatement in ``oracle/`` and,
when the OpenLB sources are present, the reference driver in ``oracle/_ref``;
both are test oracles used only by the tests and the CPU baseline.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "openlb_amd")
LIB = os.path.join(PKG, "libolbgfx.so")
HIP_SOURCES = [os.path.join(PKG, "csrc", "olbgfx.hip"), os.path.join(PKG, "csrc", "olbprobe.hip")]

HIPCC_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off (compiler-generated v_pk_* f32 forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a*b+c is fused only within one source expression, so every kernel that
    # inlines the same collision rounds it the same way (with the default
    # "fast" the fusion follows each kernel's own instruction graph); the
    # split step's dominant kernel and the masked kernel then agree bit for
    # bit in FP32, BF16 and FP64 storage
    "-ffp-contract=on",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _run(cmd: list[str], cwd: str | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def build_library() -> str:
    """Compile the HIP kernels and the C ABI into openlb_amd/libolbgfx.so."""
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), *HIP_SOURCES, "-o", tmp]
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_host_examples() -> None:
    """C++ cases written against the host mirror (openlb_amd/host); plain g++
    over the C ABI, no HIP headers"""
    for name in ("cavity3d", "channel3d"):
        src = os.path.join(PKG, "host", "examples", name + ".cpp")
        out = os.path.join(PKG, "host", "examples", name)
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
              "-I", os.path.join(PKG, "host"), src, "-L", PKG, "-lolbgfx",
              "-Wl,-rpath,$ORIGIN/../..", "-lz", "-o", out])


def build_oracle(with_reference: bool = True) -> None:
    """Compile the CPU restatement (and the reference build if possible)."""
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"])
    if with_reference and os.path.isdir("/root/reference/src"):
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_openlb_cases() -> None:
    """OpenLB's own cases, unchanged, on Platform::GPU_GFX950 (needs the
    OpenLB sources; the binaries travel in openlb_amd/platform/bin)"""
    from openlb_amd.platform import build_cases

    if os.path.isdir(os.path.join(build_cases.OLB_ROOT, "src")):
        build_cases.main([])


def build_all() -> None:
    build_library()
    build_host_examples()
    build_oracle()
    build_openlb_cases()


if __name__ == "__main__":
    build_all()
