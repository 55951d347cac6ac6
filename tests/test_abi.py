"""The C-ABI library loads and exports every entry point include/olbgfx.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    """every entry point declared by the headers under include/"""
    inc = os.path.join(ROOT, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    return sorted(set(re.findall(r"\b(olbgfx_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "olbgfx_collide_and_stream" in syms
    assert "olbgfx_lattice_create" in syms
    assert len(syms) >= 18


def test_library_exports_every_declared_symbol():
    import openlb_amd

    lib_path = openlb_amd.LIB_PATH
    if not os.path.exists(lib_path):
        pytest.fail("libolbgfx.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(lib_path)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_error_reporting_without_device():
    """Argument checks fail before any device call."""
    import openlb_amd

    lib = openlb_amd.library()
    h = ctypes.c_void_p()
    assert lib.olbgfx_lattice_create(None, ctypes.byref(h)) == 1
    assert b"null" in lib.olbgfx_last_error()
    desc = openlb_amd._LatticeDesc()
    desc.nx, desc.ny, desc.nz, desc.padding, desc.storage = 4, 4, 4, 0, 0
    assert lib.olbgfx_lattice_create(ctypes.byref(desc), ctypes.byref(h)) == 1
    assert b"padding" in lib.olbgfx_last_error()
    desc.padding, desc.storage = 1, 9
    assert lib.olbgfx_lattice_create(ctypes.byref(desc), ctypes.byref(h)) == 1
    desc.storage = 0
    desc.nx = desc.ny = desc.nz = 1300
    assert lib.olbgfx_lattice_create(ctypes.byref(desc), ctypes.byref(h)) == 1
    assert b"2^31" in lib.olbgfx_last_error()
    desc.nx = desc.ny = desc.nz = 1100          # 1.34e9 cells: below 2^31, but 5.3 GB per FP32 ring
    assert lib.olbgfx_lattice_create(ctypes.byref(desc), ctypes.byref(h)) == 1
    assert b"4 GiB" in lib.olbgfx_last_error()
    desc.storage, desc.nx, desc.ny, desc.nz = 1, 1100, 1100, 1100   # FP16 rings of 2.7 GB fit
    desc.nx = 0
    assert lib.olbgfx_lattice_create(ctypes.byref(desc), ctypes.byref(h)) == 1
    assert b"positive" in lib.olbgfx_last_error()
    out = ctypes.c_void_p()
    assert lib.olbgfx_cell_list_create(None, None, 0, ctypes.byref(out)) == 1
    assert lib.olbgfx_device_buffer_create(16, None) == 1
    assert lib.olbgfx_autotune_lattice(None, 4, None, None, None, None) == 1
