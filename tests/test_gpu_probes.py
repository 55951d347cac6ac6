"""The bandwidth probes (include/olbgfx_probe.h) that bench.py reports as the
ceilings of the step kernel's access patterns: every pattern runs and gives a
positive rate on a small lattice in each storage, and an unknown pattern is
refused by the library."""

import ctypes

import pytest

import openlb_amd as ol

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["f32", "f16", "f64"])
@pytest.mark.parametrize("pattern", sorted(ol.PROBES))
def test_probe_runs(storage, pattern):
    gbps = ol.probe_bandwidth((24, 20, 36), 1, storage, pattern, 1, 3, 2)
    assert gbps > 0.0


def test_probe_rejects_unknown_pattern():
    d = ol._LatticeDesc()
    d.nx, d.ny, d.nz = 8, 8, 8
    d.padding = 1
    d.storage = ol.STORE["f32"]
    out = ctypes.c_double()
    st = ol.library().olbgfx_probe_bandwidth(ctypes.byref(d), max(ol.PROBES.values()) + 1, 1, -1, 1, None,
                                             ctypes.byref(out))
    assert st != 0
