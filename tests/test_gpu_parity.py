"""Parity of the gfx950 kernels (through the C ABI) with the CPU oracle on the
same inputs.

Tolerances (written per test):
  * integers (cell ids, neighbour distances, ring offsets, cell counts): exact;
  * FP32 storage: max relative error of the unshifted populations
    |f_gpu - f_oracle| / |f_oracle + t_i| <= 1e-5 after N steps;
  * FP64 storage: <= 1e-12;
  * FP16/BF16 storage: the GPU and the oracle round identically (checked bit
    for bit on the conversion); after several steps the FP32 arithmetic order
    differs, so the velocity fields are compared in relative L2.
"""
import numpy as np
import pytest

from openlb_amd import (BGK, BOUNCE_BACK, FIELD_FORCE, FORCED_BGK, Dynamics, Lattice, OlbgfxError,
                        cases)
from oracle.oracle import OracleLattice, f32_to_bf16_bits, f32_to_f16_bits
from tests.helpers import core_cells, rel_err_unshifted, rel_l2, velocity

pytestmark = pytest.mark.gpu


def _pair(case, storage, steps, oracle_storage=None, block=None):
    print(f"[parity] case={case.name} storage={storage} steps={steps}", flush=True)
    gpu = Lattice(case.extent, case.padding, case.periodic, storage)
    if block:
        gpu.set_launch(block)
    case.apply(gpu)
    orc = OracleLattice(case.extent, case.padding, case.periodic, oracle_storage or storage)
    case.apply(orc)
    gpu.collide_and_stream(steps)
    gpu.synchronize()
    orc.collide_and_stream(steps)
    return gpu, orc


def test_tgv_fp32_parity():
    c = cases.tgv(32, dtype=np.float32)
    gpu, orc = _pair(c, "f32", 100)
    core = core_cells(c)
    err = rel_err_unshifted(gpu.get_populations()[:, core], orc.get_populations()[:, core])
    print("tgv32 fp32 max rel err", err)
    assert err <= 1e-5


def test_tgv_fp64_parity():
    c = cases.tgv(24)
    gpu, orc = _pair(c, "f64", 60)
    core = core_cells(c)
    err = rel_err_unshifted(gpu.get_populations()[:, core], orc.get_populations()[:, core])
    print("tgv24 fp64 max rel err", err)
    assert err <= 1e-12


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-12)])
def test_cavity_parity(storage, tol):
    """BGK bulk, BounceBack walls, BounceBackVelocity lid; every cell of the
    padded block is compared (the padding holds identical stale values)."""
    c = cases.box_cavity(20, dtype=np.float64 if storage == "f64" else np.float32)
    gpu, orc = _pair(c, storage, 150)
    err = rel_err_unshifted(gpu.get_populations(), orc.get_populations())
    print("cavity20", storage, "max rel err", err)
    assert err <= tol


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-12)])
def test_porous_trt_guo_parity(storage, tol):
    c = cases.porous_channel((48, 24, 24), force=1e-4, dtype=np.float64 if storage == "f64" else np.float32)
    gpu, orc = _pair(c, storage, 120)
    core = core_cells(c)
    err = rel_err_unshifted(gpu.get_populations()[:, core], orc.get_populations()[:, core])
    print("porous", storage, "max rel err", err)
    assert err <= tol


def test_forced_bgk_with_force_field_parity():
    """per-cell FORCE field instead of the uniform one"""
    c = cases.tgv(16)
    c.dynamics = {1: Dynamics(FORCED_BGK, omega=1.5)}
    rng = np.random.default_rng(5)
    force = (1e-5 * rng.standard_normal((c.ncells, 3))).astype(np.float64)
    gpu = Lattice(c.extent, c.padding, c.periodic, "f64")
    orc = OracleLattice(c.extent, c.padding, c.periodic, "f64")
    for L in (gpu, orc):
        c.apply(L)
        L.set_cell_field(FIELD_FORCE, force)
        L.collide_and_stream(30)
    core = core_cells(c)
    err = rel_err_unshifted(gpu.get_populations()[:, core], orc.get_populations()[:, core])
    assert err <= 1e-12


def test_integers_bit_exact():
    c = cases.box_cavity(12)
    gpu, orc = _pair(c, "f32", 7)
    np.testing.assert_array_equal(gpu.neighbor_distances(), orc.neighbor_distances())
    np.testing.assert_array_equal(gpu.ring_shifts(), orc.ring_shifts())
    assert gpu.ncells == orc.ncells == int(np.prod(c.shape))
    for xyz in [(0, 0, 0), (11, 11, 11), (-3, -3, -3), (5, 2, 9)]:
        assert gpu.cell_id(*xyz) == orc.cell_id(*xyz)


def test_statistics_parity():
    c = cases.box_cavity(16, dtype=np.float32)
    gpu = Lattice(c.extent, c.padding, c.periodic, "f32")
    orc = OracleLattice(c.extent, c.padding, c.periodic, "f32")
    for L in (gpu, orc):
        c.apply(L)
        L.collide_and_stream(20)
        L.set_statistics(True)
        L.collide_and_stream(1)
    a, b = gpu.get_statistics(), orc.get_statistics()
    assert a["nCells"] == b["nCells"] == np.count_nonzero(np.isin(c.materials, (1, 3)))
    assert abs(a["sumRho"] - b["sumRho"]) / b["sumRho"] < 1e-6
    assert abs(a["sumUSqr"] - b["sumUSqr"]) / b["sumUSqr"] < 1e-4
    assert abs(a["maxUSqr"] - b["maxUSqr"]) / b["maxUSqr"] < 1e-5


def test_uniform_and_masked_paths_agree():
    """a TGV whose cells use two identical BGK slots runs the mask path; it
    must agree with the uniform path"""
    c = cases.tgv(20, dtype=np.float32)
    a = Lattice(c.extent, c.padding, c.periodic, "f32")
    c.apply(a)
    b = Lattice(c.extent, c.padding, c.periodic, "f32")
    c.apply(b)
    mask = c.dynamics_mask().copy()
    b.set_dynamics(2, Dynamics(BGK, omega=1.0 / 0.56))
    mask[b.cell_id(3, 4, 5)] = 2
    b.set_dynamics_mask(mask)
    a.collide_and_stream(40)
    b.collide_and_stream(40)
    core = core_cells(c)
    assert rel_err_unshifted(b.get_populations()[:, core], a.get_populations()[:, core]) <= 1e-6


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-12)])
def test_smagorinsky_uniform_and_masked_match_oracle(storage, tol):
    """SmagorinskyBGKdynamics (C = 0.1, tau = 0.505) on a TGV: the uniform
    kernel and the masked kernel (a second identical slot on one cell)
    against the oracle after 60 steps"""
    from openlb_amd import SMAGORINSKY_BGK

    dtype = np.float64 if storage == "f64" else np.float32
    c = cases.tgv(16, tau=0.505, dtype=dtype)
    dyn = Dynamics(SMAGORINSKY_BGK, omega=1 / 0.505, smagorinsky=0.1)
    c.dynamics[1] = dyn
    lats = []
    for cls in (Lattice, Lattice, OracleLattice):
        lat = cls(c.extent, c.padding, c.periodic, storage)
        c.apply(lat)
        lats.append(lat)
    mask = c.dynamics_mask().copy()
    lats[1].set_dynamics(2, dyn)
    mask[lats[1].cell_id(3, 4, 5)] = 2
    lats[1].set_dynamics_mask(mask)
    for lat in lats:
        lat.collide_and_stream(60)
    core = core_cells(c)
    ref = lats[2].get_populations()[:, core]
    for lat in lats[:2]:
        assert rel_err_unshifted(lat.get_populations()[:, core], ref) <= tol


@pytest.mark.parametrize("block", [64, 128, 192])
def test_launch_shapes_give_identical_results(block):
    c = cases.tgv(18, dtype=np.float32)
    ref = Lattice(c.extent, c.padding, c.periodic, "f32")
    c.apply(ref)
    ref.collide_and_stream(10)
    alt = Lattice(c.extent, c.padding, c.periodic, "f32")
    alt.set_launch(block)
    c.apply(alt)
    alt.collide_and_stream(10)
    np.testing.assert_array_equal(alt.get_populations(), ref.get_populations())


@pytest.mark.parametrize("order,block", [(0, 64), (2, 64), (4, 128), (6, 256), (12, 64)])
def test_block_orders_give_identical_results(order, block):
    """every block order visits each cell once: the masked cavity with
    statistics on gives the same populations and statistics"""
    c = cases.box_cavity(20, 0.1, 100.0, dtype=np.float32)
    out = []
    for o in (-1, order):
        lat = Lattice(c.extent, c.padding, c.periodic, "f32")
        lat.set_launch(block)
        lat.set_block_order(o)
        c.apply(lat)
        lat.set_statistics(True)
        lat.collide_and_stream(7)
        out.append((lat.get_populations(), lat.get_statistics()))
    np.testing.assert_array_equal(out[1][0], out[0][0])
    assert out[1][1]["nCells"] == out[0][1]["nCells"]
    assert out[1][1]["maxU"] == out[0][1]["maxU"]
    assert out[1][1]["averageRho"] == pytest.approx(out[0][1]["averageRho"], rel=1e-14)


def test_block_order_outside_range_is_rejected():
    lat = Lattice((8, 8, 8), 1, (True,) * 3, "f32")
    for bad in (-2, 13):
        with pytest.raises(OlbgfxError):
            lat.set_block_order(bad)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_storage_rounding_bit_exact(storage):
    """the device conversion to FP16 / BF16 rounds exactly as the oracle
    (round to nearest even, FP16 subnormals kept)"""
    L = Lattice((4, 4, 4), padding=1, storage=storage)
    rng = np.random.default_rng(9)
    vals = np.concatenate([
        rng.standard_normal(L.ncells * 19 - 8) * 0.05,
        [1e-7, -3e-6, 2.0 ** -24, 3 * 2.0 ** -26, 1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, 0.0, -0.0],
    ]).astype(np.float32).reshape(19, L.ncells)
    L.set_populations(vals)
    got = L.get_populations()
    conv = f32_to_f16_bits if storage == "f16" else f32_to_bf16_bits
    bits = conv(vals.ravel())
    expect = bits.view(np.float16).astype(np.float32) if storage == "f16" else \
        (bits.astype(np.uint32) << 16).view(np.float32)
    np.testing.assert_array_equal(got.ravel().view(np.uint32), expect.view(np.uint32))


@pytest.mark.parametrize("storage,tol", [("f16", 2e-3), ("bf16", 2e-2)])
def test_low_precision_storage_tracks_oracle(storage, tol):
    c = cases.tgv(24, dtype=np.float32)
    gpu, orc = _pair(c, storage, 30)
    core = core_cells(c)
    ug = velocity(gpu.get_populations()[:, core])
    uo = velocity(orc.get_populations()[:, core])
    d = rel_l2(ug, uo)
    print(storage, "velocity rel L2 gpu vs oracle", d)
    assert d <= tol


def test_fp64_moments_conserve_mass():
    """FP64 reductions of the stored state: BGK conserves mass and momentum;
    in FP32 storage the per-cell rounding leaves < 1e-7 per cell."""
    c = cases.tgv(48, dtype=np.float32)
    L = Lattice(c.extent, c.padding, c.periodic, "f32")
    c.apply(L)
    m0 = L.reduce_moments()
    L.collide_and_stream(100)
    m1 = L.reduce_moments()
    n = 48 ** 3
    assert m0["nCells"] == m1["nCells"] == n
    assert abs(m1["mass"] - m0["mass"]) / n < 1e-7
    assert max(abs(v) for v in m1["momentum"]) / n < 1e-8


def test_tgv_decay_rate_on_gpu():
    n, tau, steps = 64, 0.8, 200
    c = cases.tgv(n, u0=0.005, tau=tau, dtype=np.float32)
    L = Lattice(c.extent, c.padding, c.periodic, "f32")
    c.apply(L)
    e0 = L.reduce_moments()["sumUSqr"]
    L.collide_and_stream(steps)
    e1 = L.reduce_moments()["sumUSqr"]
    nu = (tau - 0.5) / 3.0
    rate = -np.log(e1 / e0) / steps
    expect = 2.0 * nu * 3.0 * (2 * np.pi / n) ** 2
    assert abs(rate - expect) / expect < 0.01


def test_launch_shape_outside_kernel_bounds_is_rejected():
    L = Lattice((4, 4, 4), padding=1)
    for bad in (32, 320, 512, 100):
        with pytest.raises(OlbgfxError):
            L.set_launch(bad)


def test_unsupported_dynamics_is_rejected():
    L = Lattice((4, 4, 4), padding=1)
    with pytest.raises(OlbgfxError) as e:
        L.set_dynamics(1, Dynamics(42))
    assert e.value.status == 2
    with pytest.raises(OlbgfxError):
        L.set_dynamics(16, Dynamics(BOUNCE_BACK))
