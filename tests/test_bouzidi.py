"""Bouzidi interpolated bounce-back on the CPU: the setter's distance field
against the reference's own, the ray distance against the closed-form
ray-cylinder intersection, the concurrent entry list against the reference's
sequential order, and the oracle against the reference's known-answer test
(testPoiseuille3DBGK.cpp:57, forced, bouzidi, N = 11: iT = 294, velocity L2
error 0.019703 +- 1e-6)."""
import math

import numpy as np
import pytest

from openlb_amd import bouzidi
from oracle.oracle import OracleLattice
from tests.golden_runs import (bouzidi_entries, bouzidi_field, build, core_fluid, load, nonforced_bouzidi_lists,
                               padded_shape, poiseuille_wall, value_tracer_run, velocity_error_l2, assert_strain)
from tests.helpers import rel_err_unshifted

NAME = "poiseuille11_forced_bouzidi_f64"


def _origin(meta):
    return [float(v) for v in str(meta["origin"]).split(",")]


def test_distance_field_matches_reference():
    """every BOUZIDI_DISTANCE entry the reference's setter wrote, from the
    geometry and the wall cylinder alone: same set of entries, values equal
    to the last bit or two (the reference contracts a*b+c into FMAs)"""
    meta, d = load(NAME)
    mats = d["materials"].reshape(padded_shape(meta))
    q = bouzidi.distances(mats, meta["padding"], _origin(meta), meta["delta"], poiseuille_wall(meta), 2)
    ref = bouzidi_field(meta, d)
    np.testing.assert_array_equal(q >= 0, ref >= 0)
    assert np.max(np.abs(q - ref)) <= 1e-15
    # the fall-backs and the interpolated entries are both present
    on = ref >= 0
    assert np.sum(ref[on] == 0.5) > 0 and np.sum((ref[on] > 0) & (ref[on] != 0.5)) > 0


def _ray_cylinder(p, v, c, r):
    """closed form: smallest t > 0 with |(p + t v - c)_yz| = r for a cylinder
    along x through c (yz components only)"""
    py, pz = p[1] - c[1], p[2] - c[2]
    vy, vz = v[1], v[2]
    a = vy * vy + vz * vz
    b = 2 * (py * vy + pz * vz)
    cc = py * py + pz * pz - r * r
    disc = b * b - 4 * a * cc
    if a == 0 or disc < 0:
        return None
    ts = sorted(((-b - math.sqrt(disc)) / (2 * a), (-b + math.sqrt(disc)) / (2 * a)))
    ts = [t for t in ts if t > 0]
    return ts[0] if ts else None


def test_ray_distance_against_closed_form():
    """the bisection of util::distance (precision 1e-4, pitch 1/2 in units of
    the direction) finds the ray-cylinder intersection of the mathematics to
    within its precision, from points inside the cylinder (where the setter
    starts: the fluid cells) up to two lattice links from the wall"""
    rng = np.random.default_rng(7)
    cyl = bouzidi.IndicatorCylinder3D([-1.0, 0.5, 0.5], [3.0, 0.5, 0.5], 0.5)
    dx = 0.05
    checked = 0
    for _ in range(3000):
        p = [rng.uniform(0, 2), *rng.uniform(0.0, 1.0, 2)]
        i = int(rng.integers(1, 19))
        v = [-dx * float(c) for c in bouzidi.C[i]]
        if v[1] == 0 and v[2] == 0:
            continue
        if not cyl.inside(p):
            continue
        exact = _ray_cylinder(p, v, [0, 0.5, 0.5], 0.5)
        got = bouzidi.ray_distance(cyl, p, v)
        if exact is None or exact > 2:
            continue
        norm = math.sqrt(sum(c * c for c in v))
        assert got is not None
        assert abs(got / norm - exact) <= 2e-4, (p, i, got / norm, exact)
        checked += 1
    assert checked > 100


def test_entries_reproduce_the_sequential_order():
    """the concurrent entry list (solid-solid q = 1/2 pairs resolved in
    ascending cell id) gives, applied one cell after another, exactly what the
    reference's full list gives in its ascending order, on every cell"""
    meta, d = load(NAME)
    cells, q = bouzidi_entries(meta, d)
    full_cells, full_q = d["bz_cells"], d["bz_q"]
    assert cells.size == full_cells.size and (q >= 0).sum() < (full_q >= 0).sum()
    runs = []
    for cl, qq in ((full_cells, full_q), (cells, q)):
        lat = build(OracleLattice, meta, d, "f64")
        lat.set_bouzidi(cl, qq)
        lat.collide_and_stream(60)
        runs.append(lat.get_populations())
    np.testing.assert_array_equal(runs[0], runs[1])


def test_entries_reject_mutual_reads():
    """two fluid cells whose interpolations read each other's outputs cannot
    be updated concurrently"""
    shape = (6, 6, 6)
    q = np.full((19, 216), -1.0)
    a = np.ravel_multi_index((2, 2, 2), shape)
    b = np.ravel_multi_index((3, 2, 2), shape)  # a + c_10
    q[10, a] = 0.3  # a writes a[1], reads b[10]
    q[1, b] = 0.3   # b writes b[10], reads a[1]
    with pytest.raises(ValueError):
        bouzidi.entries(q, 1, shape)


def test_oracle_poiseuille_bouzidi_known_answer():
    meta, d = load(NAME)
    lat = build(OracleLattice, meta, d, "f64")
    it, energy = value_tracer_run(lat, meta["deltaT"], meta["residuum"], meta["max_steps"])
    assert it == 294 == meta["iT"]
    np.testing.assert_allclose(energy, d["energy"], rtol=1e-10)
    err = velocity_error_l2(lat, meta, d)
    assert_strain(lat, meta, d, NAME)
    print("oracle bouzidi velocity error", err)
    assert abs(err - 0.019703) <= 1e-6
    assert abs(err - meta["velocityErrorL2"]) <= 1e-9
    sel = core_fluid(meta, d)
    f1 = lat.get_populations()[:, d["fluid"][sel]]
    assert rel_err_unshifted(f1, d["f1_fluid"][:, sel]) <= 1e-12


def test_oracle_poiseuille_bouzidi_with_own_distances():
    """the same run from the distances this package computes"""
    meta, d = load(NAME)
    lat = build(OracleLattice, meta, d, "f64")
    mats = d["materials"].reshape(padded_shape(meta))
    q = bouzidi.distances(mats, meta["padding"], _origin(meta), meta["delta"], poiseuille_wall(meta), 2)
    lat.set_bouzidi(*bouzidi.entries(q, meta["padding"], padded_shape(meta)))
    it, _ = value_tracer_run(lat, meta["deltaT"], meta["residuum"], meta["max_steps"])
    assert it == 294
    assert abs(velocity_error_l2(lat, meta, d) - 0.019703) <= 1e-6


def test_circle_poiseuille_profile_closed_form():
    """CirclePoiseuille3D along x: u_x = u_max (1 - r^2 / R^2) in the yz
    distance from the axis, u_max on the axis, 0 on the wall, negative
    outside, the other components 0; the profile along y or z rotates with
    the normal"""
    prof = bouzidi.CirclePoiseuille3D([2.0, 0.5, 0.5], [2.0, 0.0, 0.0], 0.04, 0.5)
    pts = np.array([[0.3, 0.5, 0.5], [1.7, 1.0, 0.5], [0.0, 0.5, 0.0], [5.0, 0.5 + 0.3, 0.5 + 0.4],
                    [1.0, 0.5 + 0.6, 0.5], [1.0, 0.5 + 0.1, 0.5 - 0.2]])
    u = prof(pts)
    r2 = (pts[:, 1] - 0.5) ** 2 + (pts[:, 2] - 0.5) ** 2
    np.testing.assert_allclose(u[:, 0], 0.04 * (1 - r2 / 0.25), rtol=1e-14, atol=1e-16)
    assert u[0, 0] == 0.04 and abs(u[1, 0]) < 1e-16 and abs(u[3, 0]) < 1e-16 and u[4, 0] < 0
    assert np.all(u[:, 1:] == 0)
    along_y = bouzidi.CirclePoiseuille3D([0.5, 0.0, 0.5], [0.0, -1.0, 0.0], 0.04, 0.5)
    assert along_y(np.array([[0.5, 3.0, 0.5]]))[0, 1] == -0.04


@pytest.mark.parametrize("name", ["poiseuille11_nonforced_bouzidi_f64", "poiseuille21_nonforced_bouzidi_f64"])
def test_nonforced_bouzidi_lists_match_reference(name):
    """the two post-processor lists of the non-forced bouzidi rows from this
    package alone (setBouzidiBoundary's distances for the wall and the inflow
    disc, setBouzidiVelocity's coefficients with the test's Poiseuille
    profile) against the ones built from the reference's own
    BOUZIDI_DISTANCE and BOUZIDI_VELOCITY fields: the same cells, distances
    within a few ulps (the reference contracts a*b+c into FMAs) and wall
    coefficients within 4 ulps of their 0.05-0.09 magnitude.  A coefficient
    taken along c_i instead of c_opp(i), at x_b instead of the wall point,
    or from the physical instead of the lattice velocity fails here."""
    meta, d = load(name)
    (bc, bq), (vc, vq, vv) = nonforced_bouzidi_lists(meta, d)
    np.testing.assert_array_equal(bc, d["bz_cells"])
    np.testing.assert_array_equal(vc, d["bzv_cells"])
    assert np.max(np.abs(bq - d["bz_q"])) <= 5e-15
    assert np.max(np.abs(vq - d["bzv_q"])) <= 5e-15
    assert np.max(np.abs(vv - d["bzv_v"])) <= 4e-17
    assert (d["bzv_v"] != 0).sum() > 400 and np.all((vv == 0) == (d["bzv_v"] == 0))


def test_oracle_poiseuille21_nonforced_bouzidi_known_answer():
    """testPoiseuille3DBGK.cpp:64 (non-forced, bouzidi, N = 21) from the
    lists this package computes: the test's iT = 1400; velocity, pressure
    and strain-rate L2 errors within 1e-9 of the reference built here and
    within 1e-6 of the table's 0.01624583, 0.024478267 and 0.046465458;
    energies within 1e-10 and fluid populations within 1e-12 of the
    reference's run"""
    from tests.test_golden import _pressure_error_l2

    name = "poiseuille21_nonforced_bouzidi_f64"
    meta, d = load(name)
    (bc, bq), (vc, vq, vv) = nonforced_bouzidi_lists(meta, d)
    lat = build(OracleLattice, meta, d, "f64")
    lat.set_bouzidi(bc, bq)
    lat.set_bouzidi_velocity(vc, vq, vv)
    it, energy = value_tracer_run(lat, meta["deltaT"], meta["residuum"], meta["max_steps"])
    assert it == meta["iT"] == 1400
    np.testing.assert_allclose(energy, d["energy"], rtol=1e-10)
    err = velocity_error_l2(lat, meta, d)
    perr = _pressure_error_l2(lat, meta, d)
    assert_strain(lat, meta, d, name)
    print("N = 21 velocity", err, "pressure", perr)
    assert abs(err - meta["velocityErrorL2"]) <= 1e-9 and abs(perr - meta["pressureErrorL2"]) <= 1e-9
    assert abs(err - 0.01624583) <= 1e-6 and abs(perr - 0.024478267) <= 1e-6
    sel = core_fluid(meta, d)
    assert rel_err_unshifted(lat.get_populations()[:, d["fluid"][sel]], d["f1_fluid"][:, sel]) <= 1e-12


@pytest.mark.parametrize("name", ["poiseuille11_nonforced_bouzidi_f64", "poiseuille21_nonforced_bouzidi_f64"])
def test_restated_poiseuille_start_matches_reference(name):
    """the restated t = 0 state of the non-forced rows (prepareLattice's
    rho and u, equilibrium on the BGK-based cells, NoDynamics cells at rest)
    and the analytic solutions of error() against the reference's own:
    fluid, inflow and outflow populations after initialize()'s PostStream
    pass within 1e-15 (the Bouzidi wall's own cells differ where the
    concurrent list resolves mutual q = 1/2 pairs, which never reach the
    fluid), velocity, pressure and strain-rate solutions within 5e-16"""
    from tests.golden_runs import nonforced_poiseuille_state

    meta, d = load(name)
    st = nonforced_poiseuille_state(meta, d["materials"])
    np.testing.assert_array_equal(st["fluid"], d["fluid"])
    for k in ("usol_fluid", "psol_fluid", "ssol_fluid"):
        assert np.max(np.abs(st[k] - d[k])) <= 5e-16, k
    lat = build(OracleLattice, meta, {"materials": d["materials"], "normal_type": d["normal_type"],
                                      "normals": d["normals"], "cells": np.arange(st["f0"].shape[1]),
                                      "f0": st["f0"], "u0": st["u"], "rho0": st["rho"]}, "f64")
    (bc, bq), (vc, vq, vv) = nonforced_bouzidi_lists(meta, d)
    lat.set_bouzidi(bc, bq)
    lat.set_bouzidi_velocity(vc, vq, vv)
    lat.post_process()
    f = lat.get_populations()[:, d["cells"]]
    mats = d["materials"][d["cells"]]
    on = np.isin(mats, (1, 3, 4))
    assert np.max(np.abs(f[:, on] - d["f0"][:, on])) <= 1e-15


@pytest.mark.parametrize("name", ["poiseuille11_nonforced_bouzidi_f64", "poiseuille21_nonforced_bouzidi_f64",
                                  "poiseuille41_nonforced_bouzidi_light_f64"])
def test_oracle_poiseuille_nonforced_bouzidi_from_geometry(name):
    """testPoiseuille3DBGK.cpp:63-65, the non-forced bouzidi rows at N = 11,
    21 and 41, set up from the geometry alone (restated start, Bouzidi lists
    and analytic solutions): the table's iT (572, 1400, 4077), energies
    within 1e-10 of the reference built here, velocity, pressure and
    strain-rate L2 errors within 1e-9 of it and within errorFactor x 1e-6 of
    the table (errorFactor 2 at N = 41)"""
    from tests.golden_runs import KAT_TABLE, build_nonforced_bouzidi, strain_error_l2
    from tests.test_golden import _pressure_error_l2

    meta, d = load(name)
    row = KAT_TABLE[name]
    lat, sol = build_nonforced_bouzidi(OracleLattice, meta, d, "f64")
    it, energy = value_tracer_run(lat, meta["deltaT"], meta["residuum"], meta["max_steps"])
    assert it == meta["iT"] == row[0]
    np.testing.assert_allclose(energy, d["energy"], rtol=1e-10)
    errs = (velocity_error_l2(lat, meta, sol), _pressure_error_l2(lat, meta, sol), strain_error_l2(lat, meta, sol))
    print(name, "errors", errs)
    for got, ref, table in zip(errs, (meta["velocityErrorL2"], meta["pressureErrorL2"], meta["strainErrorL2"]),
                               (row[1], row[2], row[3])):
        assert abs(got - ref) <= 1e-9
        assert abs(got - table) <= row[4] * 1e-6
