"""Set-up of runs from the golden fixtures of tests/golden/ (written by the
reference's own CPU path, see tests/golden/make_fixtures.py)."""
from __future__ import annotations

import json
import os

import numpy as np

from openlb_amd import (BGK, BOUNCE_BACK, BOUNCE_BACK_VELOCITY, FIELD_RHO, FIELD_VELOCITY, FIXED_VELOCITY,
                        FORCED_BGK, FORCED_TRT, INNER_CORNER_VELOCITY, INNER_EDGE_VELOCITY, LOCAL_PRESSURE,
                        LOCAL_VELOCITY, NO_DYNAMICS, SMAGORINSKY_BGK, Dynamics)
from openlb_amd.cases import (EXTERNAL_CORNER, EXTERNAL_EDGE, FLAT, INTERNAL_CORNER, INTERNAL_EDGE, LocalPressure,
                              LocalVelocity)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name: str):
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    meta = json.loads(str(d["meta"]))
    return meta, d


def _as_value(x: float, storage: str) -> float:
    """a parameter as the reference's value type holds it"""
    return float(np.float32(x)) if storage != "f64" else float(x)


def dynamics_for(meta: dict, storage: str) -> dict:
    """material -> Dynamics of the fixture's case, parameters rounded to the
    value type the way the reference driver forms them"""
    case = meta["case"]
    if case == "cavity_lv":
        d = dynamics_for(dict(meta, case="cavity"), storage)
        om = d[1].omega
        return {1: d[1], 2: LocalVelocity(omega=om),
                3: LocalVelocity(omega=om, wall_velocity=(_as_value(meta["ulid"], storage), 0.0, 0.0))}
    if case == "cavity":
        if storage == "f64":
            nu = 0.1 * meta["res"] / 100.0
            omega = 1.0 / (3.0 * nu + 0.5)
        else:
            f = np.float32
            nu = f(0.1 * meta["res"] / 100.0)
            omega = float(f(1) / (f(3) * nu + f(0.5)))
        return {1: Dynamics(BGK, omega=omega), 2: Dynamics(BOUNCE_BACK),
                3: Dynamics(BOUNCE_BACK_VELOCITY, wall_velocity=(_as_value(meta["ulid"], storage), 0, 0),
                            wall_rho=1.0)}
    if case == "tgv":
        if "smagorinsky" in meta:
            return {1: Dynamics(SMAGORINSKY_BGK, omega=_as_value(1.0 / meta["tau"], storage),
                                smagorinsky=_as_value(meta["smagorinsky"], storage))}
        return {1: Dynamics(BGK, omega=_as_value(1.0 / meta["tau"], storage))}
    if case == "porous":
        return {1: Dynamics(FORCED_TRT, omega=_as_value(1.0 / meta["tau"], storage),
                            magic=_as_value(meta["magic"], storage),
                            force=(_as_value(meta["fx"], storage), 0.0, 0.0)),
                2: Dynamics(BOUNCE_BACK)}
    if case == "poiseuille":
        forced = Dynamics(FORCED_BGK, omega=meta["omega"], force=(meta["force"], 0.0, 0.0))
        if meta.get("boundary") in ("local", "interpolated"):
            # testPoiseuille3DBGK.cpp:118-199: the bulk dynamics on materials 1-4,
            # then LocalVelocity (2, 3) and LocalPressure (4) on top, or
            # InterpolatedVelocity and InterpolatedPressure
            from openlb_amd.cases import InterpolatedPressure, InterpolatedVelocity

            bulk = forced if meta["forced"] else Dynamics(BGK, omega=meta["omega"])
            vel, pre = ((InterpolatedVelocity, InterpolatedPressure) if meta["boundary"] == "interpolated"
                        else (LocalVelocity, LocalPressure))
            if meta["forced"]:
                return {1: bulk, 2: vel(omega=meta["omega"], base=bulk)}
            return {1: bulk, 2: vel(omega=meta["omega"], base=bulk),
                    3: vel(omega=meta["omega"], base=bulk),
                    4: pre(omega=meta["omega"], base=bulk)}
        if meta.get("boundary") == "bounceBack" and not meta.get("forced", 1):
            # testPoiseuille3DBGK.cpp:137-139, 172-197: BounceBack wall,
            # InterpolatedVelocity inflow, InterpolatedPressure outflow
            from openlb_amd.cases import InterpolatedPressure, InterpolatedVelocity

            bulk = Dynamics(BGK, omega=meta["omega"])
            return {1: bulk, 2: Dynamics(BOUNCE_BACK), 3: InterpolatedVelocity(omega=meta["omega"], base=bulk),
                    4: InterpolatedPressure(omega=meta["omega"], base=bulk)}
        if meta.get("boundary") == "bouzidi" and not meta.get("forced", 1):
            # testPoiseuille3DBGK.cpp:141-151, 168-197: NoDynamics wall and
            # inflow (their Bouzidi post-processors come with the fixture's
            # lists), InterpolatedPressure on the outflow
            from openlb_amd.cases import InterpolatedPressure

            bulk = Dynamics(BGK, omega=meta["omega"])
            return {1: bulk, 2: Dynamics(NO_DYNAMICS), 3: Dynamics(NO_DYNAMICS),
                    4: InterpolatedPressure(omega=meta["omega"], base=bulk)}
        wall = Dynamics(NO_DYNAMICS) if meta.get("boundary") == "bouzidi" else Dynamics(BOUNCE_BACK)
        return {1: forced, 2: wall}
    raise ValueError(case)


def periodic_for(meta: dict):
    return {"cavity": (False,) * 3, "cavity_lv": (False,) * 3, "tgv": (True,) * 3, "porous": (True,) * 3,
            "poiseuille": (bool(meta.get("forced", 1)), False, False)}[meta["case"]]


def build(lattice_cls, meta, d, storage: str):
    """a lattice with the fixture's geometry, dynamics and t = 0 populations"""
    ext = (meta["nx"], meta["ny"], meta["nz"])
    lat = lattice_cls(ext, meta["padding"], periodic_for(meta), storage)
    mats = d["materials"].astype(np.uint8)
    dyns = dynamics_for(meta, storage)
    lat.set_dynamics(0, Dynamics(NO_DYNAMICS))
    lut = np.zeros(256, dtype=np.uint8)
    slot = 1
    lv = {}
    for mat, dyn in sorted(dyns.items()):
        if hasattr(dyn, "shapes"):
            # one slot per boundary shape; the material's base dynamics for
            # cells boundary::set gave no normal
            shapes = dict(dyn.shapes())
            extra = dyn.extra()
            by_type = {}
            kinds = {}
            for t, kind in shapes.items():
                if kind not in kinds:
                    lat.set_dynamics(slot, Dynamics(kind, omega=dyn.omega, **extra))
                    kinds[kind] = slot
                    slot += 1
                by_type[t] = kinds[kind]
            if dyn.base is not None:
                lat.set_dynamics(slot, dyn.base)
                lut[mat] = slot
                slot += 1
            lv[mat] = by_type
        else:
            lat.set_dynamics(slot, dyn)
            lut[mat] = slot
            slot += 1
    mask = lut[mats]
    if lv:
        # the reference's own discrete normals of the boundary cells
        types = d["normal_type"]
        has_normal = d["normals"].reshape(-1, 3).any(axis=1)
        for mat, by_type in lv.items():
            on = mats == mat
            for t, sl in by_type.items():
                mask[on & (types == t) & (has_normal | (t != FLAT))] = sl
        lat.set_cell_normals(d["normals"])
    lat.set_dynamics_mask(mask)
    if "u0" in d:
        # the velocity and density the boundary momenta hold (computeRhoU at t = 0)
        u = np.zeros((lat.ncells, 3))
        u[d["cells"]] = d["u0"]
        rho = np.ones(lat.ncells)
        rho[d["cells"]] = d["rho0"]
        lat.set_cell_field(FIELD_VELOCITY, u)
        lat.set_cell_field(FIELD_RHO, rho)
    if "field_cells" in d:
        # the reference's own BOUZIDI_DISTANCE field, solid-solid pairs resolved
        lat.set_bouzidi(*bouzidi_entries(meta, d))
    f = np.zeros((19, lat.ncells))
    f[:, d["cells"]] = d["f0"]
    lat.set_populations(f)
    return lat


def padded_shape(meta) -> tuple:
    p = meta["padding"]
    return (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)


def bouzidi_field(meta, d) -> np.ndarray:
    """the reference's BOUZIDI_DISTANCE field, 19 x padded cells (-1 unset)"""
    field = np.full((19, int(np.prod(padded_shape(meta)))), -1.0)
    field[:, d["field_cells"]] = d["field_q"].T
    return field


def bouzidi_entries(meta, d):
    from openlb_amd import bouzidi

    return bouzidi.entries(bouzidi_field(meta, d), meta["padding"], padded_shape(meta))


def poiseuille_wall(meta):
    """the Bouzidi wall indicator of testPoiseuille3DBGK.cpp:130-151: the pipe
    cylinder stretched by dx/2 at both ends, and by 3 dx more for the forced
    rows"""
    from openlb_amd import bouzidi

    dx, r, lx = meta["delta"], 0.5, 2.0
    c0, c1 = [0.0, r, r], [lx, r, r]
    c0[0] -= 0.5 * dx
    c1[0] += 0.5 * dx
    if meta.get("forced", 1):
        c0[0] -= 3.0 * dx
        c1[0] += 3.0 * dx
    return bouzidi.IndicatorCylinder3D(c0, c1, r)


def origin_of(meta) -> list:
    return [float(v) for v in str(meta["origin"]).split(",")]


def nonforced_bouzidi_lists(meta, d):
    """the two post-processor lists of the non-forced bouzidi rows computed
    by this package alone from the geometry: setBouzidiBoundary's distances
    for the wall (material 2, BouzidiPostProcessor) and the inflow disc
    (material 3, BouzidiVelocityPostProcessor), both against the pipe
    cylinder (testPoiseuille3DBGK.cpp:141-151, 168-171), and
    setBouzidiVelocity's coefficients for the inflow with the test's
    CirclePoiseuille3D(origin (lx, r, r), axis x, charLatticeVelocity, r)
    (:133-136, 240).  Returns (wall cells, wall q), (inflow cells, inflow q,
    inflow coefficients)."""
    from openlb_amd import bouzidi

    p = meta["padding"]
    mats = d["materials"].reshape(padded_shape(meta))
    origin, dx = origin_of(meta), meta["delta"]
    pipe = poiseuille_wall(meta)
    q_wall = bouzidi.distances(mats, p, origin, dx, pipe, 2)
    q_inflow = bouzidi.distances(mats, p, origin, dx, pipe, 3)
    q = np.where(q_inflow >= 0, q_inflow, q_wall)
    # charLatticeVelocity = charPhysVelocity (1) / the velocity conversion factor
    profile = bouzidi.CirclePoiseuille3D([2.0, 0.5, 0.5], [1.0, 0.0, 0.0], 1.0 / meta["u_factor"], 0.5)
    v = bouzidi.wall_velocities(q, mats, p, origin, dx, profile, 3)
    wall = bouzidi.processor_entries(q, mats, p, 2)
    cells, qq = bouzidi.processor_entries(q, mats, p, 3)
    return wall, (cells, qq, np.ascontiguousarray(v[:, cells].T))


def compared_cells(meta, d) -> np.ndarray:
    """columns of f1 that are compared: the fluid for the closed cavity (its
    walls also hold values streamed in from the padding, which the reference
    keeps in rings of a different length), every core cell otherwise"""
    if meta["case"] in ("cavity", "cavity_lv"):
        return np.flatnonzero(d["materials"][d["cells"]] == 1)
    return np.arange(d["cells"].size)


def value_tracer_run(lat, deltaT: int, residuum: float, max_steps: int):
    """util::ValueTracer convergence loop of the reference's Poiseuille test
    (utilities/benchmarkUtil.hh:78-128, testPoiseuille3DBGK.cpp:353-371)"""
    values, energy = [], []
    lat.set_statistics(True)
    it = 0
    for it in range(max_steps):
        if len(values) >= deltaT:
            avg = sum(values) / len(values)
            std = np.sqrt(sum((v - avg) ** 2 for v in values) / (len(values) - 1))
            if abs(std / avg) < residuum:
                break
        lat.collide_and_stream(1)
        e = lat.get_statistics()["averageEnergy"]
        energy.append(e)
        values.append(e)
        if len(values) > deltaT:
            values.pop(0)
    else:
        it = max_steps
    return it, np.array(energy)


def core_fluid(meta, d) -> np.ndarray:
    """positions (into d["fluid"]) of the fluid cells inside the core; the
    reference keeps periodic copies of them in the padding"""
    p = meta["padding"]
    shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
    xyz = np.array(np.unravel_index(d["fluid"], shape)).T - p
    inside = np.all((xyz >= 0) & (xyz < np.array([meta["nx"], meta["ny"], meta["nz"]])), axis=1)
    return np.flatnonzero(inside)


def velocity_error_l2(lat, meta, d) -> float:
    """relative L2 error of the physical velocity over the fluid, as the test
    computes it (testPoiseuille3DBGK.cpp:262-275): u includes the F/2 shift
    of the forced momenta"""
    from tests.helpers import C

    sel = core_fluid(meta, d)
    f = lat.get_populations()[:, d["fluid"][sel]].astype(np.float64)
    rho = f.sum(0) + 1.0
    u = (C.T @ f) / rho
    u[0] += 0.5 * meta["force"]
    u_phys = u.T * meta["u_factor"]
    sol = d["usol_fluid"][sel]
    return float(np.sqrt(np.sum((sol - u_phys) ** 2) / np.sum(sol ** 2)))


# testPoiseuille3DBGK.cpp:54-66: iT, velocity, pressure and strain-rate L2
# errors and the error factor of each row (the tolerance is factor * 1e-6)
KAT_TABLE = {
    "poiseuille11_forced_bounceback_f64": (422, 0.158042, 0.0, 0.168242, 1),
    "poiseuille11_forced_local_f64": (317, 0.020342, 0.0, 0.0165907, 1),
    "poiseuille11_forced_interpolated_f64": (303, 0.016909, 0.0, 0.016009, 1),
    "poiseuille11_forced_bouzidi_f64": (294, 0.019703, 0.0, 0.0232073, 1),
    "poiseuille11_nonforced_bounceback_f64": (621, 0.1284974, 0.40862686, 0.3209910, 1),
    "poiseuille11_nonforced_local_f64": (570, 0.05043437, 0.10581659, 0.05099981, 1),
    "poiseuille11_nonforced_interpolated_f64": (570, 0.04976679, 0.10483063, 0.04959827, 1),
    "poiseuille11_nonforced_bouzidi_f64": (572, 0.056544195, 0.070186678, 0.088491618, 1),
    "poiseuille21_nonforced_bouzidi_f64": (1400, 0.01624583, 0.024478267, 0.046465458, 1),
    "poiseuille41_nonforced_bouzidi_light_f64": (4077, 0.00445271, 0.00583554, 0.02117142, 2),
}


def strain_error_l2(lat, meta, d) -> float:
    """relative L2 error of the physical strain rate over the fluid, as the
    test computes it (testPoiseuille3DBGK.cpp:277-287) through
    SuperLatticePhysStrainRate3D (functors/lattice/latticePhysStrainRate3D.hh:
    S = -Pi omega c_s^-2 / (2 rho dt)): the momenta of the fluid's dynamics
    give rho, u (with F/2 for the forced ones, momenta::Forced keeps the
    bulk stress, dynamics/momenta/interface.h:288-296) and Pi^neq
    (lbm.h:334-352, computed with that u); the analytic strain rate is the
    reference's CirclePoiseuilleStrainRate3D at each cell (fixture)"""
    from tests.helpers import C

    sel = core_fluid(meta, d)
    f = lat.get_populations()[:, d["fluid"][sel]].astype(np.float64)
    c = np.asarray(C, dtype=np.float64)
    rho = f.sum(0) + 1.0
    u = (c.T @ f) / rho
    if meta.get("forced", 1 if meta["force"] != 0 else 0):
        u[0] += 0.5 * meta["force"]
    omega = 1.0 / meta["tau"]
    s = np.empty((9, f.shape[1]))
    pairs = {(0, 0): [0], (0, 1): [1, 3], (0, 2): [2, 6], (1, 1): [4], (1, 2): [5, 7], (2, 2): [8]}
    for (a, b), outs in pairs.items():
        pi = (c[:, a] * c[:, b]) @ f - rho * u[a] * u[b]
        if a == b:
            pi -= (rho - 1.0) / 3.0
        val = -pi * omega * 3.0 / rho / 2.0 / meta["dt"]
        for k in outs:
            s[k] = val
    sol = d["ssol_fluid"][sel].T
    return float(np.sqrt(np.sum((sol - s) ** 2) / np.sum(sol ** 2)))


def assert_strain(lat, meta, d, name: str, tol_checker: float = 1e-9, table: bool = True) -> float:
    """the strain-rate column of the known-answer test for one row: within
    tol_checker of the reference built here (the fixture's strainErrorL2)
    and, where table is set, within the row's errorFactor * 1e-6 of the
    test's table (testPoiseuille3DBGK.cpp:287)"""
    s = strain_error_l2(lat, meta, d)
    row = KAT_TABLE[name]
    print(f"{name}: strain-rate L2 error {s:.9g} (reference built here {meta['strainErrorL2']:.9g}, "
          f"table {row[3]})", flush=True)
    assert abs(s - meta["strainErrorL2"]) <= tol_checker
    if table:
        assert abs(s - row[3]) <= row[4] * 1e-6
    return s


def cavity_reduced(lat, meta, d):
    """mid-plane velocity, sample populations and fluid sums of a cavity run,
    in the layout of the cavity3d_res64 fixture"""
    from tests.helpers import C

    p = meta["padding"]
    shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
    f = lat.get_populations().astype(np.float64)
    rho = f.sum(0) + 1.0
    u = (C.T @ f / rho).T.reshape(shape + (3,))
    mats = d["materials"].reshape(shape)
    fluid = (mats == 1).reshape(-1)
    zc = p + meta["res"] // 2
    mid = u[p:-p, p:-p, zc, :]
    uf = u.reshape(-1, 3)[fluid]
    sums = np.array([(rho[fluid] - 1).sum(), *uf.sum(0), (uf ** 2).sum()])
    return mid, f[:, d["sample_cells"]], sums


def plane_cells(meta: dict) -> np.ndarray:
    """CellIDs of the mid-plane z = nz/2 over the core, every `stride`-th
    cell in x and y (CellID = ((x+p) Py + (y+p)) Pz + (z+p))"""
    p, s = meta["padding"], meta["stride"]
    py, pz = meta["ny"] + 2 * p, meta["nz"] + 2 * p
    xs = np.arange(0, meta["nx"], s) + p
    ys = np.arange(0, meta["ny"], s) + p
    zc = p + meta["nz"] // 2
    return ((xs[:, None] * py + ys[None, :]) * pz + zc).reshape(-1)


def core_fluid_mask(meta: dict, materials: np.ndarray) -> np.ndarray:
    """flat mask of the core cells of material 1 (the padding of a periodic
    block holds copies that the reference also marks as fluid)"""
    p = meta["padding"]
    shape = (meta["nx"] + 2 * p, meta["ny"] + 2 * p, meta["nz"] + 2 * p)
    m = np.zeros(shape, dtype=bool)
    m[p:-p, p:-p, p:-p] = materials.reshape(shape)[p:-p, p:-p, p:-p] == 1
    return m.reshape(-1)


def _rho_u(f: np.ndarray):
    from tests.helpers import C

    f = np.asarray(f, dtype=np.float64)
    rho = f.sum(0) + 1.0
    return rho, (C.T @ f) / rho


def reduce_state(f, meta: dict, fluid: np.ndarray, sample: np.ndarray, chunk: int = 1 << 22):
    """what a full-size fixture keeps of populations f (19 x ncells, shifted
    by -t_i, any array-like): the velocity on the strided mid-plane, the
    populations of the sample cells, and FP64 sums over the core fluid
    [sum(rho-1), sum ux, sum uy, sum uz, sum u^2], formed in chunks of cells"""
    plane = plane_cells(meta)
    _, up = _rho_u(f[:, plane])
    nxs = len(range(0, meta["nx"], meta["stride"]))
    up = up.T.reshape(nxs, -1, 3)
    sums = np.zeros(5)
    n = f.shape[1]
    for a in range(0, n, chunk):
        sel = np.flatnonzero(fluid[a:a + chunk]) + a
        if sel.size == 0:
            continue
        rho, u = _rho_u(f[:, a:a + chunk][:, sel - a])
        sums += [np.sum(rho - 1.0), *u.sum(1), np.sum(u * u)]
    return up, np.asarray(f[:, sample], dtype=np.float64), sums


def local_velocity_boundary_errors(f: np.ndarray, d) -> tuple:
    """max relative errors of the unshifted populations f (19 x fixture
    cells) on the LocalVelocity boundary cells against the fixture: flat
    faces on the populations streamed in from the fluid side (those entering
    from outside are never read: the regularisation rebuilds them), edges and
    corners on all 19 (their post-processor rewrites every one)"""
    from tests.helpers import C, T

    f = np.asarray(f, dtype=np.float64)
    ref = d["f1"].astype(np.float64)
    types = d["normal_type"][d["cells"]]
    normals = d["normals"].reshape(-1, 3)[d["cells"]]
    flat_err = edge_err = 0.0
    for k in np.flatnonzero(types >= 0):
        n = normals[k]
        if types[k] == 0:
            a = int(np.flatnonzero(n)[0])
            rows = [i for i in range(19) if C[i][a] != -n[a]]
        else:
            rows = list(range(19))
        e = float(np.max(np.abs(f[rows, k] - ref[rows, k]) / np.abs(ref[rows, k] + T[rows])))
        if types[k] == 0:
            flat_err = max(flat_err, e)
        else:
            edge_err = max(edge_err, e)
    return flat_err, edge_err


def step_cavity_oracle(storage: str):
    """the step cavity of the fixture built as a Case (its discrete normals
    from the classifier, not from the fixture) and advanced by the oracle,
    with the mask of populations that do not depend on what the padding and
    the outside (material 0) cells hold: two more oracle runs refill those
    cells with different random values before every step, and a population
    is defined where the two agree.  Returns (case, fixture, core cells,
    oracle populations of the core cells, defined mask)."""
    from oracle.oracle import OracleLattice
    from openlb_amd import cases

    meta, d = load("step_cavity_res10_f64")
    dtype = np.float64 if storage == "f64" else np.float32
    c = cases.step_cavity(meta["res"], u_lid=meta["u_lid"], re=meta["re"], step=meta["step"], dtype=dtype)
    cells = d["cells"]
    garbage = ~c.core_mask().reshape(-1) | (c.materials.reshape(-1) == 0)
    runs = []
    for seed in (None, 1, 2):
        o = OracleLattice(c.extent, c.padding, c.periodic, storage)
        c.apply(o)
        if seed is None:
            o.collide_and_stream(meta["steps"])
        else:
            rng = np.random.default_rng(seed)
            for _ in range(meta["steps"]):
                f = o.get_populations()
                f[:, garbage] = rng.uniform(-0.01, 0.01, size=(19, int(garbage.sum())))
                o.set_populations(f)
                o.collide_and_stream(1)
        runs.append(np.asarray(o.get_populations()[:, cells], dtype=np.float64))
    return c, d, cells, runs[0], runs[1] == runs[2]


def nonforced_poiseuille_state(meta, materials) -> dict:
    """The t = 0 state and the analytic solutions of the non-forced bouzidi
    rows of testPoiseuille3DBGK.cpp, restated from the geometry and the unit
    converter (lx = 2, r = 1/2, Re = 10, charPhysVelocity 1, physViscosity
    1/Re):
      - prepareLattice (:218-241): rho = AnalyticalLinear3D(-p0/lx * invCs2,
        0, 0, p0 * invCs2 + 1) with p0 the lattice pressure of 4 nu u lx / r^2,
        u = CirclePoiseuille3D((lx, r, r), x, charLatticeVelocity, r) on
        materials 1-4 and 0 on material 0; iniEquilibrium sets the shifted
        second-order equilibrium on the BGK-based cells (materials 1 and 4)
        and leaves the NoDynamics cells (0, 2, 3: the outside, the Bouzidi
        wall and inflow) at f_bar = 0, whose equilibrium::None returns 0;
        the RHO and VELOCITY fields carry rho and u (defineRhoU);
      - error (:253-298) on the fluid cells: the physical Poiseuille velocity
        (maxVelocity 1), the pressure 4 nu u lx / r^2 (1 - x / lx) and
        CirclePoiseuilleStrainRate3D (functors/analytical/frameChangeF3D.hh:
        406-436).
    BlockLattice::initialize then applies the PostStream processors once
    (core/blockLattice.hh:56-67); the caller does that with post_process().
    Returns f0 (19 x cells, shifted), rho, u, fluid, usol_fluid, psol_fluid,
    ssol_fluid."""
    from openlb_amd import bouzidi
    from tests.helpers import C, T

    p = meta["padding"]
    shape = padded_shape(meta)
    n = int(np.prod(shape))
    mats = np.asarray(materials).reshape(-1)
    origin, dx = np.asarray(origin_of(meta)), meta["delta"]
    xyz = np.array(np.unravel_index(np.arange(n), shape)).T - p
    phys = origin[None, :] + xyz * dx
    lx, r, nu = 2.0, 0.5, 0.1
    u = bouzidi.CirclePoiseuille3D([lx, r, r], [1.0, 0.0, 0.0], 1.0 / meta["u_factor"], r)(phys)
    u[mats == 0] = 0.0
    p0 = 4.0 * nu * 1.0 * lx / (r * r) / meta["p_factor"]
    rho = (-p0 / lx * 3.0) * phys[:, 0] + (p0 * 3.0 + 1.0)
    c = np.asarray(C, dtype=np.float64)
    cu = u @ c.T
    usq = np.sum(u * u, axis=1)
    f0 = (rho[:, None] * T[None, :] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq[:, None]) - T[None, :]).T
    f0[:, ~np.isin(mats, (1, 4))] = 0.0
    fluid = np.flatnonzero(mats == 1)
    pf = phys[fluid]
    usol = bouzidi.CirclePoiseuille3D([lx, r, r], [1.0, 0.0, 0.0], 1.0, r)(pf)
    p0phys = 4.0 * nu * 1.0 * lx / (r * r)
    psol = (-p0phys / lx) * pf[:, 0] + p0phys
    ssol = np.zeros((fluid.size, 9))
    dudy = 1.0 * (-2.0 * (pf[:, 1] - r) / (r * r))
    dudz = 1.0 * (-2.0 * (pf[:, 2] - r) / (r * r))
    ssol[:, 1] = ssol[:, 3] = dudy / 2.0
    ssol[:, 2] = ssol[:, 6] = dudz / 2.0
    return {"f0": f0, "rho": rho, "u": u, "fluid": fluid, "usol_fluid": usol, "psol_fluid": psol,
            "ssol_fluid": ssol}


def build_nonforced_bouzidi(lattice_cls, meta, d, storage: str):
    """a non-forced bouzidi row set up from the geometry alone: the
    materials and discrete normals of `d`, the t = 0 state of
    nonforced_poiseuille_state, the Bouzidi lists of nonforced_bouzidi_lists,
    then the PostStream processors once, as BlockLattice::initialize does.
    Returns the lattice and the analytic solutions (a dict shaped like a
    fixture for the error norms)"""
    st = nonforced_poiseuille_state(meta, d["materials"])
    n = st["f0"].shape[1]
    start = {"materials": d["materials"], "normal_type": d["normal_type"], "normals": d["normals"],
             "cells": np.arange(n, dtype=np.int64), "f0": st["f0"], "u0": st["u"], "rho0": st["rho"]}
    lat = build(lattice_cls, meta, start, storage)
    (bc, bq), (vc, vq, vv) = nonforced_bouzidi_lists(meta, d)
    lat.set_bouzidi(bc, bq)
    lat.set_bouzidi_velocity(vc, vq, vv)
    lat.post_process()
    sol = {"materials": d["materials"], "fluid": st["fluid"], "usol_fluid": st["usol_fluid"],
           "psol_fluid": st["psol_fluid"], "ssol_fluid": st["ssol_fluid"]}
    return lat, sol
