"""Case set-up for the D3Q19 workloads: geometry (material numbers per cell),
dynamics per material and initial populations.

A ``Case`` is plain data; ``Case.apply(lattice)`` performs the calls an OpenLB
``prepareLattice`` / ``setInitialValues`` makes (``dynamics::set`` per
material, ``setParameter<OMEGA>``, ``defineField``, ``iniEquilibrium``) on
anything with the block-lattice methods of :class:`openlb_amd.Lattice`.

All per-cell arrays are over the padded block in CellID order, i.e. C-order
arrays of shape (nx+2p, ny+2p, nz+2p) (core/blockStructure.h:150-160).
Material 0 marks cells outside the domain (padding), as in
SuperGeometry (geometry/blockGeometry.h:59).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import (BGK, BOUNCE_BACK, BOUNCE_BACK_VELOCITY, FIELD_VELOCITY, FIXED_VELOCITY, FORCED_TRT,
               INNER_CORNER_VELOCITY, INNER_EDGE_VELOCITY, INTERPOLATED_PRESSURE, INTERPOLATED_VELOCITY,
               LOCAL_PRESSURE, LOCAL_VELOCITY, NO_DYNAMICS, Dynamics)
from .normals import EXTERNAL_CORNER, EXTERNAL_EDGE, FLAT, INTERNAL_CORNER, INTERNAL_EDGE, discrete_normals

__all__ = [
    "Case",
    "relaxation_frequency",
    "tgv",
    "box_cavity",
    "porous_channel",
    "sphere_pack",
]


def relaxation_frequency(nu: float) -> float:
    """omega = 1 / (3 nu + 1/2) (core/unitConverter.h:882 with cs^2 = 1/3)."""
    return 1.0 / (3.0 * nu + 0.5)


@dataclass
class Case:
    name: str
    extent: tuple
    padding: int
    periodic: tuple
    materials: np.ndarray                       # uint8, padded shape
    dynamics: dict = field(default_factory=dict)  # material -> Dynamics
    rho: np.ndarray | None = None               # padded, initial density
    u: np.ndarray | None = None                 # padded x 3, initial velocity
    init_materials: tuple = ()                  # materials that get iniEquilibrium
    velocity_field: np.ndarray | None = None    # optional per-cell VELOCITY
    info: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.materials.shape

    @property
    def ncells(self) -> int:
        return int(self.materials.size)

    def n_voxels(self) -> int:
        """cells with material != 0 (SuperGeometryStatistics::getNvoxel,
        geometry/superGeometryStatistics3D.hh:319-328)"""
        return int(np.count_nonzero(self.materials))

    def dynamics_slots(self):
        """material -> dynamics table slot (slot 0 stays NoDynamics).  A
        LocalVelocity or LocalPressure material takes one slot per boundary
        shape its cells have, keyed (material, discrete normal type), and one
        for its base dynamics (keyed by the material) if it has one."""
        slots = {}
        table = [Dynamics(NO_DYNAMICS)]
        normals = self.cell_normals()
        for mat in sorted(self.dynamics):
            d = self.dynamics[mat]
            if isinstance(d, _BOUNDARIES):
                on = self.materials == mat
                types = np.where(normals[1][on].any(axis=-1) | (normals[0][on] != FLAT), normals[0][on], -1)
                shapes, extra = d.shapes(), d.extra()
                by_kind = {}
                for t, kind in shapes:
                    if not np.any(types == t):
                        continue
                    if kind not in by_kind:
                        table.append(Dynamics(kind, omega=d.omega, **extra))
                        by_kind[kind] = len(table) - 1
                    slots[(mat, t)] = by_kind[kind]
                if d.base is not None:
                    table.append(d.base)
                    slots[mat] = len(table) - 1
            else:
                table.append(d)
                slots[mat] = len(table) - 1
        if len(table) > 16:
            raise ValueError(f"{len(table)} dynamics slots needed; the lattice holds 16")
        return slots, table

    def local_velocity_materials(self) -> tuple:
        """the materials with boundaries set through discrete normals
        (LocalVelocity, LocalPressure, InterpolatedVelocity,
        InterpolatedPressure)"""
        return tuple(m for m, d in sorted(self.dynamics.items()) if isinstance(d, _BOUNDARIES))

    def cell_normals(self):
        """(type, normal) per cell for the LocalVelocity materials, or None"""
        lv = self.local_velocity_materials()
        if not lv:
            return None
        if "_normals" not in self.info:
            self.info["_normals"] = discrete_normals(self.materials, lv)
        return self.info["_normals"]

    def dynamics_mask(self) -> np.ndarray:
        slots, _ = self.dynamics_slots()
        lut = np.zeros(256, dtype=np.uint8)
        for mat, slot in slots.items():
            if not isinstance(mat, tuple):
                lut[mat] = slot
        mask = lut[self.materials.reshape(-1)]
        normals = self.cell_normals()
        if normals is not None:
            # boundary::set (boundary/setBoundary.h:76-100): a classified cell
            # gets the dynamics of its shape; a cell with no discrete normal
            # (FLAT with the zero normal) keeps the material's base dynamics
            # under LocalVelocity (boundary/localVelocity3D.h:47-82) and is an
            # error under LocalPressure (boundary/localPressure3D.h:47-58)
            types = normals[0].reshape(-1)
            unset = ~normals[1].reshape(-1, 3).any(axis=1)
            mats = self.materials.reshape(-1)
            for m in self.local_velocity_materials():
                on = (mats == m) & (types >= 0)
                if isinstance(self.dynamics[m], LocalPressure) and np.any(on & ((types != FLAT) | unset)):
                    raise ValueError("No valid discrete normal found. This BC is not suited for curved walls.")
                for t in (FLAT, EXTERNAL_EDGE, EXTERNAL_CORNER, INTERNAL_EDGE, INTERNAL_CORNER):
                    if (m, t) in slots:
                        sel = on & (types == t)
                        if t == FLAT:
                            sel &= ~unset
                        mask[sel] = slots[(m, t)]
        return mask

    def apply(self, lattice) -> None:
        slots, table = self.dynamics_slots()
        for slot, dyn in enumerate(table):
            lattice.set_dynamics(slot, dyn)
        lattice.set_dynamics_mask(self.dynamics_mask())
        normals = self.cell_normals()
        if normals is not None:
            lattice.set_cell_normals(normals[1])
        if self.velocity_field is not None:
            lattice.set_cell_field(FIELD_VELOCITY, self.velocity_field)
        if self.rho is not None:
            sel = np.isin(self.materials.reshape(-1), self.init_materials).astype(np.uint8)
            lattice.ini_equilibrium(self.rho.reshape(-1), self.u.reshape(-1, 3), sel)
        if normals is not None:
            # SuperLattice::initialize runs the PostStream processors once
            # (core/blockLattice.hh:56-67)
            lattice.post_process()

    def core_mask(self) -> np.ndarray:
        p = self.padding
        m = np.zeros(self.shape, dtype=bool)
        m[p:-p, p:-p, p:-p] = True
        return m


@dataclass
class LocalVelocity:
    """boundary::set<boundary::LocalVelocity<T, D, BGKdynamics>> on a material
    (boundary/localVelocity3D.h:36-120): flat faces collide as
    CombinedRLBdynamics with RegularizedVelocityBoundaryTuple, external edges
    and corners as BGK with FixedVelocityBoundaryTuple followed by the outer
    edge and corner post-processors.  omega of the BGK part; the wall
    velocity is the VELOCITY the case defines on the material."""

    omega: float
    wall_velocity: tuple = (0.0, 0.0, 0.0)
    # dynamics of the material's cells that get no boundary dynamics (no
    # discrete normal; boundary::set leaves what defineDynamics gave them)
    base: object = None

    def shapes(self):
        return ((FLAT, LOCAL_VELOCITY), (EXTERNAL_EDGE, FIXED_VELOCITY), (EXTERNAL_CORNER, FIXED_VELOCITY),
                (INTERNAL_EDGE, INNER_EDGE_VELOCITY), (INTERNAL_CORNER, INNER_CORNER_VELOCITY))

    def extra(self):
        return {"wall_velocity": self.wall_velocity}


@dataclass
class LocalPressure:
    """boundary::set<boundary::LocalPressure<T, D, BGKdynamics>> on a material
    (boundary/localPressure3D.h:36-62): flat faces only, CombinedRLBdynamics
    with RegularizedPressureBoundaryTuple; the density is the RHO the case
    defines on the material."""

    omega: float
    rho: float = 1.0
    base: object = None

    def shapes(self):
        return ((FLAT, LOCAL_PRESSURE),)

    def extra(self):
        return {"wall_rho": self.rho}


@dataclass
class InterpolatedVelocity:
    """boundary::set<boundary::InterpolatedVelocity<T, D, BGKdynamics>>
    (boundary/interpolatedVelocity3D.h:45-100): flat faces collide as BGK
    with BasicDirichletVelocityBoundaryTuple and run PlaneFdBoundaryProcessor3D
    after streaming; edges and corners as LocalVelocity's.  Cells without a
    discrete normal keep `base`."""

    omega: float
    wall_velocity: tuple = (0.0, 0.0, 0.0)
    base: object = None

    def shapes(self):
        return ((FLAT, INTERPOLATED_VELOCITY), (EXTERNAL_EDGE, FIXED_VELOCITY), (EXTERNAL_CORNER, FIXED_VELOCITY),
                (INTERNAL_EDGE, INNER_EDGE_VELOCITY), (INTERNAL_CORNER, INNER_CORNER_VELOCITY))

    def extra(self):
        return {"wall_velocity": self.wall_velocity}


@dataclass
class InterpolatedPressure:
    """boundary::set<boundary::InterpolatedPressure<T, D, BGKdynamics>>
    (boundary/interpolatedPressure3D.h:45-85): flat faces collide as BGK
    with BasicDirichletPressureBoundaryTuple (density `rho` or the RHO
    field) and run PlaneFdBoundaryProcessor3D; every other cell keeps
    `base`."""

    omega: float
    rho: float = 1.0
    base: object = None

    def shapes(self):
        return ((FLAT, INTERPOLATED_PRESSURE),)

    def extra(self):
        return {"wall_rho": self.rho}


_BOUNDARIES = (LocalVelocity, LocalPressure, InterpolatedVelocity, InterpolatedPressure)


def _padded(extent, padding):
    return tuple(int(n) + 2 * padding for n in extent)


def tgv(n: int, u0: float = 0.04, tau: float = 0.56, padding: int = 3,
        dtype=np.float64, with_pressure: bool = True) -> Case:
    """Fully periodic Taylor-Green vortex on an n^3 box.

    u = u0 (sin x cos y cos z, -cos x sin y cos z, 0) with x = 2 pi i / n, and
    the matching pressure p = (u0^2 / 16)(cos 2x + cos 2y)(cos 2z + 2),
    rho = 1 + 3 p (cs^2 = 1/3).  BGK with omega = 1/tau everywhere
    (examples/turbulence/tgv3d/tgv3d.cpp:68-92 with the DNS bulk dynamics).
    """
    extent = (n, n, n)
    shape = _padded(extent, padding)
    mats = np.zeros(shape, dtype=np.uint8)
    mats[padding:-padding, padding:-padding, padding:-padding] = 1
    k = 2.0 * np.pi / n
    idx = (np.arange(shape[0]) - padding).astype(np.float64)
    x = (k * idx)
    sx, cx = np.sin(x).astype(dtype), np.cos(x).astype(dtype)
    c2 = np.cos(2 * x).astype(dtype)
    u = np.empty(shape + (3,), dtype=dtype)
    u[..., 0] = u0 * sx[:, None, None] * cx[None, :, None] * cx[None, None, :]
    u[..., 1] = -u0 * cx[:, None, None] * sx[None, :, None] * cx[None, None, :]
    u[..., 2] = 0
    if with_pressure:
        p = (u0 * u0 / 16.0) * (c2[:, None, None] + c2[None, :, None]) * (c2[None, None, :] + 2.0)
        rho = (1.0 + 3.0 * p).astype(dtype)
    else:
        rho = np.ones(shape, dtype=dtype)
    omega = 1.0 / tau
    return Case(
        name=f"tgv{n}",
        extent=extent,
        padding=padding,
        periodic=(True, True, True),
        materials=mats,
        dynamics={1: Dynamics(BGK, omega=omega)},
        rho=rho,
        u=u,
        init_materials=(1,),
        info={"u0": u0, "tau": tau, "nu": (tau - 0.5) / 3.0, "k": k},
    )


def box_cavity(n: int, u_lid: float = 0.1, re: float = 100.0, padding: int = 3,
               dtype=np.float64) -> Case:
    """Lid-driven cavity on an n^3 core: the outer layer is wall (material 2,
    BounceBack), the y = n-1 face without its rim is the lid (material 3,
    BounceBackVelocity with u = (u_lid, 0, 0)), the rest is fluid (material 1,
    BGK).  nu = u_lid * n / re; starts at rest with rho = 1.
    """
    extent = (n, n, n)
    shape = _padded(extent, padding)
    p = padding
    mats = np.zeros(shape, dtype=np.uint8)
    core = mats[p:-p, p:-p, p:-p]
    core[...] = 2
    core[1:-1, 1:-1, 1:-1] = 1
    core[1:-1, n - 1, 1:-1] = 3
    nu = u_lid * n / re
    omega = relaxation_frequency(nu)
    rho = np.ones(shape, dtype=dtype)
    u = np.zeros(shape + (3,), dtype=dtype)
    return Case(
        name=f"cavity{n}",
        extent=extent,
        padding=padding,
        periodic=(False, False, False),
        materials=mats,
        dynamics={
            1: Dynamics(BGK, omega=omega),
            2: Dynamics(BOUNCE_BACK),
            3: Dynamics(BOUNCE_BACK_VELOCITY, wall_velocity=(u_lid, 0.0, 0.0), wall_rho=1.0),
        },
        rho=rho,
        u=u,
        init_materials=(1, 2, 3),
        info={"nu": nu, "omega": omega, "u_lid": u_lid},
    )


class _SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014), a portable seeded generator so the
    sphere positions are the same everywhere."""

    def __init__(self, seed: int):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def sphere_pack(extent, radius: float, porosity: float, seed: int, max_tries: int = 200000):
    """Non-overlapping spheres (centres in lattice units) placed by random
    sequential addition until the solid fraction reaches 1 - porosity.
    Spheres may cross the periodic faces; overlap is tested with the minimum
    image convention (squared distance summed x, y, z in order)."""
    rng = _SplitMix64(seed)
    ext = np.asarray(extent, dtype=np.float64)
    vol = float(np.prod(ext))
    target = (1.0 - porosity) * vol
    sphere_vol = 4.0 / 3.0 * np.pi * radius ** 3
    limit = (2.0 * radius) ** 2
    cap = int(target / sphere_vol) + 2
    centres = np.empty((cap, 3))
    n = 0
    tries = 0
    while n * sphere_vol < target and tries < max_tries:
        tries += 1
        c = np.array([rng.uniform() * ext[a] for a in range(3)])
        if n:
            diff = np.abs(centres[:n] - c)
            diff = np.minimum(diff, ext - diff)
            d2 = diff[:, 0] * diff[:, 0] + diff[:, 1] * diff[:, 1] + diff[:, 2] * diff[:, 2]
            if np.any(d2 < limit):
                continue
        centres[n] = c
        n += 1
    return centres[:n].copy()


def porous_materials(extent, centres, radius: float, padding: int, dtype=np.float64) -> np.ndarray:
    """material 1 fluid, 2 solid, 0 in the padding.  A core cell is solid when
    it lies inside a sphere or one of its periodic images by the rule of
    IndicatorSphere3D (functors/analytical/indicator/indicatorF3D.hh:166-178):
    d2 < r^2 + eps*r, evaluated in the value type `dtype` with the image
    centre rounded to it, as SuperGeometry::rename(1, 2, sphere) does for
    each image."""
    T = np.dtype(dtype).type
    shape = _padded(extent, padding)
    mats = np.zeros(shape, dtype=np.uint8)
    ext = np.asarray(extent)
    if min(extent) < 2 * radius + 2:
        raise ValueError("every extent must exceed the sphere diameter")
    core = np.ones(tuple(extent), dtype=np.uint8)
    r = T(radius)
    thr = T(r * r + T(np.finfo(T).eps) * r)
    for c in np.asarray(centres, dtype=np.float64):
        for ix in (-1, 0, 1):
            for iy in (-1, 0, 1):
                for iz in (-1, 0, 1):
                    img = c + np.array([ix, iy, iz]) * ext
                    lo = np.maximum(np.floor(img - radius).astype(int), 0)
                    hi = np.minimum(np.ceil(img + radius).astype(int), ext - 1)
                    if np.any(lo > hi):
                        continue
                    cf = img.astype(T)
                    dx = (np.arange(lo[0], hi[0] + 1).astype(T) - cf[0])[:, None, None]
                    dy = (np.arange(lo[1], hi[1] + 1).astype(T) - cf[1])[None, :, None]
                    dz = (np.arange(lo[2], hi[2] + 1).astype(T) - cf[2])[None, None, :]
                    inside = (dx * dx + dy * dy) + dz * dz < thr
                    sub = core[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1]
                    sub[inside] = 2
    p = padding
    mats[p:-p, p:-p, p:-p] = core
    return mats


def porous_channel(extent=(384, 128, 128), radius: float = 6.0, porosity: float = 0.7,
                   seed: int = 20261202, force: float = 1e-6, tau: float = 0.8,
                   magic: float = 3.0 / 16.0, padding: int = 3, dtype=np.float64) -> Case:
    """Body-force driven flow through a periodic sphere pack: TRT with
    Lambda = magic and Guo forcing (ForcedTRTdynamics) in the fluid, BounceBack
    in the solid, periodic in x, y and z, starting at rest."""
    centres = sphere_pack(extent, radius, porosity, seed)
    mats = porous_materials(extent, centres, radius, padding, dtype=dtype)
    shape = mats.shape
    omega = 1.0 / tau
    rho = np.ones(shape, dtype=dtype)
    u = np.zeros(shape + (3,), dtype=dtype)
    solid = int(np.count_nonzero(mats == 2))
    return Case(
        name="porous",
        extent=tuple(extent),
        padding=padding,
        periodic=(True, True, True),
        materials=mats,
        dynamics={
            1: Dynamics(FORCED_TRT, omega=omega, magic=magic, force=(force, 0.0, 0.0)),
            2: Dynamics(BOUNCE_BACK),
        },
        rho=rho,
        u=u,
        init_materials=(1, 2),
        info={"spheres": centres, "porosity": 1.0 - solid / float(np.prod(extent)),
              "tau": tau, "magic": magic, "force": force},
    )



def ensemble_cavities(count: int = 64, nmin: int = 32, nmax: int = 128, umin: float = 0.02,
                      umax: float = 0.1, re: float = 100.0, seed: int = 20261202, padding: int = 3,
                      dtype=np.float32):
    """Independent lid-driven cavities with edge lengths drawn uniformly from
    {nmin..nmax} and lid velocities from U[umin, umax] (seeded splitmix64);
    Re is held fixed through nu = U N / Re, tau = 3 nu + 1/2."""
    rng = _SplitMix64(seed)
    out = []
    for _ in range(count):
        n = nmin + int(rng.next() % (nmax - nmin + 1))
        u = umin + (umax - umin) * rng.uniform()
        out.append(box_cavity(n, u_lid=u, re=re, padding=padding, dtype=dtype))
    return out


def cavity3d_local_velocity(res: int, u_lid: float = 0.1, re: float = 100.0, padding: int = 3,
                            dtype=np.float64) -> Case:
    """cavity3dBenchmark in its default configuration (LID_DRIVEN without
    LID_DRIVEN_BOUNCE_BACK, cavity3d.cpp:111-124): the geometry of cavity3d
    with LocalVelocity boundaries on the walls (u = 0) and on the lid
    (u = (u_lid, 0, 0))."""
    c = cavity3d(res, u_lid=u_lid, re=re, padding=padding, dtype=dtype)
    omega = c.info["omega"]
    c.name = f"cavity3d_lv_res{res}"
    c.dynamics[2] = LocalVelocity(omega=omega)
    c.dynamics[3] = LocalVelocity(omega=omega, wall_velocity=(u_lid, 0.0, 0.0))
    return c


def step_cavity(res: int, u_lid: float = 0.1, re: float = 100.0, step: int | None = None, padding: int = 3,
                dtype=np.float64) -> Case:
    """The LocalVelocity cavity of cavity3d_local_velocity with a solid cube
    in the corner at the origin: core cells with x, y, z all <= `step`
    (default res // 2) are solid, wall (material 2) where a fluid cell is
    among their 26 neighbours and outside (0) elsewhere.  The cube's three
    edges inside the fluid are internal edges, its tip an internal corner,
    so boundary::set gives the walls every LocalVelocity shape.  Wall cells
    without a discrete normal are BounceBack."""
    c = cavity3d_local_velocity(res, u_lid=u_lid, re=re, padding=padding, dtype=dtype)
    s = res // 2 if step is None else int(step)
    if not 1 <= s <= res - 3:
        raise ValueError("the step must leave fluid on every side")
    p = padding
    core = c.materials[p:-p, p:-p, p:-p]
    fluid = core == 1
    fluid[: s + 1, : s + 1, : s + 1] = False
    block = np.zeros_like(fluid)
    block[: s + 1, : s + 1, : s + 1] = True
    near = np.zeros_like(fluid)
    padded = np.pad(fluid, 1)
    n = core.shape
    for dx in (0, 1, 2):
        for dy in (0, 1, 2):
            for dz in (0, 1, 2):
                near |= padded[dx:dx + n[0], dy:dy + n[1], dz:dz + n[2]]
    inner = block.copy()
    inner[0, :, :] = inner[:, 0, :] = inner[:, :, 0] = False
    core[inner & near] = 2
    core[inner & ~near] = 0
    c.name = f"step_cavity_res{res}_step{s}"
    omega = c.info["omega"]
    c.dynamics[2] = LocalVelocity(omega=omega, base=Dynamics(BOUNCE_BACK))
    c.dynamics[3] = LocalVelocity(omega=omega, wall_velocity=(u_lid, 0.0, 0.0), base=Dynamics(BOUNCE_BACK))
    c.info["step"] = s
    c.info.pop("_normals", None)
    return c


def cavity3d(res: int, u_lid: float = 0.1, re: float = 100.0, padding: int = 3,
             dtype=np.float64) -> Case:
    """The lid-driven cavity of examples/laminar/cavity3dBenchmark at
    RESOLUTION res: (res+1)^3 nodes, wall shell (material 2), lid face y = res
    without its rim (material 3), fluid inside (material 1)
    (cavity3d.cpp:44-85), with BounceBack walls and a BounceBackVelocity lid
    moving at u_lid; nu = u_lid * res / re (res = 64, u = 0.1, Re = 100 give
    nu = 0.064, tau = 0.692)."""
    c = box_cavity(res + 1, u_lid=u_lid, re=re, padding=padding, dtype=dtype)
    nu = u_lid * res / re
    omega = relaxation_frequency(nu)
    c.name = f"cavity3d_res{res}"
    c.dynamics[1] = Dynamics(BGK, omega=omega)
    c.info.update({"nu": nu, "omega": omega, "res": res})
    return c


def channel3d(nx: int, ny: int, nz: int, u_in: float = 0.05, tau: float = 0.8, padding: int = 3,
              dtype=np.float64, boundary: str = "local") -> Case:
    """A straight channel along x on an nx x ny x nz core: BounceBack walls
    (material 2) on the y and z faces, the x = 0 face without its rim a
    LocalVelocity inflow at (u_in, 0, 0) (material 3), the x = nx-1 face
    without its rim a LocalPressure outflow at rho = 1 (material 4), BGK fluid
    (material 1) starting at rest.  boundary="interpolated" makes the inflow
    InterpolatedVelocity and the outflow InterpolatedPressure."""
    if boundary not in ("local", "interpolated"):
        raise ValueError("boundary is local or interpolated")
    extent = (nx, ny, nz)
    shape = _padded(extent, padding)
    p = padding
    mats = np.zeros(shape, dtype=np.uint8)
    core = mats[p:-p, p:-p, p:-p]
    core[...] = 2
    core[:, 1:-1, 1:-1] = 1
    core[0, 1:-1, 1:-1] = 3
    core[-1, 1:-1, 1:-1] = 4
    omega = 1.0 / tau
    return Case(
        name=f"channel{nx}x{ny}x{nz}",
        extent=extent,
        padding=padding,
        periodic=(False, False, False),
        materials=mats,
        dynamics={1: Dynamics(BGK, omega=omega), 2: Dynamics(BOUNCE_BACK),
                  3: (InterpolatedVelocity if boundary == "interpolated" else LocalVelocity)(
                      omega=omega, wall_velocity=(u_in, 0.0, 0.0)),
                  4: (InterpolatedPressure if boundary == "interpolated" else LocalPressure)(omega=omega, rho=1.0)},
        rho=np.ones(shape, dtype=dtype),
        u=np.zeros(shape + (3,), dtype=dtype),
        init_materials=(1, 2, 3, 4),
        info={"omega": omega, "u_in": u_in},
    )
