"""Bouzidi interpolated bounce-back: the host side of setBouzidiBoundary.

setBouzidiBoundary<T, DESCRIPTOR, BouzidiPostProcessor>(lattice, geometry,
solidMaterial, indicator) (boundary/setBouzidiBoundary.h:193-330) walks every
core cell of the solid material and, for each direction c_i whose neighbour
x_b = x_s + c_i is inside the block, stores in x_b's BOUZIDI_DISTANCE field
the normed distance q along -c_i from x_b to the indicator's surface:
  - x_b a bulk (fluid) cell: q = dist / (dx |c_i|) from the indicator's ray
    distance, 0.5 when none is found or it falls outside [0, 1], and 0.5 as
    well when the cell beyond x_b (x_b + c_i) is not bulk;
  - x_b neither bulk nor material 0: q = 0.5.
Every non-padding cell that received an entry runs BouzidiPostProcessor after
streaming.  This module computes that field for the indicators a case passes
(IndicatorCylinder3D here, with the generic bisection ray distance of
functors/analytical/indicator/indicatorBase.h:35-93) and turns it into the
entry list the GPU platform applies concurrently (see entries()).
"""
from __future__ import annotations

import math

import numpy as np

Q = 19
C = np.array([
    [0, 0, 0],
    [-1, 0, 0], [0, -1, 0], [0, 0, -1],
    [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
    [-1, 0, 1], [0, -1, -1], [0, -1, 1],
    [1, 0, 0], [0, 1, 0], [0, 0, 1],
    [1, 1, 0], [1, -1, 0], [1, 0, 1],
    [1, 0, -1], [0, 1, 1], [0, 1, -1]], dtype=np.int64)
OPP = np.array([0] + [i + 9 for i in range(1, 10)] + [i - 9 for i in range(10, 19)])

_EPS = 3.0 * np.finfo(np.float64).eps
_MIN = np.finfo(np.float64).tiny / np.finfo(np.float64).eps


def eps_relative(a: float) -> float:
    """util::epsRelative (utilities/floatingPointRelationalOperators.h:278-296)"""
    if a == 0.0 or (-np.finfo(np.float64).eps < a < np.finfo(np.float64).eps):
        return _MIN
    return _EPS * abs(a)


def _dot(a, b) -> float:
    s = 0.0
    for k in range(3):
        s += a[k] * b[k]
    return s


class IndicatorCylinder3D:
    """IndicatorCylinder3D(center1, center2, radius)
    (functors/analytical/indicator/indicatorF3D.hh:312-418), FP64"""

    def __init__(self, center1, center2, radius: float):
        self.c1 = [float(v) for v in center1]
        self.c2 = [float(v) for v in center2]
        self.ba = [self.c2[k] - self.c1[k] for k in range(3)]
        baba = self.ba[0] * self.ba[0] + self.ba[1] * self.ba[1] + self.ba[2] * self.ba[2]
        self.radius2 = float(radius) * float(radius)
        self.length = math.sqrt(baba)
        K = [self.ba[k] / self.length for k in range(3)]
        if self.c2[1] == self.c1[1] and self.c2[0] == self.c1[0]:
            # util::approxEqual of the two centres' x and y
            I, J = [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]
        else:
            normi = math.sqrt(K[1] * K[1] + K[0] * K[0])
            I = [-K[1] / normi, K[0] / normi, 0.0]
            J = [K[1] * I[2] - K[2] * I[1], K[2] * I[0] - K[0] * I[2], K[0] * I[1] - K[1] * I[0]]
        r = math.sqrt(self.radius2)
        self.lo = [0.0] * 3
        self.hi = [0.0] * 3
        for k in range(3):
            s = math.sqrt(I[k] * I[k] + J[k] * J[k]) * r
            self.hi[k] = self.c1[k] + s + max(K[k] * self.length, 0.0)
            self.lo[k] = self.c1[k] - s + min(K[k] * self.length, 0.0)
        self._eps_max = max(eps_relative(self.hi[0]), max(eps_relative(self.hi[1]), eps_relative(self.hi[2])))
        self._eps_rad = eps_relative(math.sqrt(self.radius2))

    def inside(self, p) -> bool:
        """operator() (indicatorF3D.hh:356-379)"""
        pa = [p[k] - self.c1[k] for k in range(3)]
        d = [self.c2[k] - self.c1[k] for k in range(3)]
        h = (pa[0] * d[0] + pa[1] * d[1] + pa[2] * d[2]) / (d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        if h < 0.0 - self._eps_max or h > 1.0 + self._eps_max:
            return False
        closest = [self.c1[k] + h * d[k] for k in range(3)]
        r = [p[k] - closest[k] for k in range(3)]
        return _dot(r, r) < self.radius2 + self._eps_rad

    def inside_box(self, p) -> bool:
        """IndicatorF3D::isInsideBox (indicatorBaseF3D.hh:305-308)"""
        return all(self.lo[k] <= p[k] for k in range(3)) and all(p[k] <= self.hi[k] for k in range(3))


def ray_distance(indicator, origin, direction, precision: float = 0.0001, pitch: float = 0.5):
    """util::distance with the indicator's operator() and bounding box
    (functors/analytical/indicator/indicatorBase.h:35-93; the precision and
    pitch of IndicatorF3D::distance, indicatorBaseF3D.hh:80-87): the distance
    along `direction` to the first change of inside/outside, or None"""
    step = 10 * precision
    p = [origin[k] + step * direction[k] for k in range(3)]
    a = indicator.inside(p)
    p = [origin[k] - step * direction[k] for k in range(3)]
    b = indicator.inside(p)
    if a == (not b):
        return 0.0
    origin_value = indicator.inside(origin)
    cur = list(origin)
    value = origin_value
    while value == origin_value and indicator.inside_box(cur):
        cur = [cur[k] + direction[k] for k in range(3)]
        value = indicator.inside(cur)
    if not indicator.inside_box(cur) and not origin_value:
        return None
    while pitch >= precision:
        if not indicator.inside_box(cur) and origin_value:
            cur = [cur[k] - pitch * direction[k] for k in range(3)]
        else:
            value = indicator.inside(cur)
            if value == origin_value:
                cur = [cur[k] + pitch * direction[k] for k in range(3)]
            else:
                cur = [cur[k] - pitch * direction[k] for k in range(3)]
        pitch *= 0.5
    r = [cur[k] - origin[k] for k in range(3)]
    return math.sqrt(_dot(r, r))


def distances(materials: np.ndarray, padding: int, origin, delta: float, indicator, solid: int,
              bulk=(1,)) -> np.ndarray:
    """the BOUZIDI_DISTANCE field (19 x padded cells, -1 where unset) that
    setBouzidiBoundary (setBouzidiBoundary.h:229-330) writes for the solid
    material `solid`, the bulk materials `bulk` and an indicator with
    inside(p), inside_box(p)"""
    m = np.asarray(materials)
    P = np.array(m.shape)
    q = np.full((Q, m.size), -1.0)
    is_bulk = np.isin(m, np.asarray(bulk))
    flat = np.arange(m.size).reshape(m.shape)

    def material(r):
        # BlockGeometry::getMaterial: 0 outside the padded block
        return int(m[tuple(r)]) if np.all(r >= 0) and np.all(r < P) else 0

    # forSpatialLocations covers the padding as well (core/blockStructure.h:236-258)
    for sp in np.argwhere(m == solid):
        for i in range(1, Q):
            bp = sp + C[i]
            if np.any(bp < 0) or np.any(bp >= P):
                continue                    # BlockGeometry::isInside
            b = bp - padding                # latticeR of x_b relative to the core
            cell = flat[tuple(bp)]
            o = OPP[i]
            if is_bulk[tuple(bp)]:
                phys = [origin[k] + float(b[k]) * delta for k in range(3)]
                direction = [-delta * float(C[i][k]) for k in range(3)]
                norm = delta * math.sqrt(float(np.dot(C[i], C[i])))
                d = ray_distance(indicator, phys, direction)
                qi = 0.5 if d is None else d / norm
                if qi < 0 or qi > 1:
                    qi = 0.5
                q[o, cell] = qi if material(bp + C[i]) in bulk else 0.5
            elif material(bp) != 0:
                q[o, cell] = 0.5
    return q


def entries(q: np.ndarray, padding: int, shape):
    """the cells that run BouzidiPostProcessor and their 19 distances, ready
    for Lattice.set_bouzidi: non-padding cells with an entry, ascending id.

    The reference applies the post-processors one cell after another in
    ascending cell id (core/platform/cpu/sisd/operator.h:368-386); the GPU
    applies them all at once.  The two differ only where one cell's update
    reads a population another cell's update writes, which the setter
    produces only for q = 0.5 entries between two solid cells (a pure copy
    X[k] <- Y[opp k] met by Y[opp k] <- X[k]).  In ascending order the lower
    id takes the other's value and the higher id's copy then writes back the
    value it already holds, so the higher id's entry is dropped here."""
    P = tuple(shape)
    count = int(np.prod(P))
    core = np.zeros(P, dtype=bool)
    core[padding:P[0] - padding, padding:P[1] - padding, padding:P[2] - padding] = True
    has = (q >= 0).any(axis=0) & core.reshape(-1)
    cells = np.flatnonzero(has)
    qq = q[:, cells].copy()
    listed = np.full(count, -1, dtype=np.int64)
    listed[cells] = np.arange(cells.size)
    step = np.array([P[1] * P[2], P[2], 1], dtype=np.int64)
    d = C @ step
    for e, x in enumerate(cells):
        for j in range(1, Q):
            if qq[j, e] < 0:
                continue
            # x writes x[opp j] reading y[j], y = x + c_j, which y writes when q_y[opp j] >= 0
            y = x + d[j]
            ey = listed[y]
            if ey >= 0 and qq[OPP[j], ey] >= 0:
                if qq[j, e] != 0.5 or qq[OPP[j], ey] != 0.5:
                    raise ValueError("two Bouzidi cells read each other's updates")
                if y < x:
                    qq[j, e] = -1.0
            # for q > 1/2 x also reads z[opp j], z = x - c_j, which z writes when q_z[j] >= 0
            if qq[j, e] > 0.5:
                ez = listed[x - d[j]]
                if ez >= 0 and qq[j, ez] >= 0:
                    raise ValueError("two Bouzidi cells read each other's updates")
    keep = (qq >= 0).any(axis=0)
    return cells[keep].astype(np.int64), np.ascontiguousarray(qq[:, keep].T)


class CirclePoiseuille3D:
    """CirclePoiseuille3D(center, normal, maxVelocity, radius)
    (functors/analytical/frameChangeF3D.hh:353-356), i.e. CirclePowerLaw3D
    with n = 1 (:166-169, 237-243), FP64: velocity component d is
    maxVelocity * normal_d * (1 - (r_d / radius)^2), r_d the distance from
    the centre in the two coordinates other than d (a pipe along a
    coordinate axis); the normal is normalised."""

    def __init__(self, center, normal, max_velocity: float, radius: float, scale: float = 1.0):
        n = np.asarray(normal, dtype=np.float64)
        self.center = np.asarray(center, dtype=np.float64)
        self.normal = n / math.sqrt(float(n @ n))
        self.max_velocity = float(max_velocity)
        self.radius = float(radius)
        self.scale = float(scale)

    def __call__(self, x: np.ndarray) -> np.ndarray:
        """velocities (..., 3) at points x (..., 3)"""
        x = np.asarray(x, dtype=np.float64)
        dx = x - self.center
        out = np.empty_like(x)
        for d in range(3):
            a, b = [k for k in range(3) if k != d]
            r = np.sqrt(dx[..., a] * dx[..., a] + dx[..., b] * dx[..., b])
            out[..., d] = (self.scale * self.max_velocity * self.normal[d]) * (1.0 - (r / self.radius) ** 2)
        return out


def wall_velocities(q: np.ndarray, materials: np.ndarray, padding: int, origin, delta: float, velocity,
                    solid: int) -> np.ndarray:
    """the BOUZIDI_VELOCITY field (19 x padded cells, 0 where unset) that
    setBouzidiVelocity (boundary/setBouzidiBoundary.h:380-420) writes for the
    cells of material `solid` and the analytical velocity `velocity`
    (a callable on points (..., 3)): for every cell x_s of the material,
    padding included, and direction c_i with x_b = x_s + c_i inside the block
    and a distance q_b[opp i] >= 0, the wall point is
    x_b + q_b[opp i] dx c_opp(i) and the coefficient c_opp(i) . u(wall point)"""
    m = np.asarray(materials)
    P = np.array(m.shape)
    v = np.zeros_like(q)
    solids = np.argwhere(m == solid)
    for i in range(1, Q):
        o = OPP[i]
        bp = solids + C[i]
        inside = np.all((bp >= 0) & (bp < P), axis=1)
        bp = bp[inside]
        cell = np.ravel_multi_index(bp.T, m.shape)
        qb = q[o, cell]
        on = qb >= 0
        bp, cell, qb = bp[on], cell[on], qb[on]
        phys = np.asarray(origin, dtype=np.float64)[None, :] + (bp - padding).astype(np.float64) * delta
        wall = phys + (qb * delta)[:, None] * C[o].astype(np.float64)[None, :]
        u = velocity(wall)
        v[o, cell] = u @ C[o].astype(np.float64)
    return v


def processor_entries(q: np.ndarray, materials: np.ndarray, padding: int, solid: int):
    """the cells that run the post-processor setBouzidiBoundary attached for
    the solid material `solid` (setBouzidiBoundary.h:290-330): a core cell
    runs it when one of its distances points at a cell of that material,
    and then applies all of its distances (the processor reads the whole
    BOUZIDI_DISTANCE field of its cell); as entries() returns them"""
    m = np.asarray(materials)
    shape = m.shape
    count = int(np.prod(shape))
    mats = m.reshape(-1)
    step = np.array([shape[1] * shape[2], shape[2], 1], dtype=np.int64)
    d = C @ step
    owner = np.zeros(count, dtype=bool)
    for j in range(1, Q):
        cells = np.flatnonzero(q[j] >= 0)
        target = cells + d[j]
        ok = (target >= 0) & (target < count)
        owner[cells[ok][mats[target[ok]] == solid]] = True
    return entries(np.where(owner[None, :], q, -1.0), padding, shape)
