"""Discrete normals of boundary cells: the classification boundary::set
makes before it picks a boundary's dynamics and post-processors
(boundary/setBoundary.h:76-100 calling computeBoundaryTypeAndNormal,
geometry/discreteNormals.hh:176-1190).

A cell is classified by which of its 26 neighbours are fluid (F), outside
(O, material 0) or neither (B).  The classification is a fixed list of
neighbourhood patterns tried in a fixed order: the first flat, external
corner, internal corner or external edge pattern that matches decides
(later matches of those kinds are ignored), and any internal-edge pattern
that matches afterwards replaces it, the last one winning.  A cell matching
nothing is returned as FLAT with the zero normal, as the reference returns
it; boundary::set then leaves the cell's dynamics alone for LocalVelocity
and refuses it for LocalPressure.

Only cells at least two cells inside the padded block are classified
(BlockStructure::getNeighborhoodRadius >= 2 in setBoundary.h:86-87); the
patterns look one cell further.
"""
from __future__ import annotations

import numpy as np

__all__ = ["FLAT", "EXTERNAL_CORNER", "INTERNAL_CORNER", "EXTERNAL_EDGE", "INTERNAL_EDGE", "RULES",
           "discrete_normals"]

# DiscreteNormalType (geometry/discreteNormals.h:33-39)
FLAT, EXTERNAL_CORNER, INTERNAL_CORNER, EXTERNAL_EDGE, INTERNAL_EDGE = 0, 1, 2, 3, 4

# (type, outward normal, pattern).  A pattern is a list of neighbour tests
# "Kxyz": K is F (fluid), B (neither fluid nor outside) or N (not fluid);
# x, y, z in {+, -, 0} give the offset.  Every pattern also needs the cell
# itself to be B.  The order is the reference's order of tests.
RULES = (
    # flat faces: the four in-plane neighbours are B, fluid on one side
    (FLAT, (-1, 0, 0), "F+00 B0+0 B0-0 B00+ B00-"),
    (FLAT, (1, 0, 0), "F-00 B0+0 B0-0 B00+ B00-"),
    (FLAT, (0, -1, 0), "F0+0 B+00 B-00 B00+ B00-"),
    (FLAT, (0, 1, 0), "F0-0 B+00 B-00 B00+ B00-"),
    (FLAT, (0, 0, -1), "F00+ B+00 B-00 B0+0 B0-0"),
    (FLAT, (0, 0, 1), "F00- B+00 B-00 B0+0 B0-0"),
    # external corners: fluid on the diagonal, the six cells between are B
    (EXTERNAL_CORNER, (-1, -1, -1), "F+++ B++0 B+0+ B+00 B0++ B0+0 B00+"),
    (EXTERNAL_CORNER, (-1, 1, -1), "F+-+ B+-0 B+0+ B+00 B0-+ B0-0 B00+"),
    (EXTERNAL_CORNER, (-1, 1, 1), "F+-- B+-0 B+0- B+00 B0-- B0-0 B00-"),
    (EXTERNAL_CORNER, (-1, -1, 1), "F++- B++0 B+0- B+00 B0+- B0+0 B00-"),
    (EXTERNAL_CORNER, (1, 1, 1), "F--- B--0 B-0- B-00 B0-- B0-0 B00-"),
    (EXTERNAL_CORNER, (1, -1, 1), "F-+- B-+0 B-0- B-00 B0+- B0+0 B00-"),
    (EXTERNAL_CORNER, (1, -1, -1), "F-++ B-+0 B-0+ B-00 B0++ B0+0 B00+"),
    (EXTERNAL_CORNER, (1, 1, -1), "F--+ B--0 B-0+ B-00 B0-+ B0-0 B00+"),
    # internal corners: fluid on three faces, B on the three opposite ones
    (INTERNAL_CORNER, (1, 1, 1), "F-00 F0-0 F00- B+00 B0+0 B00+"),
    (INTERNAL_CORNER, (1, -1, 1), "F-00 F0+0 F00- B+00 B0-0 B00+"),
    (INTERNAL_CORNER, (1, -1, -1), "F-00 F0+0 F00+ B+00 B0-0 B00-"),
    (INTERNAL_CORNER, (1, 1, -1), "F-00 F0-0 F00+ B+00 B0+0 B00-"),
    (INTERNAL_CORNER, (-1, 1, 1), "F+00 F0-0 F00- B-00 B0+0 B00+"),
    (INTERNAL_CORNER, (-1, -1, 1), "F+00 F0+0 F00- B-00 B0-0 B00+"),
    (INTERNAL_CORNER, (-1, 1, -1), "F+00 F0-0 F00+ B-00 B0+0 B00-"),
    (INTERNAL_CORNER, (-1, -1, -1), "F+00 F0+0 F00+ B-00 B0-0 B00-"),
    # external edges: fluid on one in-plane diagonal (the tests are not
    # symmetric between the edge directions; kept as the reference has them)
    (EXTERNAL_EDGE, (0, 1, -1), "F0-+ B+00 B-00 B0-0 B00+ N+0+ N-0+"),
    (EXTERNAL_EDGE, (0, -1, -1), "F0++ B+00 B-00 B0+0 B00+ N+0+ N-0+"),
    (EXTERNAL_EDGE, (0, -1, 1), "F0+- B+00 B-00 B00- N+0- N-0- N0+0"),
    (EXTERNAL_EDGE, (0, 1, 1), "F0-- B+00 B-00 B00- N+0- N-0- N0-0"),
    (EXTERNAL_EDGE, (-1, 0, -1), "F+0+ B+00 B0+0 B0-0 B00+ N-00 N00-"),
    (EXTERNAL_EDGE, (1, 0, -1), "F-0+ B-00 B0+0 B0-0 B00+ N+00 N00-"),
    (EXTERNAL_EDGE, (-1, 0, 1), "F+0- B+00 B0+0 B0-0 B00- N-00 N00+"),
    (EXTERNAL_EDGE, (1, 0, 1), "F-0- B-00 B0+0 B0-0 B00- N+00 N00+"),
    (EXTERNAL_EDGE, (-1, -1, 0), "F++0 B0+0 B00+ B00- N+00"),
    (EXTERNAL_EDGE, (1, -1, 0), "F-+0 B0+0 B00+ B00- N-00"),
    (EXTERNAL_EDGE, (1, 1, 0), "F--0 B0-0 B00+ B00- N-00"),
    (EXTERNAL_EDGE, (-1, 1, 0), "F+-0 B0-0 B00+ B00- N+00"),
    # internal edges (these replace an earlier match)
    (INTERNAL_EDGE, (0, -1, -1), "F0+0 F00+ B+-0 B+00 B--0 B-00 B0-0 B00-"),
    (INTERNAL_EDGE, (0, 1, -1), "F0-0 F00+ B++0 B+00 B-+0 B-00 B0+0 B00-"),
    (INTERNAL_EDGE, (0, -1, 1), "F0+0 F00- B+-0 B+00 B--0 B-00 B0-0 B00+"),
    (INTERNAL_EDGE, (0, 1, 1), "F0-0 F00- B++0 B+00 B-+0 B-00 B0+0 B00+"),
    (INTERNAL_EDGE, (1, 0, 1), "F-00 F00- B+00 B0++ B0+0 B0-+ B0-0 B00+"),
    (INTERNAL_EDGE, (1, 0, -1), "F-00 F00+ B+00 B0+- B0+0 B0-- B0-0 B00-"),
    (INTERNAL_EDGE, (-1, 0, 1), "F+00 F00- B-00 B0++ B0+0 B0-+ B0-0 B00+"),
    (INTERNAL_EDGE, (-1, 0, -1), "F+00 F00+ B-00 B0+- B0+0 B0-- B0-0 B00-"),
    (INTERNAL_EDGE, (1, 1, 0), "F--0 F-00 F0-0 B0+0 B00+ B00-"),
    (INTERNAL_EDGE, (-1, 1, 0), "F+-0 F+00 F0-0 B0+0 B00+ B00-"),
    (INTERNAL_EDGE, (1, -1, 0), "F-+0 F-00 F0+0 B0-0 B00+ B00-"),
    (INTERNAL_EDGE, (-1, -1, 0), "F++0 F+00 F0+0 B0-0 B00+ B00-"),
)

_SIGN = {"+": 1, "-": -1, "0": 0}


def _parse(pattern: str):
    return [(t[0], tuple(_SIGN[c] for c in t[1:])) for t in pattern.split()]


_PARSED = tuple((t, n, _parse(p)) for t, n, p in RULES)


def discrete_normals(materials: np.ndarray, boundary, fluid: int = 1, outside: int = 0,
                     with_rule: bool = False):
    """Type and outward normal of every cell whose material is in `boundary`
    (fluid = material `fluid`, outside = material `outside`).

    materials: the padded block, C-order (x, y, z).  Returns (types, normals):
    int8 type per cell, -1 for cells that are not classified; int8 normal x 3
    per cell.  with_rule=True also returns the index into RULES of the
    pattern that decided each cell (-1 where none did)."""
    m = np.asarray(materials)
    if m.ndim != 3:
        raise ValueError("materials must be a 3D array over the padded block")
    shape = m.shape
    types = np.full(shape, -1, dtype=np.int8)
    normals = np.zeros(shape + (3,), dtype=np.int8)
    rule = np.full(shape, -1, dtype=np.int16)
    if min(shape) < 5:
        return (types, normals, rule) if with_rule else (types, normals)
    isf = m == fluid
    iso = m == outside
    sel = np.zeros(shape, dtype=bool)
    inner = (slice(2, -2),) * 3
    sel[inner] = np.isin(m[inner], list(boundary))
    # views over the cells two or more inside the block, shifted by d
    view = (slice(2, shape[0] - 2), slice(2, shape[1] - 2), slice(2, shape[2] - 2))

    def shifted(arr, d):
        return arr[2 + d[0]:shape[0] - 2 + d[0], 2 + d[1]:shape[1] - 2 + d[1], 2 + d[2]:shape[2] - 2 + d[2]]

    def test(kind, d):
        f = shifted(isf, d)
        if kind == "F":
            return f
        if kind == "N":
            return ~f
        return ~f & ~shifted(iso, d)

    s = sel[view]
    cand = s & test("B", (0, 0, 0))
    t_v = types[view]
    n_v = normals[view]
    r_v = rule[view]
    t_v[s] = FLAT
    decided = np.zeros_like(s)
    for k, (t, n, terms) in enumerate(_PARSED):
        hit = cand.copy()
        for kind, d in terms:
            hit &= test(kind, d)
            if not hit.any():
                break
        if t != INTERNAL_EDGE:
            hit &= ~decided
            decided |= hit
        t_v[hit] = t
        n_v[hit] = n
        r_v[hit] = k
    return (types, normals, rule) if with_rule else (types, normals)
