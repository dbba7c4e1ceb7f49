"""The kernels' segment/triangle intersection (pt_seg_tri: CGAL's case analysis evaluated
without its branches, one set of orientation predicates for every case) against the
oracle's restatement of CGAL intersection(Triangle_3, Segment_3) with the case branches
(oracle/ptref.c seg_tri): result and intersection point bit-exact, on generic segments
and on every degenerate case the branches distinguish (an end on the triangle's plane,
the segment through an edge or a vertex, coplanar segments, zero-length segments)."""
import ctypes as C

import numpy as np
import pytest

from pentrack_amd import _native


@pytest.fixture(scope="module")
def both(oracle_lib):
    return _native.lib(), oracle_lib.load("fp64")


def run(lib, fn, a, b, c, p, q):
    d = lambda v: np.ascontiguousarray(v, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))
    pt = np.full(3, np.nan)
    r = getattr(lib, fn)(d(a), d(b), d(c), d(p), d(q), d(pt))
    return r, pt


def check(both, a, b, c, p, q):
    L, O = both
    rg, pg = run(L, "pt_seg_tri", a, b, c, p, q)
    ro, po = run(O, "ptref_seg_tri", a, b, c, p, q)
    assert rg == ro, (a, b, c, p, q, rg, ro)
    if rg == 1:
        assert np.array_equal(pg, po), (a, b, c, p, q, pg, po)
    return rg


def test_generic_segments(both):
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(20000):
        a, b, c = rng.uniform(-1, 1, (3, 3))
        p, q = rng.uniform(-1.5, 1.5, (2, 3))
        hits += check(both, a, b, c, p, q) == 1
    assert hits > 500   # the hit branch is exercised


def test_degenerate_cases(both):
    """Triangle in z = 0 with integer vertices; segment ends on the plane, through an edge,
    through a vertex, coplanar (overlapping or not) and of zero length, in both
    directions and for both orientations of the triangle."""
    a, b, c = np.array([0., 0., 0.]), np.array([4., 0., 0.]), np.array([0., 4., 0.])
    ends = [np.array(v, float) for v in
            [(1, 1, 0), (1, 1, 1), (1, 1, -1), (2, 0, 0), (2, 0, 1), (2, 0, -1), (0, 0, 0), (0, 0, 1), (0, 0, -2),
             (2, 2, 0), (2, 2, 1), (2, 2, -1), (5, 5, 0), (5, 5, 1), (5, 5, -1), (-1, 1, 0), (-1, 1, 3), (1, -1, -3),
             (6, -2, 0), (-2, 6, 0), (3, 3, 0), (1, 1, 2), (0.5, 0.25, 0)]]
    seen = set()
    for tri in ((a, b, c), (a, c, b)):
        for p in ends:
            for q in ends:
                seen.add(check(both, *tri, p, q))
    assert seen == {-1, 0, 1}


def test_random_ends_on_the_plane(both):
    """Ends exactly on the plane of an axis-aligned triangle (orientation exactly 0),
    the other end above, below or on it."""
    rng = np.random.default_rng(9)
    a, b, c = np.array([-1., -1., 0.5]), np.array([2., -1., 0.5]), np.array([-1., 2., 0.5])
    for _ in range(5000):
        p = np.array([*rng.uniform(-2, 3, 2), 0.5])
        q = np.array([*rng.uniform(-2, 3, 2), rng.choice([0.5, 1.0, 0.0, 0.5 + 2 ** -40])])
        if rng.random() < 0.5:
            p, q = q, p
        check(both, a, b, c, p, q)
