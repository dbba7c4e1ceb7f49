"""Oracle geometry: CGAL-semantics segment/triangle tests, collision records,
point-in-solid parity, RotateVector and the segment/bbox test.

Pinned against exact rational arithmetic (fractions.Fraction), closed forms for
axis-aligned faces and brute-force sampling.
"""
import ctypes as C
from fractions import Fraction as Fr

import numpy as np
import pytest

from pentrack_amd.geometry import box_mesh, hollow_box_mesh

MATS = np.zeros((2, 9))


@pytest.fixture(scope="module")
def P(oracle_lib):
    return oracle_lib


def exact_orient(p, q, r, s):
    p, q, r, s = ([Fr(float(v)) for v in x] for x in (p, q, r, s))
    m = [[q[i] - p[i], r[i] - p[i], s[i] - p[i]] for i in range(3)]
    det = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
           + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))
    return (det > 0) - (det < 0)


def exact_hit(a, b, c, p, q):
    """Exact segment/triangle intersection point (None if none), generic position."""
    o1, o2 = exact_orient(a, b, c, p), exact_orient(a, b, c, q)
    if o1 == o2 or o1 == 0 or o2 == 0:
        return None
    u, w = (p, q) if o1 > 0 else (q, p)
    if exact_orient(u, w, a, b) > 0 or exact_orient(u, w, b, c) > 0 or exact_orient(u, w, c, a) > 0:
        return None
    A, B, Cc, Pp, Q = ([Fr(float(v)) for v in x] for x in (a, b, c, p, q))
    n = [(B[1] - A[1]) * (Cc[2] - A[2]) - (B[2] - A[2]) * (Cc[1] - A[1]),
         (B[2] - A[2]) * (Cc[0] - A[0]) - (B[0] - A[0]) * (Cc[2] - A[2]),
         (B[0] - A[0]) * (Cc[1] - A[1]) - (B[1] - A[1]) * (Cc[0] - A[0])]
    d = [Q[i] - Pp[i] for i in range(3)]
    t = sum(n[i] * (A[i] - Pp[i]) for i in range(3)) / sum(n[i] * d[i] for i in range(3))
    return [float(Pp[i] + t * d[i]) for i in range(3)], float(t)


def one_triangle(P, tri):
    return P.single_geometry(np.asarray(tri, float).reshape(1, 9), ([0], [2], [0]), MATS)


def test_segment_triangle_matches_exact_arithmetic(P):
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(1500):
        tri = rng.uniform(-1, 1, (3, 3))
        p, q = rng.uniform(-1.5, 1.5, (2, 3))
        g = one_triangle(P, tri)
        n, rec = P.collisions(g, p, q)
        ex = exact_hit(tri[0], tri[1], tri[2], p, q)
        assert (n == 1) == (ex is not None)
        if ex is not None:
            hits += 1
            assert abs(rec["s"][0] - ex[1]) < 1e-12
    assert hits > 20


def test_axis_aligned_face_closed_form(P):
    cube = box_mesh((0, 0, 0), (1, 1, 1)).astype(np.float64)
    g = P.single_geometry(cube.reshape(-1, 9), (np.zeros(12, np.int32), [2], [0]), MATS)
    p, q = np.array([0.5, 0.25, 0.75]), np.array([1.5, 0.5, 0.75])
    n, rec = P.collisions(g, p, q)
    assert n == 1
    assert np.array_equal(rec["normal"][0], [1.0, 0.0, 0.0])  # outward
    assert rec["s"][0] == 0.5
    assert rec["distnormal"][0] == 1.0  # (q-p).n > 0: leaving
    n, rec = P.collisions(g, q, p)
    assert n == 1 and rec["distnormal"][0] == -1.0  # entering
    # parallel segment inside: no hit
    assert P.collisions(g, [0.2, 0.2, 0.2], [0.8, 0.2, 0.2])[0] == 0


def test_collision_order_s_then_id_descending(P):
    # two nested boxes: a segment leaving both has two hits ordered by s
    inner = box_mesh((0.25, 0.25, 0.25), (0.75, 0.75, 0.75)).astype(np.float64)
    outer = box_mesh((0, 0, 0), (1, 1, 1)).astype(np.float64)
    tris = np.concatenate([outer, inner]).reshape(-1, 9)
    ts = np.array([0] * 12 + [1] * 12, np.int32)
    g = P.single_geometry(tris, (ts, [2, 3], [0, 0]), MATS)
    n, rec = P.collisions(g, [0.5, 0.3, 0.6], [1.5, 0.3, 0.6])
    assert n == 2 and list(rec["ID"]) == [3, 2] and rec["s"][0] < rec["s"][1]
    # a segment through the shared diagonal of a face's two triangles hits both
    n, rec = P.collisions(g, [0.5, 0.5, 0.5], [1.5, 0.5, 0.5])
    assert n == 4 and list(rec["ID"]) == [3, 3, 2, 2] and list(rec["tri"]) == sorted(rec["tri"][:2]) + sorted(rec["tri"][2:])
    # coincident faces: same s, higher ID first
    same = np.concatenate([outer, outer]).reshape(-1, 9)
    g2 = P.single_geometry(same, (ts, [2, 5], [0, 0]), MATS)
    n, rec = P.collisions(g2, [0.5, 0.3, 0.6], [1.5, 0.3, 0.6])
    assert n == 2 and rec["s"][0] == rec["s"][1] and list(rec["ID"]) == [5, 2]


def test_coplanar_overlap_is_reported(P):
    cube = box_mesh((0, 0, 0), (1, 1, 1)).astype(np.float64)
    g = P.single_geometry(cube.reshape(-1, 9), (np.zeros(12, np.int32), [2], [0]), MATS)
    # a segment lying in the face x = 1 (the reference throws: stopID -6)
    n, _ = P.collisions(g, [1.0, 0.2, 0.3], [1.0, 0.7, 0.4])
    assert n == -1
    # in the plane of the face but outside the triangle: no error, no hit
    assert P.collisions(g, [1.0, 2.2, 0.3], [1.0, 2.7, 0.4])[0] == 0


def test_get_solids_parity_and_priority(P):
    outer = hollow_box_mesh((0, 0, 0), (1, 1, 1), 0.01).astype(np.float64)
    inner = box_mesh((0.4, 0.4, 0.4), (0.6, 0.6, 0.6)).astype(np.float64)
    tris = np.concatenate([outer, inner]).reshape(-1, 9)
    ts = np.array([0] * len(outer) + [1] * len(inner), np.int32)
    ign_off = np.array([0, 0, 1], np.int32)
    g = P.single_geometry(tris, (ts, [2, 3], [0, 0]), MATS, ign=(ign_off, [[1.0, 2.0]]))
    assert P.get_solids(g, 0.0, [0.3, 0.2, 0.3]) == ([1], [0])          # inside the shell cavity: default only
    assert P.get_solids(g, 0.0, [0.3, 0.2, -0.005])[0] == [1, 2]        # inside the shell wall
    assert P.get_solids(g, 0.0, [0.52, 0.47, 0.5]) == ([1, 3], [0, 0])    # inside the inner box
    assert P.get_solids(g, 1.5, [0.52, 0.47, 0.5]) == ([1, 3], [0, 1])    # ... during its ignore interval


def test_rotate_vector_orthonormal_frame(P):
    L = P.load()
    rng = np.random.default_rng(3)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    for _ in range(200):
        n = rng.normal(size=3)
        x = rng.normal(size=3)
        # local z maps to n/|n|
        v = np.array([0.0, 0.0, 1.0])
        L.ptref_rotate_vector(dp(v), dp(n), dp(x))
        assert np.allclose(v, n / np.linalg.norm(n), atol=1e-15)
        # local x maps to the normalised projection of x on the plane normal to n
        v = np.array([1.0, 0.0, 0.0])
        L.ptref_rotate_vector(dp(v), dp(n), dp(x))
        xp = x - (x @ n) / (n @ n) * n
        assert np.allclose(v, xp / np.linalg.norm(xp), atol=1e-14)
        # lengths are preserved
        w = rng.normal(size=3)
        w0 = w.copy()
        L.ptref_rotate_vector(dp(w), dp(n), dp(x))
        assert abs(np.linalg.norm(w) - np.linalg.norm(w0)) < 1e-14 * np.linalg.norm(w0)
    # x parallel to n falls back to Plane_3::base1: n = (0,0,1) -> (1,0,0)
    v = np.array([1.0, 0.0, 0.0])
    L.ptref_rotate_vector(dp(v), dp(np.array([0.0, 0.0, 2.0])), dp(np.array([0.0, 0.0, 5.0])))
    assert np.array_equal(v, [1.0, 0.0, 0.0])


def test_segment_bbox_against_sampling(P):
    L = P.load()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rng = np.random.default_rng(4)
    bb = np.array([0.0, 0.0, 0.0, 1.0, 2.0, 0.5])
    for _ in range(2000):
        p, q = rng.uniform(-1, 2.5, (2, 3))
        t = np.linspace(0, 1, 4001)[:, None]
        pts = p + t * (q - p)
        inside = np.all((pts >= bb[:3]) & (pts <= bb[3:]), axis=1).any()
        got = L.ptref_segment_bbox(dp(p), dp(q), dp(bb))
        if inside:
            assert got == 1
        # sampling can miss a grazing intersection; never report one that is far
        if got == 1 and not inside:
            d = np.maximum(np.maximum(bb[:3] - pts, pts - bb[3:]), 0).max(axis=1).min()
            assert d < 2e-3
