"""Clearance grid (host side of the collision-query skip): every cell's bound is a
lower bound of the distance from any point of the cell to the nearest triangle.

Pinned by brute force: random points of random cells against the exact
point-triangle distance over all triangles (numpy, dense sampling of each
triangle is avoided by the closed-form projection onto plane, edges and vertices).
"""
import ctypes as C

import numpy as np

from pentrack_amd import _native
from pentrack_amd.geometry import box_mesh, cylinder_mesh


def seg_dist(p, a, b):
    ab = b - a
    t = np.clip(((p - a) @ ab) / (ab @ ab), 0, 1)
    return np.linalg.norm(p - (a + t[:, None] * ab), axis=1)


def tri_dist(p, T):
    a, b, c = T
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n)
    d = (p - a) @ n
    q = p - d[:, None] * n
    # barycentric inside test of the projection
    v0, v1, v2 = b - a, c - a, q - a
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    d20, d21 = v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    inside = (v >= 0) & (w >= 0) & (v + w <= 1)
    e = np.minimum(np.minimum(seg_dist(p, a, b), seg_dist(p, b, c)), seg_dist(p, c, a))
    return np.where(inside, np.abs(d), e)


def clearance(tris, cells):
    L = _native.lib()
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    gn = (C.c_int32 * 3)()
    lo = (C.c_double * 3)()
    ih = (C.c_double * 3)()
    tp = t.ctypes.data_as(C.POINTER(C.c_float))
    assert L.pt_clearance_grid(len(t), tp, cells, None, gn, lo, ih) == 0
    out = np.zeros(gn[0] * gn[1] * gn[2], np.float32)
    assert L.pt_clearance_grid(len(t), tp, cells, out.ctypes.data_as(C.POINTER(C.c_float)), gn, lo, ih) == 0
    return out.reshape(tuple(gn)), np.array(lo[:]), np.array(ih[:])


def check(tris, cells, npts, seed):
    grid, lo, ih = clearance(tris, cells)
    T = np.asarray(tris, np.float32).astype(np.float64).reshape(-1, 3, 3)
    rng = np.random.default_rng(seed)
    idx = np.stack([rng.integers(0, g, npts) for g in grid.shape], 1)
    p = lo + (idx + rng.uniform(0, 1, (npts, 3))) / ih
    d = np.full(npts, np.inf)
    for tri in T:
        d = np.minimum(d, tri_dist(p, tri))
    bound = grid[idx[:, 0], idx[:, 1], idx[:, 2]].astype(np.float64)
    assert np.all(bound <= d), (bound - d).max()
    return bound, d


def test_clearance_is_lower_bound_box():
    bound, d = check(box_mesh((0, 0, 0), (1, 1, 1)), 32, 4000, 0)
    # and it is not trivially zero: cells away from the walls carry most of their distance
    far = d > 0.1
    assert np.all(bound[far] >= d[far] - np.sqrt(3) / 32 - 1e-6)  # at most one cell diagonal lost


def test_clearance_is_lower_bound_cylinder():
    tris = cylinder_mesh(0.25, 0.5, 60, 8, 4)
    bound, d = check(tris, 48, 3000, 1)
    assert (bound > 0).mean() > 0.5
