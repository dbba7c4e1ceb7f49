"""GPU: the intersection counts the tracking kernels take for chord pieces
(pt_geometry_segment_hits), which come from the clearance grid, from the candidate list of
the piece's clearance cell, from the BVH subtree of the cell's triangles or from the whole
BVH: every way gives the whole BVH's count, and the counts equal the oracle's
TGeometry::GetCollisions restatement (src/trianglemesh.cpp:127-160)."""
import numpy as np
import pytest

from pentrack_amd import queries, workloads

pytestmark = pytest.mark.gpu


def segments_near_walls(rng, n, radius=0.25, height=0.5):
    """Segments of 0 to 4 cm starting within 3 cm of the cylinder's side or caps (some of
    them leaving it), in random directions."""
    phi = rng.uniform(0, 2 * np.pi, n)
    r = radius - rng.uniform(-0.005, 0.03, n)
    z = rng.uniform(0, height, n)
    cap = rng.random(n) < 0.3
    r = np.where(cap, radius * np.sqrt(rng.random(n)), r)
    z = np.where(cap, np.where(rng.random(n) < 0.5, rng.uniform(-0.005, 0.03, n), height - rng.uniform(-0.005, 0.03, n)), z)
    p = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    L = rng.uniform(0, 0.04, n) ** 1.5 / 0.04 ** 0.5
    return p, p + d * L[:, None]


def test_cell_lists_count_as_the_bvh(gpu):
    g = workloads.cylinder_geometry()
    rng = np.random.default_rng(11)
    p, q = segments_near_walls(rng, 200_000)
    print("case: cylinder 19600 triangles, 200000 segments near the walls", flush=True)
    way = queries.segment_hits(g, p, q, mode=2)
    fast = queries.segment_hits(g, p, q, mode=0)
    bvh = queries.segment_hits(g, p, q, mode=1)
    assert np.array_equal(fast, bvh)
    frac = np.bincount(way, minlength=4) / len(way)
    print(f"ways: grid {frac[0]:.3f} cell list {frac[1]:.3f} whole BVH {frac[2]:.3f} cell subtree {frac[3]:.3f}; "
          f"segments with hits {np.mean(bvh > 0):.3f}", flush=True)
    assert frac[1] > 0.15 and frac[0] > 0.05 and frac[2] > 0.0 and frac[3] > 0.05
    assert np.count_nonzero((way == 3) & (bvh > 0)) > 1000
    assert np.count_nonzero((way == 1) & (bvh > 0)) > 1000


def test_segment_counts_match_oracle(gpu, oracle_lib):
    w = workloads.cylinder(n=10, seed=2, n_grid=8)
    rng = np.random.default_rng(12)
    p, q = segments_near_walls(rng, 3000)
    print("case: cylinder, 3000 segments against the oracle", flush=True)
    fast = queries.segment_hits(w.geometry, p, q, mode=0)
    P = oracle_lib.Problem(w.geometry, w.fields)
    for i in range(len(p)):
        n, _ = oracle_lib.collisions(P, p[i], q[i])
        assert fast[i] == (n if n >= 0 else fast[i]) and (n >= 0) == (fast[i] >= 0), (i, n, fast[i])
