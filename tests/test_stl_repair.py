"""The polygon-soup repair of the STL front end (PMP::repair_polygon_soup and
PMP::orient_polygon_soup as TTriangleMesh::ReadFile calls them, src/trianglemesh.cpp:63-64):
known answers on small soups for each step, and the identity on the reference's own meshes
(every mesh under in/ and test/ is closed, manifold and consistently oriented; the fixtures
here are test/HollowUnitCube.STL, in/guide_experiment.STL and
in/storagevolume_60degfeeder_bellow.STL)."""
import os

import numpy as np
import pytest

from pentrack_amd.geometry import box_mesh, read_stl, repair_soup, write_stl

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def normals(t):
    t = np.asarray(t, np.float64)
    return np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])


def outward(t):
    """Signed volume of the closed mesh (> 0: normals point outward)."""
    t = np.asarray(t, np.float64)
    return np.einsum("ij,ij->i", t[:, 0], np.cross(t[:, 1], t[:, 2])).sum() / 6


def test_identity_on_reference_meshes():
    for f in ("HollowUnitCube.STL", "guide_experiment.STL", "storagevolume_60degfeeder_bellow.STL"):
        _, raw = read_stl(os.path.join(GOLDEN, f), repair=False)
        raw = np.where(np.abs(raw) < 1e-8, np.float32(0), raw)
        rep, src = repair_soup(raw)
        assert np.array_equal(rep, raw) and np.array_equal(src, np.arange(len(raw))), f


def test_flipped_face_is_reoriented():
    box = box_mesh((0, 0, 0), (1, 2, 3))
    soup = box.copy()
    soup[5] = soup[5][::-1]
    rep, src = repair_soup(soup)
    assert np.array_equal(rep, box) and np.array_equal(src, np.arange(12))
    assert outward(rep) == pytest.approx(6.0)


def test_first_face_sets_the_orientation():
    """The soup is made consistent with its first triangle, not made outward: flipping the
    first face of a closed box turns every other face around."""
    box = box_mesh((0, 0, 0), (1, 1, 1))
    soup = box.copy()
    soup[0] = soup[0][::-1]
    rep, _ = repair_soup(soup)
    assert outward(rep) == pytest.approx(-1.0)
    assert np.allclose(normals(rep), -normals(box))


def test_components_are_oriented_separately():
    a, b = box_mesh((0, 0, 0), (1, 1, 1)), box_mesh((3, 0, 0), (4, 1, 1))
    b_in = b[:, ::-1, :]                     # second box entirely inward
    soup = np.concatenate([a, b_in])
    soup[7] = soup[7][::-1]                  # and one face of the first flipped
    soup[14] = soup[14][::-1]                # and one face of the second
    rep, _ = repair_soup(soup)
    assert outward(rep[:12]) == pytest.approx(1.0) and outward(rep[12:]) == pytest.approx(-1.0)
    assert np.array_equal(rep[:12], a) and np.array_equal(rep[12:], b_in)


def test_degenerate_faces_removed_in_order():
    box = box_mesh((0, 0, 0), (1, 1, 1))
    v = box[2, 0]
    degen = np.array([[v, v, box[2, 2]], [box[3, 1], box[4, 2], box[3, 1]]], np.float32)
    soup = np.concatenate([box[:4], degen[:1], box[4:9], degen[1:], box[9:]])
    rep, src = repair_soup(soup)
    assert np.array_equal(rep, box)
    assert list(src) == [0, 1, 2, 3] + list(range(5, 10)) + list(range(11, 14))
    # collinear but distinct vertices are not removed by the repair
    col = np.array([[[0, 0, 0], [1, 0, 0], [2, 0, 0]]], np.float32)
    rep2, _ = repair_soup(np.concatenate([box, col]))
    assert len(rep2) == 13


def test_duplicate_faces_removed_by_swap():
    """merge_duplicate_polygons_in_polygon_soup keeps the first of each group and swaps the
    others to the end: [A, A', B, C, D] becomes [A, D, B, C]."""
    box = box_mesh((0, 0, 0), (1, 1, 1))
    A, B, C, D = box[0], box[1], box[2], box[3]
    soup = np.stack([A, A[::-1], B, C, D])   # A' has A's vertices in the other orientation
    rep, src = repair_soup(soup)
    assert list(src) == [0, 4, 2, 3]
    assert np.array_equal(rep[0], A) and np.array_equal(rep[1], D)
    # a duplicate already at the end is simply cut off
    rep, src = repair_soup(np.stack([A, B, C, A]))
    assert list(src) == [0, 1, 2]


def test_nonmanifold_edge_is_not_crossed():
    """Three triangles on one edge: the edge is marked and the walk does not cross it, so a
    wrongly oriented third fin stays as it is while its manifold neighbours follow."""
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1]], np.float32)
    f0 = p[[0, 1, 2]]
    f1 = p[[1, 0, 3]]          # consistent with f0 across edge 0-1
    f2 = p[[0, 1, 4]]          # a fin on the same edge, same direction as f0
    rep, _ = repair_soup(np.stack([f0, f1, f2]))
    assert np.array_equal(rep[0], f0) and np.array_equal(rep[1], f1) and np.array_equal(rep[2], f2)


def test_read_stl_applies_the_repair(tmp_path):
    box = box_mesh((0, 0, 0), (1, 1, 1))
    soup = box.copy()
    soup[9] = soup[9][::-1]
    path = os.path.join(tmp_path, "flip.stl")
    write_stl(path, soup)
    _, tris = read_stl(path)
    assert np.array_equal(tris, box)
    _, raw = read_stl(path, repair=False)
    assert np.array_equal(raw, soup)
