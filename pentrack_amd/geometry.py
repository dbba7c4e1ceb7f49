"""TGeometry for the MI355X engine: materials, solids and triangle meshes.

Mirrors the reference's world model (include/geometry.h:20-155,
src/geometry.cpp:10-175, src/trianglemesh.cpp:17-130): materials are the
nine numbers of `material`, solids carry an ID, a material and ignore
intervals, the default solid (ID 1) has no mesh, and meshes are loaded in
lexicographic GEOMETRY key order.  The device BVH is built by
pt_geometry_create in the native library.
"""
import ctypes as C
import struct
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from . import _native

MATERIAL_FIELDS = ("FermiReal", "FermiImag", "DiffProb", "SpinflipProb", "RMSRoughness",
                   "CorrelLength", "InternalBField", "ModifiedLambertProb", "LossPerBounce")


@dataclass
class Material:
    """struct material (include/geometry.h:20-31); FermiReal/FermiImag in neV."""
    name: str
    FermiReal: float = 0.0
    FermiImag: float = 0.0
    DiffProb: float = 0.0
    SpinflipProb: float = 0.0
    RMSRoughness: float = 0.0
    CorrelLength: float = 0.0
    InternalBField: float = 0.0
    ModifiedLambertProb: float = 0.0
    LossPerBounce: float = 0.0

    def __post_init__(self):
        # range checks of operator>>(istream, material), src/geometry.cpp:10-52
        if not 0 <= self.DiffProb <= 1:
            raise ValueError(f"diffuse-reflection probability outside 0..1 for material {self.name}")
        if not 0 <= self.SpinflipProb <= 1:
            raise ValueError(f"spin-flip probability outside 0..1 for material {self.name}")
        if not 0 <= self.LossPerBounce <= 1:
            raise ValueError(f"loss-per-bounce probability outside 0..1 for material {self.name}")
        models = (self.DiffProb != 0) + (self.RMSRoughness != 0 or self.CorrelLength != 0) + (self.ModifiedLambertProb != 0)
        if models > 1:
            raise ValueError(f"more than one diffuse-reflection model set for material {self.name}")

    def as_array(self):
        return np.array([getattr(self, f) for f in MATERIAL_FIELDS], dtype=np.float64)


@dataclass
class Solid:
    """struct solid (include/geometry.h:36-51)."""
    ID: int
    material: str
    triangles: np.ndarray = None          # [n][3][3] float32, STL vertex order
    ignoretimes: List[Tuple[float, float]] = field(default_factory=list)
    name: str = ""

    def is_ignored(self, t):
        return any(a <= t < b for a, b in self.ignoretimes)


def repair_soup(tris):
    """The polygon-soup repair TTriangleMesh::ReadFile applies to every STL file
    (src/trianglemesh.cpp:63-64): PMP::repair_polygon_soup then PMP::orient_polygon_soup of
    the vendored CGAL 4.14.1, restated for triangles.

    tris: [n][3][3] float32 after the 1e-8 vertex snap.  Returns (triangles, source index of
    each kept triangle).  The steps, in CGAL's order:
      * merge_duplicate_points_in_polygon_soup (repair_polygon_soup.h:511-582): vertices are
        identified by exact coordinates;
      * simplify_polygons_in_polygon_soup and remove_invalid_polygons_in_polygon_soup
        (:123-243, :339-372): a triangle with two identical vertices collapses below three
        vertices and is erased, keeping the order of the others (triangles have no pinched
        form, :246-337);
      * merge_duplicate_polygons_in_polygon_soup (:906-1003): triangles on the same three
        vertices, in either orientation, are duplicates; the first of a group stays, the
        others are swapped one by one to the end of the list and cut off.  The order in which
        several groups are visited is that of an unordered_map in CGAL (parity unpinned with
        more than one group); here the group with the largest first index goes first;
      * remove_isolated_points_in_polygon_soup changes no triangle;
      * Polygon_soup_orienter::orient (orient_polygon_soup.h:188-284): from the first
        triangle not yet oriented, a depth-first walk over edges shared by at most two
        triangles; a neighbour that runs the shared edge in the same direction is reversed,
        one that runs it the other way is taken as it is; an edge met by two already oriented
        triangles of the same direction is not crossed.  The first triangle of each connected
        component keeps its orientation (the soup is made consistent, not outward).  The
        duplication of singular vertices that follows changes no triangle.
    """
    t = np.asarray(tris, dtype=np.float32).reshape(-1, 3, 3)
    n = len(t)
    _, vid = np.unique(t.reshape(-1, 3), axis=0, return_inverse=True)
    F = vid.reshape(n, 3).astype(np.int64)
    keep = (F[:, 0] != F[:, 1]) & (F[:, 1] != F[:, 2]) & (F[:, 0] != F[:, 2])
    src = np.nonzero(keep)[0]
    F = [list(f) for f in F[keep]]
    # duplicate triangles
    groups = {}
    for k, f in enumerate(F):
        groups.setdefault(tuple(sorted(f)), []).append(k)
    dups = [g for g in groups.values() if len(g) > 1]
    if dups:
        m = len(F)
        pos = list(range(m))          # position of each polygon id
        at = list(range(m))           # polygon id at each position
        order = list(F)
        srcl = list(src)
        swap = m - 1
        for g in sorted(dups, key=lambda g: g[0], reverse=True):      # largest first index first
            for pid in g[1:]:
                p = pos[pid]
                other = at[swap]
                pos[other], pos[pid] = p, swap
                at[p], at[swap] = other, pid
                order[swap], order[p] = order[p], order[swap]
                srcl[swap], srcl[p] = srcl[p], srcl[swap]
                swap -= 1
        F, src = order[:swap + 1], np.asarray(srcl[:swap + 1])
    # orientation
    m = len(F)
    edges = {}
    for k, f in enumerate(F):
        for j in range(3):
            edges.setdefault((f[j], f[(j + 1) % 3]), set()).add(k)
    marked = set()
    for f in F:
        for j in range(3):
            a, b = f[j], f[(j + 1) % 3]
            if len(edges.get((a, b), ())) + len(edges.get((b, a), ())) > 2:
                marked.add((min(a, b), max(a, b)))
    oriented = [False] * m
    for start in range(m):
        if oriented[start]:
            continue
        oriented[start] = True
        stack = [start]
        while stack:
            cur = stack.pop()
            for j in range(3):
                i1, i2 = F[cur][j], F[cur][(j + 1) % 3]
                if (min(i1, i2), max(i1, i2)) in marked:
                    continue
                same = edges[(i1, i2)]
                if len(same) > 1:
                    idx = min(same)
                    if idx == cur:
                        idx = sorted(same)[1]
                    if oriented[idx]:
                        marked.add((min(i1, i2), max(i1, i2)))
                        continue
                    f = F[idx]
                    for q in range(3):
                        e = (f[q], f[(q + 1) % 3])
                        edges[e].discard(idx)
                        if not edges[e]:
                            del edges[e]
                    F[idx] = f[::-1]
                    f = F[idx]
                    for q in range(3):
                        edges.setdefault((f[q], f[(q + 1) % 3]), set()).add(idx)
                    oriented[idx] = True
                    stack.append(idx)
                else:
                    other = edges.get((i2, i1))
                    if other:
                        idx = min(other)
                        if not oriented[idx]:
                            oriented[idx] = True
                            stack.append(idx)
    # triangles from the (possibly reversed) vertex cycles: a reversed triangle (a, b, c)
    # is (c, b, a) in CGAL's std::reverse
    out = np.empty((m, 3, 3), dtype=np.float32)
    uniq = np.unique(t.reshape(-1, 3), axis=0)
    for k, f in enumerate(F):
        out[k] = uniq[f]
    return out, np.asarray(src, dtype=np.int64)


def read_stl(path, repair=True):
    """Binary STL reader (TTriangleMesh::ReadFile, src/trianglemesh.cpp:17-64): the 1e-8 vertex
    snap and, with repair, the polygon-soup repair and orientation of repair_soup.

    Returns (header name, [n][3][3] float32 vertices); normals in the file are ignored.
    """
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < 84:
        raise ValueError(f"{path}: not a binary STL file")
    header = data[:80].decode("latin-1").rstrip(" ")
    (count,) = struct.unpack("<I", data[80:84])
    if count == 0:
        raise ValueError(f"{path} contains no triangles")
    rec = np.dtype([("n", "<f4", (3,)), ("v", "<f4", (3, 3)), ("attr", "<u2")])
    if 84 + count * rec.itemsize > len(data):
        raise ValueError(f"{path} should contain {count} triangles but is truncated")
    arr = np.frombuffer(data, dtype=rec, count=count, offset=84)
    tris = np.ascontiguousarray(arr["v"], dtype=np.float32)
    if repair:
        tris = np.where(np.abs(tris) < 1e-8, np.float32(0), tris)
        tris, _ = repair_soup(tris)
    return header, tris


def write_stl(path, tris, name="pentrack_amd"):
    tris = np.asarray(tris, dtype=np.float32).reshape(-1, 3, 3)
    rec = np.zeros(len(tris), dtype=[("n", "<f4", (3,)), ("v", "<f4", (3, 3)), ("attr", "<u2")])
    rec["v"] = tris
    with open(path, "wb") as f:
        f.write(name.encode("latin-1")[:80].ljust(80, b" "))
        f.write(struct.pack("<I", len(tris)))
        f.write(rec.tobytes())


def snap(tris):
    """Vertex snap |v| < 1e-8 -> 0 (src/trianglemesh.cpp:42), as float64."""
    t = np.asarray(tris, dtype=np.float32).astype(np.float64)
    t[np.abs(t) < 1e-8] = 0.0
    return t


# ---------------------------------------------------------------------------
# mesh generators (outward normals: counter-clockwise seen from outside)
# ---------------------------------------------------------------------------

def box_mesh(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """12-triangle axis-aligned box with outward normals."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    P = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], dtype=np.float64)
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (1, 2, 6, 5), (0, 4, 7, 3)]
    tris = []
    for a, b, c, d in quads:
        tris.append([P[a], P[b], P[c]])
        tris.append([P[a], P[c], P[d]])
    return np.array(tris, dtype=np.float32)


def hollow_box_mesh(inner_lo=(0, 0, 0), inner_hi=(1, 1, 1), wall=0.001):
    """A closed shell: outer box with outward normals plus inner box with inward normals (24 triangles)."""
    outer = box_mesh(tuple(v - wall for v in inner_lo), tuple(v + wall for v in inner_hi))
    inner = box_mesh(inner_lo, inner_hi)[:, ::-1, :]
    return np.concatenate([outer, inner]).astype(np.float32)


def cylinder_mesh(radius, height, n_phi, n_z, n_r, z0=0.0):
    """Closed cylinder (axis z, base at z0) with outward normals.

    Side: n_phi x n_z quads; caps: n_r concentric rings (innermost a fan).
    Triangle count: 2*n_phi*n_z + 2*(n_phi + 2*n_phi*(n_r-1)).
    """
    phi = np.linspace(0.0, 2 * np.pi, n_phi, endpoint=False)
    c, s = np.cos(phi), np.sin(phi)
    zs = np.linspace(z0, z0 + height, n_z + 1)
    tris = []
    for k in range(n_z):
        for i in range(n_phi):
            j = (i + 1) % n_phi
            a = (radius * c[i], radius * s[i], zs[k])
            b = (radius * c[j], radius * s[j], zs[k])
            cc = (radius * c[j], radius * s[j], zs[k + 1])
            d = (radius * c[i], radius * s[i], zs[k + 1])
            tris.append([a, b, cc])
            tris.append([a, cc, d])
    rs = np.linspace(0.0, radius, n_r + 1)
    for zc, up in ((z0, False), (z0 + height, True)):
        for r in range(n_r):
            for i in range(n_phi):
                j = (i + 1) % n_phi
                p_in_i = (rs[r] * c[i], rs[r] * s[i], zc)
                p_in_j = (rs[r] * c[j], rs[r] * s[j], zc)
                p_out_i = (rs[r + 1] * c[i], rs[r + 1] * s[i], zc)
                p_out_j = (rs[r + 1] * c[j], rs[r + 1] * s[j], zc)
                if r == 0:
                    t = [p_in_i, p_out_i, p_out_j]
                    tris.append(t if up else t[::-1])
                else:
                    t1 = [p_in_i, p_out_i, p_out_j]
                    t2 = [p_in_i, p_out_j, p_in_j]
                    tris.append(t1 if up else t1[::-1])
                    tris.append(t2 if up else t2[::-1])
    return np.array(tris, dtype=np.float32)


class TGeometry:
    """World model: materials, the default solid and meshed solids.

    solids are kept in lexicographic order of their IDs as strings, the order in
    which the reference iterates its GEOMETRY config section (include/config.h:32,
    src/geometry.cpp:115).
    """

    def __init__(self, materials, default_material, solids):
        self.materials = {m.name: m for m in materials}
        if default_material not in self.materials:
            raise ValueError(f"Material {default_material} used but not defined!")
        self.default_material = default_material
        ids = [s.ID for s in solids]
        if len(set(ids)) != len(ids):
            raise ValueError("You defined solids with identical ID! IDs have to be unique!")
        for s in solids:
            if s.ID <= 1:
                raise ValueError("solid IDs of meshes must be > 1 (ID 1 is the default solid)")
            if s.material not in self.materials:
                raise ValueError(f"Material {s.material} used but not defined!")
            if s.triangles is None or len(s.triangles) == 0:
                raise ValueError(f"solid {s.ID} has no triangles")
        self.solids = sorted(solids, key=lambda s: str(s.ID))
        self._handle = None

    # flattened arrays, shared by the native library and the oracle
    def arrays(self):
        names = list(self.materials)
        mats = np.stack([self.materials[n].as_array() for n in names])
        tri = np.concatenate([np.asarray(s.triangles, dtype=np.float32).reshape(-1, 3, 3) for s in self.solids])
        tri_solid = np.concatenate([np.full(len(s.triangles), k, dtype=np.int32) for k, s in enumerate(self.solids)])
        solid_id = np.array([s.ID for s in self.solids], dtype=np.uint32)
        solid_mat = np.array([names.index(s.material) for s in self.solids], dtype=np.int32)
        ign_off = np.zeros(len(self.solids) + 1, dtype=np.int32)
        ign = []
        for k, s in enumerate(self.solids):
            ign.extend(s.ignoretimes)
            ign_off[k + 1] = len(ign)
        ign = np.array(ign, dtype=np.float64).reshape(-1, 2)
        return dict(tri=np.ascontiguousarray(tri), tri_solid=tri_solid, solid_id=solid_id, solid_mat=solid_mat,
                    ign_off=ign_off, ign=np.ascontiguousarray(ign), default_mat=names.index(self.default_material),
                    mats=np.ascontiguousarray(mats))

    def n_triangles(self):
        return sum(len(s.triangles) for s in self.solids)

    def device(self, leaf_size=4):
        """Native handle (uploads the BVH to the GPU); one per BVH leaf size, kept until the
        geometry is released, so batches built on any of them stay valid."""
        if self._handle is None:
            self._handle = {}
        if leaf_size in self._handle:
            return self._handle[leaf_size]
        L = _native.lib()
        a = self.arrays()
        self._keep = a
        h = C.c_void_p()
        rc = L.pt_geometry_create(
            len(a["tri"]), a["tri"].ctypes.data_as(C.POINTER(C.c_float)),
            a["tri_solid"].ctypes.data_as(C.POINTER(C.c_int32)), len(a["solid_id"]),
            a["solid_id"].ctypes.data_as(C.POINTER(C.c_uint32)),
            a["solid_mat"].ctypes.data_as(C.POINTER(C.c_int32)),
            a["ign_off"].ctypes.data_as(C.POINTER(C.c_int32)),
            a["ign"].ctypes.data_as(C.POINTER(C.c_double)), a["default_mat"], len(a["mats"]),
            a["mats"].ctypes.data_as(C.POINTER(C.c_double)), leaf_size, C.byref(h))
        _native.check(rc, "pt_geometry_create")
        self._handle[leaf_size] = h
        return h

    def __del__(self):
        try:
            if self._handle and _native._lib is not None:
                for h in self._handle.values():
                    _native._lib.pt_geometry_destroy(h)
        except Exception:
            pass
