"""Field, solid and segment queries on the GPU through the C ABI (pt_fields_eval,
pt_geometry_solid_at, pt_geometry_segment_hits)."""
import ctypes as C

import numpy as np

from . import _native


def fields_eval(fields, pts, t, stream=None, grad=True, variant=None, cooperative=False):
    """TFieldManager::BField / EField at pts [n][3], times t (scalar or [n]): B [n][3],
    dB [n][3][3], V [n], E [n][3].  grad False: the value-only BField(x, y, z, t, B) the
    reference uses for energies and logs, dB returned as None.  variant (PT_VARIANT_FP64 or
    PT_VARIANT_MIXED) and cooperative select the tables and the coefficient gathers the
    tracking kernels use (pt_fields_eval_tables)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    n = len(pts)
    t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (n,)))
    B, dB, V, E = np.zeros((n, 3)), np.zeros((n, 3, 3)), np.zeros(n), np.zeros((n, 3))
    d = _native.dptr
    if variant is None and not cooperative:
        _native.check(_native.lib().pt_fields_eval(fields.device(), stream, n, d(pts), d(t), d(B),
                                                   d(dB) if grad else None, d(V), d(E)), "pt_fields_eval")
    else:
        v = _native.PT_VARIANT_FP64 if variant is None else variant
        dev = fields.device(with_mixed=True) if v == _native.PT_VARIANT_MIXED else fields.device()
        _native.check(_native.lib().pt_fields_eval_tables(dev, stream, n, d(pts), d(t), v, 1 if cooperative else 0, d(B),
                                                          d(dB) if grad else None, d(V), d(E)), "pt_fields_eval_tables")
    return B, dB if grad else None, V, E


def solid_at(geometry, pts, t, stream=None, leaf_size=4):
    """TGeometry::GetSolid at pts [n][3] and times t: solid IDs [n] (1 = default solid)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    n = len(pts)
    t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (n,)))
    out = np.zeros(n, np.int32)
    d = _native.dptr
    _native.check(_native.lib().pt_geometry_solid_at(geometry.device(leaf_size), stream, n, d(pts), d(t),
                                                     out.ctypes.data_as(C.POINTER(C.c_int32))), "pt_geometry_solid_at")
    return out


def segment_hits(geometry, p, q, mode=0, stream=None, leaf_size=4):
    """Intersection counts of the segments p[i] -> q[i] ([n][3] each) with the mesh, as
    TTracker::CheckHit takes them from TGeometry::GetCollisions (-1 coplanar overlap, -2 more
    than 16 intersections).  mode 0 as the tracking kernels count, 1 from the BVH alone, 2
    which way mode 0 took (0 clearance grid, 1 cell candidate list, 2 BVH)."""
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    if p.shape != q.shape:
        raise ValueError("p and q must have the same shape")
    out = np.zeros(len(p), np.int32)
    d = _native.dptr
    _native.check(_native.lib().pt_geometry_segment_hits(geometry.device(leaf_size), stream, len(p), d(p), d(q), mode,
                                                         out.ctypes.data_as(C.POINTER(C.c_int32))),
                  "pt_geometry_segment_hits")
    return out
