"""Field and solid queries on the GPU through the C ABI (pt_fields_eval, pt_geometry_solid_at)."""
import ctypes as C

import numpy as np

from . import _native


def fields_eval(fields, pts, t, stream=None, grad=True):
    """TFieldManager::BField / EField at pts [n][3], times t (scalar or [n]): B [n][3],
    dB [n][3][3], V [n], E [n][3].  grad False: the value-only BField(x, y, z, t, B) the
    reference uses for energies and logs, dB returned as None."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    n = len(pts)
    t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (n,)))
    B, dB, V, E = np.zeros((n, 3)), np.zeros((n, 3, 3)), np.zeros(n), np.zeros((n, 3))
    d = _native.dptr
    _native.check(_native.lib().pt_fields_eval(fields.device(), stream, n, d(pts), d(t), d(B), d(dB) if grad else None,
                                               d(V), d(E)), "pt_fields_eval")
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
