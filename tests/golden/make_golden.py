"""Writes tests/golden/reference_islands.npz: input/output vectors of the reference's
own Boost-free numerical code (ALGLIB 3.15.0 spline1dgriddiffcubic, libtricubic
tricubic_get_coeff / tricubic_eval, the TabField3::PreInterpol call sequence on them,
and TabField's ALGLIB spline2dbuildbicubicv with spline2dcalc / spline2ddiff), computed
with oracle/_ref/libref_islands.so built by oracle/ref_build.sh from the sources under
/root/reference.  Run here, where the reference is mounted:

    bash oracle/ref_build.sh && python tests/golden/make_golden.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_islands.so")


def load():
    L = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    L.ref_spline_diff.argtypes = [C.c_int, dp, dp, dp]
    L.ref_tricubic_coeff.argtypes = [dp, dp]
    L.ref_tricubic_eval.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
    L.ref_tricubic_eval.restype = C.c_double
    L.ref_preinterpol.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp]
    L.ref_spline2d.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_int, dp, dp, C.c_int, dp, dp, dp]
    return L


def spline2d(L, x, y, values, px, py, diff):
    """ALGLIB bicubic spline of values[len(y)][len(x)] at the points: F (and dF/dx, dF/dy)."""
    n = len(px)
    F, Fx, Fy = np.zeros(n), np.zeros(n), np.zeros(n)
    L.ref_spline2d(len(x), len(y), p(x), p(y), p(values), n, p(px), p(py), int(diff),
                   F.ctypes.data_as(C.POINTER(C.c_double)), Fx.ctypes.data_as(C.POINTER(C.c_double)),
                   Fy.ctypes.data_as(C.POINTER(C.c_double)))
    return np.stack([F, Fx, Fy]) if diff else F


def p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def spline_diff(L, x, y):
    d = np.zeros(len(x))
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    L.ref_spline_diff(len(x), p(x), p(y), d.ctypes.data_as(C.POINTER(C.c_double)))
    return d


def tricubic_coeff(L, corner):
    a = np.zeros(64)
    c = np.ascontiguousarray(corner, np.float64)
    L.ref_tricubic_coeff(p(c), a.ctypes.data_as(C.POINTER(C.c_double)))
    return a


def preinterpol(L, xg, yg, zg, values):
    nx, ny, nz = len(xg), len(yg), len(zg)
    coef = np.zeros(((nx - 1) * (ny - 1) * (nz - 1), 64))
    v = np.ascontiguousarray(values, np.float64)
    L.ref_preinterpol(nx, ny, nz, p(xg), p(yg), p(zg), p(v), coef.ctypes.data_as(C.POINTER(C.c_double)))
    return coef


def main():
    L = load()
    rng = np.random.default_rng(20261203)
    out = {}
    # spline derivatives on non-uniform grids of several lengths (n = 2 and 3 hit the end cases)
    for n in (2, 3, 4, 7, 12):
        x = np.cumsum(rng.uniform(0.1, 1.0, n))
        y = rng.normal(size=n)
        out[f"spline_x_{n}"], out[f"spline_y_{n}"], out[f"spline_d_{n}"] = x, y, spline_diff(L, x, y)
    # tricubic coefficients from random corner data, and evaluations at random points
    corners = rng.normal(size=(40, 64))
    out["tric_corner"] = corners
    out["tric_coef"] = np.stack([tricubic_coeff(L, c) for c in corners])
    pts = rng.uniform(0, 1, size=(40, 3))
    ders = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    out["tric_pts"] = pts
    out["tric_eval"] = np.array([[L.ref_tricubic_eval(p(out["tric_coef"][k]), *pts[k], *d) for d in ders]
                                 for k in range(len(pts))])
    # PreInterpol on a small non-uniform 3-D grid
    xg = np.cumsum(rng.uniform(0.05, 0.2, 6)) - 0.3
    yg = np.cumsum(rng.uniform(0.05, 0.2, 5))
    zg = np.cumsum(rng.uniform(0.05, 0.2, 4)) + 0.1
    vals = rng.normal(size=(6, 5, 4))
    out["pre_x"], out["pre_y"], out["pre_z"], out["pre_values"] = xg, yg, zg, vals
    out["pre_coef"] = preinterpol(L, xg, yg, zg, vals)
    # bicubic splines of 2D maps on non-uniform grids (m = 2 and n = 2 hit the end cases):
    # random points, the grid nodes themselves and points just outside the grid
    for tag, (m, n) in (("a", (7, 5)), ("b", (2, 4)), ("c", (3, 2))):
        x = np.cumsum(rng.uniform(0.05, 0.3, m)) - 0.1
        y = np.cumsum(rng.uniform(0.05, 0.3, n)) + 0.2
        v = rng.normal(size=(n, m))
        X, Y = np.meshgrid(x, y)
        px = np.concatenate([rng.uniform(x[0], x[-1], 60), X.ravel(), [x[0] - 0.01, x[-1] + 0.01]])
        py = np.concatenate([rng.uniform(y[0], y[-1], 60), Y.ravel(), [y[-1] + 0.01, y[0] - 0.01]])
        out[f"s2d_x_{tag}"], out[f"s2d_y_{tag}"], out[f"s2d_v_{tag}"] = x, y, v
        out[f"s2d_px_{tag}"], out[f"s2d_py_{tag}"] = px, py
        out[f"s2d_calc_{tag}"] = spline2d(L, x, y, v, px, py, 0)
        out[f"s2d_diff_{tag}"] = spline2d(L, x, y, v, px, py, 1)
    np.savez_compressed(os.path.join(HERE, "reference_islands.npz"), **out)
    print("wrote", os.path.join(HERE, "reference_islands.npz"))


if __name__ == "__main__":
    main()
