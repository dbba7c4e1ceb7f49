"""Tricubic field maps and field containers.

Pinned by exact rational evaluation of the Horner form, by the 64x64
Lekien-Marsden matrix inverting the corner-constraint matrix exactly, by exact
reproduction of fields that are quadratic along each axis (the parabolic end
conditions of the spline derivatives are exact for them), and by the on-axis
closed form of a current loop.
"""
import ctypes as C
from fractions import Fraction as Fr
from itertools import product

import numpy as np
import pytest

from pentrack_amd import _native
from pentrack_amd.fields import TFieldManager, FieldContainer, loop_field, tabfield3, uniform_efield
from pentrack_amd.constants import MU0


@pytest.fixture(scope="module")
def P(oracle_lib):
    return oracle_lib


def test_horner_matches_exact_polynomial(P):
    L = P.load()
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = rng.normal(size=64)
        x, y, z = rng.uniform(0, 1, 3)
        ap = a.ctypes.data_as(C.POINTER(C.c_double))
        for d in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]:
            got = L.ptref_tricubic_eval(ap, x, y, z, *d)
            exact = Fr(0)
            for i, j, k in product(range(4), repeat=3):
                if i < d[0] or j < d[1] or k < d[2]:
                    continue
                c = Fr(float(a[i + 4 * j + 16 * k]))
                c *= (i if d[0] else 1) * (j if d[1] else 1) * (k if d[2] else 1)
                c *= Fr(float(x)) ** (i - d[0]) * Fr(float(y)) ** (j - d[1]) * Fr(float(z)) ** (k - d[2])
                exact += c
            assert abs(got - float(exact)) <= 1e-14 * (1 + abs(float(exact)))


def test_tricubic_matrix_inverts_corner_constraints():
    L = _native.lib()
    A = np.zeros(64 * 64, np.int32)
    L.pt_tricubic_matrix(A.ctypes.data_as(C.POINTER(C.c_int32)))
    A = A.reshape(64, 64).astype(np.int64)
    ders = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]

    def dmono(p, m, x):
        if p < m:
            return 0
        c = 1
        for q in range(m):
            c *= p - q
        return c * (x ** (p - m))

    M = np.zeros((64, 64), np.int64)
    for d, (dx, dy, dz) in enumerate(ders):
        for c in range(8):
            cx, cy, cz = c & 1, (c >> 1) & 1, (c >> 2) & 1
            for i, j, k in product(range(4), repeat=3):
                M[8 * d + c, i + 4 * j + 16 * k] = dmono(i, dx, cx) * dmono(j, dy, cy) * dmono(k, dz, cz)
    assert np.array_equal(A @ M, np.eye(64, dtype=np.int64))


def quad_field(x, y, z):
    # quadratic along each axis (tri-quadratic), exactly reproduced by the map
    return 0.3 + 1.5 * x - 0.7 * y * y + 0.25 * x * z + 0.9 * z * z * y - 0.2 * x * x * y * z


def test_map_reproduces_triquadratic_fields_and_gradients(P):
    xg = np.linspace(-0.2, 0.3, 7)
    yg = np.array([0.0, 0.05, 0.13, 0.2, 0.31, 0.4])   # non-uniform grid
    zg = np.linspace(0.1, 0.5, 5)
    X, Y, Z = np.meshgrid(xg, yg, zg, indexing="ij")
    B = np.stack([quad_field(X, Y, Z), -quad_field(Y, Z, X), quad_field(Z, X, Y)])
    fm = TFieldManager([tabfield3(xg, yg, zg, B=B)])
    prob = P.Problem(_dummy_geometry(), fm)
    rng = np.random.default_rng(1)
    pts = np.stack([rng.uniform(-0.2, 0.3, 500), rng.uniform(0, 0.4, 500), rng.uniform(0.1, 0.5, 500)], 1)
    Bg, dB, V, E = P.field_eval(prob, pts)
    x, y, z = pts.T
    ref = np.stack([quad_field(x, y, z), -quad_field(y, z, x), quad_field(z, x, y)], 1)
    assert np.abs(Bg - ref).max() < 1e-12
    h = 1e-6
    for a in range(3):
        dp = pts.copy(); dp[:, a] += h
        dm = pts.copy(); dm[:, a] -= h
        fd = (np.stack([quad_field(*dp.T), -quad_field(dp[:, 1], dp[:, 2], dp[:, 0]), quad_field(dp[:, 2], dp[:, 0], dp[:, 1])], 1)
              - np.stack([quad_field(*dm.T), -quad_field(dm[:, 1], dm[:, 2], dm[:, 0]), quad_field(dm[:, 2], dm[:, 0], dm[:, 1])], 1)) / (2 * h)
        assert np.abs(dB[:, :, a] - fd).max() < 1e-8
    # outside the grid the map contributes nothing (TabField3::Interpolate returns early)
    Bo, dBo, _, _ = P.field_eval(prob, np.array([[0.3, 0.2, 0.3], [-0.25, 0.2, 0.3], [0.0, 0.41, 0.3]]))
    assert np.all(Bo == 0) and np.all(dBo == 0)
    # the upper grid edge is outside (std::upper_bound), the lower edge inside
    Bl, _, _, _ = P.field_eval(prob, np.array([[-0.2, 0.0, 0.1]]))
    assert np.abs(Bl[0] - [quad_field(-0.2, 0.0, 0.1), -quad_field(0.0, 0.1, -0.2), quad_field(0.1, -0.2, 0.0)]).max() < 1e-13


def _dummy_geometry():
    from pentrack_amd.geometry import Material, Solid, TGeometry, box_mesh
    return TGeometry([Material("v")], "v", [Solid(2, "v", box_mesh())])


def test_container_scale_bounds_and_taper(P):
    xg = yg = zg = np.linspace(0, 1, 6)
    X, Y, Z = np.meshgrid(xg, yg, zg, indexing="ij")
    B = np.stack([np.ones_like(X), 2 * np.ones_like(X), -np.ones_like(X)])
    box = (0.9, 0.1, 0.9, 0.1, 0.9, 0.1)
    fm = TFieldManager([tabfield3(xg, yg, zg, B=B, bscale=2.0, box=box, bw=0.1)])
    prob = P.Problem(_dummy_geometry(), fm)
    Bc, dB, _, _ = P.field_eval(prob, np.array([[0.5, 0.5, 0.5], [0.05, 0.5, 0.5], [0.1, 0.5, 0.5], [0.15, 0.5, 0.5]]))
    assert np.array_equal(Bc[0], [2.0, 4.0, -2.0])                 # scaled, away from the boundary
    assert np.all(Bc[1] == 0)                                       # outside the box
    assert np.all(Bc[2] == 0)                                       # smooth step is 0 at the box edge
    s = lambda u: ((6 * u - 15) * u + 10) * u ** 3
    assert np.allclose(Bc[3], np.array([2.0, 4.0, -2.0]) * s(0.5), rtol=1e-14)
    # uniform E field with V = Ez*z (TEDMStaticEField), B scale "0"
    prob2 = P.Problem(_dummy_geometry(), TFieldManager([uniform_efield(1.0, -2.0, 1e5)]))
    B2, _, V2, E2 = P.field_eval(prob2, np.array([[0.1, 0.2, 0.3]]))
    assert np.all(B2 == 0) and np.array_equal(E2[0], [1.0, -2.0, 1e5]) and V2[0] == 1e5 * 0.3


def test_loop_field_on_axis_closed_form():
    R, I = 0.3, 2.0
    z = np.linspace(-1, 1, 11)
    pts = np.stack([np.zeros_like(z), np.zeros_like(z), z + 0.1], 1)
    B = loop_field(pts, R, 0.1, I)
    ref = MU0 * I * R * R / (2 * (R * R + z * z) ** 1.5)
    assert np.allclose(B[:, 2], ref, rtol=1e-12) and np.allclose(B[:, :2], 0)
    # off axis: divergence-free (finite differences)
    p = np.array([[0.11, -0.07, 0.25]])
    h = 1e-5
    div = 0.0
    for a in range(3):
        e = np.zeros(3); e[a] = h
        div += (loop_field(p + e, R, 0.1, I)[0, a] - loop_field(p - e, R, 0.1, I)[0, a]) / (2 * h)
    assert abs(div) < 1e-8
