"""TFieldManager for the MI355X engine.

Field containers mirror TFieldContainer (include/field.h, src/field.cpp):
a field, B and E scale factors -- constants or exprtk formulas of the time t
(TFieldScaler, src/field.cpp:7-45; compiled by pt_formula.cpp) -- and an
optional boundary box with a smooth taper.  Supported fields on the tracking
path:
  * TabField3 -- tricubic interpolation of a 3D table (src/field_3d.cpp);
    the 64 coefficients per cell come from pt_tricubic_precompute (host C++)
  * TEDMStaticEField -- uniform E with V = Ez*z (src/edmfields.cpp:175-187)
  * TConductorField -- finite straight wire (src/conductor.cpp)
  * TExponentialFieldX, TLinearFieldZ, TB0GradZ, TB0GradX2, TB0GradXY, TB0_XY
    (src/analyticFields.cpp)
  * TabField -- axisymmetric r-z maps with ALGLIB's bicubic splines (src/field_2d.cpp);
    the per-cell Hermite data come from pt_bicubic_precompute (host C++)
Coil fields used to fill the synthetic tables are computed here on the host.
"""
import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native
from .constants import MU0


@dataclass
class FieldContainer:
    type: int
    bscale: float = 1.0
    escale: float = 1.0
    box: Sequence[float] = (0, 0, 0, 0, 0, 0)   # xmax xmin ymax ymin zmax zmin (no bounds by default)
    bw: float = 0.0
    # tricubic
    grid: Optional[tuple] = None                 # (xg, yg, zg)
    has_b: int = 0
    has_v: int = 0
    coef: Optional[np.ndarray] = None            # [cells][ncomp][64]
    # uniform E
    e: Sequence[float] = (0.0, 0.0, 0.0)
    # analytic fields and conductors
    params: Sequence[float] = ()
    # exprtk formulas of t replacing bscale / escale
    bformula: Optional[str] = None
    eformula: Optional[str] = None
    # 2D maps: the samples [ncomp][nz][nr] the coefficients were made from
    values: Optional[np.ndarray] = None
    # tricubic, mixed-precision tables: bound on the FP32 remainder's gradient error of a
    # cell component [field unit / m]; components above it keep FP64 (0: library default)
    mixed_tol: float = 0.0

    def __post_init__(self):
        if len(self.params) > 8:
            raise ValueError("at most 8 field parameters")
        b = self.box
        # TFieldBoundaryBox constructor (src/field.cpp:116-125)
        if self.bw < 0 or self.bw > (b[0] - b[1]) / 2 or self.bw > (b[2] - b[3]) / 2 or self.bw > (b[4] - b[5]) / 2:
            raise ValueError("You defined a field with an invalid boundary. Check field parameters.")


def tabfield3(xg, yg, zg, B=None, V=None, bscale=1.0, escale=0.0, box=(0, 0, 0, 0, 0, 0), bw=0.0, threads=None,
              precompute="host"):
    """TabField3 container from samples on a grid.

    B: [3][nx][ny][nz] (T) or None; V: [nx][ny][nz] (V) or None.
    Grid coordinates are made sorted and unique as in TabField3::TabField3
    (src/field_3d.cpp:417-426).  precompute "host": coefficients now, on the host
    (pt_tricubic_precompute); "gpu": the samples are kept and pre-interpolated on the GPU
    when the field manager is uploaded (bit-identical coefficients).
    """
    xg, yg, zg = (np.unique(np.asarray(g, dtype=np.float64)) for g in (xg, yg, zg))
    comps = []
    if B is not None:
        B = np.asarray(B, dtype=np.float64).reshape(3, len(xg), len(yg), len(zg))
        comps.extend([B[0], B[1], B[2]])
    if V is not None:
        comps.append(np.asarray(V, dtype=np.float64).reshape(len(xg), len(yg), len(zg)))
    if not comps:
        raise ValueError("tabfield3 needs B or V samples")
    values = np.ascontiguousarray(np.stack(comps))
    if precompute not in ("host", "gpu"):
        raise ValueError("precompute is 'host' or 'gpu'")
    coef = tricubic_coefficients(xg, yg, zg, values, threads=threads) if precompute == "host" else None
    bs, bf = _scale(bscale)
    es, ef = _scale(escale)
    return FieldContainer(type=_native.PT_FIELD_TRICUBIC, bscale=bs, escale=es, bformula=bf, eformula=ef,
                          box=tuple(box), bw=bw, grid=(xg, yg, zg), has_b=int(B is not None),
                          has_v=int(V is not None), coef=coef, values=None if coef is not None else values)


def tricubic_coefficients_gpu(xg, yg, zg, values):
    """TabField3::PreInterpol on the GPU (pt_tricubic_precompute_gpu): [cells][ncomp][64],
    the same layout and values as tricubic_coefficients."""
    L = _native.lib()
    xg, yg, zg = (np.ascontiguousarray(g, dtype=np.float64) for g in (xg, yg, zg))
    values = np.ascontiguousarray(values, dtype=np.float64)
    ncomp = values.shape[0]
    coef = np.empty(((len(xg) - 1) * (len(yg) - 1) * (len(zg) - 1), ncomp, 64), dtype=np.float64)
    d = _native.dptr
    _native.check(L.pt_tricubic_precompute_gpu(len(xg), len(yg), len(zg), d(xg), d(yg), d(zg), ncomp, d(values),
                                               d(coef)), "pt_tricubic_precompute_gpu")
    return coef


TAB2D_B = ("Br", "Bphi", "Bz")
TAB2D_V = ("V", "Er", "Ephi", "Ez")


def tabfield2(rg, zg, bscale=1.0, escale=1.0, **cols):
    """TabField container (src/field_2d.cpp) of an axisymmetric map: samples on an r-z grid
    (rg, zg strictly increasing), each column given as [len(zg)][len(rg)] by name (Br,
    Bphi, Bz in T; V in V or Er, Ephi, Ez in V/m).  A tabulated E replaces the potential
    (src/field_2d.cpp:262-284).  No bounds box: TFieldContainer(field, Bscale, Escale)
    (src/field_2d.cpp:68)."""
    rg, zg = np.asarray(rg, dtype=np.float64), np.asarray(zg, dtype=np.float64)
    unknown = set(cols) - set(TAB2D_B + TAB2D_V)
    if unknown:
        raise ValueError(f"unknown 2D table columns {sorted(unknown)}")
    if any(cols.get(k) is not None for k in ("Er", "Ephi", "Ez")):
        cols = dict(cols, V=None)
    names = [k for k in TAB2D_B + TAB2D_V if cols.get(k) is not None]
    if not names:
        raise ValueError("tabfield2 needs at least one column")
    has_b = sum(1 << i for i, k in enumerate(TAB2D_B) if cols.get(k) is not None)
    has_v = sum(1 << i for i, k in enumerate(TAB2D_V) if cols.get(k) is not None)
    values = np.ascontiguousarray(np.stack([np.asarray(cols[k], dtype=np.float64).reshape(len(zg), len(rg))
                                            for k in names]))
    m, n = len(rg), len(zg)
    coef = np.empty(((n - 1) * (m - 1), len(names), 16))
    rc = _native.lib().pt_bicubic_precompute(m, n, _native.dptr(np.ascontiguousarray(rg)),
                                             _native.dptr(np.ascontiguousarray(zg)), len(names),
                                             _native.dptr(values), _native.dptr(coef))
    _native.check(rc, "pt_bicubic_precompute")
    bs, bf = _scale(bscale)
    es, ef = _scale(escale)
    return FieldContainer(type=_native.PT_FIELD_TAB2D, bscale=bs, escale=es, bformula=bf, eformula=ef,
                          grid=(rg, None, zg), has_b=has_b, has_v=has_v, coef=coef, values=values)


def tricubic_coefficients(xg, yg, zg, values, threads=None):
    """TabField3::PreInterpol via the native host routine: [cells][ncomp][64]."""
    L = _native.lib()
    nx, ny, nz = len(xg), len(yg), len(zg)
    values = np.ascontiguousarray(values, dtype=np.float64)
    ncomp = values.shape[0]
    coef = np.empty(((nx - 1) * (ny - 1) * (nz - 1), ncomp, 64), dtype=np.float64)
    xg, yg, zg = (np.ascontiguousarray(g, dtype=np.float64) for g in (xg, yg, zg))
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    rc = L.pt_tricubic_precompute(nx, ny, nz, _native.dptr(xg), _native.dptr(yg), _native.dptr(zg), ncomp,
                                  _native.dptr(values), _native.dptr(coef), threads)
    _native.check(rc, "pt_tricubic_precompute")
    return coef


def _scale(s):
    """A scale given as a number or as an exprtk formula of t."""
    if isinstance(s, str):
        return 1.0, s
    return float(s), None


def uniform_efield(ex, ey, ez, escale=1.0):
    """TEDMStaticEField in a container with B scale "0" (src/fields.cpp:100-104)."""
    es, ef = _scale(escale)
    return FieldContainer(type=_native.PT_FIELD_UNIFORM_E, bscale=0.0, escale=es, eformula=ef, e=(ex, ey, ez))


def conductor(I, p1, p2, bscale=1.0):
    """TConductorField from p1 to p2 carrying I [A], in a container without bounds and
    E scale "1" (src/fields.cpp:42-46, include/field.h:280-281)."""
    if tuple(p1) == tuple(p2):
        raise ValueError("You have defined a conductor with length zero! Aborting...")
    bs, bf = _scale(bscale)
    return FieldContainer(type=_native.PT_FIELD_CONDUCTOR, bscale=bs, bformula=bf, escale=1.0,
                          params=(float(I),) + tuple(float(v) for v in p1) + tuple(float(v) for v in p2))


_ANALYTIC = {
    "ExponentialFieldX": (_native.PT_FIELD_EXPONENTIAL_X, 5),
    "LinearFieldZ": (_native.PT_FIELD_LINEAR_Z, 2),
    "B0GradZ": (_native.PT_FIELD_B0_GRAD_Z, 3),
    "B0GradX2": (_native.PT_FIELD_B0_GRAD_X2, 4),
    "B0GradXY": (_native.PT_FIELD_B0_GRAD_XY, 3),
    "B0_XY": (_native.PT_FIELD_B0_XY, 2),
}


def analytic_field(name, params, box=(0, 0, 0, 0, 0, 0), bscale=1.0):
    """One of the analytic B fields of src/analyticFields.cpp in a container with the
    given bounds (xmax xmin ymax ymin zmax zmin), no taper and E scale "0"
    (src/fields.cpp:106-134)."""
    ty, npar = _ANALYTIC[name]
    if len(params) != npar:
        raise ValueError(f"{name} takes {npar} parameters")
    bs, bf = _scale(bscale)
    return FieldContainer(type=ty, bscale=bs, bformula=bf, escale=0.0, box=tuple(box), bw=0.0,
                          params=tuple(float(v) for v in params))


def loop_field(points, radius, z0, current):
    """B of a circular current loop coaxial with z (Biot-Savart, elliptic integrals).

    points [n][3] -> B [n][3] in T.
    """
    from scipy.special import ellipe, ellipk
    p = np.asarray(points, dtype=np.float64)
    x, y, z = p[:, 0], p[:, 1], p[:, 2] - z0
    rho = np.hypot(x, y)
    a = radius
    q = (a + rho) ** 2 + z ** 2
    m = 4 * a * rho / q
    K = ellipk(m)
    E = ellipe(m)
    d = (a - rho) ** 2 + z ** 2
    pref = MU0 * current / (2 * np.pi)
    Bz = pref / np.sqrt(q) * (K + (a * a - rho * rho - z * z) / d * E)
    with np.errstate(divide="ignore", invalid="ignore"):
        Brho = np.where(rho > 1e-12, pref * z / (rho * np.sqrt(q)) * (-K + (a * a + rho * rho + z * z) / d * E), 0.0)
        cx = np.where(rho > 1e-12, x / np.where(rho > 0, rho, 1), 0.0)
        cy = np.where(rho > 1e-12, y / np.where(rho > 0, rho, 1), 0.0)
    return np.stack([Brho * cx, Brho * cy, Bz], axis=1)


def coil_set(seed, n=8, radius=0.30, zmin=-0.05, zmax=0.55):
    """8 coaxial loops with seed-derived currents of alternating sign."""
    rng = np.random.default_rng(seed)
    zs = np.linspace(zmin, zmax, n)
    currents = rng.uniform(0.5, 1.5, n) * np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return [(radius, float(z), float(I)) for z, I in zip(zs, currents)]


def sample_coils(coils, xg, yg, zg, scale=1.0):
    X, Y, Z = np.meshgrid(xg, yg, zg, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    B = np.zeros_like(pts)
    for a, z0, I in coils:
        B += loop_field(pts, a, z0, I * scale)
    return B.T.reshape(3, len(xg), len(yg), len(zg))


class TFieldManager:
    """Ordered list of field containers, summed in order (src/fields.cpp:139-182)."""

    def __init__(self, containers=()):
        self.containers = list(containers)
        self._handle = None

    def descs(self):
        ds = (_native.FieldDesc * max(1, len(self.containers)))()
        keep = []
        for k, c in enumerate(self.containers):
            d = ds[k]
            d.type = c.type
            d.bscale = c.bscale
            d.escale = c.escale
            for i in range(6):
                d.box[i] = c.box[i]
            d.bw = c.bw
            for i in range(3):
                d.e[i] = c.e[i]
            for i, v in enumerate(c.params):
                d.p[i] = v
            for attr, form in (("bscale_formula", c.bformula), ("escale_formula", c.eformula)):
                if form is not None:
                    b = C.create_string_buffer(form.encode())
                    keep.append(b)
                    setattr(d, attr, C.cast(b, C.c_char_p))
            if c.type == _native.PT_FIELD_TRICUBIC:
                xg, yg, zg = (np.ascontiguousarray(g, dtype=np.float64) for g in c.grid)
                keep.extend([xg, yg, zg])
                d.nx, d.ny, d.nz = len(xg), len(yg), len(zg)
                d.xg, d.yg, d.zg = _native.dptr(xg), _native.dptr(yg), _native.dptr(zg)
                d.has_b, d.has_v = c.has_b, c.has_v
                d.mixed_tol = c.mixed_tol
                if c.coef is not None:
                    coef = np.ascontiguousarray(c.coef, dtype=np.float64)
                    keep.append(coef)
                    d.coef = _native.dptr(coef)
                else:                      # samples only: pre-interpolated on the GPU
                    vals = np.ascontiguousarray(c.values, dtype=np.float64)
                    keep.append(vals)
                    d.values = _native.dptr(vals)
            elif c.type == _native.PT_FIELD_TAB2D:
                rg, zg = (np.ascontiguousarray(c.grid[i], dtype=np.float64) for i in (0, 2))
                coef = np.ascontiguousarray(c.coef, dtype=np.float64)
                keep.extend([rg, zg, coef])
                d.nx, d.ny, d.nz = len(rg), 1, len(zg)
                d.xg, d.zg = _native.dptr(rg), _native.dptr(zg)
                d.has_b, d.has_v = c.has_b, c.has_v
                d.coef = _native.dptr(coef)
        return ds, keep

    def device(self, with_mixed=False):
        if self._handle is not None and (self._handle[1] or not with_mixed):
            return self._handle[0]
        L = _native.lib()
        ds, keep = self.descs()
        h = C.c_void_p()
        rc = L.pt_fields_create(len(self.containers), ds, int(bool(with_mixed)), C.byref(h))
        _native.check(rc, "pt_fields_create")
        if self._handle is not None:
            # batches may still use the old tables: keep them until the manager is released
            self._retired = getattr(self, "_retired", []) + [self._handle[0]]
        self._handle = (h, bool(with_mixed))
        return h

    def mixed_info(self):
        """(blocks, blocks kept in FP64) of the mixed-precision tables on the device."""
        if self._handle is None or not self._handle[1]:
            return 0, 0
        b, f = C.c_int64(), C.c_int64()
        _native.check(_native.lib().pt_fields_mixed_info(self._handle[0], C.byref(b), C.byref(f)),
                      "pt_fields_mixed_info")
        return b.value, f.value

    def __del__(self):
        try:
            if self._handle is not None and _native._lib is not None:
                _native._lib.pt_fields_destroy(self._handle[0])
                for h in getattr(self, "_retired", []):
                    _native._lib.pt_fields_destroy(h)
        except Exception:
            pass
