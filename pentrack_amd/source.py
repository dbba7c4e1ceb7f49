"""Particle sources of the SOURCE section (TParticleSource and its volume and surface
sources, src/source.cpp:21-230, include/source.h:22-392).

Volume sources: `boxvolume`, `cylvolume` and `STLvolume` (points uniform in the STL
solid's bounding box, kept where the +z ray from the point crosses the mesh an odd
number of times, TTriangleMesh::InSolid, src/trianglemesh.cpp:167-171), optionally with
phase-space weighting (TVolumeSource::CreateParticle, src/source.cpp:126-202: total
energies drawn from the spectrum, start points accepted with probability
sqrt(H - V) / sqrt(H - Vmin), Vmin from 100000 random points).  Surface sources:
`STLsurface` and `cylsurface` (TSurfaceSource::CreateParticle, src/source.cpp:99-124):
points on the geometry's walls (mesh by area, triangle by area, uniform in the triangle,
include/trianglemesh.h:210-231) inside the source volume, moved 1e-8 m off the wall,
velocities cosine-distributed about the wall normal with the `Enormal` boost.

The energy spectrum and the angular distributions are the reference's piecewise linear
distributions of formulas of x (parse_distribution, include/mc.h:310-335) built by
pt_distribution_table; inside-solid tests and potentials run on the GPU
(pt_geometry_solid_at, pt_fields_eval).  Draws come from a numpy generator seeded with
the run's seed, not from the reference's mt19937_64, so a run reproduces the reference's
distributions, not its individual particles.

Differences: the reference's `cylsurface` never reads its parameters (all six stay 0,
include/source.h:286-287, and no point is ever accepted); here `parameters` gives
rmin rmax phimin phimax zmin zmax (angles in degrees, as for cylvolume).  Surface
normals are those of the triangles as stored in the STL file (the reference orients
meshes with CGAL when it reads them).
"""
import ctypes as C
import math
import os

import numpy as np

from . import _native
from .config import ConfigError, SPECIES_BY_NAME, _floats, deg
from .particle import make_states

REFLECT_TOLERANCE = 1e-8
VOLUME_MODES = ("boxvolume", "cylvolume", "STLvolume")
SURFACE_MODES = ("STLsurface", "cylsurface")


class PiecewiseLinear:
    """std::piecewise_linear_distribution<double> of a formula of x on [lo, hi]."""

    def __init__(self, formula, lo, hi):
        L = _native.lib()
        n = C.c_int()
        f = formula.strip().encode()
        _native.check(L.pt_distribution_table(f, lo, hi, 0, None, None, None, None, None, C.byref(n)),
                      "pt_distribution_table")
        ni = n.value
        self.x, raw, self.den = np.zeros(ni + 1), np.zeros(ni + 1), np.zeros(ni + 1)
        self.cp, self.m = np.zeros(ni), np.zeros(ni)
        d = _native.dptr
        _native.check(L.pt_distribution_table(f, lo, hi, ni + 1, d(self.x), d(raw), d(self.den), d(self.cp), d(self.m),
                                              C.byref(n)), "pt_distribution_table")

    def sample_from(self, p):
        """libstdc++ operator() for variates p in [0, 1)."""
        i = np.searchsorted(self.cp, p, side="left")
        pref = np.where(i > 0, self.cp[np.maximum(i - 1, 0)], 0.0)
        a = 0.5 * self.m[i]
        b = self.den[i]
        cm = p - pref
        with np.errstate(divide="ignore", invalid="ignore"):
            quad = 0.5 * (np.sqrt(b * b + 4.0 * a * cm) - b) / a
        return self.x[i] + np.where(a == 0, cm / b, quad)

    def __call__(self, rng, n):
        return self.sample_from(rng.random(n))

    def min(self):
        return self.x[0]

    def max(self):
        return self.x[-1]


def rotate_vectors(v, n):
    """RotateVector(v, n) without a reference direction (src/globals.cpp:40-64): the
    frame (base1 of the plane normal to n, n x base1, n) applied to v; rows of v, n."""
    v = np.asarray(v, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    a, b, c = n[:, 0], n[:, 1], n[:, 2]
    x = np.zeros_like(n)
    # CGAL Plane_3::base1
    ca, cb, cc = a == 0, (a != 0) & (b == 0), (a != 0) & (b != 0) & (c == 0)
    rest = ~(ca | cb | cc)
    x[ca] = [1, 0, 0]
    x[cb] = [0, 1, 0]
    x[cc] = [0, 0, 1]
    A, Bb, Cc = np.abs(a), np.abs(b), np.abs(c)
    s1 = rest & (A <= Bb) & (A <= Cc)
    s2 = rest & ~s1 & (Bb <= A) & (Bb <= Cc)
    s3 = rest & ~s1 & ~s2
    x[s1] = np.stack([0 * a[s1], -c[s1], b[s1]], 1)
    x[s2] = np.stack([-c[s2], 0 * a[s2], a[s2]], 1)
    x[s3] = np.stack([-b[s3], a[s3], 0 * a[s3]], 1)
    x = x / np.sqrt((x * x).sum(1))[:, None]
    z = n / np.sqrt((n * n).sum(1))[:, None]
    y = np.cross(z, x)
    return x * v[:, :1] + y * v[:, 1:2] + z * v[:, 2:3]


def _inside(tris, pts):
    """TTriangleMesh::InSolid on the GPU: the solid query of a geometry holding only `tris`."""
    from .geometry import Material, Solid, TGeometry
    from .queries import solid_at
    g = TGeometry([Material("vacuum")], "vacuum", [Solid(2, "vacuum", tris, [], "source")])
    return solid_at(g, pts, np.zeros(len(pts))) == 2


class ParticleSource:
    """TParticleSource (src/source.cpp:21-79): species, spectrum, angles, polarisation and
    the start-time distribution of the SOURCE section."""

    def __init__(self, cfg, base_dir="."):
        sc = cfg["SOURCE"]
        self.base_dir = base_dir
        get = self._get = lambda k, d="": sc.get(k, d).strip()
        self.mode = get("sourcemode").split()[0] if get("sourcemode") else ""
        self.particle = get("particle").split()[0] if get("particle") else ""
        if self.particle not in SPECIES_BY_NAME:
            raise ConfigError("Could not create particle " + self.particle)
        self.kind = SPECIES_BY_NAME[self.particle]
        self._num = num = lambda k, d=0.0: _floats(get(k))[0] if _floats(get(k)) else d
        self.active = num("ActiveTime")
        self.pulse_gap, self.pulse_width = num("pulseGap"), num("pulseWidth")
        self.polarization = num("polarization")
        self.spectrum = PiecewiseLinear(get("spectrum") or "1", num("Emin"), num("Emax"))
        self.phi_v = PiecewiseLinear(get("phi_v") or "1", deg(num("phi_v_min")), deg(num("phi_v_max", 360.0)))
        self.theta_v = PiecewiseLinear(get("theta_v") or "sin(x)", deg(num("theta_v_min")), deg(num("theta_v_max", 180.0)))
        self.par = _floats(get("parameters"))
        # start-time distribution (src/source.cpp:49-78)
        if self.active == 0:
            self.t_int, self.t_w = [0.0, math.nextafter(0.0, math.inf)], [1.0]
        elif self.pulse_gap > 0:
            iv, w = [], []
            i = 0.0
            while i < self.active:
                iv.append(i)
                if self.pulse_width == 0:
                    iv.append(math.nextafter(i, math.inf))
                    w.append(1.0)
                elif i + self.pulse_width > self.active:
                    iv.append(self.active)
                    w.append((self.active - i) / self.pulse_width)
                else:
                    iv.append(i + self.pulse_width)
                    w.append(1.0)
                w.append(0.0)
                i += self.pulse_gap
            w.pop()
            self.t_int, self.t_w = iv, w
        else:
            self.t_int, self.t_w = [0.0, self.active], [1.0]

    def _need_par(self, k):
        if len(self.par) < k:
            raise ConfigError(f"source {self.mode} needs {k} parameters")
        return self.par[:k]

    def _stl(self):
        from .geometry import read_stl, snap
        f = self._get("STLfile").split()
        if not f:
            raise ConfigError(f"source {self.mode} needs an STLfile")
        _, tris = read_stl(os.path.join(self.base_dir, f[0]))
        return np.asarray(snap(tris), dtype=np.float32)

    def _times(self, rng, n):
        """std::piecewise_constant_distribution over the intervals."""
        x = np.asarray(self.t_int)
        w = np.asarray(self.t_w) * np.diff(x)
        cp = np.cumsum(w) / w.sum()
        p = rng.random(n)
        i = np.minimum(np.searchsorted(cp, p, side="right"), len(cp) - 1)
        pref = np.where(i > 0, cp[np.maximum(i - 1, 0)], 0.0)
        frac = np.where(cp[i] > pref, (p - pref) / (cp[i] - pref), 0.0)
        return x[i] + frac * (x[i + 1] - x[i])

    def _pol(self, rng, n):
        """polarization_distribution: +1 with probability (1 + polarization) / 2."""
        return np.where(rng.random(n) < 0.5 * (self.polarization + 1.0), 1.0, -1.0)


class VolumeSource(ParticleSource):
    """TCuboidVolumeSource, TCylindricalVolumeSource, TSTLVolumeSource."""

    def __init__(self, cfg, base_dir="."):
        super().__init__(cfg, base_dir)
        if self.mode not in VOLUME_MODES:
            raise ConfigError(f"sourcemode {self.mode!r} is not a volume source ({', '.join(VOLUME_MODES)})")
        self.weighting = int(self._num("PhaseSpaceWeighting")) != 0
        if self.mode == "STLvolume":
            self.tris = self._stl()
            v = self.tris.reshape(-1, 3).astype(np.float64)
            self.bbox = (v.min(0), v.max(0))
        else:
            self.par = self._need_par(6)
        self.min_pot = math.inf

    def points(self, rng, n):
        """RandomPointInSourceVolume for n points."""
        if self.mode == "boxvolume":
            xmin, xmax, ymin, ymax, zmin, zmax = self.par
            return np.stack([xmin + rng.random(n) * (xmax - xmin), ymin + rng.random(n) * (ymax - ymin),
                             zmin + rng.random(n) * (zmax - zmin)], axis=1)
        if self.mode == "cylvolume":
            rmin, rmax, phimin, phimax, zmin, zmax = self.par
            r = np.sqrt(rng.random(n) * (rmax * rmax - rmin * rmin) + rmin * rmin)   # linear_distribution(rmin, rmax)
            phr = deg(phimin) + rng.random(n) * (deg(phimax) - deg(phimin))
            return np.stack([r * np.cos(phr), r * np.sin(phr), zmin + rng.random(n) * (zmax - zmin)], axis=1)
        lo, hi = self.bbox                        # TTriangleMesh::RandomPointInVolume
        out = np.zeros((0, 3))
        while len(out) < n:
            k = max(2 * (n - len(out)), 1024)
            p = lo + rng.random((k, 3)) * (hi - lo)
            out = np.concatenate([out, p[_inside(self.tris, p)]])
        return out[:n]

    def _potential(self, geometry, fields, t, pos, pol):
        """TParticle::GetInitialTotalEnergy of a particle at rest: its potential energy with
        the material potential of the solid it starts in."""
        from .logger import potential_energy
        from .queries import fields_eval, solid_at
        n = len(pos)
        y = np.zeros((n, 9))
        y[:, :3] = pos
        y[:, 7] = pol
        B, _, V, _ = fields_eval(fields, pos, t, grad=False)
        sid = solid_at(geometry, pos, t)
        names = list(geometry.materials)
        mats = np.stack([m.as_array() for m in geometry.materials.values()])
        id_to_mat = {1: names.index(geometry.default_material)}
        for s in geometry.solids:
            id_to_mat[s.ID] = names.index(s.material)
        ms = np.array([id_to_mat.get(int(i), id_to_mat[1]) for i in sid])
        return potential_energy(self.kind, y, B, V, mats, ms)

    def sample(self, n, rng, geometry=None, fields=None):
        """n particles: (t0 [n], y0 [n][9], spin0 [n][5]) as the constructors set them up."""
        if not self.weighting:
            t0 = self._times(rng, n)
            pos = self.points(rng, n)
            pol = self._pol(rng, n)
            E = self.spectrum(rng, n)
        else:
            if geometry is None or fields is None:
                raise ConfigError("PhaseSpaceWeighting needs the geometry and the fields")
            if self.min_pot == math.inf:        # TVolumeSource::FindPotentialMinimum, 100000 points
                N = 100000
                self.min_pot = float(self._potential(geometry, fields, self._times(rng, N), self.points(rng, N),
                                                     self._pol(rng, N)).min())
                if self.min_pot > self.spectrum.max():
                    raise ConfigError(f"Error: your chosen spectrum is below the minimal potential energy in the "
                                      f"source volume ({self.spectrum.max()} eV < {self.min_pot} eV). Exiting!")
            H = np.empty(n)
            todo = np.arange(n)
            while len(todo):                    # total energies above the potential minimum
                h = self.spectrum(rng, len(todo))
                ok = h >= self.min_pot
                H[todo[ok]] = h[ok]
                todo = todo[~ok]
            t0, pos, pol, E = np.empty(n), np.empty((n, 3)), np.empty(n), np.empty(n)
            todo = np.arange(n)
            while len(todo):                    # start points accepted with sqrt(H - V) / sqrt(H - Vmin)
                k = len(todo)
                tt, pp, pl = self._times(rng, k), self.points(rng, k), self._pol(rng, k)
                V = self._potential(geometry, fields, tt, pp, pl)
                Hk = H[todo]
                u = rng.random(k) * np.sqrt(Hk - self.min_pot)
                with np.errstate(invalid="ignore"):
                    ok = (Hk >= V) & (np.sqrt(np.maximum(Hk - V, 0)) > u)
                sel = todo[ok]
                t0[sel], pos[sel], pol[sel], E[sel] = tt[ok], pp[ok], pl[ok], Hk[ok] - V[ok]
                todo = todo[~ok]
        phi = self.phi_v(rng, n)
        theta = self.theta_v(rng, n)
        _, y0, spin0 = make_states(self.kind, pos, E, phi, theta, pol)
        return t0, y0, spin0


class SurfaceSource(ParticleSource):
    """TSTLSurfaceSource, TCylindricalSurfaceSource: particles start on the geometry's walls."""

    def __init__(self, cfg, base_dir="."):
        super().__init__(cfg, base_dir)
        if self.mode not in SURFACE_MODES:
            raise ConfigError(f"sourcemode {self.mode!r} is not a surface source ({', '.join(SURFACE_MODES)})")
        self.Enormal = self._num("Enormal")
        if self.mode == "STLsurface":
            self.tris = self._stl()
            v = self.tris.reshape(-1, 3).astype(np.float64)
            self.bbox = (v.min(0), v.max(0))
        else:
            rmin, rmax, phimin, phimax, zmin, zmax = self._need_par(6)
            self.cyl = (rmin, rmax, deg(phimin), deg(phimax), zmin, zmax)
            self.bbox = (np.array([-rmax, -rmax, zmin]), np.array([rmax, rmax, zmax]))

    def _in_source(self, p):
        if self.mode == "STLsurface":
            return _inside(self.tris, p)
        rmin, rmax, phimin, phimax, zmin, zmax = self.cyl
        r = np.hypot(p[:, 0], p[:, 1])
        phi = np.arctan2(p[:, 1], p[:, 0])
        return (r > rmin) & (r < rmax) & (phi > phimin) & (phi < phimax) & (p[:, 2] > zmin) & (p[:, 2] < zmax)

    def points(self, rng, n, geometry):
        """TTriangleMesh::RandomPointOnSurface until inside the source volume: (points, normals)."""
        lo, hi = self.bbox
        meshes = []
        for s in geometry.solids:
            t = np.asarray(s.triangles, dtype=np.float64)
            v = t.reshape(-1, 3)
            if (v.min(0) <= hi).all() and (v.max(0) >= lo).all():     # bounding boxes intersect
                cr = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
                area = 0.5 * np.sqrt((cr * cr).sum(1))
                meshes.append((t, cr, area))
            else:
                meshes.append(None)
        total = np.array([m[2].sum() if m is not None else 0.0 for m in meshes])
        if total.sum() == 0:
            raise ConfigError("no wall of the geometry lies in the surface source's bounding box")
        pts, nrm = np.zeros((0, 3)), np.zeros((0, 3))
        while len(pts) < n:
            k = max(2 * (n - len(pts)), 1024)
            mi = rng.choice(len(meshes), size=k, p=total / total.sum())
            P, N = np.empty((k, 3)), np.empty((k, 3))
            for j in np.unique(mi):
                t, cr, area = meshes[j]
                sel = np.flatnonzero(mi == j)
                ti = rng.choice(len(t), size=len(sel), p=area / area.sum())
                a, b = rng.random(len(sel)), rng.random(len(sel))
                fold = a + b > 1
                a[fold], b[fold] = 1 - a[fold], 1 - b[fold]
                tt = t[ti]
                P[sel] = tt[:, 0] + a[:, None] * (tt[:, 1] - tt[:, 0]) + b[:, None] * (tt[:, 2] - tt[:, 0])
                c = cr[ti]
                N[sel] = c / np.sqrt((c * c).sum(1))[:, None]
            ok = self._in_source(P)
            pts, nrm = np.concatenate([pts, P[ok]]), np.concatenate([nrm, N[ok]])
        return pts[:n], nrm[:n]

    def sample(self, n, rng, geometry=None, fields=None):
        if geometry is None:
            raise ConfigError("a surface source needs the geometry")
        p, nv = self.points(rng, n, geometry)
        p = p + nv * REFLECT_TOLERANCE
        E = self.spectrum(rng, n)
        phi = rng.random(n) * (2 * math.pi)
        theta = np.arccos(np.sqrt(rng.random(n) * (math.cos(0.5 * math.pi) ** 2 - 1.0) + 1.0))   # sincos_distribution
        if self.Enormal > 0:
            vn = np.sqrt(E * np.cos(theta) ** 2 + self.Enormal)
            vt = np.sqrt(E) * np.sin(theta)
            theta = np.arctan2(vt, vn)
            E = vn * vn + vt * vt
        v = np.stack([np.cos(phi) * np.sin(theta), np.sin(phi) * np.sin(theta), np.cos(theta)], 1)
        v = rotate_vectors(v, nv)
        phi = np.arctan2(v[:, 1], v[:, 0])
        theta = np.arccos(np.clip(v[:, 2], -1.0, 1.0))
        pol = self._pol(rng, n)
        t0 = self._times(rng, n)
        _, y0, spin0 = make_states(self.kind, p, E, phi, theta, pol)
        return t0, y0, spin0


def create_source(cfg, base_dir="."):
    """CreateParticleSource (src/source.cpp:204-230)."""
    sc = cfg["SOURCE"]
    mode = sc.get("sourcemode", "").split()[:1]
    mode = mode[0] if mode else ""
    if mode in VOLUME_MODES:
        return VolumeSource(cfg, base_dir)
    if mode in SURFACE_MODES:
        return SurfaceSource(cfg, base_dir)
    raise ConfigError(f"Could not load source {mode}!")
