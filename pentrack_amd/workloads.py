"""Synthetic workloads of the tracking path (inputs shared by tests, smoke and bench).

cube        1 m cube (12 triangles) as solid 2 of material "vacuum" inside a default
            solid of Fermi potential 335 neV; specular walls; gravity only.
cylinder    closed cylinder R = 0.25 m, H = 0.5 m (about 20k triangles); 2% Lambert,
            loss 1e-4 per bounce; B from a tricubic map of 8 coaxial loops.
Particles are created like a volume source: uniform positions inside, kinetic
energy uniform in [0, Emax], isotropic directions (src/source.cpp:157-202 draws
the same quantities).  Host-side sampling uses numpy's PCG64 seeded per
workload; the tracker's own random draws use the per-particle streams.
"""
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .constants import NEUTRON
from .fields import TFieldManager, coil_set, sample_coils, tabfield3
from .geometry import Material, Solid, TGeometry, box_mesh, cylinder_mesh
from .particle import SpeciesConfig, make_states


@dataclass
class Workload:
    name: str
    geometry: TGeometry
    fields: TFieldManager
    species: List[SpeciesConfig]
    tmax: float
    species_idx: np.ndarray
    t0: np.ndarray
    y0: np.ndarray
    spin0: np.ndarray
    pid: np.ndarray
    seed: int
    meta: dict = field(default_factory=dict)
    loss: Optional[np.ndarray] = None       # per-particle loss per bounce (None: the materials')

    @property
    def n(self):
        return len(self.t0)


def cube_geometry(fermi=335.0, diffprob=0.0, loss=0.0):
    mats = [Material("wall", FermiReal=fermi), Material("vacuum", DiffProb=diffprob, LossPerBounce=loss)]
    return TGeometry(mats, "wall", [Solid(2, "vacuum", box_mesh((0, 0, 0), (1, 1, 1)))])


def cylinder_geometry(radius=0.25, height=0.5, n_phi=200, n_z=40, n_r=5, fermi=335.0, diffprob=0.02, loss=1e-4):
    mats = [Material("wall", FermiReal=fermi), Material("storage", DiffProb=diffprob, LossPerBounce=loss)]
    return TGeometry(mats, "wall", [Solid(2, "storage", cylinder_mesh(radius, height, n_phi, n_z, n_r))])


def neutron_source(rng, n, inside, lo, hi, Emax, pol=None):
    pos = np.empty((n, 3))
    k = 0
    while k < n:
        p = rng.uniform(lo, hi, size=(max(2 * (n - k), 16), 3))
        p = p[inside(p)]
        take = min(len(p), n - k)
        pos[k:k + take] = p[:take]
        k += take
    E = rng.uniform(0.0, Emax, n)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    theta = np.arccos(rng.uniform(-1.0, 1.0, n))
    if pol is None:
        pol = np.where(rng.uniform(size=n) < 0.5, 1.0, -1.0)
    else:
        pol = np.full(n, float(pol))
    return make_states(NEUTRON, pos, E, phi, theta, pol)


def coil_map(seed, n_grid, lo, hi, Bmax=1.0, threads=None):
    """n_grid^3 tricubic B map of the 8-loop coil set over [lo, hi] plus one cell of margin,
    scaled so that max |B| on the grid is Bmax."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    h = (hi - lo) / (n_grid - 3)
    axes = [lo[a] - h[a] + h[a] * np.arange(n_grid) for a in range(3)]
    coils = coil_set(seed)
    B = sample_coils(coils, *axes)
    scale = Bmax / np.sqrt((B ** 2).sum(axis=0)).max()
    B *= scale
    return tabfield3(*axes, B=B, threads=threads), dict(coils=coils, scale=scale, axes=axes)


def cube(n=10000, seed=1, tmax=10.0, Emax=300e-9, with_field=False, n_grid=32, Bmax=0.1):
    """Neutrons in a 1 m cube under gravity with specular walls (335 neV); with_field adds an
    n_grid^3 coil-field map over the cube."""
    rng = np.random.default_rng(seed)
    geom = cube_geometry()
    eps = 1e-6
    t0, y0, spin0 = neutron_source(rng, n, lambda p: np.ones(len(p), bool), eps, 1 - eps, Emax)
    containers = []
    meta = {}
    if with_field:
        fmap, meta = coil_map(seed, n_grid, (0, 0, 0), (1, 1, 1), Bmax=Bmax)
        containers.append(fmap)
    sp = [SpeciesConfig(kind=NEUTRON, tau=0.0, tmax=9e99, lmax=9e99)]
    return Workload("cube", geom, TFieldManager(containers), sp, tmax, np.zeros(n, np.int32), t0, y0, spin0,
                    np.arange(n, dtype=np.uint64), seed, meta)


def cylinder(n=1000000, seed=2, tmax=200.0, Emax=250e-9, n_grid=64, Bmax=1.0, n_phi=200, n_z=40, n_r=5):
    """Cylinder storage volume (R 0.25 m, H 0.5 m) with a 64^3 coil-field map."""
    rng = np.random.default_rng(seed)
    R, H = 0.25, 0.5
    geom = cylinder_geometry(R, H, n_phi, n_z, n_r)
    # inscribed polygon radius, so sampled points are inside the faceted mesh
    r_in = R * np.cos(np.pi / n_phi) * (1 - 1e-6)
    inside = lambda p: np.hypot(p[:, 0], p[:, 1]) < r_in
    t0, y0, spin0 = neutron_source(rng, n, inside, (-R, -R, 1e-6), (R, R, H - 1e-6), Emax, pol=1.0)
    fmap, meta = coil_map(seed, n_grid, (-R, -R, 0.0), (R, R, H), Bmax=Bmax)
    sp = [SpeciesConfig(kind=NEUTRON, tau=0.0, tmax=9e99, lmax=9e99)]
    return Workload("cylinder", geom, TFieldManager([fmap]), sp, tmax, np.zeros(n, np.int32), t0, y0, spin0,
                    np.arange(n, dtype=np.uint64), seed, meta)


def adjoint_cube(n=100000, seed=5, n_grid=64, nmaps=8, Bmax=0.5, Emax=300e-9, dt=1e-4, threads=None):
    """Differentiation workload: neutrons in the 1 m cube, B = sum_k s_k B_k with one
    tricubic map per coaxial loop (loops centred on the cube axis), RK4 step dt.
    Returns (AdjointModel, y0 [n][6], s0 [nmaps])."""
    from .adjoint import AdjointModel
    from .fields import coil_set, loop_field, tricubic_coefficients
    rng = np.random.default_rng(seed)
    lo, hi = np.zeros(3), np.ones(3)
    h = (hi - lo) / (n_grid - 3)
    axes = [lo[a] - h[a] + h[a] * np.arange(n_grid) for a in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([X.ravel() - 0.5, Y.ravel() - 0.5, Z.ravel()], axis=1)
    coils = [(0.6, -0.05 + 1.1 * k / max(nmaps - 1, 1), I) for k, (_, _, I) in enumerate(coil_set(seed, nmaps))]
    Bk = [loop_field(pts, a, z0, I).T.reshape(3, n_grid, n_grid, n_grid) for a, z0, I in coils]
    scale = Bmax / np.sqrt((sum(Bk) ** 2).sum(axis=0)).max()
    coefs = [tricubic_coefficients(*axes, np.ascontiguousarray(B * scale), threads=threads) for B in Bk]
    model = AdjointModel(grid=tuple(axes), coefs=coefs, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), dt=dt)
    _, y9, _ = neutron_source(rng, n, lambda p: np.ones(len(p), bool), 1e-3, 1 - 1e-3, Emax, pol=1.0)
    return model, np.ascontiguousarray(y9[:, :6]), np.ones(nmaps)


def uneven(n=4000000, seed=4, n_tris=50000, tmax=200.0, n_grid=64, Bmax=1.0, E_field=1e5):
    """Mixed species with very uneven lifetimes: 90% neutrons with
    exponential decay (tau = 880 s), 8% protons (1-800 eV), 2% electrons (1-700 keV) in
    a closed cylinder tessellated to about n_tris triangles, B from the coil map plus a
    uniform E = E_field V/m along z; species proper-time limits 3e-3 s (protons) and
    1e-5 s (electrons) as in in/config.in."""
    from .constants import ELECTRON, PROTON
    from .fields import uniform_efield
    rng = np.random.default_rng(seed)
    R, H = 0.25, 0.5
    # side quads 2 n_phi n_z plus caps ~ 2 n_phi (2 n_r - 1): n_z ~ 0.6 n_phi, n_r = 5
    n_phi = max(12, int(np.sqrt(n_tris / 1.2 / 1.0 + 81) - 9))
    n_z = max(4, int(0.6 * n_phi))
    geom = cylinder_geometry(R, H, n_phi, n_z, 5)
    r_in = R * np.cos(np.pi / n_phi) * (1 - 1e-6)
    inside = lambda p: np.hypot(p[:, 0], p[:, 1]) < r_in
    kinds = np.where(rng.uniform(size=n) < 0.9, NEUTRON, np.where(rng.uniform(size=n) < 0.8, PROTON, ELECTRON))
    kinds = kinds.astype(np.int32)
    _, y0, spin0 = neutron_source(rng, n, inside, (-R, -R, 1e-6), (R, R, H - 1e-6), 250e-9, pol=1.0)
    pos = y0[:, :3]
    E = np.where(kinds == PROTON, rng.uniform(1, 800, n), np.where(kinds == ELECTRON, rng.uniform(1e3, 7e5, n),
                                                                  rng.uniform(0, 250e-9, n)))
    phi = rng.uniform(0, 2 * np.pi, n)
    theta = np.arccos(rng.uniform(-1, 1, n))
    y = np.empty((n, 9))
    t0 = np.zeros(n)
    for k in (NEUTRON, PROTON, ELECTRON):
        sel = kinds == k
        if sel.any():
            _, yk, _ = make_states(k, pos[sel], E[sel], phi[sel], theta[sel], 1.0)
            y[sel] = yk
    fmap, meta = coil_map(seed, n_grid, (-R, -R, 0.0), (R, R, H), Bmax=Bmax)
    fm = TFieldManager([fmap, uniform_efield(0.0, 0.0, E_field)])
    species = [SpeciesConfig(NEUTRON, tau=880.0, tmax=9e99), SpeciesConfig(PROTON, tmax=3e-3),
               SpeciesConfig(ELECTRON, tmax=1e-5)]
    meta = dict(meta, n_phi=n_phi, n_z=n_z)
    return Workload("uneven", geom, fm, species, tmax, kinds, t0, y, spin0, np.arange(n, dtype=np.uint64), seed,
                    meta)


# the reference's own meshes used by the uneven-lifetime workloads (binary STL, metres)
REFERENCE_MESHES = {"guide": "guide_experiment.STL",                    # in/guide_experiment.STL, 1032 tris
                    "storage": "storagevolume_60degfeeder_bellow.STL"}  # in/storagevolume_..., 61468 tris
MESH_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden")


def vessel_geometry(mesh, wall, margin=0.05):
    """A vessel whose inside is the default solid (vacuum, ID 1), where charged particles
    fly (TProton/TElectron::OnStep absorb them inside any solid with ID > 1,
    src/proton.cpp:25-33): solid 2 of the wall material is the shell between the closed
    mesh (reversed, so its normals face the vacuum) and a box around it, as
    test/HollowUnitCube.STL is built."""
    mesh = np.asarray(mesh, dtype=np.float32)
    lo, hi = mesh.reshape(-1, 3).min(0) - margin, mesh.reshape(-1, 3).max(0) + margin
    shell = np.concatenate([box_mesh(tuple(lo), tuple(hi)), mesh[:, ::-1, :]]).astype(np.float32)
    return TGeometry([Material("vacuum"), wall], "vacuum", [Solid(2, wall.name, shell)])


def uneven_vessel(mesh="storage", n=4000000, seed=4, tmax=200.0, n_grid=64, Bmax=1.0, E_field=1e5, n_tris=200000,
                  loss=(1e-5, 1e-3)):
    """Config 4: mixed species with very uneven lifetimes in a vessel of the mesh `mesh`
    ("guide", "storage": the reference's in/guide_experiment.STL and
    in/storagevolume_60degfeeder_bellow.STL; "cylinder": a synthetic closed cylinder of about
    n_tris triangles): 90% neutrons with exponential decay (tau = 880 s) and a loss per
    bounce drawn per particle from U(loss) (a build extension, Workload.loss), 8% protons
    (1-800 eV) and 2% electrons (1-700 keV), in a uniform E = E_field V/m along z plus the
    coil B map (n_grid^3 over the vessel); species proper-time limits 3e-3 s (protons) and
    1e-5 s (electrons) as in in/config.in.  Walls: 335 neV, 2% Lambert.  Start points are
    uniform inside the mesh, found with the GPU's TGeometry::GetSolid (needs a GPU)."""
    from .constants import ELECTRON, PROTON
    from .fields import uniform_efield
    from .geometry import read_stl
    from .queries import solid_at
    rng = np.random.default_rng(seed)
    if mesh in REFERENCE_MESHES:
        _, tris = read_stl(os.path.join(MESH_DIR, REFERENCE_MESHES[mesh]))
    elif mesh == "cylinder":
        n_phi = max(12, int(np.sqrt(n_tris / 1.2 + 81) - 9))
        tris = cylinder_mesh(0.25, 0.5, n_phi, max(4, int(0.6 * n_phi)), 5)
    else:
        raise ValueError(f"unknown mesh {mesh}")
    wall = Material("wall", FermiReal=335.0, DiffProb=0.02)
    geom = vessel_geometry(tris, wall)
    lo, hi = tris.reshape(-1, 3).min(0).astype(float), tris.reshape(-1, 3).max(0).astype(float)
    pos = np.empty((n, 3))
    k = 0
    while k < n:
        p = rng.uniform(lo, hi, size=(max(2 * (n - k), 1024), 3))
        p = p[solid_at(geom, p, 0.0) == 1]
        take = min(len(p), n - k)
        pos[k:k + take] = p[:take]
        k += take
    kinds = np.where(rng.uniform(size=n) < 0.9, NEUTRON, np.where(rng.uniform(size=n) < 0.8, PROTON, ELECTRON))
    kinds = kinds.astype(np.int32)
    E = np.where(kinds == PROTON, rng.uniform(1, 800, n), np.where(kinds == ELECTRON, rng.uniform(1e3, 7e5, n),
                                                                  rng.uniform(0, 250e-9, n)))
    phi = rng.uniform(0, 2 * np.pi, n)
    theta = np.arccos(rng.uniform(-1, 1, n))
    y = np.empty((n, 9))
    spin0 = np.zeros((n, 5))
    for kd in (NEUTRON, PROTON, ELECTRON):
        sel = kinds == kd
        if sel.any():
            _, yk, _ = make_states(kd, pos[sel], E[sel], phi[sel], theta[sel], 1.0)
            y[sel] = yk
    fmap, meta = coil_map(seed, n_grid, lo, hi, Bmax=Bmax)
    fm = TFieldManager([fmap, uniform_efield(0.0, 0.0, E_field)])
    species = [SpeciesConfig(NEUTRON, tau=880.0, tmax=9e99), SpeciesConfig(PROTON, tmax=3e-3),
               SpeciesConfig(ELECTRON, tmax=1e-5)]
    lpb = np.where(kinds == NEUTRON, rng.uniform(loss[0], loss[1], n), 0.0)
    meta = dict(meta, mesh=mesh, mesh_triangles=len(tris))
    return Workload(f"uneven_{mesh}", geom, fm, species, tmax, kinds, np.zeros(n), y, spin0,
                    np.arange(n, dtype=np.uint64), seed, meta, loss=lpb)


def analytic(n=300, seed=6, tmax=0.1, n_grid=16):
    """Neutrons, protons and electrons in a 1 m box under every analytic field type:
    a finite wire (12.5 kA, outside the cube) switched on by a time ramp, the six analytic
    B fields of src/analyticFields.cpp with bounds and time-dependent scales, a coil map
    with a ramped scale and a tapered boundary, and a uniform E whose scale steps in time."""
    from .constants import ELECTRON, PROTON
    from .fields import analytic_field, conductor, uniform_efield
    from .geometry import hollow_box_mesh
    rng = np.random.default_rng(seed)
    # a 1 mm shell (solid 2, 335 neV, 5% diffuse) around the unit cube; the inside is the
    # default solid, where charged particles fly until they enter the shell
    mats = [Material("vacuum"), Material("wall", FermiReal=335.0, DiffProb=0.05)]
    geom = TGeometry(mats, "vacuum", [Solid(2, "wall", hollow_box_mesh((0, 0, 0), (1, 1, 1), 0.001))])
    kinds = np.where(np.arange(n) % 3 == 0, PROTON, np.where(np.arange(n) % 3 == 1, ELECTRON, NEUTRON))
    kinds = kinds.astype(np.int32)
    _, y0, spin0 = neutron_source(rng, n, lambda p: np.ones(len(p), bool), 1e-3, 1 - 1e-3, 300e-9)
    E = np.where(kinds == PROTON, rng.uniform(1, 800, n), np.where(kinds == ELECTRON, rng.uniform(1e3, 7e5, n),
                                                                  rng.uniform(0, 300e-9, n)))
    phi = rng.uniform(0, 2 * np.pi, n)
    theta = np.arccos(rng.uniform(-1, 1, n))
    y = y0.copy()
    for k in (NEUTRON, PROTON, ELECTRON):
        sel = kinds == k
        _, yk, _ = make_states(k, y0[sel, :3], E[sel], phi[sel], theta[sel], y0[sel, 7])
        y[sel] = yk
    fmap, meta = coil_map(seed, n_grid, (0, 0, 0), (1, 1, 1), Bmax=0.05)
    fmap.bformula = "t < 0.02 ? 0.2 : (t < 0.06 ? 0.2 + 20*(t - 0.02) : 1)"
    fmap.box, fmap.bw = (1.0, 0.0, 1.0, 0.0, 1.0, 0.0), 0.1
    containers = [
        conductor(12500.0, (1.5, 0.5, -1.0), (1.5, 0.5, 2.0), bscale="t < 0.01 ? 0 : (t < 0.03 ? 50*(t-0.01) : 1)"),
        analytic_field("B0GradZ", (2e-3, -1e-3, 1e-3), box=(1, 0, 1, 0, 1, 0), bscale="0.5 + 0.5*cos(2*pi*t/0.05)"),
        analytic_field("ExponentialFieldX", (5e-3, 1.0, -0.5, 1e-4, 2e-4), box=(3, -3, 1, -1, 1.5, -1)),
        analytic_field("LinearFieldZ", (2e-3, 1e-3), box=(2, -2, 2, -2, 2, -2), bscale="t/0.1"),
        analytic_field("B0GradX2", (1e-3, 2e-4, -3e-4, 1e-3), box=(0.9, 0.1, 0.9, 0.1, 0.9, 0.1)),
        analytic_field("B0GradXY", (1e-3, -2e-4, 5e-4)),
        analytic_field("B0_XY", (1e-3, 2e-3), box=(1, 0, 1, 0, 0.7, 0.3), bscale="t > 0.05"),
        fmap,
        uniform_efield(0.0, 1e4, 1e5, escale="t < 0.02 ? 1 : 0.5"),
    ]
    species = [SpeciesConfig(NEUTRON, tau=0.0, tmax=9e99), SpeciesConfig(PROTON, tmax=3e-3),
               SpeciesConfig(ELECTRON, tmax=1e-5)]
    t0 = rng.uniform(0.0, 0.06, n)     # start times across the ramps of the scale formulas
    return Workload("analytic", geom, TFieldManager(containers), species, tmax, kinds, t0, y, spin0,
                    np.arange(n, dtype=np.uint64), seed, meta)
