"""Chord splitting (src/tracking.cpp:96-123) and the collision bisection
(TTracker::iterate_collision, src/tracking.cpp:228-265) of the oracle pinned by closed
forms of free fall, independently of the GPU:

* chord pieces: a step from x1 to x is cut while the sagitta bound
  sqrt((l^2 - d^2) / 4) of the remaining arc (path length l, chord d) exceeds
  MAX_TRACK_DEVIATION = 1 mm (include/particle.h:20), the next piece ending at
  x1 + (x - x1) / ceil(bound / 1 mm); the number of pieces (Nstep, one DoStep per piece,
  src/particle.cpp:205) equals the count replayed from the closed-form parabola and arc
  length at the oracle's step ends;
* bisection: the bracket ends before the wall within REFLECT_TOLERANCE = 1e-8 m
  (include/trianglemesh.h:39) of the closed-form crossing point, at a time before the
  closed-form crossing time; with two walls in one chord the first half is searched first,
  so the earlier wall is the one hit."""
import numpy as np

from oracle import ptref
from pentrack_amd.constants import GRAVCONST, NEUTRON
from pentrack_amd.fields import TFieldManager
from pentrack_amd.geometry import Material, Solid, TGeometry, box_mesh, hollow_box_mesh
from pentrack_amd.particle import SpeciesConfig
from pentrack_amd.workloads import Workload

MAXDEV = 1e-3


def parabola(p0, v0, t):
    t = np.asarray(t, float)[..., None]
    p = p0 + v0 * t
    p[..., 2] -= 0.5 * GRAVCONST * t[..., 0] ** 2
    return p


def arc_length(v0, t):
    """Closed-form path length of free fall from 0 to t: |v| = sqrt(a^2 + u^2), u = vz - g t."""
    a = np.hypot(v0[0], v0[1])

    def F(u):
        if a == 0:
            return 0.5 * u * abs(u)
        return 0.5 * (u * np.sqrt(u * u + a * a) + a * a * np.arcsinh(u / a))
    return (F(v0[2]) - F(v0[2] - GRAVCONST * t)) / GRAVCONST


def workload(geom, pos, vel, tmax, t0=0.0):
    n = len(pos)
    y0 = np.zeros((n, 9))
    y0[:, :3], y0[:, 3:6], y0[:, 7] = pos, vel, 1.0
    sp = [SpeciesConfig(kind=NEUTRON, tau=0.0, tmax=9e99, lmax=9e99)]
    return Workload("freefall", geom, TFieldManager([]), sp, tmax, np.zeros(n, np.int32), np.full(n, t0), y0,
                    np.zeros((n, 5)), np.arange(n, dtype=np.uint64), 1)


def test_chord_piece_count_free_fall(oracle_lib):
    """Neutrons far from any wall for 0.6 s: Nstep equals the piece count of the reference's
    splitting rule evaluated on the exact parabola at the oracle's step ends."""
    rng = np.random.default_rng(12)
    n = 120
    geom = TGeometry([Material("vacuum"), Material("wall", FermiReal=335.0)], "vacuum",
                     [Solid(2, "wall", box_mesh((-50, -50, -50), (50, 50, 50)))])
    pos = np.zeros((n, 3))
    speed = rng.uniform(0.3, 6.0, n)
    d = rng.normal(size=(n, 3))
    vel = d / np.linalg.norm(d, axis=1)[:, None] * speed[:, None]
    vel[:5, :2] = 0.0                                    # vertical downward throws: straight,
    vel[:5, 2] = -np.abs(vel[:5, 2])                     # one piece per step
    w = workload(geom, pos, vel, 0.6)
    out, rec, cnt = ptref.track_tracklog(w, 1e-6, 4096)
    assert (out["stopID"] == -1).all() and (cnt < 4096).all()
    checked = 0
    for k in range(n):
        ends = rec["t"][k, :cnt[k]]
        assert ends[-1] == 0.6
        pieces, ambiguous = 0, False
        x1 = 0.0
        # the oracle's path length differs from the closed form by its stepper's error at
        # the step ends and by the dense output's interpolation error at interior piece
        # ends (which dominates at the long steps of free fall): lerr from the step ends,
        # scaled up, bounds the error of l below
        lerr = np.abs(rec["y"][k, :cnt[k], 8] - np.array([arc_length(vel[k], x) for x in ends]))
        dl = 20 * lerr.max() + 1e-8
        for j, x in enumerate(ends):
            while x1 < x:
                l = arc_length(vel[k], x) - arc_length(vel[k], x1)
                dd = np.linalg.norm(parabola(pos[k], vel[k], x) - parabola(pos[k], vel[k], x1))
                dev2 = 0.25 * (l * l - dd * dd)
                x2 = x
                # skip particles whose count depends on the path-length error dl:
                # d(dev2) = l dl / 2, so q = sqrt(dev2) / 1 mm moves by l dl / (4 sqrt(dev2) 1 mm)
                q = np.sqrt(max(dev2, 0.0)) / MAXDEV
                dq = dl * l / (4 * max(np.sqrt(max(dev2, 0.0)), 1e-12) * MAXDEV)
                ambiguous |= q > 0.5 and abs(q - np.rint(q)) < dq + 1e-9
                if dev2 > MAXDEV * MAXDEV:
                    x2 = x1 + (x - x1) / np.ceil(q)
                pieces += 1
                x1 = x2
        if ambiguous:
            continue
        assert pieces == out["Nstep"][k], (k, pieces, out["Nstep"][k], len(ends))
        checked += 1
        if k < 5:
            assert pieces == len(ends)                     # vertical: no step is split
    assert checked >= 0.75 * n, checked
    assert out["Nstep"][5:].sum() > 2 * cnt[5:].sum()      # the splits mattered


def shell_geometry():
    """1 mm shell of 335 neV material around the unit cube (solid 2); inside, vacuum."""
    return TGeometry([Material("vacuum"), Material("wall", FermiReal=335.0)], "vacuum",
                     [Solid(2, "wall", hollow_box_mesh((0, 0, 0), (1, 1, 1), 0.001))])


def crossing_time(z0, vz0, zw):
    """First t > 0 with z0 + vz0 t - g t^2 / 2 = zw."""
    return (vz0 + np.sqrt(vz0 * vz0 + 2 * GRAVCONST * (z0 - zw))) / GRAVCONST


def test_bisection_bracket_before_the_floor(oracle_lib):
    """Falling onto the floor z = 0 of the shell: the logged hit point (the bracket end
    before the wall, where the reflection is placed) lies within 1e-8 m of the closed-form
    crossing point, on the near side, at a time before the crossing time."""
    rng = np.random.default_rng(5)
    n = 40
    pos = np.column_stack([rng.uniform(0.45, 0.55, n), rng.uniform(0.45, 0.55, n), rng.uniform(0.2, 0.8, n)])
    vel = np.column_stack([rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n), rng.uniform(-2.5, 1.0, n)])
    ts = crossing_time(pos[:, 2], vel[:, 2], 0.0)
    w = workload(shell_geometry(), pos, vel, float(ts.max()) + 1e-3)
    out, hits = ptref.track(w, "fp64", hit_cap=4)
    for k in range(n):
        h = hits[k, 0]
        assert out["Nhit"][k] >= 1 and h["entering"] == 2 and h["leaving"] == 1
        pstar = parabola(pos[k], vel[k], ts[k])
        assert h["t"] < ts[k]
        assert np.linalg.norm(h["pos"] - pstar) <= 1e-8 * (1 + 1e-6)
        assert 0 <= h["pos"][2] <= 1e-8
        # the bracket is not much tighter than the tolerance either: the search stopped by
        # the 1e-8 m rule, halving at most once past it
        assert np.linalg.norm(h["pos"] - pstar) > 1e-8 / 2 ** 12
        vz = vel[k, 2] - GRAVCONST * ts[k]
        assert (ts[k] - h["t"]) * abs(vz) <= 1e-8 * (1 + 1e-6)


def test_bisection_searches_the_first_half_first(oracle_lib):
    """Vertical falls: every chord is straight, so one chord piece of a step spans both the
    inner floor z = 0 and the outer floor z = -1 mm of the shell.  The bisection looks in
    the first half of an interval first, so it brackets the earlier crossing (z = 0); a
    second-half-first search would end at z = -1 mm."""
    rng = np.random.default_rng(6)
    n = 30
    pos = np.column_stack([np.full(n, 0.5), np.full(n, 0.5), rng.uniform(0.3, 0.9, n)])
    vel = np.column_stack([np.zeros(n), np.zeros(n), rng.uniform(-3.0, -1.0, n)])
    ts = crossing_time(pos[:, 2], vel[:, 2], 0.0)
    tb = crossing_time(pos[:, 2], vel[:, 2], -0.001)
    w = workload(shell_geometry(), pos, vel, float(ts.max()) + 1e-4)
    out, hits = ptref.track(w, "fp64", hit_cap=4)
    # both walls lie on one chord piece: the step over the floor is longer than the 1 mm
    # between them (steps here are >= 10 ms at >= 1 m/s)
    assert (out["n_acc"] < 200).all()
    spans = 0
    for k in range(n):
        h = hits[k, 0]
        assert h["t"] < ts[k] < tb[k]
        assert abs(h["pos"][2]) <= 1e-8 and h["pos"][0] == 0.5 and h["pos"][1] == 0.5
        spans += 1
    assert spans == n
