"""GPU: the reference's own geometry test and the tracker's failure paths, bit for bit
against the FP64 oracle.

- test/GeometryTest/config.in as the reference runs it (test/GeometryTest/RunTest.sh):
  neutrons stored in a hollow steel cube that holds unit cubes switching material at
  t = 5 s (ignore intervals 5-10 and 0-5).  Its README asks for no geometry errors at the
  end of the simulation; the command line must report no -6 or -7.
- A chord lying in the plane of a triangle it overlaps: the reference's CGAL intersection
  is a segment and TTriangleMesh::Collision throws (src/trianglemesh.cpp:147), caught as
  ID_CGAL_ERROR (-6).
- Leaving a solid the particle is not in (a mesh turned inside out, entered from
  outside): the inconsistent-crossing branch of TTracker::DoHit (src/tracking.cpp:310-320),
  ID_GEOMETRY_ERROR (-7).
- A neutron at rest: the first step 10 * MAX_TRACK_DEVIATION / |v| is infinite
  (src/tracking.cpp:64).
- pt_request_quit (the reference's `quit`, include/globals.h:33) between and during runs,
  then resumed: the same results as an uninterrupted run.
- A plain C program (pentrack_amd/csrc/pt_capi_example.c) that calls pt_track_batch
  through include/pentrack_gpu.h with no Python involved.
"""
import ctypes as C
import os
import subprocess
import sys
import threading
import time

import numpy as np
import pytest

from pentrack_amd import _native
from pentrack_amd.constants import NEUTRON
from pentrack_amd.fields import TFieldManager
from pentrack_amd.geometry import Material, Solid, TGeometry, box_mesh
from pentrack_amd.particle import SpeciesConfig, make_states
from pentrack_amd.tracker import Batch, LaunchConfig
from pentrack_amd.workloads import Workload, cube
from tests.test_gpu_parity import assert_bit_exact

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEOMETRY_TEST = os.path.join(ROOT, "tests", "golden", "GeometryTest", "config.in")
HITS = 1024


def run_gpu(w, hit_cap=HITS, cfg=None):
    b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid,
              hit_cap=hit_cap)
    assert b.run_to_completion(cfg or LaunchConfig(), max_launches=100000)
    return b.download()


def geometry_test_workload(seed):
    """The particles `python -m pentrack_amd 0 config.in out seed` tracks for GeometryTest."""
    from pentrack_amd import config as cfgmod
    from pentrack_amd.source import create_source
    cfg = cfgmod.TConfig(GEOMETRY_TEST)
    base = os.path.dirname(GEOMETRY_TEST)
    g = cfgmod.global_settings(cfg)
    geometry, fields = cfgmod.load_geometry(cfg, base), cfgmod.load_fields(cfg, base)
    source = create_source(cfg, base)
    t0, y0, spin0 = source.sample(g.simcount, np.random.default_rng(seed), geometry, fields)
    species = [cfgmod.species_config(cfg, n) for n in ("neutron", "proton", "electron")]
    n = g.simcount
    return Workload("GeometryTest", geometry, fields, species, g.simtime, np.full(n, source.kind, np.int32), t0, y0,
                    spin0, np.arange(1, n + 1, dtype=np.uint64), seed)


def stop_counts(line):
    """'neutron: -4: 12, -1: 9000' -> {-4: 12, -1: 9000}"""
    body = line.split(":", 1)[1]
    return {int(k): int(v) for k, v in (item.split(":") for item in body.split(","))}


def test_geometry_test_command_line_reports_no_geometry_errors(gpu, tmp_path):
    seed = 20240
    print("case: GeometryTest through python -m pentrack_amd, simcount 10000, simtime 10 s", flush=True)
    r = subprocess.run([sys.executable, "-m", "pentrack_amd", "0", GEOMETRY_TEST, str(tmp_path / "out"), str(seed)],
                       cwd=ROOT, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("neutron: ")]
    assert len(lines) == 1, r.stdout[-2000:]
    counts = stop_counts(lines[0])
    assert sum(counts.values()) == 10000
    assert counts.get(-6, 0) == 0 and counts.get(-7, 0) == 0, counts
    assert not set(counts) & {-3, -8, -9}, counts
    # the command line's counts are those of the same particles tracked in this process
    w = geometry_test_workload(seed)
    g, _ = run_gpu(w, hit_cap=0)
    ids, cnt = np.unique(g["stopID"], return_counts=True)
    assert counts == {int(i): int(c) for i, c in zip(ids, cnt)}
    # the detector cubes switched on at 5 s and the steel absorb most neutrons within 10 s
    assert set(counts) <= {-4, -2, -1, 1, 2} and counts.get(1, 0) > 5000 and counts.get(-1, 0) > 0


def test_geometry_test_bit_exact(gpu, oracle_lib):
    w = geometry_test_workload(20240)
    sel = np.arange(2000)
    sub = Workload("GeometryTest[:2000]", w.geometry, w.fields, w.species, w.tmax, w.species_idx[sel], w.t0[sel],
                   w.y0[sel], w.spin0[sel], w.pid[sel], w.seed)
    print("case: GeometryTest first 2000 particles, 10 s, full hit logs", flush=True)
    g, gh = run_gpu(sub)
    o, oh = oracle_lib.track(sub, "fp64", hit_cap=HITS, threads=8)
    assert_bit_exact(g, gh, o, oh)
    assert not np.isin(g["stopID"], [-3, -6, -7, -8, -9]).any()
    # the switch at t = 5 s (src/geometry.cpp:145,155; src/tracking.cpp:297): a solid keeps
    # the ignore state it had when the particle entered it, so the detector cubes (IDs 5, 6,
    # ignored 0-5 s) are entered only after 5 s and the default-material cubes (IDs 2, 3,
    # ignored 5-10 s) only before; particles inside 2 or 3 at 5 s still leave them later
    logged = np.arange(HITS)[None, :] < g["n_hits"][:, None]
    t, ent, lea = gh["t"][logged], gh["entering"][logged], gh["leaving"][logged]
    assert np.isin(ent, [5, 6]).any() and np.isin(ent, [2, 3]).any()
    assert (t[np.isin(ent, [5, 6]) | np.isin(lea, [5, 6])] >= 5.0).all()
    assert (t[np.isin(ent, [2, 3])] < 5.0).all()
    assert (t[np.isin(lea, [2, 3])] >= 5.0).any()


def vessel(extra, mats=()):
    """A 1 m cube of 'vacuum' (ID 2) inside the default solid 'wall' (335 neV), plus solids."""
    materials = [Material("wall", FermiReal=335.0), Material("vacuum")] + list(mats)
    return TGeometry(materials, "wall", [Solid(2, "vacuum", box_mesh((0, 0, 0), (1, 1, 1)))] + list(extra))


def neutrons(pos, vel, tmax, seed):
    pos = np.asarray(pos, float)
    n = len(pos)
    t0, y0, spin0 = make_states(NEUTRON, pos, np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n))
    y0[:, 3:6] = vel
    return Workload("scenario", None, TFieldManager([]), [SpeciesConfig(kind=NEUTRON)], tmax, np.zeros(n, np.int32),
                    t0, y0, spin0, np.arange(n, dtype=np.uint64), seed)


def test_coplanar_chord_stops_as_cgal_error(gpu, oracle_lib):
    """A fin (solid 3, x in [0.5, 0.6]) whose face x = 0.5 holds neutrons moving with vx = 0
    exactly at x = 0.5: every chord over that face lies in its plane."""
    fin = Solid(3, "vacuum", box_mesh((0.5, 0.3, 0.1), (0.6, 0.7, 0.9)))
    rng = np.random.default_rng(7)
    m = 64
    pos = np.column_stack([np.full(m, 0.5), rng.uniform(0.02, 0.25, m), rng.uniform(0.3, 0.7, m)])
    vel = np.column_stack([np.zeros(m), rng.uniform(1.0, 4.0, m), rng.uniform(-0.5, 0.5, m)])
    # and neutrons elsewhere in the vessel, which the fin reflects or lets through
    pos = np.concatenate([pos, np.column_stack([rng.uniform(0.05, 0.45, m), rng.uniform(0.05, 0.95, m),
                                                rng.uniform(0.05, 0.95, m)])])
    vel = np.concatenate([vel, rng.normal(0, 2.0, (m, 3))])
    w = neutrons(pos, vel, 1.0, 7)
    w.geometry = vessel([fin])
    print("case: coplanar chord, 128 neutrons", flush=True)
    g, gh = run_gpu(w)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS)
    assert_bit_exact(g, gh, o, oh)
    assert (g["stopID"][:m] == _native.ID_CGAL_ERROR).all(), g["stopID"][:m]
    assert (g["stopID"][m:] != _native.ID_CGAL_ERROR).mean() > 0.9


def test_inverted_mesh_stops_as_geometry_error(gpu, oracle_lib):
    """Solid 3 with every triangle turned around: its normals point inward, so entering it
    from outside looks like leaving a solid the particle is not in."""
    inner = box_mesh((0.3, 0.3, 0.3), (0.7, 0.7, 0.7))[:, ::-1, :].copy()
    rng = np.random.default_rng(8)
    n = 128
    pos = rng.uniform(0.02, 0.98, (4 * n, 3))
    pos = pos[(np.abs(pos - 0.5) > 0.22).any(axis=1)][:n]
    vel = rng.normal(0, 2.5, (n, 3))
    w = neutrons(pos, vel, 2.0, 8)
    w.geometry = vessel([Solid(3, "vacuum", inner)])
    print("case: inverted inner mesh, 128 neutrons", flush=True)
    g, gh = run_gpu(w)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS)
    assert_bit_exact(g, gh, o, oh)
    assert (g["stopID"] == _native.ID_GEOMETRY_ERROR).mean() > 0.5
    assert np.isin(g["stopID"], [_native.ID_GEOMETRY_ERROR, _native.ID_NOT_FINISH]).all()


def test_neutron_at_rest(gpu, oracle_lib):
    """E = 0: the initial step 10 * MAX_TRACK_DEVIATION / |v| is infinite, the first trial
    step's state is not finite and odeint's max-norm of the error (which passes over NaN)
    accepts it.  The engine stops such a particle as a stepper failure (ID_ODEINT_ERROR) with
    the state of the step start; the reference would carry the NaN state on.  GPU and oracle
    agree bit for bit."""
    rng = np.random.default_rng(9)
    n = 16
    pos = rng.uniform(0.1, 0.9, (n, 3))
    vel = np.zeros((n, 3))
    vel[8:, 2] = rng.uniform(0.5, 2.0, 8)       # and, for contrast, vertical motion only
    w = neutrons(pos, vel, 1.0, 9)
    w.geometry = vessel([])
    print("case: neutrons at rest", flush=True)
    g, gh = run_gpu(w)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS)
    assert_bit_exact(g, gh, o, oh)
    assert (g["stopID"][:8] == _native.ID_ODEINT_ERROR).all() and (g["t"][:8] == w.t0[:8]).all()
    assert np.array_equal(g["y"][:8], w.y0[:8]) and (g["Nstep"][:8] == 0).all()
    assert (g["stopID"][8:] == _native.ID_NOT_FINISH).all()


def test_quit_between_runs_then_resume(gpu):
    """With the quit flag set, pt_batch_run advances no particle; cleared, the run resumes
    and ends as an uninterrupted one, bit for bit."""
    w = cube(n=2000, seed=12, tmax=2.0)
    ref, ref_h = run_gpu(w, hit_cap=HITS)
    cfg = LaunchConfig(steps_per_launch=5)
    b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid, hit_cap=HITS)
    L = _native.lib()
    print("case: quit between runs, cube n=2000", flush=True)
    try:
        for _ in range(3):
            b.run(cfg)
        before, before_h = b.download()
        active = b.active()
        assert 0 < active == (before["stopID"] == 0).sum()
        L.pt_request_quit(1)
        for _ in range(3):
            b.run(cfg)
        after, after_h = b.download()
        for f in ("t", "y", "stopID", "Nstep", "Nhit", "n_acc", "n_rej"):
            assert np.array_equal(before[f], after[f]), f
        assert np.array_equal(before["t"], after["t"]) and b.active() == active
    finally:
        L.pt_request_quit(0)
    assert b.run_to_completion(cfg, max_launches=100000)
    g, gh = b.download()
    assert_bit_exact(g, gh, ref, ref_h)


def track_batch(w, cfg, quit_after=None):
    """pt_track_batch on the workload; with quit_after, pt_request_quit(1) from another
    thread that many seconds into the call.  Returns (out, seconds)."""
    L = _native.lib()
    n = w.n
    sp = (_native.Species * 1)(w.species[0].native())
    out = np.zeros(n, dtype=_native.OUT_DTYPE)
    c = cfg.native()
    keep = [np.ascontiguousarray(w.species_idx, np.int32), np.ascontiguousarray(w.t0), np.ascontiguousarray(w.y0),
            np.ascontiguousarray(w.spin0), np.ascontiguousarray(w.pid, np.uint64)]
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    args = (None, w.geometry.device(), w.fields.device(), 1, sp, w.tmax, n, keep[0].ctypes.data_as(C.POINTER(C.c_int32)),
            dp(keep[1]), dp(keep[2]), dp(keep[3]), w.seed, keep[4].ctypes.data_as(C.POINTER(C.c_uint64)),
            out.ctypes.data_as(C.c_void_p), None, 0, C.byref(c))
    timer = threading.Timer(quit_after, L.pt_request_quit, args=(1,)) if quit_after else None
    t = time.perf_counter()
    if timer:
        timer.start()
    try:
        rc = L.pt_track_batch(*args)
    finally:
        if timer:
            timer.cancel()
        L.pt_request_quit(0)
    _native.check(rc, "pt_track_batch")
    return out, time.perf_counter() - t


def test_quit_during_track_batch(gpu):
    """pt_request_quit from another thread while pt_track_batch loops over its launches: the
    call returns early; particles it finished equal the uninterrupted run's, the others are
    left unfinished (stopID 0, as the reference leaves the particles it abandons)."""
    w = cube(n=20000, seed=13, tmax=20.0)
    w.t0 = np.random.default_rng(13).uniform(0.0, 20.0, w.n)    # a spread of finishing times
    cfg = LaunchConfig(steps_per_launch=1)
    ref, seconds = track_batch(w, cfg)
    assert (ref["stopID"] != 0).all()
    print(f"case: quit during pt_track_batch, cube n=20000, {seconds:.2f} s uninterrupted", flush=True)
    out, _ = track_batch(w, cfg, quit_after=0.3 * seconds)
    done = out["stopID"] != 0
    print(f"quit left {int((~done).sum())} of {w.n} particles unfinished", flush=True)
    assert done.any() and (~done).any()
    for f in ("t", "y", "stopID", "Nhit", "Nstep", "n_acc", "n_rej", "Hmax"):
        assert np.array_equal(out[f][done], ref[f][done]), f
    # the unfinished particles were stopped on the way
    assert (out["t"][~done] < ref["t"][~done]).all()


def read_capi_output(path):
    raw = open(path, "rb").read()
    assert raw[:4] == b"PTCX"
    n, n_tri, cap, _ = np.frombuffer(raw, np.int32, 4, 4)
    seed = int(np.frombuffer(raw, np.uint64, 1, 20)[0])
    tmax = float(np.frombuffer(raw, np.float64, 1, 28)[0])
    off = 36

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off)
        off += a.nbytes
        return a
    tri = take(np.float32, n_tri * 9).reshape(n_tri, 3, 3)
    mats = take(np.float64, 18).reshape(2, 9)
    t0 = take(np.float64, n).copy()
    y0 = take(np.float64, n * 9).reshape(n, 9).copy()
    spin0 = take(np.float64, n * 5).reshape(n, 5).copy()
    pid = take(np.uint64, n).copy()
    out = take(_native.OUT_DTYPE, n).copy()
    hits = take(_native.HIT_DTYPE, n * cap).reshape(n, cap).copy()
    assert off == len(raw)
    return dict(n=n, cap=cap, seed=seed, tmax=tmax, tri=tri, mats=mats, t0=t0, y0=y0, spin0=spin0, pid=pid, out=out,
                hits=hits)


def test_c_program_through_the_header(gpu, oracle_lib, tmp_path):
    from pentrack_amd.build import C_EXAMPLE
    assert os.path.exists(C_EXAMPLE), "pt_capi_example is built by build()"
    path = str(tmp_path / "capi.bin")
    print("case: pt_capi_example 3000 neutrons", flush=True)
    r = subprocess.run([C_EXAMPLE, "3000", "99", path], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "tracked 3000 neutrons" in r.stdout
    c = read_capi_output(path)
    wall = c["mats"][0]
    vac = c["mats"][1]
    assert wall[0] == 335.0 and vac[2] == 0.1
    geom = TGeometry([Material("wall", FermiReal=335.0), Material("vacuum", DiffProb=0.1)], "wall",
                     [Solid(2, "vacuum", c["tri"])])
    assert np.array_equal(c["tri"], box_mesh((0, 0, 0), (1, 1, 1)))
    n = c["n"]
    w = Workload("capi", geom, TFieldManager([]), [SpeciesConfig(kind=NEUTRON)], c["tmax"], np.zeros(n, np.int32),
                 c["t0"], c["y0"], c["spin0"], c["pid"], c["seed"])
    o, oh = oracle_lib.track(w, "fp64", hit_cap=c["cap"], threads=8)
    assert_bit_exact(c["out"], c["hits"], o, oh)
    # and the Python route over the same library gives the same
    g, gh = run_gpu(w, hit_cap=c["cap"])
    assert_bit_exact(g, gh, o, oh)
    assert (c["out"]["Nhit"] > 0).mean() > 0.9
