"""GPU: snapshots (TLogger::PrintSnapshot, src/logger.cpp:128-140, src/tracking.cpp:126):
the stepper's dense-output state at the first snapshot time of each step, with the
particle's counters, bit for bit against the oracle, and on a free-fall parabola."""
import numpy as np
import pytest

from pentrack_amd import workloads
from pentrack_amd.constants import NEUTRON
from pentrack_amd.tracker import Batch, LaunchConfig

pytestmark = pytest.mark.gpu

# unsorted, a duplicate, and two in one step with the later time first: find_if over the
# configured order (src/logger.cpp:134) records 1.0 + 1e-9 there, not 1.0
TIMES = [2.5, 0.1, 0.5, 1.0 + 1e-9, 1.0, 0.5]
SORTED_TIMES = [0.1, 0.5, 1.0, 1.0 + 1e-9, 2.5]


def run_snapshots(w, times, cap, cfg=None):
    b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid)
    b.set_snapshots(times, cap)
    assert b.run_to_completion(cfg or LaunchConfig(), max_launches=100000)
    out, _ = b.download()
    return out, *b.snapshots()


@pytest.mark.parametrize("persistent", [0, 1])
def test_snapshots_match_oracle(gpu, oracle_lib, persistent):
    w = workloads.cylinder(n=1500, seed=7, tmax=3.0, n_grid=24)
    print("case: snapshots cylinder n=1500 persistent", persistent, flush=True)
    out, rec, cnt = run_snapshots(w, TIMES, 4, LaunchConfig(persistent=persistent))
    o, orec, ocnt = oracle_lib.track_snapshots(w, TIMES, 4, threads=8)
    assert np.array_equal(out["y"], o["y"]) and np.array_equal(cnt, ocnt)
    assert cnt.max() <= 4 and (cnt > 0).mean() > 0.9
    for k in range(w.n):
        a, b = rec[k, :cnt[k]], orec[k, :ocnt[k]]
        for f in ("t", "y", "spin", "Hmax", "noflipprob", "stopID", "Nhit", "Nstep", "Nspinflip"):
            assert np.array_equal(a[f], b[f]), (k, f)
    # at most one record per step: 1.0 and 1.0 + 1e-9 fall into one step of almost every
    # particle, where the configured order picks 1.0 + 1e-9 (slots past a particle's count
    # hold no record)
    t = np.concatenate([rec["t"][k, :min(cnt[k], 4)] for k in range(w.n)])
    assert (np.isin(t, [0.1, 0.5, 1.0, 1.0 + 1e-9, 2.5])).all()
    assert (t == 1.0 + 1e-9).sum() > 10 * (t == 1.0).sum()
    # a sorted list is searched by bisection and gives find_if's answer for that order
    out2, rec2, cnt2 = run_snapshots(w, SORTED_TIMES, 4, LaunchConfig(persistent=persistent))
    o2, orec2, ocnt2 = oracle_lib.track_snapshots(w, SORTED_TIMES, 4, threads=8)
    assert np.array_equal(cnt2, ocnt2)
    for k in range(w.n):
        m = min(cnt2[k], 4)
        assert np.array_equal(rec2["t"][k, :m], orec2["t"][k, :m]) and np.array_equal(rec2["y"][k, :m], orec2["y"][k, :m])


def test_snapshots_on_a_parabola(gpu):
    """No walls in reach: the snapshot state of a falling neutron is the parabola."""
    w = workloads.cube(n=200, seed=3, tmax=0.2)
    w.y0[:, :3] = [0.5, 0.5, 0.9]
    w.y0[:, 3:6] = np.random.default_rng(1).uniform(-0.5, 0.5, (200, 3))
    w.y0[:, 8] = 0
    w.t0[:] = 0
    times = [0.05, 0.1, 0.15]
    print("case: snapshots parabola", flush=True)
    out, rec, cnt = run_snapshots(w, times, 3)
    assert (cnt == 3).all()
    g = 9.80665
    for j, t in enumerate(times):
        r = rec[:, j]
        assert (r["t"] == t).all()
        want_z = w.y0[:, 2] + w.y0[:, 5] * t - 0.5 * g * t * t
        assert np.allclose(r["y"][:, 2], want_z, rtol=0, atol=1e-9)
        assert np.allclose(r["y"][:, 0], w.y0[:, 0] + w.y0[:, 3] * t, rtol=0, atol=1e-9)


@pytest.mark.parametrize("persistent", [0, 1])
def test_track_log_matches_oracle(gpu, oracle_lib, persistent):
    """TLogger::PrintTrack records (src/tracking.cpp:140): step ends where the trajectory
    length crossed a multiple of the interval, with the current solid, bit for bit."""
    w = workloads.cylinder(n=1000, seed=8, tmax=1.0, n_grid=24)
    b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid)
    b.set_track(0.05, 64)
    print("case: track log cylinder n=1000 persistent", persistent, flush=True)
    assert b.run_to_completion(LaunchConfig(persistent=persistent), max_launches=100000)
    out, _ = b.download()
    rec, cnt = b.track()
    o, orec, ocnt = oracle_lib.track_tracklog(w, 0.05, 64, threads=8)
    assert np.array_equal(out["y"], o["y"]) and np.array_equal(cnt, ocnt)
    assert (cnt > 3).mean() > 0.9                    # counts go on past the 64 records kept
    kept = np.minimum(cnt, 64)
    for k in range(w.n):
        assert np.array_equal(rec[k, :kept[k]], orec[k, :kept[k]]), k
    # consecutive records are at least one interval apart in trajectory length
    for k in range(0, w.n, 50):
        L = rec["y"][k, :kept[k], 8]
        assert (np.floor(L / 0.05)[1:] > np.floor(L / 0.05)[:-1]).all()
