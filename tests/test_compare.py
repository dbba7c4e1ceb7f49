"""pentrack_amd.compare on hand-made runs: records keyed by their time (not their position
in the list), particles that stop before the horizon compared on their path only, stop
reasons and hit sequences required equal."""
import numpy as np
import pytest

from pentrack_amd import _native, compare
from pentrack_amd.compare import deviation, snapshot_times


def run(ends, stops, tends, snaps, hits):
    """ends [n][3], stops [n], tends [n]; snaps: per particle a list of (t, position);
    hits: per particle a list of triangle indices."""
    n = len(ends)
    out = np.zeros(n, dtype=_native.OUT_DTYPE)
    out["t"] = tends
    out["stopID"] = stops
    out["y"][:, :3] = ends
    cap = max(len(s) for s in snaps)
    rec = np.zeros((n, cap), dtype=_native.SNAP_DTYPE)
    cnt = np.zeros(n, np.int32)
    for k, s in enumerate(snaps):
        for r, (t, pos) in enumerate(s):
            rec[k, r]["t"] = t
            rec[k, r]["y"][:3] = pos
        cnt[k] = len(s)
    hcap = max(1, max(len(h) for h in hits))
    hr = np.zeros((n, hcap), dtype=_native.HIT_DTYPE)
    for k, h in enumerate(hits):
        out["n_hits"][k] = len(h)
        for r, tri in enumerate(h):
            hr[k, r]["t"] = 0.001 * (r + 1)
            hr[k, r]["tri"] = tri
    return out, hr, rec, cnt


def test_snapshot_times():
    t = snapshot_times(0.25, (0.5, 1.0))
    assert t[0] == 0.01 and 0.25 in t and 0.5 in t and t[-1] == 1.0 and np.all(np.diff(t) > 0)
    assert np.count_nonzero(t <= 0.25) == 25


def test_records_keyed_by_time_and_stopped_particles():
    times = np.array([0.1, 0.2, 0.3])
    z = np.zeros(3)
    # particle 0 reaches the horizon 0.3 (end states there); run b did not record 0.1, so
    # index alignment would compare its 0.2 record with run a's 0.1 record
    # particle 1 stops at 0.15 in both runs (different end points, same reason)
    a = run([z, [1., 0, 0]], [-1, -2], [0.3, 0.150], [[(0.1, z), (0.2, [0.0, 0, 0])], [(0.1, z)]], [[1, 2], [3]])
    b = run([[5e-7, 0, 0], [1.001, 0, 0]], [-1, -2], [0.3, 0.151], [[(0.2, [4e-7, 0, 0])], [(0.1, [1e-7, 0, 0])]],
            [[1, 2], [3]])
    r = deviation(a, b, times, 0.3)
    assert r["flying_at_horizon"] == 1 and r["stopped_before"] == 1 and r["snapshots_compared"] == 2
    assert np.isclose(r["trajectory_position_max_m"], 5e-7) and np.isclose(r["end_position_max_m"], 5e-7)
    assert np.isclose(r["stopped_end_position_max_m"], 1e-3) and np.isclose(r["stopped_stop_time_difference_max_s"], 1e-3)
    assert r["same_stop_reasons"] == 1.0 and r["same_hit_sequences"] == 1.0 and r["condition_holds"]
    # a different hit sequence, a different stop reason or a larger deviation fails
    c = run([[2e-6, 0, 0], [1., 0, 0]], [-1, -3], [0.3, 0.150], [[(0.1, z), (0.2, z)], [(0.1, z)]], [[1, 5], [3]])
    r2 = deviation(a, c, times, 0.3)
    assert r2["same_hit_sequences"] == 0.5 and r2["same_stop_reasons"] == 0.5
    assert np.isclose(r2["trajectory_position_max_m"], 2e-6) and not r2["condition_holds"]
    with pytest.raises(ValueError):
        deviation(a, b, times, 0.25)
