"""Deviation of one tracking run from another on the same particles (the mixed-precision
kernel against the FP64 kernel, or the reference arithmetic against FP64).

Both runs track the particles to tmax = the horizon, so a particle that reaches it ends
with the stepper's dense-output state at exactly that time, and both record snapshots at
the same sorted times (TLogger::PrintSnapshot: the first snapshot time inside each step,
src/logger.cpp:134, so each run records a subset that depends on its step ends).  A
particle's trajectory deviation is the largest position difference over its end state
at the horizon and the snapshot times both runs recorded.  A particle that leaves the
geometry or stops earlier is compared on its path up to its last common snapshot: its end
state is where a step of the adaptive stepper ended (outside the bounding box for an
escape, src/tracking.cpp:188-191), and step ends differ between two arithmetics by up to a
step length, so end states of stopped particles are reported separately (with their
stop-time differences) and not held to the position bound.
"""
import numpy as np


def snapshot_times(horizon, longer=(), step=0.01, longer_step=0.05):
    """Snapshot times: every `step` seconds up to the stated horizon, then every
    `longer_step` seconds up to the longest of the longer horizons (all horizons included)."""
    t = list(np.arange(step, horizon + 0.5 * step, step))
    hi = max([horizon] + list(longer))
    t += list(np.arange(horizon + longer_step, hi + 0.5 * longer_step, longer_step))
    t += [horizon] + list(longer)
    return np.unique(np.round(np.asarray(t, dtype=np.float64), 12))


def _by_time(rec, cnt, times):
    """Snapshot states [n][len(times)][6] (position, velocity) keyed by their recorded
    time, NaN where the particle has no record at that time."""
    n = len(cnt)
    Y = np.full((n, len(times), 6), np.nan)
    for r in range(rec.shape[1] if rec.ndim == 2 else 0):
        valid = cnt > r
        if not valid.any():
            break
        t = rec[:, r]["t"]
        j = np.searchsorted(times, t)
        jj = np.clip(j, 0, len(times) - 1)
        ok = valid & (j < len(times)) & (times[jj] == t)
        Y[ok, jj[ok]] = rec[ok, r]["y"][:, :6]
    return Y


def deviation(run_a, run_b, times, horizon, bound=1e-6):
    """run_a, run_b: (out, hits, snapshot records, snapshot counts) of two runs of the same
    particles tracked to tmax = horizon with the snapshot times `times`.  Returns the
    comparison of the two runs up to the horizon."""
    (g, gh, gr, gc), (m, mh, mr, mc) = run_a, run_b
    n = len(g)
    if (g["t"] > horizon).any() or (m["t"] > horizon).any():
        raise ValueError("both runs must be tracked to tmax = horizon")
    times = np.asarray(times, dtype=np.float64)
    nj = int(np.count_nonzero(times <= horizon))
    A = _by_time(gr, gc, times)[:, :nj]
    B = _by_time(mr, mc, times)[:, :nj]
    both = ~np.isnan(A[:, :, 0]) & ~np.isnan(B[:, :, 0])
    ex = np.where(both, np.linalg.norm(A[:, :, :3] - B[:, :, :3], axis=2), 0.0)
    ev = np.where(both, np.linalg.norm(A[:, :, 3:6] - B[:, :, 3:6], axis=2), 0.0)
    flying = (g["t"] == horizon) & (m["t"] == horizon)
    stopped = (g["t"] < horizon) & (m["t"] < horizon)
    end_x = np.where(flying, np.linalg.norm(g["y"][:, :3] - m["y"][:, :3], axis=1), 0.0)
    end_v = np.where(flying, np.linalg.norm(g["y"][:, 3:6] - m["y"][:, 3:6], axis=1), 0.0)
    dx = np.maximum(ex.max(axis=1, initial=0.0), end_x)
    dv = np.maximum(ev.max(axis=1, initial=0.0), end_v)
    # both reached the horizon, or both stopped before it for the same reason
    same_stop = flying | (stopped & (g["stopID"] == m["stopID"]))
    same_hits = np.zeros(n, bool)
    for k in range(n):
        x, y = gh[k, :g["n_hits"][k]], mh[k, :m["n_hits"][k]]
        same_hits[k] = len(x) == len(y) and np.array_equal(x["tri"], y["tri"])
    res = {"horizon_s": float(horizon), "particles": n, "flying_at_horizon": int(flying.sum()),
           "stopped_before": int(stopped.sum()), "same_stop_reasons": float(same_stop.mean()),
           "same_hit_sequences": float(same_hits.mean()),
           "snapshots_compared": int(both.sum()),
           "trajectory_position_max_m": float(dx.max()), "trajectory_position_p99_m": float(np.percentile(dx, 99)),
           "trajectory_velocity_max_m_s": float(dv.max()), "trajectory_velocity_p99_m_s": float(np.percentile(dv, 99)),
           "end_position_max_m": float(end_x[flying].max()) if flying.any() else 0.0,
           "end_position_p99_m": float(np.percentile(end_x[flying], 99)) if flying.any() else 0.0,
           "end_velocity_max_m_s": float(end_v[flying].max()) if flying.any() else 0.0}
    if stopped.any():
        se = np.linalg.norm(g["y"][stopped, :3] - m["y"][stopped, :3], axis=1)
        res["stopped_end_position_max_m"] = float(se.max())
        res["stopped_stop_time_difference_max_s"] = float(np.abs(g["t"][stopped] - m["t"][stopped]).max())
    res["condition_holds"] = bool(same_stop.all() and same_hits.all() and dx.max() <= bound)
    return res
