"""Launch-configuration autotuner for the tracker.

Times the tracker on a sample of a workload over a grid of launch shapes -- kernel
(per-particle grid or persistent work-queue slots), block size or persistent grid,
accepted steps per run call, step/collision rounds per run call, BVH leaf size -- and
returns the fastest.  Fewer rounds per call leave a particle's second or third wall
collision of the call to the next call instead of running short, latency-bound rounds;
the score is accepted steps per second, so that trade is measured, not assumed.  Results are
cached per device and workload signature in a JSON file, so a production run tunes
once.  Timing uses HIP events on the tracker's stream; every trial starts from the
same state so trials are comparable.
"""
import hashlib
import itertools
import json
import os
import time

import numpy as np

from .tracker import Batch, LaunchConfig

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "pentrack_amd", "autotune.json")


def _signature(w, n):
    h = hashlib.sha1()
    h.update(f"{w.name}:{w.geometry.n_triangles()}:{len(w.fields.containers)}:{w.seed}:{n}".encode())
    h.update(np.bincount(w.species_idx[:n], minlength=3).tobytes())
    return h.hexdigest()[:16]


def _device_name():
    import torch
    return torch.cuda.get_device_name(0)


def tune(w, n_sample=200_000, calls=3, blocks=(64, 128, 256), grids=(1, 2, 4), steps=(8, 16, 32, 64),
         leaf_sizes=(2, 4, 8), rounds=(2, 4), variant=0, cache=DEFAULT_CACHE, stream=None, verbose=False):
    """Best LaunchConfig for workload `w` (evaluated on its first n_sample particles).

    Each trial creates a fresh batch from the same initial states, runs `calls` untimed
    warm-up calls, then `calls` timed calls, and scores accepted RK steps per second.
    """
    import torch
    n = min(n_sample, w.n)
    key = f"{_device_name()}:{_signature(w, n)}:v{variant}:r{'-'.join(map(str, rounds))}"
    if cache and os.path.exists(cache):
        try:
            entry = json.load(open(cache)).get(key)
            if entry:
                return LaunchConfig(**entry["cfg"]), entry
        except (OSError, ValueError):
            pass
    dev = torch.device("cuda:0")
    props = torch.cuda.get_device_properties(dev)
    cus = props.multi_processor_count
    s = stream or torch.cuda.current_stream(dev).cuda_stream
    sel = slice(0, n)
    trials = []
    shapes = [dict(persistent=0, block=b) for b in blocks] + [dict(persistent=1, grid=g * cus) for g in grids]
    for leaf in leaf_sizes:
        for shape, k, nr in itertools.product(shapes, steps, rounds):
            cfg = LaunchConfig(variant=variant, steps_per_launch=k, rounds=nr, leaf_size=leaf, **shape)
            b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx[sel], w.t0[sel], w.y0[sel],
                      w.spin0[sel], w.seed, w.pid[sel], with_mixed=(variant == 1), leaf_size=leaf)
            if getattr(w, "loss", None) is not None:
                b.set_loss(w.loss[sel])
            for _ in range(calls):
                b.run(cfg, s)
            torch.cuda.synchronize()
            a0, _ = b.step_counts(s)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(calls):
                b.run(cfg, s)
            e1.record()
            torch.cuda.synchronize()
            a1, _ = b.step_counts(s)
            ms = e0.elapsed_time(e1)
            rate = (a1 - a0) / (ms / 1e3) if ms > 0 else 0.0
            trials.append((rate, cfg))
            if verbose:
                print(f"[autotune] leaf {leaf} {shape} steps {k} rounds {nr}: {rate:.3e} steps/s")
            del b
    best_rate, best = max(trials, key=lambda t: t[0])
    entry = {"cfg": {k: getattr(best, k) for k in ("variant", "block", "steps_per_launch", "persistent", "grid",
                                                   "rounds", "leaf_size")},
             "rate": best_rate, "trials": len(trials), "time": time.time(), "n_sample": n}
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        try:
            data = json.load(open(cache)) if os.path.exists(cache) else {}
        except (OSError, ValueError):
            data = {}
        data[key] = entry
        json.dump(data, open(cache, "w"), indent=1)
    return best, entry
