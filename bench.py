"""Benchmark of the tracking hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 1e6 neutrons in a
closed cylinder (R = 0.25 m, H = 0.5 m, 19,600 triangles), 2% Lambert walls with
1e-4 loss per bounce, B from a 64^3 FP64 tricubic map of 8 coaxial loops
(|B| <= 1 T), low-field seekers, energies uniform in 0-250 neV.

A bench step is one launch of the tracker over the whole resident batch,
advancing every live particle by `--rk` accepted dopri5 steps.  `value` is
accepted RK steps per second over the timed steps (inputs resident in HBM).
The CPU baseline is the oracle (the C restatement of the reference algorithm)
on a bounded sample of the same workload, on all host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
BYTES_PER_FIELD_EVAL = {"fp64": 3 * 64 * 8, "mixed": 3 * 64 * 4}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=1_000_000)
    p.add_argument("--rk", type=int, default=16, help="accepted RK steps per particle per bench step")
    p.add_argument("--grid", type=int, default=64)
    p.add_argument("--block", type=int, default=256)
    p.add_argument("--variants", default="fp64,mixed")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def cpu_baseline(w, budget_s):
    """Oracle on the first particles of the workload, all host cores, scaled to ~budget_s."""
    from oracle import ptref
    from pentrack_amd.workloads import Workload
    cores = min(16, os.cpu_count() or 1)
    P = ptref.Problem(w.geometry, w.fields)

    def sub(m, tmax):
        return Workload(w.name, w.geometry, w.fields, w.species, tmax, w.species_idx[:m], w.t0[:m], w.y0[:m],
                        w.spin0[:m], w.pid[:m], w.seed)

    m, tmax = 4 * cores, 1.0
    t = time.perf_counter()
    out, _ = ptref.track(sub(m, tmax), "fp64", threads=cores, problem=P)
    dt = time.perf_counter() - t
    scale = max(1.0, min(50.0, budget_s / max(dt, 1e-3)))
    tmax2 = min(w.tmax, tmax * scale)
    t = time.perf_counter()
    out, _ = ptref.track(sub(m, tmax2), "fp64", threads=cores, problem=P)
    dt = time.perf_counter() - t
    acc = int(out["n_acc"].sum())
    fields_per_step = float(out["n_field"].sum()) / max(acc, 1)
    flops = None
    return dict(value=acc / dt, unit="RK steps/s", cores=cores, kind="port",
                sample=f"{m} particles of the workload tracked for {tmax2:g} s of flight "
                       f"({acc} accepted steps in {dt:.1f} s)",
                fields_per_step=fields_per_step, rhs_per_step=float(out["n_rhs"].sum()) / max(acc, 1))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the tracking path shards trivially "
              "by particle index but multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)
    import torch
    from pentrack_amd import _native, workloads
    from pentrack_amd.tracker import Batch, LaunchConfig

    _native.lib()
    dev = torch.device("cuda:0")
    torch.cuda.init()
    stream = torch.cuda.current_stream(dev)
    sptr = stream.cuda_stream

    t_build = time.perf_counter()
    w = workloads.cylinder(n=a.n, seed=2, tmax=200.0, n_grid=a.grid)
    t_build = time.perf_counter() - t_build
    print(f"[bench] workload built in {t_build:.1f} s: n={w.n} tris={w.geometry.n_triangles()} grid={a.grid}^3",
          file=sys.stderr, flush=True)

    results = {}
    for variant in [v for v in a.variants.split(",") if v]:
        vid = _native.PT_VARIANT_MIXED if variant == "mixed" else _native.PT_VARIANT_FP64
        b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid,
                  with_fp32=(variant == "mixed"))
        cfg = LaunchConfig(variant=vid, block=a.block, steps_per_launch=a.rk)
        for _ in range(a.warmup):
            b.run(cfg, sptr)
        torch.cuda.synchronize()
        acc0, rej0 = b.step_counts(sptr)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(a.steps):
            b.run(cfg, sptr)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = ev0.elapsed_time(ev1)
        acc1, rej1 = b.step_counts(sptr)
        active = b.active(sptr)
        steps = acc1 - acc0
        results[variant] = dict(steps=steps, rejected=rej1 - rej0, gpu_ms=gpu_ms, wall_s=wall,
                                rate=steps / (gpu_ms / 1e3), active=active)
        print(f"[bench] {variant}: {steps} accepted steps in {gpu_ms:.1f} ms -> {steps / (gpu_ms / 1e3):.3e} steps/s "
              f"(active {active}/{w.n})", file=sys.stderr, flush=True)
        del b

    cpu = None if a.no_cpu else cpu_baseline(w, a.cpu_seconds)
    head = "fp64" if "fp64" in results else next(iter(results))
    r = results[head]
    fields_per_step = cpu["fields_per_step"] if cpu else 8.0
    bytes_per_step = fields_per_step * BYTES_PER_FIELD_EVAL[head]
    launch_ms = r["gpu_ms"] / a.steps
    achieved = bytes_per_step * (r["steps"] / a.steps) / (launch_ms / 1e3) / 1e9
    line = {
        "metric": "RK steps/s on 1 MI355X (cylinder, 1e6 neutrons, 64^3 tricubic B map)",
        "value": r["rate"],
        "unit": "RK steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": launch_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cylinder R=0.25 H=0.5 (19600 tris), 1e6 neutrons, 64^3 FP64 tricubic B map of 8 loops, "
                               "2% Lambert, loss 1e-4, dopri5 rtol=atol=1e-9",
                   "n_particles": w.n, "rk_steps_per_bench_step": a.rk, "variant": head, "block": a.block},
        "variants": {k: {"rk_steps_per_s": v["rate"], "gpu_ms": v["gpu_ms"], "accepted": v["steps"],
                         "rejected": v["rejected"]} for k, v in results.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                     "per_step_bytes": bytes_per_step, "field_evals_per_step": fields_per_step},
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
