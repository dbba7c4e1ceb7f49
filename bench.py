"""Benchmark of the tracking hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 1e6 neutrons in a
closed cylinder (R = 0.25 m, H = 0.5 m, 19,600 triangles), 2% Lambert walls with
1e-4 loss per bounce, B from a 64^3 FP64 tricubic map of 8 coaxial loops
(|B| <= 1 T), low-field seekers, energies uniform in 0-250 neV, tmax 200 s.

A bench step is one `pt_batch_run` call over the whole resident batch: it
advances every live particle by up to `--rk` accepted dopri5 steps (64) in `--rounds`
step-kernel/collision-kernel rounds (2): a particle whose third wall collision would
fall in the call keeps its remaining steps for the next call, which skips the short,
latency-bound rounds that resolve the last few collisions of a call.  A variant's
rate is accepted RK steps per second of GPU time over the timed steps (inputs
resident in HBM; HIP events on the library's stream).  Per-kernel times come from
HIP events the library records around each step-kernel and collision-kernel launch.

`value` is the rate of the fastest variant whose accuracy condition holds: the FP64
variants are bit-exact with the FP64 oracle (the GPU tests); a mixed-precision variant
counts only if the error leg holds the metric's condition -- end positions within 1e-6 m
of FP64 and identical full hit sequences for every particle -- over the stated horizon.

Roofline of the dominant kernel (that variant's step kernel): algorithmic bytes =
field-map cell evaluations made by the step kernel (device counters) x the bytes one
evaluation gathers (3 x 512 B FP64; mixed: the 896 B cell record of a 3-component map,
128 B of FP64 constant and linear terms and 3 x 256 B of FP32 lines, plus 512 B for a
component kept in FP64), divided by the step kernel's average launch duration, against HBM's 8 TB/s, and
against the measured ceiling of the same random gather in the step kernels' cooperative
form (pt_gather_bandwidth_coop; the per-lane form of pt_gather_bandwidth beside it).  `traffic`
is the measured HBM bytes per step-kernel launch from the committed rocprofv3 counter
summary of that kernel, when present.

The CPU baseline is the oracle (the C restatement of the reference algorithm) on a
bounded sample of the same workload, one thread per host core the process may use, in
its FP64 mode (the GPU's arithmetic) and its reference-semantics mode (the reference's
long double expressions and glibc).
"""
import argparse
import json
import os
import platform
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
VARIANTS = {
    "fp64": dict(mixed=False, persistent=0, graph=0),
    "fp64_persistent": dict(mixed=False, persistent=1, graph=0),
    "fp64_persistent_graph": dict(mixed=False, persistent=1, graph=1),
    "mixed": dict(mixed=True, persistent=0, graph=0),
    "mixed_persistent": dict(mixed=True, persistent=1, graph=0),
    "mixed_persistent_graph": dict(mixed=True, persistent=1, graph=1),
    # work list handed out in Morton order of the particle positions
    "mixed_persistent_spatial": dict(mixed=True, persistent=1, graph=0, spatial=1),
    "fp64_persistent_spatial": dict(mixed=False, persistent=1, graph=0, spatial=1),
    # Morton-ordered work list split into one chunk per XCD
    "fp64_persistent_xcd": dict(mixed=False, persistent=1, graph=0, spatial=2),
    "fp64_persistent_xcd_graph": dict(mixed=False, persistent=1, graph=1, spatial=2),
    "mixed_persistent_xcd": dict(mixed=True, persistent=1, graph=0, spatial=2),
    # collision queue ordered by candidate count
    "fp64_persistent_corder": dict(mixed=False, persistent=1, graph=0, corder=1),
    "fp64_corder": dict(mixed=False, persistent=0, graph=0, corder=1),
}
# graph replays carry no per-kernel events: the roofline uses the same kernels timed eagerly
PROFILED_FOR = {"fp64_persistent_graph": "fp64_persistent", "mixed_persistent_graph": "mixed_persistent",
                "fp64_persistent_xcd_graph": "fp64_persistent_xcd"}
KERNEL_NAMES = {"fp64": "step_kernel_fp64", "mixed": "step_kernel_mixed",
                "fp64_persistent": "step_kernel_persistent_fp64",
                "fp64_persistent_spatial": "step_kernel_persistent_fp64_spatial",
                "fp64_persistent_xcd": "step_kernel_persistent_fp64_xcd",
                "mixed_persistent": "step_kernel_persistent_mixed",
                "mixed_persistent_spatial": "step_kernel_persistent_mixed_spatial"}
MIX_CELL, FP64_BLOCK = 896, 512


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=1_000_000)
    p.add_argument("--rk", type=int, default=64, help="accepted RK steps per particle per bench step")
    p.add_argument("--grid", type=int, default=64)
    p.add_argument("--block", type=int, default=256)
    p.add_argument("--rounds", type=int, default=2, help="step/collision rounds per run call (0: library default, 4)")
    p.add_argument("--variants", default="fp64,fp64_persistent,fp64_persistent_graph,mixed,mixed_persistent,"
                                         "mixed_persistent_graph")
    p.add_argument("--no-mixed-error", action="store_true")
    p.add_argument("--err-n", type=int, default=10_000, help="mixed-precision error leg: neutrons")
    p.add_argument("--err-tmax", type=float, default=0.25, help="mixed-precision error leg: stated horizon [s]")
    p.add_argument("--err-longer", default="0.5,1.0", help="longer horizons whose deviation is reported too")
    p.add_argument("--err-ref-n", type=int, default=1000, help="particles of the oracle fp64-vs-refsem comparison")
    p.add_argument("--err-speed-n", type=int, default=1_000_000,
                   help="neutrons of the FP64-vs-mixed speed comparison on the error leg's map (0: skip)")
    p.add_argument("--no-uneven", action="store_true")
    p.add_argument("--uneven-n", type=int, default=4_000_000)
    p.add_argument("--uneven-meshes", default="guide,storage,cylinder")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-ceiling", action="store_true")
    p.add_argument("--adj-n", type=int, default=100_000, help="differentiable leg: neutrons")
    p.add_argument("--adj-steps", type=int, default=2000, help="differentiable leg: RK4 steps (dt = 1e-4 s)")
    p.add_argument("--adj-chunk", type=int, default=0,
                   help="differentiable leg: steps per chunk, a multiple of 256 (0: one pass; config 5's 2e5 "
                        "steps: 20224)")
    p.add_argument("--no-adjoint", action="store_true")
    return p.parse_args()


def host_cpu():
    """(threads this process may use, CPU model name)."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return max(1, cores), model


def cpu_baseline(w, budget_s):
    """Oracle on the first particles of the workload over 2 s of flight, one worker thread
    per host core the process may use, sized to about budget_s per mode."""
    from oracle import ptref
    from pentrack_amd.workloads import Workload
    cores, model = host_cpu()
    P = ptref.Problem(w.geometry, w.fields)
    tmax = 2.0

    def sub(m):
        return Workload(w.name, w.geometry, w.fields, w.species, tmax, w.species_idx[:m], w.t0[:m], w.y0[:m],
                        w.spin0[:m], w.pid[:m], w.seed)

    res = {}
    for mode in ("fp64", "refsem"):
        m = 4 * cores
        for target in (0.25 * budget_s, budget_s):   # two calibration passes, then the sample
            t = time.perf_counter()
            ptref.track(sub(m), mode, threads=cores, problem=P)
            dt = time.perf_counter() - t
            m = int(min(w.n, max(m, m * target / max(dt, 1e-3))))
        t = time.perf_counter()
        out, _ = ptref.track(sub(m), mode, threads=cores, problem=P)
        dt = time.perf_counter() - t
        acc = int(out["n_acc"].sum())
        res[mode] = dict(value=acc / dt, m=m, seconds=dt, accepted=acc, particles_per_s=m / dt)
        print(f"[bench] cpu baseline {mode}: {acc / dt:.3e} RK steps/s on {cores} threads ({m} particles, "
              f"{dt:.1f} s)", file=sys.stderr, flush=True)
    f = res["fp64"]
    return dict(value=f["value"], unit="RK steps/s", cores=cores, kind="port",
                sample=f"first {f['m']} particles of the workload tracked for {tmax:g} s of flight by the oracle "
                       f"(fp64 mode, the GPU's arithmetic): {f['accepted']} accepted steps in {f['seconds']:.1f} s "
                       f"on {cores} threads of a {model}",
                cpu_model=model, nproc=os.cpu_count(), particles_per_s=f["particles_per_s"],
                refsem={"value": res["refsem"]["value"], "particles": res["refsem"]["m"],
                        "seconds": res["refsem"]["seconds"], "particles_per_s": res["refsem"]["particles_per_s"]})


def mixed_error_leg(a, sptr):
    """Mixed precision against FP64 on config 3's workload (1 m cube, 128^3 coil map,
    |B| <= 1 T): every particle, full hit sequences, trajectory deviation at snapshots every
    10 ms up to the stated horizon (pentrack_amd.compare).  The FP64 kernel is bit-exact
    with the FP64 oracle (GPU tests), so it stands for the oracle here.  Deviations at longer
    horizons are reported beside the stated one, with the deviation of the oracle's
    reference-semantics mode (the reference's own long double arithmetic) from its FP64 mode
    on a sample, the floor any change of arithmetic is measured against."""
    import torch
    from oracle import ptref
    from pentrack_amd import _native, compare, workloads
    from pentrack_amd.tracker import Batch, LaunchConfig
    from pentrack_amd.workloads import Workload
    longer = [float(v) for v in a.err_longer.split(",") if v]
    times = compare.snapshot_times(a.err_tmax, longer)
    horizons = [a.err_tmax] + longer
    w = workloads.cube(n=a.err_n, seed=3, tmax=max(horizons), with_field=True, n_grid=128, Bmax=1.0)
    cap = 1024
    rows = []
    for tj in horizons:
        # both arithmetics tracked to tmax = the horizon: end states there, snapshots before it
        tt = times[times <= tj]
        out = {}
        for name, vid in (("fp64", _native.PT_VARIANT_FP64), ("mixed", _native.PT_VARIANT_MIXED)):
            b = Batch(w.geometry, w.fields, w.species, tj, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid,
                      hit_cap=cap, with_mixed=(name == "mixed"))
            b.set_snapshots(tt, len(tt))
            b.run_to_completion(LaunchConfig(variant=vid, steps_per_launch=256, rounds=8), max_launches=100000,
                                stream=sptr)
            o, h = b.download(sptr)
            rec, cnt = b.snapshots(sptr)
            assert (o["n_hits"] <= cap).all()
            out[name] = (o, h, rec, cnt)
            del b
        rows.append(compare.deviation(out["fp64"], out["mixed"], tt, tj))
    blocks, fp64_blocks = w.fields.mixed_info()
    # the reference's own arithmetic against the FP64 oracle, on a sample
    r = min(a.err_ref_n, w.n)
    sw = Workload(w.name, w.geometry, w.fields, w.species, w.tmax, w.species_idx[:r], w.t0[:r], w.y0[:r],
                  w.spin0[:r], w.pid[:r], w.seed)
    cores, _ = host_cpu()
    refrows = []
    for tj in horizons:
        tt = times[times <= tj]
        sw.tmax = tj
        ref = {mode: ptref.track_snapshots(sw, tt, len(tt), mode=mode, threads=cores) for mode in ("fp64", "refsem")}
        nh = np.zeros((r, 0), _native.HIT_DTYPE)
        d = compare.deviation((ref["fp64"][0], nh) + tuple(ref["fp64"][1:]), (ref["refsem"][0], nh) + tuple(ref["refsem"][1:]),
                              tt, tj)
        refrows.append({"horizon_s": tj, "particles": r, "trajectory_position_max_m": d["trajectory_position_max_m"],
                        "trajectory_position_p99_m": d["trajectory_position_p99_m"]})
    holds = rows[0]["condition_holds"]
    # speed of the two arithmetics on this workload's map at the benchmark's scale
    speed = {}
    if a.err_speed_n > 0:
        ws = workloads.cube(n=a.err_speed_n, seed=31, tmax=50.0, with_field=False)
        from pentrack_amd.fields import TFieldManager
        ws.fields = TFieldManager(w.fields.containers)
        for name, vid, pers in (("fp64", _native.PT_VARIANT_FP64, 0), ("fp64_persistent", _native.PT_VARIANT_FP64, 1),
                                ("mixed_persistent", _native.PT_VARIANT_MIXED, 1)):
            b = Batch(ws.geometry, ws.fields, ws.species, ws.tmax, ws.species_idx, ws.t0, ws.y0, ws.spin0, ws.seed,
                      ws.pid, with_mixed=(vid == _native.PT_VARIANT_MIXED))
            cfg = LaunchConfig(variant=vid, steps_per_launch=a.rk, persistent=pers)
            for _ in range(2):
                b.run(cfg, sptr)
            torch.cuda.synchronize()
            a0, _ = b.step_counts(sptr)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(5):
                b.run(cfg, sptr)
            e1.record()
            torch.cuda.synchronize()
            a1, _ = b.step_counts(sptr)
            speed[name] = (a1 - a0) / (e0.elapsed_time(e1) / 1e3)
            del b
        # the metric's comparison (mixed + persistent against the plain FP64 kernel), and the
        # same persistent kernel in both arithmetics
        speed["mixed_persistent_over_plain_fp64"] = speed["mixed_persistent"] / speed["fp64"]
        speed["mixed_over_fp64"] = speed["mixed_persistent"] / speed["fp64_persistent"]
        speed["particles"] = ws.n
    res = {"workload": f"cube 1 m, {w.n} neutrons, 128^3 coil map (|B| <= 1 T), mixed tables with "
                       f"{fp64_blocks}/{blocks} blocks kept in FP64",
           "stated_horizon_s": a.err_tmax, "condition_holds": bool(holds), "at_horizons": rows,
           "reference_arithmetic_vs_fp64": {"particles": r, "at_horizons": refrows},
           "fp64_block_fraction": fp64_blocks / max(blocks, 1),
           "rk_steps_per_s_on_this_map": speed}
    print(f"[bench] mixed-precision error: {res}", file=sys.stderr, flush=True)
    return res


def uneven_leg(a, sptr):
    """Config 4: mixed species with very uneven lifetimes in vessels of the reference's
    1032-triangle guide and 61468-triangle storage volume and of a synthetic 200k-triangle
    cylinder, a loss per bounce per neutron: the launch shape the autotuner picks (persistent
    work-queue kernel or per-particle grid, steps per call, BVH leaf size) against the default
    per-particle grid."""
    import torch
    from pentrack_amd import autotune, workloads
    from pentrack_amd.tracker import Batch, LaunchConfig
    res = {}
    for mesh in [m for m in a.uneven_meshes.split(",") if m]:
        w = workloads.uneven_vessel(mesh, n=a.uneven_n, seed=4, n_grid=a.grid, n_tris=200_000)
        best, entry = autotune.tune(w, n_sample=min(200_000, w.n), calls=2, blocks=(256,), grids=(2, 4),
                                    steps=(16, 64), leaf_sizes=(2, 4), rounds=(2, 4), cache=None, stream=sptr)
        r = {"workload": f"{mesh} vessel ({w.meta['mesh_triangles']} mesh triangles, {w.geometry.n_triangles()} "
                         f"with the shell box), {w.n} particles: 90% neutrons tau=880 s and loss per bounce "
                         "U(1e-5, 1e-3), 8% protons 1-800 eV, 2% electrons 1-700 keV, uniform E 1e5 V/m + "
                         f"{a.grid}^3 B map",
             "tuned_config": {k: getattr(best, k) for k in ("block", "steps_per_launch", "persistent", "grid",
                                                            "rounds", "leaf_size")},
             "autotune_trials": entry.get("trials")}
        for name, cfg in (("default", LaunchConfig(steps_per_launch=16)), ("autotuned", best)):
            b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid,
                      leaf_size=cfg.leaf_size)
            b.set_loss(w.loss)
            b.run(cfg, sptr)
            torch.cuda.synchronize()
            a0, _ = b.step_counts(sptr)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(4):
                b.run(cfg, sptr)
            e1.record()
            torch.cuda.synchronize()
            a1, _ = b.step_counts(sptr)
            ms = e0.elapsed_time(e1)
            r[name] = {"rk_steps_per_s": (a1 - a0) / (ms / 1e3), "gpu_ms": ms, "particles_active": b.active(sptr)}
            del b
        r["autotuned_over_default"] = r["autotuned"]["rk_steps_per_s"] / r["default"]["rk_steps_per_s"]
        res[mesh] = r
        print(f"[bench] uneven lifetimes, {mesh}: {r}", file=sys.stderr, flush=True)
        del w
    return res


def adjoint_leg(a, sptr):
    """Differentiable RK4 path: cube, B = sum of 8 coil maps (64^3 each), dt = 1e-4 s;
    forward with checkpoints every 256 steps, then the adjoint of J = <E_kin> + <z>.
    With --adj-chunk the horizon runs in chunks of that many steps (config 5's 2e5 steps;
    pentrack_amd.adjoint.rk4_forward_chunks / rk4_adjoint_chunks), the GPU work timed as a
    whole (the synchronisations between chunks only print progress)."""
    import torch
    from pentrack_amd import workloads
    from pentrack_amd.adjoint import rk4_adjoint_chunks, rk4_forward, rk4_forward_chunks
    model, y0, s0 = workloads.adjoint_cube(n=a.adj_n, seed=5, n_grid=a.grid, nmaps=8, dt=1e-4)
    dev = torch.device("cuda:0")
    y0t = torch.tensor(y0, device=dev)
    rk4_forward(model, s0, y0t[:256], 16)          # warm-up (module load)
    torch.cuda.synchronize()
    chunk = a.adj_chunk if 0 < a.adj_chunk < a.adj_steps else a.adj_steps
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def report(kind):
        def f(j, nc, m):
            if nc > 1:
                torch.cuda.synchronize()
                print(f"[bench] adjoint leg: {kind} chunk {j + 1}/{nc} ({m} steps) done", file=sys.stderr, flush=True)
        return f
    ev[0].record()
    yT, cks, lens = rk4_forward_chunks(model, s0, y0t, a.adj_steps, chunk, 256, report("forward"))
    ev[1].record()
    lamT = torch.zeros_like(yT)
    lamT[:, 3:] = model.mass * yT[:, 3:] / len(yT)
    lamT[:, 2] = 1.0 / len(yT)
    ev[2].record()
    lam, gs = rk4_adjoint_chunks(model, s0, cks, lens, lamT, 256, report("adjoint"))
    ev[3].record()
    torch.cuda.synchronize()
    fwd, adj = ev[0].elapsed_time(ev[1]), ev[2].elapsed_time(ev[3])
    steps = a.adj_n * a.adj_steps
    finite = bool(torch.isfinite(gs).all().item() and torch.isfinite(lam).all().item())
    print(f"[bench] adjoint leg: forward {fwd:.1f} ms, adjoint {adj:.1f} ms (ratio {adj / fwd:.2f}), "
          f"{steps / (fwd / 1e3):.3e} RK4 steps/s forward", file=sys.stderr, flush=True)
    return {"workload": f"cube, {a.adj_n} neutrons, B = sum of 8 coil maps ({a.grid}^3 each), RK4 dt=1e-4 s, "
                        f"{a.adj_steps} steps ({a.adj_steps * 1e-4:g} s) in {len(lens)} chunk(s), checkpoints "
                        "every 256, specular cube walls as reflection maps at the steps that crossed them, "
                        "J = <E_kin> + <z>",
            "forward_ms": fwd, "adjoint_ms": adj, "adjoint_over_forward": adj / fwd,
            "rk4_steps_per_s_forward": steps / (fwd / 1e3), "gradients_finite": finite,
            "dJ_ds": gs.sum(0).cpu().tolist()}


def read_profile(kernel, key):
    """A figure of `kernel` from its committed rocprofv3 summary under profiles/, if any."""
    try:
        txt = open(os.path.join(HERE, "profiles", f"traffic_{kernel}_cylinder_1e6.txt")).read()
    except OSError:
        return None
    m = re.search(r"^" + key + r":\s*([0-9.eE+]+)", txt, re.M)
    return float(m.group(1)) if m else None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); particles are independent and the "
              "path would shard by particle index, but multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from pentrack_amd import _native, workloads
    from pentrack_amd.tracker import Batch, LaunchConfig

    _native.lib()
    dev = torch.device("cuda:0")
    torch.cuda.init()
    stream = torch.cuda.Stream(dev)          # graph capture needs a non-default stream
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream

    t_build = time.perf_counter()
    w = workloads.cylinder(n=a.n, seed=2, tmax=200.0, n_grid=a.grid)
    t_build = time.perf_counter() - t_build
    print(f"[bench] workload built in {t_build:.1f} s: n={w.n} tris={w.geometry.n_triangles()} grid={a.grid}^3",
          file=sys.stderr, flush=True)

    results = {}
    for variant in [v for v in a.variants.split(",") if v]:
        if variant not in VARIANTS:
            raise SystemExit(f"unknown variant {variant}; known: {', '.join(VARIANTS)}")
        kw = VARIANTS[variant]
        vid = _native.PT_VARIANT_MIXED if kw["mixed"] else _native.PT_VARIANT_FP64
        torch.cuda.synchronize()
        t_up = time.perf_counter()
        b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid,
                  with_mixed=kw["mixed"])
        torch.cuda.synchronize()
        t_up = time.perf_counter() - t_up   # the particle state's allocation and upload from host arrays
        cfg = LaunchConfig(variant=vid, block=a.block, steps_per_launch=a.rk, rounds=a.rounds,
                           persistent=kw["persistent"], graph=kw["graph"], spatial_order=kw.get("spatial", 0),
                           collision_order=kw.get("corder", 0))
        for _ in range(a.warmup):
            b.run(cfg, sptr)
        torch.cuda.synchronize()
        acc0, rej0, nf0 = b.step_counts(sptr, with_fields=True)
        b.set_profiling(not kw["graph"])
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
        kt = b.kernel_times(sptr)
        acc1, rej1, nf1 = b.step_counts(sptr, with_fields=True)
        active = b.active(sptr)
        steps = acc1 - acc0
        t_dl = time.perf_counter()
        b.download(sptr)                    # end states and counters back to host arrays
        t_dl = time.perf_counter() - t_dl
        results[variant] = dict(steps=steps, rejected=rej1 - rej0, field_evals=nf1 - nf0, gpu_ms=gpu_ms,
                                wall_s=wall, rate=steps / (gpu_ms / 1e3), active=active, kernels=kt,
                                mixed=kw["mixed"], upload_s=t_up, download_s=t_dl)
        print(f"[bench] {variant}: {steps} accepted steps in {gpu_ms:.1f} ms -> {steps / (gpu_ms / 1e3):.3e} steps/s "
              f"(active {active}/{w.n}); step kernel {kt['step'][0]:.1f} ms / {kt['step'][1]} launches, "
              f"collision kernel {kt['collision'][0]:.1f} ms / {kt['collision'][1]} launches",
              file=sys.stderr, flush=True)
        del b
    blocks, fp64_blocks = w.fields.mixed_info()
    f64frac = fp64_blocks / max(blocks, 1)

    mixed_err = None if a.no_mixed_error else mixed_error_leg(a, sptr)
    uneven_res = None if a.no_uneven else uneven_leg(a, sptr)
    adj = None if a.no_adjoint else adjoint_leg(a, sptr)
    ceiling = None
    if not a.no_ceiling:
        from pentrack_amd.profiling import gather_ceiling, gather_ceiling_coop
        table = w.fields.containers[0].coef.nbytes
        gg, cc = gather_ceiling(table)
        gcoop = gather_ceiling_coop(table)
        # the same gather from tables that an XCD's L2 (4 MB) and the Infinity Cache (256 MB) hold
        gl2, _ = gather_ceiling(2 << 20)
        gmall, _ = gather_ceiling(128 << 20)
        ceiling = {"table_bytes": table, "random_1536B_gather_gbs": gg, "cooperative_1536B_gather_gbs": gcoop,
                   "stream_copy_gbs": cc, "random_1536B_gather_gbs_2MB_table_l2": gl2,
                   "random_1536B_gather_gbs_128MB_table_mall": gmall}
        print(f"[bench] gather ceiling: {ceiling}", file=sys.stderr, flush=True)
    cpu = None if a.no_cpu else cpu_baseline(w, a.cpu_seconds)

    mixed_ok = mixed_err is not None and mixed_err["condition_holds"]
    valid = [v for v in results if not results[v]["mixed"] or mixed_ok]
    head = max(valid or results, key=lambda v: results[v]["rate"])
    prof = PROFILED_FOR.get(head, head)
    prof = prof if prof in results else head
    r, rp = results[head], results[prof]
    step_ms, step_launches = rp["kernels"]["step"]
    per_launch_ms = step_ms / max(step_launches, 1)
    # tables that keep every block in FP64 run the FP64 kernels
    bpe = MIX_CELL + 3 * f64frac * FP64_BLOCK if rp["mixed"] and f64frac < 1 else 3 * FP64_BLOCK
    bytes_per_launch = rp["field_evals"] * bpe / max(step_launches, 1)
    achieved = bytes_per_launch / (per_launch_ms / 1e3) / 1e9 if per_launch_ms > 0 else 0.0
    kernel_name = KERNEL_NAMES.get(prof, prof)
    traffic = read_profile(kernel_name, "traffic_bytes_per_launch")
    flops = read_profile(kernel_name, "fp64_flops_per_launch")
    fp64_rates = [v["rate"] for k, v in results.items() if not v["mixed"]]
    mixed_rates = {k: v["rate"] for k, v in results.items() if v["mixed"]}
    line = {
        "metric": "RK steps/s on 1 MI355X: mixed+persistent vs plain FP64 @<=1e-6 m (cylinder, 1e6 neutrons, "
                  "64^3 tricubic B map)",
        "value": r["rate"],
        "unit": "RK steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["gpu_ms"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state and control, f64 linear + f32 remainder field interpolation" if r["mixed"] else "f64",
        "data": "synthetic",
        "config": {"workload": "cylinder R=0.25 H=0.5 (19600 tris), 1e6 neutrons, 64^3 tricubic B map of 8 loops, "
                               "2% Lambert, loss 1e-4, dopri5 rtol=atol=1e-9, tmax 200 s",
                   "n_particles": w.n, "rk_steps_per_bench_step": a.rk, "rounds_per_bench_step": a.rounds, "variant": head, "block": a.block},
        "variants": {k: {"rk_steps_per_s": v["rate"], "gpu_ms": v["gpu_ms"], "accepted": v["steps"],
                         "rejected": v["rejected"], "particles_active": v["active"],
                         "step_kernel_ms_per_launch": v["kernels"]["step"][0] / max(v["kernels"]["step"][1], 1),
                         "collision_kernel_ms_per_launch":
                             v["kernels"]["collision"][0] / max(v["kernels"]["collision"][1], 1),
                         "counts_for_value": (not v["mixed"]) or mixed_ok}
                     for k, v in results.items()},
        "mixed_over_best_fp64": {k: v / max(fp64_rates) for k, v in mixed_rates.items()} if fp64_rates else None,
        "mixed_persistent_over_plain_fp64": (results["mixed_persistent"]["rate"] / results["fp64"]["rate"]
                                             if "mixed_persistent" in results and "fp64" in results else None),
        "mixed_tables_fp64_block_fraction": f64frac,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel": kernel_name, "kernel_ms_per_launch": per_launch_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "field_evals_per_accepted_step": rp["field_evals"] / max(rp["steps"], 1),
                     "bytes_per_field_eval": bpe,
                     "gather_ceiling_gbs": ceiling["cooperative_1536B_gather_gbs"] if ceiling else None,
                     "frac_of_gather_ceiling": achieved / ceiling["cooperative_1536B_gather_gbs"] if ceiling else None,
                     "fp64_tflops": flops / (per_launch_ms * 1e9) if flops else None,
                     "fp64_frac": flops / (per_launch_ms * 1e9) / PEAK_FP64_TFLOPS if flops else None},
        "gather_ceiling": ceiling,
        # pt_track_batch's form: host arrays in, one run call, host arrays out (never `value`)
        "with_host_transfers": {"upload_s": r["upload_s"], "download_s": r["download_s"],
                                "rk_steps_per_s_one_call": (r["steps"] / a.steps) /
                                (r["upload_s"] + r["gpu_ms"] / a.steps / 1e3 + r["download_s"])},
        "cpu_baseline": cpu,
        "mixed_precision_error": mixed_err,
        "uneven_lifetimes": uneven_res,
        "adjoint": adj,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
