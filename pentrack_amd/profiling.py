"""HBM traffic of the step kernel from rocprofv3 counters.

`python -m pentrack_amd.profiling run KERNEL` is the program profiled: a plain
device copy of known size (the calibration: FETCH_SIZE and WRITE_SIZE of a kernel
that reads and writes exactly COPY_BYTES), then the benchmark workload (1e6
neutrons, cylinder, 64^3 map) advanced by a few run calls of 16 RK steps with the
step kernel KERNEL (step_kernel_fp64, step_kernel_mixed,
step_kernel_persistent_mixed, or step_kernel_persistent_mixed_spatial: the persistent
kernel with its work list in Morton order of the particle positions).  Profile it twice, once per counter:

    rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pmc_fetch -o run -- python3 -m pentrack_amd.profiling run KERNEL
    rocprofv3 --pmc WRITE_SIZE -d gpurun_out/pmc_write -o run -- python3 -m pentrack_amd.profiling run KERNEL

`python -m pentrack_amd.profiling summarize KERNEL FETCH_DB WRITE_DB` prints the
summary: counter values (KiB per dispatch) of the copy and of the step kernel,
the scale factor (bytes the copy moved / bytes the counter reports), the step
kernel's scaled HBM bytes per launch and its algorithmic bytes per launch.
"""
import sqlite3
import sys

COPY_BYTES = 1 << 30
KERNELS = {
    # name: (variant, persistent, bytes per field evaluation, dispatch-name pattern, spatial order)
    # (the kernels' template argument is the mode: 0 FP64 tables, 1 FP32 tables)
    "step_kernel_fp64": (0, 0, 1536, "step_kernel<0>", 0),
    "step_kernel_mixed": (1, 0, 768, "step_kernel<1>", 0),
    "step_kernel_persistent_mixed": (1, 1, 768, "step_kernel_persistent<1>", 0),
    "step_kernel_persistent_mixed_spatial": (1, 1, 768, "step_kernel_persistent<1>", 1),
}


def run(kernel="step_kernel_fp64"):
    import torch
    from pentrack_amd import _native, workloads
    from pentrack_amd.tracker import Batch, LaunchConfig

    _native.lib()
    src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1.0)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    w = workloads.cylinder(n=1_000_000, seed=2, tmax=200.0, n_grid=64)
    variant, persistent, bpe, _, spatial = KERNELS[kernel]
    b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid,
              with_mixed=bool(variant))
    stream = torch.cuda.current_stream().cuda_stream
    cfg = LaunchConfig(variant=variant, persistent=persistent, steps_per_launch=16, spatial_order=spatial)
    b.set_profiling(True)
    for _ in range(3):
        b.run(cfg, stream)
    torch.cuda.synchronize()
    acc, rej, nf = b.step_counts(stream, with_fields=True)
    kt = b.kernel_times(stream)
    launches = kt["step"][1]
    print(f"profiling run done: {acc} accepted steps, {nf} step-kernel field evaluations over {launches} "
          f"step-kernel launches = {nf * bpe / max(launches, 1):.6e} algorithmic bytes per launch", flush=True)


def _per_dispatch(db, counter):
    """{dispatch_id: (kernel_name, value in KiB)} for one counter of a rocpd database."""
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, kernel_name, sum(value) from counters_collection "
                       "where counter_name = ? group by dispatch_id", (counter,)).fetchall()
    return {d: (k, float(v)) for d, k, v in rows}


def summarize(kernel, fetch_db, write_db, algorithmic=None):
    f = _per_dispatch(fetch_db, "FETCH_SIZE")
    w = _per_dispatch(write_db, "WRITE_SIZE")
    kib = COPY_BYTES / 1024
    # the calibration copies: device-to-device copies that wrote COPY_BYTES
    copies = [d for d in f if "copy" in f[d][0].lower() and d in w and w[d][1] > 0.9 * kib]
    pattern = KERNELS[kernel][3]
    step = [d for d in f if pattern in f[d][0]]
    if not copies or not step:
        raise RuntimeError("calibration copy or step kernel not found")
    fc = sum(f[d][1] for d in copies) / len(copies)
    wc = sum(w[d][1] for d in copies) / len(copies)
    fscale, wscale = kib / fc, kib / wc
    fs = sum(f[d][1] for d in step) / len(step)
    ws = sum(w[d][1] for d in step if d in w) / len(step)
    fetch, write = fs * fscale * 1024, ws * wscale * 1024
    print(f"# HBM traffic per launch of {kernel}, cylinder 1e6 neutrons, 64^3 map, 16 RK steps per call")
    print("commands:")
    print(f"  rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pmc_fetch -o run -- python3 -m pentrack_amd.profiling run {kernel}")
    print(f"  rocprofv3 --pmc WRITE_SIZE -d gpurun_out/pmc_write -o run -- python3 -m pentrack_amd.profiling run {kernel}")
    print(f"  python3 -m pentrack_amd.profiling summarize {kernel} {fetch_db} {write_db}")
    print()
    print(f"calibration: {len(copies)} device copies of {COPY_BYTES} B ({f[copies[0]][0]})")
    print(f"  FETCH_SIZE {fc:.1f} KiB, WRITE_SIZE {wc:.1f} KiB per copy; {kib:.0f} KiB moved each way")
    print(f"  scale factors: fetch {fscale:.4f}, write {wscale:.4f}")
    print(f"step kernel: {f[step[0]][0]} ({len(step)} dispatches)")
    print(f"  raw per dispatch: FETCH_SIZE {fs:.1f} KiB, WRITE_SIZE {ws:.1f} KiB")
    print(f"  fetch_bytes_per_launch: {fetch:.6e}")
    print(f"  write_bytes_per_launch: {write:.6e}")
    if algorithmic:
        print(f"algorithmic_bytes_per_launch (field gathers, from the run's device counters): {algorithmic:.6e}")
    print(f"traffic_bytes_per_launch: {fetch + write:.6e}")


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "run":
        run(*sys.argv[2:3])
    elif len(sys.argv) in (5, 6) and sys.argv[1] == "summarize":
        summarize(sys.argv[2], sys.argv[3], sys.argv[4], float(sys.argv[5]) if len(sys.argv) == 6 else None)
    else:
        print(__doc__)
        sys.exit(2)
