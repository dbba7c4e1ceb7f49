"""Measurements of the step kernel from rocprofv3: HBM traffic, FP64 instructions, kernel table.

`python -m pentrack_amd.profiling run KERNEL` is the program profiled: a plain
device copy of known size (the calibration: FETCH_SIZE and WRITE_SIZE of a kernel
that reads and writes exactly COPY_BYTES), then the benchmark workload (1e6
neutrons, cylinder, 64^3 map) advanced by a few run calls of 64 RK steps in 2 rounds with the
step kernel KERNEL (step_kernel_fp64, step_kernel_persistent_fp64, step_kernel_mixed,
step_kernel_persistent_mixed, or step_kernel_persistent_mixed_spatial: the persistent
kernel with its work list in Morton order of the particle positions).  Profile it three
times, once per counter set (a pass holds at most 4 TCC and 8 SQ counters):

    rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pmc_fetch -o run -- python3 -m pentrack_amd.profiling run KERNEL
    rocprofv3 --pmc WRITE_SIZE -d gpurun_out/pmc_write -o run -- python3 -m pentrack_amd.profiling run KERNEL
    rocprofv3 --pmc SQ_INSTS_VALU_ADD_F64 SQ_INSTS_VALU_MUL_F64 SQ_INSTS_VALU_FMA_F64 SQ_INSTS_VALU_TRANS_F64 \
        -d gpurun_out/pmc_flops -o run -- python3 -m pentrack_amd.profiling run KERNEL

`python -m pentrack_amd.profiling summarize KERNEL FETCH_DB WRITE_DB [ALGORITHMIC [FLOPS_DB]]`
prints the summary: counter values (KiB per dispatch) of the copy and of the step kernel,
the scale factor (bytes the copy moved / bytes the counter reports), the step kernel's
scaled HBM bytes per launch, its algorithmic bytes per launch and its issued FP64 lane
operations per launch.  `python -m pentrack_amd.profiling trace CSV TITLE COMMAND` prints
the kernel table of a `rocprofv3 --kernel-trace --output-format csv` run.
"""
import ctypes as C
import sqlite3
import sys

COPY_BYTES = 1 << 30
# the launch shape of the benchmark's run calls (bench.py --rk / --rounds defaults)
RK_STEPS, ROUNDS = 64, 2
KERNELS = {
    # name: (variant, persistent, dispatch-name pattern, spatial order)
    # (the kernels' template argument is the mode: 0 FP64 tables, 1 mixed-precision tables)
    "step_kernel_fp64": (0, 0, "step_kernel<0>", 0),
    "step_kernel_persistent_fp64": (0, 1, "step_kernel_persistent<0>", 0),
    "step_kernel_mixed": (1, 0, "step_kernel<1>", 0),
    "step_kernel_persistent_mixed": (1, 1, "step_kernel_persistent<1>", 0),
    "step_kernel_persistent_mixed_spatial": (1, 1, "step_kernel_persistent<1>", 1),
    # Morton-ordered work list split into one chunk per XCD
    "step_kernel_persistent_fp64_xcd": (0, 1, "step_kernel_persistent<0>", 2),
}
# the FP64 VALU instruction counters of a third pass (wave-level instruction counts)
FLOP_COUNTERS = ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_TRANS_F64")


def gather_ceiling(table_bytes, block_bytes=1536, threads=1 << 20, gathers=8, reps=5):
    """(random-gather GB/s, streaming-copy GB/s) of pt_gather_bandwidth for a table of
    table_bytes: lanes reading their own block_bytes block (a cell's three components)
    from random blocks, as the tricubic evaluation does."""
    from . import _native
    g, c = C.c_double(), C.c_double()
    _native.check(_native.lib().pt_gather_bandwidth(int(table_bytes), int(block_bytes), int(threads), int(gathers),
                                                    int(reps), C.byref(g), C.byref(c)), "pt_gather_bandwidth")
    return g.value, c.value


def gather_ceiling_coop(table_bytes, threads=1 << 20, gathers=8, reps=5):
    """GB/s of pt_gather_bandwidth_coop: the same random 1536-byte block gather in the step
    kernels' cooperative form (half lines loaded by four lanes, passed through LDS) at
    their occupancy."""
    from . import _native
    g = C.c_double()
    _native.check(_native.lib().pt_gather_bandwidth_coop(int(table_bytes), int(threads), int(gathers), int(reps),
                                                         C.byref(g)), "pt_gather_bandwidth_coop")
    return g.value


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
    variant, persistent, _, spatial = KERNELS[kernel]
    b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid,
              with_mixed=bool(variant))
    blocks, fp64 = w.fields.mixed_info()
    # bytes one field evaluation gathers: 3 blocks of 512 B (FP64), or the 896 B cell record
    # plus 512 B for each component kept in FP64 (mixed)
    bpe = 3 * 512 if not variant or fp64 == blocks else 896 + 3 * 512 * fp64 / max(blocks, 1)
    stream = torch.cuda.current_stream().cuda_stream
    cfg = LaunchConfig(variant=variant, persistent=persistent, steps_per_launch=RK_STEPS, rounds=ROUNDS,
                       spatial_order=spatial)
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


def _flops_per_dispatch(db, pattern):
    """Issued FP64 lane operations per dispatch of the kernels matching pattern: 64 lanes x
    (ADD + MUL + TRANS + 2 FMA) wave instructions (inactive lanes of a wave included, so an
    upper bound of the useful FP64 FLOPs)."""
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, kernel_name, counter_name, sum(value) from counters_collection "
                       "group by dispatch_id, counter_name").fetchall()
    per = {}
    for d, k, c, v in rows:
        if pattern in k and c in FLOP_COUNTERS:
            per.setdefault(d, {})[c] = float(v)
    w = {"SQ_INSTS_VALU_ADD_F64": 1, "SQ_INSTS_VALU_MUL_F64": 1, "SQ_INSTS_VALU_FMA_F64": 2,
         "SQ_INSTS_VALU_TRANS_F64": 1}
    return [64 * sum(w[c] * v for c, v in cs.items()) for cs in per.values()], per


def summarize(kernel, fetch_db, write_db, algorithmic=None, flops_db=None):
    f = _per_dispatch(fetch_db, "FETCH_SIZE")
    w = _per_dispatch(write_db, "WRITE_SIZE")
    kib = COPY_BYTES / 1024
    # the calibration copies: device-to-device copies that wrote COPY_BYTES
    copies = [d for d in f if "copy" in f[d][0].lower() and d in w and w[d][1] > 0.9 * kib]
    pattern = KERNELS[kernel][2]
    step = [d for d in f if pattern in f[d][0]]
    if not copies or not step:
        raise RuntimeError("calibration copy or step kernel not found")
    fc = sum(f[d][1] for d in copies) / len(copies)
    wc = sum(w[d][1] for d in copies) / len(copies)
    fscale, wscale = kib / fc, kib / wc
    fs = sum(f[d][1] for d in step) / len(step)
    ws = sum(w[d][1] for d in step if d in w) / len(step)
    fetch, write = fs * fscale * 1024, ws * wscale * 1024
    print(f"# HBM traffic per launch of {kernel}, cylinder 1e6 neutrons, 64^3 map, {RK_STEPS} RK steps in "
          f"{ROUNDS} rounds per call")
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
    if flops_db:
        fl, per = _flops_per_dispatch(flops_db, pattern)
        if fl:
            one = next(iter(per.values()))
            print(f"FP64 instructions (third pass, {' '.join(FLOP_COUNTERS)}), first dispatch: "
                  + ", ".join(f"{c} {v:.4e}" for c, v in sorted(one.items())))
            print(f"  rocprofv3 --pmc {' '.join(FLOP_COUNTERS)} -d gpurun_out/pmc_flops -o run -- "
                  f"python3 -m pentrack_amd.profiling run {kernel}")
            print(f"fp64_flops_per_launch: {sum(fl) / len(fl):.6e}")


# counter passes of the step and collision kernels (each within one pass's per-block limits)
COUNTER_PASSES = {
    "sq": "SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR "
          "SQ_INSTS_LDS",
    "sq_wait": "SQ_ACTIVE_INST_ANY SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_IFETCH "
               "SQ_INSTS_BRANCH SQ_INSTS_SMEM",
    "mem": "TA_TA_BUSY_sum TA_ADDR_STALLED_BY_TC_CYCLES_sum TD_TD_BUSY_sum TD_TC_STALL_sum TCP_TCC_READ_REQ_sum "
           "TCP_TCC_WRITE_REQ_sum TCP_PENDING_STALL_CYCLES_sum TCP_TOTAL_CACHE_ACCESSES_sum TCC_HIT_sum TCC_MISS_sum "
           "TCC_EA0_RDREQ_sum GRBM_GUI_ACTIVE",
}


def summarize_counters(kernel, dbs):
    """Per-dispatch averages of the counter passes (COUNTER_PASSES, one rocpd database each)
    for the step kernel of `kernel` and the collision kernel of the same mode, with the
    ratios the measurements quote."""
    pattern = KERNELS[kernel][2]
    mode = pattern[pattern.index("<"):]
    kernels = {"step": pattern, "collision": "collision_kernel" + mode}
    vals = {k: {} for k in kernels}
    dur = {k: [] for k in kernels}
    for db in dbs:
        con = sqlite3.connect(db)
        for k, pat in kernels.items():
            rows = con.execute("select counter_name, avg(v), avg(d) from (select dispatch_id, counter_name, "
                               "sum(value) as v, max(end - start) as d from counters_collection where kernel_name "
                               "like ? group by dispatch_id, counter_name) group by counter_name",
                               (f"%{pat}%",)).fetchall()
            for c, v, d in rows:
                vals[k][c] = float(v)
                dur[k].append(float(d))
    print(f"# rocprofv3 --pmc counter passes: {kernels['step']} and {kernels['collision']}, cylinder 1e6 "
          f"neutrons, 64^3 map, {RK_STEPS} RK steps in {ROUNDS} rounds per call (averages per dispatch)")
    print("commands:")
    for name, counters in COUNTER_PASSES.items():
        print(f"  rocprofv3 --pmc {counters} -d gpurun_out/pmc_{name} -o run -- python3 -m pentrack_amd.profiling "
              f"run {kernel}")
    print(f"  python3 -m pentrack_amd.profiling counters {kernel} " +
          " ".join(f"gpurun_out/pmc_{n}/run_results.db" for n in COUNTER_PASSES))
    for k in kernels:
        v = vals[k]
        print()
        print(f"{kernels[k]}: {sum(dur[k]) / max(len(dur[k]), 1) / 1e6:.2f} ms per dispatch")
        for c in sorted(v):
            print(f"  {c}: {v[c]:.4e}")
        g = v.get
        if g("SQ_ACTIVE_INST_VALU") and g("SQ_THREAD_CYCLES_VALU"):
            print(f"  active lanes per VALU instruction (THREAD_CYCLES_VALU / ACTIVE_INST_VALU, 64 = full): "
                  f"{g('SQ_THREAD_CYCLES_VALU') / g('SQ_ACTIVE_INST_VALU'):.1f}")
        if g("SQ_WAVE_CYCLES"):
            for c in ("SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_ANY"):
                if g(c):
                    print(f"  {c} / SQ_WAVE_CYCLES: {g(c) / g('SQ_WAVE_CYCLES'):.3f}")
        if g("TCC_HIT_sum") is not None and g("TCC_MISS_sum"):
            print(f"  L2 hit rate: {g('TCC_HIT_sum') / (g('TCC_HIT_sum') + g('TCC_MISS_sum')):.3f}")
        if g("GRBM_GUI_ACTIVE") and g("TD_TD_BUSY_sum"):
            # GRBM_GUI_ACTIVE sums the 8 XCDs; the TA/TD/TCP sums run over the 256 CUs
            per_cu = g("GRBM_GUI_ACTIVE") / 8
            for c in ("TD_TD_BUSY_sum", "TD_TC_STALL_sum", "TA_TA_BUSY_sum", "TCP_PENDING_STALL_CYCLES_sum"):
                if g(c):
                    print(f"  {c} per CU / active cycles: {g(c) / 256 / per_cu:.3f}")


def summarize_trace(csv_path, title, command):
    """The kernel table of a rocprofv3 --kernel-trace --output-format csv run: calls, average
    and total duration, registers, scratch and LDS of every kernel, slowest first."""
    import csv

    def col(row, *names):
        for n in names:
            if n in row and row[n] != "":
                return row[n]
        return ""
    stats = {}
    with open(csv_path) as fh:
        for row in csv.DictReader(fh):
            name = col(row, "Kernel_Name", "Kernel-Name", "KernelName")
            t = (int(col(row, "End_Timestamp", "End-Timestamp")) - int(col(row, "Start_Timestamp", "Start-Timestamp")))
            e = stats.setdefault(name, {"calls": 0, "ns": 0, "row": row})
            e["calls"] += 1
            e["ns"] += t
    print(f"# rocprofv3 --kernel-trace: {title}")
    print(f"command: {command}")
    print()
    print("kernel | calls | avg ms | total ms | arch VGPR | AGPR | SGPR | scratch B/lane | LDS B | block | grid")
    for name, e in sorted(stats.items(), key=lambda kv: -kv[1]["ns"]):
        r = e["row"]
        print(" | ".join([name, str(e["calls"]), f"{e['ns'] / e['calls'] / 1e6:.3f}", f"{e['ns'] / 1e6:.2f}",
                          col(r, "Arch_VGPR_Count", "VGPR_Count"), col(r, "Accum_VGPR_Count"), col(r, "SGPR_Count"),
                          col(r, "Scratch_Size", "Private_Segment_Size"), col(r, "LDS_Block_Size", "LDS_Size",
                                                                               "Group_Segment_Size"),
                          col(r, "Workgroup_Size_X", "Workgroup_Size"), col(r, "Grid_Size_X", "Grid_Size")]))


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "run":
        run(*sys.argv[2:3])
    elif len(sys.argv) in (5, 6, 7) and sys.argv[1] == "summarize":
        summarize(sys.argv[2], sys.argv[3], sys.argv[4], float(sys.argv[5]) if len(sys.argv) >= 6 else None,
                  sys.argv[6] if len(sys.argv) == 7 else None)
    elif len(sys.argv) >= 4 and sys.argv[1] == "counters":
        summarize_counters(sys.argv[2], sys.argv[3:])
    elif len(sys.argv) == 5 and sys.argv[1] == "trace":
        summarize_trace(sys.argv[2], sys.argv[3], sys.argv[4])
    else:
        print(__doc__)
        sys.exit(2)
