"""Build libpentrack_gpu.so in-tree with hipcc for gfx950.

Every source is compiled to its own object in parallel; the tracking kernels
(pt_track.hip) are compiled three times per kernel mode (-DPT_MODE_TU=m: step,
persistent-step and collision kernels) plus once for
the mode-independent kernels and launchers, then everything is linked.
"""
import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpentrack_gpu.so")
OBJ = os.path.join(HERE, "_obj")

N_MODES = 4
SOURCES = ["pt_adjoint.hip", "pt_precompute.hip", "pt_membench.hip", "pt_host.cpp", "pt_formula.cpp"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # bit-reproducible FP64: no contraction, no reassociation
    "-ffp-contract=off", "-fno-fast-math",
    # packed FP32 instructions disabled (hipcc only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _units():
    """(object name, source, extra flags) of every translation unit."""
    units = [("pt_track_common.o", "pt_track.hip", [])]
    units += [(f"pt_track_m{m}{'scp'[k]}.o", "pt_track.hip", [f"-DPT_MODE_TU={m}", f"-DPT_KERNEL_SET={k}"])
              for m in reversed(range(N_MODES)) for k in (0, 2, 1)]
    units += [(os.path.splitext(s)[0] + ".o", s, []) for s in SOURCES]
    return units


def build(verbose=False, jobs=None, out=OUT, defines=()):
    """Compile and link the library to `out`; `defines` are extra -D flags (launch-shape
    experiments, e.g. PT_STEP_MIN_WAVES=1), built in their own object folder."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    obj_dir = OBJ if not defines else OBJ + "_" + "_".join(d.replace("=", "") for d in defines)
    os.makedirs(obj_dir, exist_ok=True)
    extra_d = [f"-D{d}" for d in defines]
    units = _units()

    def compile_one(u):
        obj, src, extra = u
        cmd = [hipcc] + FLAGS + extra_d + extra + ["-c", "-x", "hip", os.path.join(CSRC, src),
                                                  "-o", os.path.join(obj_dir, obj)]
        t = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True)
        if verbose:
            print(f"  {obj}: {time.perf_counter() - t:.0f} s", flush=True)
        return obj, r

    jobs = jobs or min(len(units), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, units))
    failed = [(o, r) for o, r in results if r.returncode != 0]
    for o, r in failed:
        sys.stderr.write(f"--- {o}\n" + r.stdout + r.stderr)
    if failed:
        raise RuntimeError("hipcc failed building libpentrack_gpu.so: " + ", ".join(o for o, _ in failed))
    link = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + \
        [os.path.join(obj_dir, u[0]) for u in units]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libpentrack_gpu.so failed")
    os.replace(out + ".tmp", out)
    if out == OUT:
        build_c_example(verbose)
    if verbose:
        print("built", out)
    return out


C_EXAMPLE = os.path.join(HERE, "pt_capi_example")


def build_c_example(verbose=False):
    """The plain C caller of the library (csrc/pt_capi_example.c), linked against the
    in-tree libpentrack_gpu.so with an $ORIGIN run path."""
    cc = os.environ.get("CC", "gcc")
    cmd = [cc, "-O2", "-std=c11", "-Wall", "-Wextra", "-o", C_EXAMPLE + ".tmp", os.path.join(CSRC, "pt_capi_example.c"),
           "-L" + HERE, "-lpentrack_gpu", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building pt_capi_example failed")
    os.replace(C_EXAMPLE + ".tmp", C_EXAMPLE)
    if verbose:
        print("built", C_EXAMPLE)


if __name__ == "__main__":
    build(verbose=True)
