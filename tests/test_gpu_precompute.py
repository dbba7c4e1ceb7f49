"""GPU: TabField3::PreInterpol on the device (pt_tricubic_precompute_gpu, and maps handed
to pt_fields_create as samples) against the host pre-interpolation, bit for bit."""
import time

import numpy as np
import pytest

from pentrack_amd import queries
from pentrack_amd.fields import TFieldManager, tabfield3, tricubic_coefficients, tricubic_coefficients_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 2, 5), (7, 9, 4), (17, 12, 23)])
def test_gpu_coefficients_equal_host(gpu, shape):
    rng = np.random.default_rng(sum(shape))
    grids = [np.cumsum(rng.uniform(0.01, 0.3, n)) - 0.4 for n in shape]
    values = rng.normal(size=(4,) + shape)
    print("case: precompute shape", shape, flush=True)
    h = tricubic_coefficients(*grids, values)
    g = tricubic_coefficients_gpu(*grids, values)
    assert np.array_equal(g, h)


def test_maps_uploaded_as_samples(gpu, oracle_lib):
    """A manager whose maps are pre-interpolated on upload gives the same fields (FP64 and
    the FP32 copy) as one built from host coefficients."""
    rng = np.random.default_rng(3)
    xg, yg, zg = np.linspace(-0.5, 0.5, 21), np.cumsum(rng.uniform(0.02, 0.06, 18)), np.linspace(0, 1, 25)
    B = rng.normal(size=(3, 21, 18, 25)) * 0.01
    V = rng.normal(size=(21, 18, 25))
    host = TFieldManager([tabfield3(xg, yg, zg, B=B, V=V, escale=1.0, box=(0.5, -0.5, yg[-1], yg[0], 1, 0), bw=0.1)])
    dev = TFieldManager([tabfield3(xg, yg, zg, B=B, V=V, escale=1.0, box=(0.5, -0.5, yg[-1], yg[0], 1, 0), bw=0.1,
                                   precompute="gpu")])
    assert dev.containers[0].coef is None
    pts = np.stack([rng.uniform(-0.6, 0.6, 2000), rng.uniform(0, 1, 2000), rng.uniform(-0.1, 1.1, 2000)], 1)
    print("case: maps uploaded as samples", flush=True)
    a = queries.fields_eval(host, pts, 0.0)
    b = queries.fields_eval(dev, pts, 0.0)
    assert all(np.array_equal(u, v) for u, v in zip(a, b))
    o = oracle_lib.field_eval(dev, pts, 0.0, "fp64")
    assert all(np.array_equal(u, v) for u, v in zip(a, o))
    # the FP32 copy made on the device equals the host's rounding of the coefficients
    host.device(with_mixed=True)
    dev.device(with_mixed=True)
    from pentrack_amd.workloads import Workload
    from pentrack_amd.particle import SpeciesConfig
    from pentrack_amd.constants import NEUTRON
    from pentrack_amd.tracker import LaunchConfig
    from pentrack_amd import _native
    from tests.test_gpu_parity import run_gpu
    w0 = oracle_free_fall(rng)
    ws = [Workload("m", w0.geometry, f, [SpeciesConfig(NEUTRON, tmax=9e99)], 0.05, w0.species_idx, w0.t0, w0.y0,
                   w0.spin0, w0.pid, 5) for f in (host, dev)]
    cfg = LaunchConfig(variant=_native.PT_VARIANT_MIXED)
    ga, _ = run_gpu(ws[0], cfg=cfg, with_mixed=True)
    gb, _ = run_gpu(ws[1], cfg=cfg, with_mixed=True)
    assert np.array_equal(ga["y"], gb["y"]) and np.array_equal(ga["Nstep"], gb["Nstep"])


def oracle_free_fall(rng):
    from pentrack_amd import workloads
    return workloads.cube(n=500, seed=5, tmax=0.05)


def test_large_table_timing(gpu):
    """64^3 grid, B and V: host (16 threads) and GPU pre-interpolation agree; times printed."""
    rng = np.random.default_rng(1)
    grids = [np.linspace(0, 1, 64)] * 3
    values = rng.normal(size=(4, 64, 64, 64))
    tricubic_coefficients_gpu(*[g[:4] for g in grids], values[:, :4, :4, :4])     # warm-up
    t = time.perf_counter()
    g = tricubic_coefficients_gpu(*grids, values)
    tg = time.perf_counter() - t
    t = time.perf_counter()
    h = tricubic_coefficients(*grids, values, threads=16)
    th = time.perf_counter() - t
    print(f"precompute 64^3 x 4: gpu {tg * 1e3:.1f} ms (incl. transfers), host {th * 1e3:.1f} ms", flush=True)
    assert np.array_equal(g, h)
