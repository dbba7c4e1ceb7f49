"""GPU: BMT spin tracking (IntegrateSpin inside spin-tracking windows) against the
oracle, bit for bit, and on the Larmor precession."""
import numpy as np
import pytest

from pentrack_amd import workloads
from pentrack_amd.particle import SpeciesConfig
from pentrack_amd.workloads import Workload
from tests.test_gpu_parity import FIELDS, HITS, assert_bit_exact, run_gpu
from tests.test_oracle_bmt import GAMMA_N, falling_workload, larmor_workload

pytestmark = pytest.mark.gpu


def test_larmor_precession_gpu(gpu, oracle_lib):
    w = larmor_workload(1e-6, (0.0, 1.0), 0.05, n=64)
    print("case: BMT Larmor n=64", flush=True)
    g, _ = run_gpu(w, hit_cap=0)
    o, _ = oracle_lib.track(w, "fp64")
    for f in FIELDS:
        assert np.array_equal(g[f], o[f]), f
    wl = -GAMMA_N * 1e-6
    assert np.allclose(g["spin"][:, 0], np.cos(wl * 0.05), atol=1e-8)


def test_bmt_with_walls_and_map(gpu, oracle_lib):
    """Cylinder with the coil map and wall hits; spin tracking in two windows and Bmax
    below part of the map: the step and collision paths both hand spin steps over."""
    # |B| <= 20 uT keeps the Larmor period above 1.7 ms: a few hundred spin steps per window
    base = workloads.cylinder(n=300, seed=9, tmax=0.3, n_grid=24, Bmax=2e-5)
    sp = [SpeciesConfig(kind=0, tau=0.0, tmax=9e99, lmax=9e99, Bmax=1e-5, spintimes=(0.0, 0.1, 0.2, 0.25))]
    spin0 = base.spin0.copy()
    spin0[:, 0] = 1.0
    w = Workload("bmt", base.geometry, base.fields, sp, base.tmax, base.species_idx, base.t0, base.y0, spin0,
                 base.pid, base.seed)
    print("case: BMT cylinder n=400", flush=True)
    g, gh = run_gpu(w)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS, threads=8)
    assert_bit_exact(g, gh, o, oh)
    assert g["Nhit"].sum() > 100 and (g["spin"][:, 3] > 0).mean() > 0.5      # spin integrated in the windows


@pytest.mark.parametrize("name,params", [("B0GradZ", (0.0, 1e-5, 1e-6)), ("ExponentialFieldX", (2e-6, 3.0, 0.0, 1e-6, 0.0))])
def test_interpolated_axis_gpu(gpu, oracle_lib, name, params):
    """interpolatefields 1: the precession axis from splines over 11 points of each step."""
    w = falling_workload(1, name, params, n=64)
    print(f"case: BMT interpolatefields {name} n=64", flush=True)
    g, _ = run_gpu(w, hit_cap=0)
    o, _ = oracle_lib.track(w, "fp64")
    for f in FIELDS:
        assert np.array_equal(g[f], o[f]), f
    assert (g["spin"][:, 4] > 1).all()


@pytest.mark.parametrize("which", ["larmor", "cylinder"])
def test_spin_log_matches_oracle(gpu, oracle_lib, which):
    """TLogger::PrintSpin rows recorded inside bmt_integrate, bit for bit with the oracle."""
    from pentrack_amd.tracker import Batch, LaunchConfig
    if which == "larmor":
        w, iv = larmor_workload(1e-6, (0.0, 1.0), 0.05, n=64), 2e-3
    else:
        base = workloads.cylinder(n=200, seed=9, tmax=0.15, n_grid=24, Bmax=2e-5)
        sp = [SpeciesConfig(kind=0, tau=0.0, tmax=9e99, lmax=9e99, Bmax=1e-5, spintimes=(0.0, 0.1))]
        spin0 = base.spin0.copy()
        spin0[:, 0] = 1.0
        w = Workload("bmt", base.geometry, base.fields, sp, base.tmax, base.species_idx, base.t0, base.y0, spin0,
                     base.pid, base.seed)
        iv = 5e-3
    b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid)
    b.set_spinlog(iv, 128)
    print(f"case: spin log {which} n={w.n}", flush=True)
    assert b.run_to_completion(LaunchConfig(), max_launches=100000)
    out, _ = b.download()
    rec, cnt = b.spinlog()
    o, orec, ocnt = oracle_lib.track_spinlog(w, iv, 128, threads=8)
    assert np.array_equal(out["spin"], o["spin"]) and np.array_equal(cnt, ocnt)
    assert cnt.sum() > 10 * w.n
    kept = np.minimum(cnt, 128)
    for k in range(w.n):
        assert np.array_equal(rec[k, :kept[k]], orec[k, :kept[k]]), k
