"""The step kernels' wave-cooperative coefficient gathers against per-lane loads of the same
FP64 table (pt_fields_eval_tables): B and its gradient bit-identical, on the 128^3 coil map
of the mixed-precision workload.  Points partly outside the map leave lanes of a wave
inactive, and the point count is not a multiple of the wave size, so the cooperative
gathers run over partial waves too.  The mixed-precision tables (FP64 linear part plus FP32
remainder, FP64 blocks where the remainder's bound exceeds the tolerance; 128-byte cell
records) are read per lane by the step kernels: evaluated through the same entry point they
agree with the FP64 tables within the remainder's bound."""
import numpy as np
import pytest

from pentrack_amd import _native, workloads
from pentrack_amd.queries import fields_eval

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cube128():
    w = workloads.cube(n=16, seed=3, tmax=0.25, with_field=True, n_grid=128, Bmax=1.0)
    rng = np.random.default_rng(11)
    pts = rng.uniform(-0.06, 1.06, (100003, 3))
    t = np.zeros(len(pts))
    return w.fields, pts, t


def test_fp64_cooperative_equals_per_lane(gpu, cube128):
    f, pts, t = cube128
    print("case: FP64 tables, per-lane vs cooperative gathers, n=100003", flush=True)
    B0, dB0, _, _ = fields_eval(f, pts, t, variant=_native.PT_VARIANT_FP64, cooperative=False)
    B1, dB1, _, _ = fields_eval(f, pts, t, variant=_native.PT_VARIANT_FP64, cooperative=True)
    Bp, dBp, _, _ = fields_eval(f, pts, t)
    assert np.array_equal(B0, Bp) and np.array_equal(dB0, dBp)
    assert np.array_equal(B0, B1) and np.array_equal(dB0, dB1)
    assert np.count_nonzero(np.abs(B0).sum(axis=1) == 0) > 1000    # lanes outside the map


def test_mixed_tables(gpu, cube128):
    f, pts, t = cube128
    print("case: mixed tables against FP64 tables, n=100003", flush=True)
    B0, dB0, _, _ = fields_eval(f, pts, t, variant=_native.PT_VARIANT_MIXED, cooperative=False)
    B1, dB1, _, _ = fields_eval(f, pts, t, variant=_native.PT_VARIANT_MIXED, cooperative=True)
    assert np.array_equal(B0, B1) and np.array_equal(dB0, dB1)
    blocks, fp64 = f.mixed_info()
    assert 0.05 * blocks < fp64 < 0.5 * blocks          # both kinds of component are exercised
    Bd, dBd, _, _ = fields_eval(f, pts, t)
    assert not np.array_equal(B0, Bd)                   # the mixed tables are in use
    assert np.abs(B0 - Bd).max() <= 1e-6 * np.abs(Bd).max()
    # value-only evaluations (the dense-output and energy paths) agree as well
    V0 = fields_eval(f, pts, t, grad=False, variant=_native.PT_VARIANT_MIXED, cooperative=False)[0]
    V1 = fields_eval(f, pts, t, grad=False, variant=_native.PT_VARIANT_MIXED, cooperative=True)[0]
    assert np.array_equal(V0, V1)
