"""GPU: axisymmetric 2D maps (TabField) through the C ABI against the oracle, and the
reference's IntegrationTest (test/IntegrationTest/config.in: a neutron in the 0.5 + 0.5 T/m
vertical gradient of the OPERA2D and OPERA3D fixture tables, hitting the hollow unit cube)
tracked on the GPU and by the oracle, with the closed-form parabola of
test/IntegrationTest/showintegrationresult.cxx."""
import os
import shutil

import numpy as np
import pytest

from pentrack_amd import config as C
from pentrack_amd import queries
from pentrack_amd.fields import TFieldManager, tabfield2, uniform_efield
from pentrack_amd.source import VolumeSource
from pentrack_amd.workloads import Workload

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def synthetic_fields(rng):
    m, n = 23, 17
    rg = np.concatenate([[0.0], np.cumsum(rng.uniform(0.02, 0.05, m - 1))])
    zg = np.linspace(-0.4, 0.6, n)
    R, Z = np.meshgrid(rg, zg)
    a = tabfield2(rg, zg, bscale="1 + 0.5*t", escale=3.0, Br=R * Z + R ** 2 + 0.01 * np.sin(9 * Z),
                  Bphi=R * (1 - Z) / 2, Bz=Z ** 2 - R / 3 + 0.02 * np.cos(7 * R), V=R ** 2 * Z + np.sin(5 * Z))
    b = tabfield2(rg * 0.7, zg, Bz=np.cos(3 * R) + Z, Er=3 * R + Z, Ephi=0 * R + 0.5, Ez=R - 2 * Z)
    return TFieldManager([a, b, uniform_efield(0.0, 10.0, 20.0)])


def test_field_queries_match_oracle(gpu, oracle_lib):
    rng = np.random.default_rng(11)
    fm = synthetic_fields(rng)
    pts = np.concatenate([rng.uniform(-0.9, 0.9, (3000, 3)), np.zeros((1, 3)), [[0.0, 0.0, 0.3]],
                          [[1e-9, -1e-9, 0.1]]])
    t = rng.uniform(0, 1, len(pts))
    print("case: 2D-map field queries n=%d" % len(pts), flush=True)
    for grad in (True, False):
        B, dB, V, E = queries.fields_eval(fm, pts, t, grad=grad)
        oB, odB, oV, oE = oracle_lib.field_eval(fm, pts, t, "fp64", grad=grad)
        assert np.array_equal(B, oB) and np.array_equal(V, oV) and np.array_equal(E, oE), grad
        if grad:
            assert np.array_equal(dB, odB)
    assert np.abs(B).max() > 0.1 and (np.abs(B).sum(1) == 0).sum() > 100     # in and out of range


def integration_test_config(tmp):
    """test/IntegrationTest/config.in with the fixture files (2D and 3D gradient tables,
    test/HollowUnitCube.STL) and the materials it uses (in/materials.in:15,23)."""
    for name in ("VerticalLinearGradientField2D.tab", "VerticalLinearGradientField3D.tab.gz", "HollowUnitCube.STL"):
        shutil.copy(os.path.join(GOLDEN, name), tmp)
    text = """[GLOBAL]
simtype 1
simcount 400
simtime 10
secondaries 0
[MATERIALS]
default		0			0			0				0	0	0	0	0   0
UCNdet		0.1			1			0				0	0	0	0	0   0
[GEOMETRY]
1	ignored				default
2	HollowUnitCube.STL		UCNdet
[SOURCE]
sourcemode boxvolume
parameters 0.1	0.1	0.5	0.5		0.01	0.01
particle neutron
ActiveTime 0
PhaseSpaceWeighting 0
Emin 10e-9
Emax 150e-9
spectrum 1
phi_v_min 0
phi_v_max 0
phi_v 1
theta_v_min 5
theta_v_max 5
theta_v 1
polarization 1
[FIELDS]
1 OPERA2D VerticalLinearGradientField2D.tab	0.00005	0	0.01
2 OPERA3D VerticalLinearGradientField3D.tab.gz	0.00005	0	0	0.01
[PARTICLES]
tau 0
tmax 9e99
lmax 9e99
Bmax 0.1
flipspin 0
[neutron]
tau 0
[proton]
tmax 3e-3
[electron]
tmax 1e-5
"""
    path = os.path.join(tmp, "config.in")
    with open(path, "w") as f:
        f.write(text)
    return path


def test_reference_integration_test(gpu, oracle_lib, tmp_path):
    cfgpath = integration_test_config(str(tmp_path))
    cfg = C.TConfig(cfgpath)
    geometry, fields = C.load_geometry(cfg, str(tmp_path)), C.load_fields(cfg)
    n = 400
    t0, y0, spin0 = VolumeSource(cfg).sample(n, np.random.default_rng(21))
    species = [C.species_config(cfg, s) for s in ("neutron", "proton", "electron")]
    w = Workload("integration", geometry, fields, species, 10.0, np.zeros(n, np.int32), t0, y0, spin0,
                 np.arange(1, n + 1, dtype=np.uint64), 21)
    from tests.test_gpu_parity import FIELDS, run_gpu
    print("case: IntegrationTest n=400", flush=True)
    g, _ = run_gpu(w, hit_cap=4)
    o, _ = oracle_lib.track(w, "fp64", hit_cap=4, threads=8)
    for f in FIELDS:
        assert np.array_equal(g[f], o[f]), f
    # every neutron ends on the detector; x at the hit follows the closed-form parabola with
    # a = -g - mu |dB/dz| / m = -15.5754762729 m/s^2 (showintegrationresult.cxx:6)
    assert (g["stopID"] > 0).all()
    a = -15.5754762729
    xs, zs, vx, vz = y0[:, 0], y0[:, 2], y0[:, 3], y0[:, 5]
    ye = g["y"]
    thit = (-vz - np.sqrt(vz * vz - 2 * a * (zs - ye[:, 2]))) / a
    resid = ye[:, 0] - xs - vx * thit
    assert np.abs(resid).max() < 1e-6, np.abs(resid).max()
