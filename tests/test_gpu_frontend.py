"""GPU: field and solid queries through the C ABI against the oracle, and the
PENTrack-style command line (config.in -> source -> tracking -> end logs)."""
import os

import numpy as np
import pytest

from pentrack_amd import queries, workloads

pytestmark = pytest.mark.gpu


def test_field_and_solid_queries_match_oracle(gpu, oracle_lib):
    w = workloads.analytic(n=10, seed=3)
    rng = np.random.default_rng(4)
    pts = rng.uniform(-0.2, 1.2, (4000, 3))
    t = rng.uniform(0, 0.1, 4000)
    print("case: field queries n=4000", flush=True)
    B, dB, V, E = queries.fields_eval(w.fields, pts, t)
    oB, odB, oV, oE = oracle_lib.field_eval(w.fields, pts, t, "fp64")
    for a, b in ((B, oB), (dB, odB), (V, oV), (E, oE)):
        assert np.array_equal(a, b)
    print("case: solid queries n=4000", flush=True)
    sid = queries.solid_at(w.geometry, pts, t)
    P = oracle_lib.Problem(w.geometry, w.fields)
    for i in range(0, 4000, 7):
        ids, ign = oracle_lib.get_solids(P, t[i], pts[i])
        live = [int(s) for s, g in zip(ids, ign) if not g]
        assert sid[i] == (max(live) if live else 1), i


def test_cell_search_on_non_uniform_grid_matches_oracle(gpu, oracle_lib):
    """Points whose uniform-spacing cell guess misses (a strongly non-uniform grid), points
    on interior nodes and on both grid edges, outside the grid and NaN: the GPU cell search
    falls back to its binary search and must return the oracle's cells, so B, dB, V and E
    are bit-exact."""
    from pentrack_amd.fields import TFieldManager, tabfield3
    xg = np.array([-0.2, -0.19, -0.15, 0.0, 0.02, 0.25, 0.3])
    yg = np.array([0.0, 0.05, 0.13, 0.2, 0.31, 0.4])
    zg = np.array([0.1, 0.11, 0.12, 0.4, 0.45, 0.5])
    X, Y, Z = np.meshgrid(xg, yg, zg, indexing="ij")
    B = np.stack([np.sin(3 * X) * Y, np.cos(2 * Y) + Z * X, X * Y * Z + 0.1])
    V = X * X - Y * Z
    fm = TFieldManager([tabfield3(xg, yg, zg, B=B, V=V, escale=1.0)])
    rng = np.random.default_rng(11)
    inner = np.stack([rng.uniform(-0.2, 0.3, 3000), rng.uniform(0.0, 0.4, 3000), rng.uniform(0.1, 0.5, 3000)], 1)
    nodes = np.stack(np.meshgrid(xg, yg, zg, indexing="ij"), -1).reshape(-1, 3)
    edges = np.array([[0.3, 0.2, 0.3], [-0.2, 0.0, 0.1], [-0.2000001, 0.2, 0.3], [0.0, 0.4, 0.3],
                      [np.nan, 0.2, 0.3], [0.1, np.nan, 0.3], [5.0, 0.2, 0.3], [-5.0, -5.0, -5.0]])
    pts = np.concatenate([inner, nodes, edges])
    print(f"case: non-uniform grid field queries n={len(pts)}", flush=True)
    got = queries.fields_eval(fm, pts, 0.0)
    ref = oracle_lib.field_eval(fm, pts, 0.0, "fp64")
    for a, b in zip(got, ref):
        assert np.array_equal(a, b, equal_nan=True)


def read_log(path):
    lines = open(path).read().splitlines()
    titles = lines[0].split()
    vals = np.array([[float(x) for x in l.split()] for l in lines[1:]])
    return titles, vals


def test_command_line_run(gpu, oracle_lib, tmp_path):
    from tests.test_config import write_config
    from pentrack_amd.__main__ import main
    from pentrack_amd import config as C
    from pentrack_amd.source import VolumeSource
    from pentrack_amd.workloads import Workload
    cfgpath = write_config(str(tmp_path))
    out = str(tmp_path / "out")
    print("case: command line simcount 200", flush=True)
    assert main(["3", cfgpath, out, "77"]) == 0
    titles, nlog = read_log(os.path.join(out, "000000000003neutronend.out"))
    assert len(nlog) == 200 and titles[:3] == ["jobnumber", "particle", "tstart"]
    # the same particles through the oracle
    cfg = C.TConfig(cfgpath)
    geometry, fields = C.load_geometry(cfg, str(tmp_path)), C.load_fields(cfg)
    t0, y0, spin0 = VolumeSource(cfg).sample(200, np.random.default_rng(77))
    species = [C.species_config(cfg, n) for n in ("neutron", "proton", "electron")]
    w = Workload("cli", geometry, fields, species, 0.5, np.zeros(200, np.int32), t0, y0, spin0,
                 np.arange(1, 201, dtype=np.uint64), 77)
    o, _ = oracle_lib.track(w, "fp64", threads=8)
    col = {t: k for k, t in enumerate(titles)}
    assert np.array_equal(nlog[:, col["stopID"]], o["stopID"])
    assert np.array_equal(nlog[:, col["Nhit"]], o["Nhit"])
    assert np.allclose(nlog[:, col["xend"]], o["y"][:, 0], rtol=1e-14, atol=1e-300)
    assert np.allclose(nlog[:, col["Hmax"]], o["Hmax"], rtol=1e-14)
    # hit log: every hit of every neutron, particles in order, as the oracle records them
    htitles, hlog = read_log(os.path.join(out, "000000000003neutronhit.out"))
    assert htitles == ["jobnumber", "particle", "t", "x", "y", "z", "v1x", "v1y", "v1z", "pol1", "v2x", "v2y", "v2z",
                       "pol2", "nx", "ny", "nz", "solid1", "solid2"]
    o2, oh = oracle_lib.track(w, "fp64", hit_cap=256, threads=8)
    assert (o2["n_hits"] <= 256).all()
    rows = np.concatenate([oh[k, :o2["n_hits"][k]] for k in range(200)])
    assert len(hlog) == len(rows) == o["Nhit"].sum()
    hc = {t: k for k, t in enumerate(htitles)}
    assert np.array_equal(hlog[:, hc["particle"]], np.repeat(np.arange(1, 201), o2["n_hits"]))
    for name, want in (("t", rows["t"]), ("x", rows["pos"][:, 0]), ("z", rows["pos"][:, 2]),
                       ("v2y", rows["v2"][:, 1]), ("nz", rows["normal"][:, 2]), ("pol2", rows["pol2"])):
        assert np.allclose(hlog[:, hc[name]], want, rtol=1e-14, atol=1e-300), name
    assert np.array_equal(hlog[:, hc["solid2"]], rows["entering"])
    dec = int((o["stopID"] == -4).sum())
    if dec:
        _, plog = read_log(os.path.join(out, "000000000003protonend.out"))
        _, elog = read_log(os.path.join(out, "000000000003electronend.out"))
        assert len(plog) == dec and len(elog) == dec


def test_command_line_spin_log(gpu, oracle_lib, tmp_path):
    """spinlog 1 on the command line: the spin log file holds the rows the oracle records
    (src/logger.cpp:237-284), with B at each row's position and time, filtered by a
    spinlogfilter formula and with a formula column (src/logger.cpp:286-308)."""
    from tests.test_config import CONFIG, write_config
    from pentrack_amd.__main__ import main
    from pentrack_amd import config as C
    from pentrack_amd.source import VolumeSource
    from pentrack_amd.workloads import Workload
    from oracle import ptref
    cfgpath = write_config(str(tmp_path))
    text = CONFIG.replace("secondaries 1", "secondaries 0").replace("simcount 200", "simcount 20")
    text = text.replace("[neutron]\ntau 0.05\n", "[neutron]\ntau 0.05\nspintimes 0 0.02\nspinlog 1\nspinloginterval 1e-3\n"
                        "spinlogfilter late\nspinlogvars jobnumber particle t x y z Sx Sy Sz Wx Wy Wz Bx By Bz Sperp\n")
    text = text.replace("[FORMULAS]\n", "[FORMULAS]\nlate t >= 0.005 and z > 0\nSperp sqrt(Sx^2 + Sy^2)\n")
    open(cfgpath, "w").write(text)
    out = str(tmp_path / "out")
    print("case: command line spin log simcount 20", flush=True)
    assert main(["3", cfgpath, out, "77"]) == 0
    titles, slog = read_log(os.path.join(out, "000000000003neutronspin.out"))
    assert titles == ["jobnumber", "particle", "t", "x", "y", "z", "Sx", "Sy", "Sz", "Wx", "Wy", "Wz", "Bx", "By", "Bz",
                      "Sperp"]
    cfg = C.TConfig(cfgpath)
    geometry, fields = C.load_geometry(cfg, str(tmp_path)), C.load_fields(cfg)
    t0, y0, spin0 = VolumeSource(cfg).sample(20, np.random.default_rng(77))
    species = [C.species_config(cfg, n) for n in ("neutron", "proton", "electron")]
    w = Workload("cli", geometry, fields, species, 0.5, np.zeros(20, np.int32), t0, y0, spin0,
                 np.arange(1, 21, dtype=np.uint64), 77)
    o, rec, cnt = oracle_lib.track_spinlog(w, 1e-3, 4096, threads=8)
    assert (cnt <= 4096).all()
    rows = np.concatenate([rec[k, :cnt[k]] for k in range(20)])
    num = np.repeat(np.arange(1, 21), cnt)
    keep = (rows["t"] >= 0.005) & (rows["pos"][:, 2] > 0)
    assert 20 * 5 < keep.sum() < len(rows)
    rows, num = rows[keep], num[keep]
    c = {t: k for k, t in enumerate(titles)}
    assert len(slog) == len(rows) and np.array_equal(slog[:, c["particle"]], num)
    assert np.allclose(slog[:, c["Sperp"]], np.hypot(rows["spin"][:, 0], rows["spin"][:, 1]), rtol=1e-14)
    for name, want in (("t", rows["t"]), ("x", rows["pos"][:, 0]), ("z", rows["pos"][:, 2]), ("Sx", rows["spin"][:, 0]),
                       ("Sz", rows["spin"][:, 2]), ("Wz", rows["W"][:, 2])):
        assert np.allclose(slog[:, c[name]], want, rtol=1e-14, atol=1e-300), name
    B, _, _, _ = ptref.field_eval(fields, np.ascontiguousarray(rows["pos"]), np.ascontiguousarray(rows["t"]), grad=False)
    assert np.allclose(slog[:, c["Bz"]], B[:, 2], rtol=1e-14, atol=1e-300)
