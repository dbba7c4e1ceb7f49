"""Field-table readers (OPERA3D, 3Dtable, COMSOL; plain, gzip, bzip2) against the
reference's own table fixtures and closed forms, evaluated with the oracle."""
import bz2
import gzip
import os
import shutil

import numpy as np
import pytest

from pentrack_amd import config as C
from pentrack_amd import tables
from pentrack_amd.fields import tabfield3

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fields_of(text, base):
    return C.load_fields(C.TConfig(text="[FIELDS]\n" + text + "\n"), base)


def plain(tmp_path, name):
    """Decompressed copy of a fixture in tmp_path."""
    src = os.path.join(GOLDEN, name)
    opener = {".gz": gzip.open, ".bz2": bz2.open}.get(os.path.splitext(name)[1], open)
    dst = tmp_path / os.path.splitext(name)[0] if opener is not open else tmp_path / name
    with opener(src, "rb") as f, open(dst, "wb") as g:
        g.write(f.read())
    return str(dst)


def test_opera3d_vertical_gradient_closed_form(oracle_lib, tmp_path):
    """test/VerticalLinearGradientField3D.tab: 6x6x6 points 0..100 cm, BZ = V = 100 z[cm]
    (CRLF lines).  Tricubic interpolation of linear data is exact: with B scale -1e-4 and
    E scale -1 (test/2DFieldTest/config.in:19) Bz = -z [T], dBz/dz = -1, V = -1e4 z."""
    shutil.copy(os.path.join(GOLDEN, "VerticalLinearGradientField3D.tab.gz"), tmp_path)
    f = fields_of("3 OPERA3D VerticalLinearGradientField3D.tab.gz -0.0001 -1 0 0.01", str(tmp_path))
    c = f.containers[0]
    assert tuple(len(g) for g in c.grid) == (6, 6, 6) and c.has_b and c.has_v
    assert c.box == (1.0, 0.0, 1.0, 0.0, 1.0, 0.0) and c.bw == 0.0
    rng = np.random.default_rng(1)
    pts = rng.uniform(0.0, 0.999, (500, 3))
    B, dB, V, E = oracle_lib.field_eval(f, pts, 0.0, "fp64")
    z = pts[:, 2]
    assert np.allclose(B[:, 2], -z, rtol=1e-12, atol=1e-13) and np.abs(B[:, :2]).max() == 0
    assert np.allclose(dB[:, 2, 2], -1.0, rtol=1e-12) and np.abs(dB[:, 2, :2]).max() < 1e-12
    assert np.allclose(V, -1e4 * z, rtol=1e-12, atol=1e-9)
    assert np.allclose(E[:, 2], 1e4, rtol=1e-12) and np.abs(E[:, :2]).max() < 1e-8
    # outside the table's extent the container holds no field
    B2, _, V2, _ = oracle_lib.field_eval(f, np.array([[0.5, 0.5, 1.0], [-1e-9, 0.5, 0.5]]), 0.0, "fp64")
    assert not B2.any() and not V2.any()


def test_opera3d_compression_and_3dtable_units(tmp_path):
    """Plain, gzip and bzip2 files give the same table; the old 3Dtable type is OPERA3D
    with the scales multiplied by 1e-4 (B) and 100 (E) and lengths in cm."""
    p = plain(tmp_path, "VerticalLinearGradientField3D.tab.gz")
    with open(p, "rb") as f:
        raw = f.read()
    with bz2.open(p + ".bz2", "wb") as g:
        g.write(raw)
    with gzip.open(p + ".gz", "wb") as g:
        g.write(raw)
    cs = [fields_of(f"1 OPERA3D {os.path.basename(n)} 0.0001 100 0 0.01", str(tmp_path)).containers[0]
          for n in (p, p + ".bz2", p + ".gz")]
    old = fields_of(f"1 3Dtable {os.path.basename(p)} 1 1 0", str(tmp_path)).containers[0]
    for c in cs[1:] + [old]:
        assert np.array_equal(c.coef, cs[0].coef)
        assert all(np.array_equal(a, b) for a, b in zip(c.grid, cs[0].grid))
    assert cs[0].bformula == "0.0001" and old.bformula == "(1)*0.0001" and old.eformula == "(1)*100"


def write_opera3d(path, xg, yg, zg, cols, rows_order=None, header_sizes=None, sep="  "):
    """An OPERA3D text file: sizes, column descriptions, ' 0', then rows x y z cols...
    cols: {name: [nx][ny][nz]} in the order BX, BY, BZ, V."""
    X, Y, Z = np.meshgrid(xg, yg, zg, indexing="ij")
    names = [k for k in ("BX", "BY", "BZ", "V") if k in cols]
    data = np.stack([X.ravel(), Y.ravel(), Z.ravel()] + [np.asarray(cols[k]).ravel() for k in names], 1)
    if rows_order is not None:
        data = data[rows_order]
    sizes = header_sizes or (len(xg), len(yg), len(zg))
    with open(path, "w") as f:
        f.write("%11d%11d%11d  2\n 1 X [METRE]\n 2 Y [METRE]\n 3 Z [METRE]\n" % sizes)
        for i, k in enumerate(names):
            f.write(" %d %s [UNIT]\n" % (i + 4, k))
        f.write(" 0\n")
        for r in data:
            f.write(sep.join("%.17g" % v for v in r) + "\n")
    return data


def test_opera3d_missing_components_scatter_and_errors(oracle_lib, tmp_path):
    rng = np.random.default_rng(2)
    xg, yg, zg = np.array([0, 0.1, 0.25, 0.3]), np.array([-0.2, 0.0, 0.2]), np.array([1.0, 1.5, 1.75, 2.0, 2.5])
    By = rng.normal(size=(4, 3, 5))
    V = rng.normal(size=(4, 3, 5))
    # rows in random order: samples go to their grid points by coordinate
    write_opera3d(tmp_path / "a.table", xg, yg, zg, {"BY": By, "V": V}, rows_order=rng.permutation(60))
    c = fields_of("1 OPERA3D a.table 2 3 0.05 1", str(tmp_path)).containers[0]
    direct = tabfield3(xg, yg, zg, B=np.stack([np.zeros_like(By), By, np.zeros_like(By)]), V=V,
                       bscale="2", escale="3", box=(0.3, 0.0, 0.2, -0.2, 2.5, 1.0), bw=0.05)
    assert np.array_equal(c.coef, direct.coef) and c.box == direct.box and c.bw == 0.05
    pts = np.stack([rng.uniform(0, 0.3, 300), rng.uniform(-0.2, 0.2, 300), rng.uniform(1, 2.5, 300)], 1)
    a = oracle_lib.field_eval(fields_of("1 OPERA3D a.table 2 3 0.05 1", str(tmp_path)), pts, 0.0, "fp64")
    b = oracle_lib.field_eval(C.TFieldManager([direct]), pts, 0.0, "fp64")
    assert all(np.array_equal(u, v) for u, v in zip(a, b))
    # the table's values at its grid points (tricubic reproduces the samples there)
    i = (1, 1, 3)
    Bn, _, Vn, _ = oracle_lib.field_eval(C.TFieldManager([tabfield3(xg, yg, zg, B=np.stack([0 * By, By, 0 * By]), V=V,
                                                                    escale=1.0)]),
                                         np.array([[xg[i[0]], yg[i[1]], zg[i[2]]]]), 0.0, "fp64")
    assert Bn[0, 1] == By[i] and Vn[0] == V[i]
    # header size that does not match the rows
    write_opera3d(tmp_path / "b.table", xg, yg, zg, {"BZ": By}, header_sizes=(4, 3, 6))
    with pytest.raises(C.ConfigError, match="header says the size is 72, actually it is 60"):
        fields_of("1 OPERA3D b.table 1 1 0 1", str(tmp_path))
    # column description not closed by a line starting " 0"
    (tmp_path / "c.table").write_text("2 2 2\n 1 X\n 2 Y\n 3 Z\n 4 BX\n 5 BY\n")
    with pytest.raises(C.ConfigError, match="corrupt"):
        fields_of("1 OPERA3D c.table 1 1 0 1", str(tmp_path))
    with pytest.raises(C.ConfigError, match="Could not open"):
        fields_of("1 OPERA3D missing.table 1 1 0 1", str(tmp_path))
    with pytest.raises(C.ConfigError, match="required parameters"):
        fields_of("1 OPERA3D a.table 1 1 0", str(tmp_path))
    with pytest.raises(ValueError, match="invalid boundary"):
        fields_of("1 OPERA3D a.table 1 1 0.25 1", str(tmp_path))      # wider than half the table


def test_comsol_fixture_values_at_nodes(oracle_lib, tmp_path):
    """test/comsolFieldTest/comsolField.txt (a COMSOL arrow export, 1791 nodes with '%'
    header lines): the table reproduces every node inside its bounds exactly, and grid
    points without a node hold zero field."""
    shutil.copy(os.path.join(GOLDEN, "comsolField.txt.bz2"), tmp_path)
    f = fields_of("1 COMSOL comsolField.txt.bz2 1 0 1", str(tmp_path))
    c = f.containers[0]
    assert not c.has_v and c.eformula == "0" and c.bformula == "1"
    with bz2.open(os.path.join(GOLDEN, "comsolField.txt.bz2"), "rt") as g:
        rows = np.array([[float(v) for v in l.split()] for l in g if l[:1] not in "%#"])
    assert len(rows) == 1791
    xg, yg, zg = c.grid
    assert (xg == np.unique(rows[:, 0])).all() and (zg == np.unique(rows[:, 2])).all()
    inside = (rows[:, 0] < xg[-1]) & (rows[:, 1] < yg[-1]) & (rows[:, 2] < zg[-1])
    B, _, _, _ = oracle_lib.field_eval(f, rows[inside, :3], 0.0, "fp64")
    assert np.array_equal(B, rows[inside, 3:])
    n_grid = len(xg) * len(yg) * len(zg)
    assert n_grid > 1791                       # some grid points have no node: zero field there
    # the same rows with commas and tabs, and a length conversion
    with open(tmp_path / "c.txt", "w") as g:
        g.write("# x,y,z,Bx,By,Bz\n")
        for r in rows[:200]:
            g.write("%r,%r\t%r, %r,%r,%r\n" % tuple(map(float, r)))
    with open(tmp_path / "s.txt", "w") as g:
        for r in rows[:200]:
            g.write("%r %r %r %r %r %r\n" % tuple(map(float, r)))
    a = fields_of("1 COMSOL c.txt 1 0 100", str(tmp_path)).containers[0]
    b = fields_of("1 COMSOL s.txt 1 0 100", str(tmp_path)).containers[0]
    assert np.array_equal(a.coef, b.coef) and np.array_equal(a.grid[0], np.unique(rows[:200, 0]) * 100)
    (tmp_path / "bad.txt").write_text("1 2 3 4 5 6\n1 2 3 4 5\n")
    with pytest.raises(C.ConfigError, match="Error reading line 2"):
        fields_of("1 COMSOL bad.txt 1 0 1", str(tmp_path))
    (tmp_path / "lead.txt").write_text(" 1 2 3 4 5 6\n")                 # a leading blank is an empty column
    with pytest.raises(C.ConfigError, match="Error reading line 1"):
        fields_of("1 COMSOL lead.txt 1 0 1", str(tmp_path))


def test_stod_and_stream_numbers():
    assert tables._stod("1.5e3") == 1500.0 and tables._stod("6\r") == 6.0 and tables._stod("2x") == 2.0
    assert tables._stod("1_0") == 1.0                                     # no digit separators in C
    with pytest.raises(ValueError):
        tables._stod("x1")
    v = tables._to_doubles(["1", "-.5", "2E-3", "abc", "4"])
    assert list(v) == [1.0, -0.5, 0.002]                                  # reading stops at the first non-number
