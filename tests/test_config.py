"""config.in front end: TConfig semantics (src/config.cpp:14-53), materials and solids
(src/geometry.cpp:10-137), the FIELDS section (src/fields.cpp:19-136), particle
settings (src/main.cpp:242-249), volume sources (src/source.cpp) and the text end
log format (src/logger.cpp:314-334)."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from pentrack_amd import _native
from pentrack_amd import config as C
from pentrack_amd.geometry import box_mesh, hollow_box_mesh, write_stl
from pentrack_amd.logger import _fmt

CONFIG = """
# test configuration
[GLOBAL]
simtype 1
simcount 200       # particles
simtime 0.5
secondaries 1

[MATERIALS]
#name FermiReal FermiImag Lambert Spinflip RMS Corr B_int ModLambert Loss
vacuum  0 0 0 0 0 0
wall    335 0.01 0.05 0 0 0 0 0 1e-4
steel   183 0.0852 0 1e-5 0 0

[GEOMETRY]
1   ignored   vacuum
2   shell.STL wall
10  flap.STL  steel   0-0.1 0.3-1e9

[FORMULAS]
ramp  t < 0.1 ? 0 : 1

[FIELDS]
5 Conductor 12500 1.5 0.5 -1 1.5 0.5 2 ramp
7 EDMStaticEField 0 0 1e5 1
8 B0GradZ 1e-3 -1e-3 1e-3 1 0 1 0 1 0 0.5 + 0.5*t

[SOURCE]
sourcemode boxvolume
parameters 0.1 0.9 0.1 0.9 0.1 0.6
particle neutron
ActiveTime 0
Emin 100e-9
Emax 300e-9
spectrum 1.96616e39*x^5 - 0.00204264e36*x^4 + 0.834378e27*x^3 - 167.958e18*x^2 + 16674.8e9*x - 639317
phi_v_min 0
phi_v_max 360
phi_v 1
theta_v_min 0
theta_v_max 180
theta_v sin(x)
polarization 1

[PARTICLES]
tau 0
tmax 9e99
lmax 9e99
endlog 1
Bmax 0.1

[neutron]
tau 0.05
hitlog 1

[proton]
tmax 3e-3
"""


def write_config(d):
    path = os.path.join(d, "config.in")
    open(path, "w").write(CONFIG)
    write_stl(os.path.join(d, "shell.STL"), hollow_box_mesh((0, 0, 0), (1, 1, 1), 0.001), "shell")
    write_stl(os.path.join(d, "flap.STL"), box_mesh((0.4, 0.4, 0.7), (0.6, 0.6, 0.75)), "flap")
    return path


def test_tconfig_semantics():
    cfg = C.TConfig(text=CONFIG)
    assert cfg["GLOBAL"]["simcount"].split() == ["200"]          # comment cut
    assert [k for k, _ in cfg.items("GEOMETRY")] == ["1", "10", "2"]   # std::map order
    assert C.global_settings(cfg) == C.GlobalSettings(1, 200, 0.5, 1)
    with pytest.raises(C.ConfigError, match="ambiguous"):
        C.TConfig(text="[A]\nx 1\nx 2\n")
    c2 = C.TConfig(text="[A]\nx 1\n[/A]\ny 2\n[B]\n# c\nz 3 # d\n")
    assert "y" not in c2["A"] and c2["B"]["z"].strip() == "3"
    # PARTICLES defaults merged under each species' own keys
    assert C.species_config(cfg, "neutron").tau == 0.05 and C.species_config(cfg, "proton").tau == 0.0
    assert C.species_config(cfg, "proton").tmax == 3e-3 and C.species_config(cfg, "electron").Bmax == 0.1


def test_materials_geometry_and_fields():
    with tempfile.TemporaryDirectory() as d:
        cfg = C.TConfig(write_config(d))
        g = C.load_geometry(cfg, d)
        assert g.default_material == "vacuum"
        assert [s.ID for s in g.solids] == [10, 2]                    # lexicographic IDs
        assert g.solids[0].ignoretimes == [(0.0, 0.1), (0.3, 1e9)]
        w = g.materials["wall"]
        assert (w.FermiReal, w.FermiImag, w.DiffProb, w.LossPerBounce) == (335.0, 0.01, 0.05, 1e-4)
        assert g.materials["steel"].LossPerBounce == 0.0             # optional trailing values
        f = C.load_fields(cfg)
        types = [c.type for c in f.containers]
        assert types == [_native.PT_FIELD_CONDUCTOR, _native.PT_FIELD_UNIFORM_E, _native.PT_FIELD_B0_GRAD_Z]
        assert f.containers[0].bformula == "t < 0.1 ? 0 : 1" and tuple(f.containers[0].params) == (12500, 1.5, 0.5, -1, 1.5, 0.5, 2)
        assert f.containers[2].box == (1, 0, 1, 0, 1, 0) and f.containers[2].bformula == "0.5"
    with pytest.raises(ValueError):
        C.load_materials(C.TConfig(text="[MATERIALS]\nx 1 0 1.5 0 0 0\n"))
    with pytest.raises(ValueError):
        C.load_materials(C.TConfig(text="[MATERIALS]\nx 1 0 0.5 0 1e-9 1e-8\n"))
    with pytest.raises(C.ConfigError, match="not supported"):
        C.load_fields(C.TConfig(text="[FIELDS]\n1 HarmonicExpandedBField 0 0 0 1\n"))


@pytest.mark.skipif(not os.path.exists("/root/reference/in/config.in"), reason="reference inputs absent")
def test_reference_config_file_parses():
    cfg = C.TConfig("/root/reference/in/config.in")
    cfg.read("/root/reference/in/materials.in")
    assert C.global_settings(cfg).simcount > 0
    mats = C.load_materials(cfg)
    assert len(mats) > 10 and all(m.name for m in mats)
    nc = C.species_config(cfg, "neutron")
    assert nc.tau > 800 and nc.spintimes == (500.0, 700.0)            # BMT between 500 and 700 s


def test_volume_source_distributions():
    from pentrack_amd.source import VolumeSource
    cfg = C.TConfig(text=CONFIG)
    src = VolumeSource(cfg)
    rng = np.random.default_rng(1)
    t0, y0, spin0 = src.sample(20000, rng)
    assert np.all(t0 <= 5e-324) and np.all(y0[:, 7] == 1.0)         # ActiveTime 0: [0, nextafter(0)); polarization 1
    assert y0[:, 0].min() >= 0.1 and y0[:, 0].max() <= 0.9 and y0[:, 2].max() <= 0.6
    from pentrack_amd.particle import kinetic_energy
    E = kinetic_energy(0, y0[:, 3:6])
    assert E.min() >= 100e-9 * (1 - 1e-9) and E.max() <= 300e-9 * (1 + 1e-9)
    v = y0[:, 3:6] / np.linalg.norm(y0[:, 3:6], axis=1)[:, None]
    assert np.abs(v.mean(0)).max() < 0.02                            # isotropic (theta ~ sin)
    # the energy spectrum follows the formula: compare histogram with the density
    p = src.spectrum
    hist, edges = np.histogram(E, bins=20, range=(100e-9, 300e-9), density=True)
    mid = 0.5 * (edges[1:] + edges[:-1])
    dens = np.interp(mid, p.x, p.den)
    assert np.abs(hist - dens).max() < 0.1 * dens.max()


CPP_FMT = r"""
#include <iostream>
#include <iterator>
#include <limits>
#include <vector>
int main() {
    std::vector<double> v{0.0, -0.0, 1.0, 1e-9, 123456789.123456789, 3.141592653589793, -2.5e-300, 1e300,
                          0.1, 2.0/3.0, 7e15, 1234567890123456.0};
    std::cout << std::setprecision(std::numeric_limits<double>::digits10);
    std::copy(v.begin(), v.end(), std::ostream_iterator<double>(std::cout, " "));
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_end_log_number_format_matches_iostream():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "f.cpp"), os.path.join(d, "f")
        open(src, "w").write(CPP_FMT.replace("#include <limits>", "#include <limits>\n#include <iomanip>"))
        subprocess.run(["g++", src, "-o", exe], check=True)
        ref = subprocess.run([exe], capture_output=True, check=True, text=True).stdout
    vals = [0.0, -0.0, 1.0, 1e-9, 123456789.123456789, 3.141592653589793, -2.5e-300, 1e300, 0.1, 2.0 / 3.0, 7e15,
            1234567890123456.0]
    assert "".join(_fmt(v) + " " for v in vals) == ref


def test_hit_log_rows_and_format(tmp_path):
    """TLogger::PrintHit variables (src/logger.cpp:198-232): one row per logged hit,
    particles in order; a particle's hits beyond the record capacity are dropped."""
    from pentrack_amd.logger import DEFAULT_HITLOGVARS, hit_values, write_log
    hits = np.zeros((3, 2), dtype=_native.HIT_DTYPE)
    hits["t"] = [[0.1, 0.2], [0.3, 0.0], [0.5, 0.6]]
    hits["pos"][..., 2] = hits["t"] * 10
    hits["entering"] = [[2, 1], [10, 0], [2, 2]]
    v = hit_values(hits, np.array([2, 1, 5]), np.array([7.0, 8.0, 9.0]), jobnumber=4)
    assert list(v["particle"]) == [7, 7, 8, 9, 9] and list(v["t"]) == [0.1, 0.2, 0.3, 0.5, 0.6]
    assert list(v["z"]) == [1.0, 2.0, 3.0, 5.0, 6.0] and list(v["solid2"]) == [2, 1, 10, 2, 2]
    p = write_log(str(tmp_path), 4, "neutron", "hit", DEFAULT_HITLOGVARS, v)
    lines = open(p).read().split("\n")
    assert os.path.basename(p) == "000000000004neutronhit.out"
    assert lines[0] == " ".join(DEFAULT_HITLOGVARS) + " " and lines[1].startswith("4 7 0.1 0 0 1 ")


def test_log_truncated_on_first_write_of_a_run(tmp_path):
    """TTextLogger::DoLog (src/logger.cpp:316-329) truncates a log when it first opens it in a
    run and appends afterwards: a second write in the same run adds rows after one header, a
    new run (the process's set of opened logs cleared) starts the file over."""
    from pentrack_amd import logger
    v = {"a": np.array([1.0, 2.0]), "b": np.array([3.0, 4.0])}
    p = logger.write_log(str(tmp_path), 1, "neutron", "end", ["a", "b"], v)
    logger.write_log(str(tmp_path), 1, "neutron", "end", ["a", "b"], v)
    assert open(p).read().split("\n")[:-1] == ["a b ", "1 3 ", "2 4 ", "1 3 ", "2 4 "]
    logger._opened.clear()
    logger.write_log(str(tmp_path), 1, "neutron", "end", ["a", "b"], {"a": np.array([5.0]), "b": np.array([6.0])})
    assert open(p).read().split("\n")[:-1] == ["a b ", "5 6 "]


def test_rotate_vectors_match_oracle(oracle_lib):
    """RotateVector without a reference direction (src/globals.cpp:40-64), as the surface
    sources use it: CGAL's base1 cases (zero components) and general normals, bit for bit."""
    from pentrack_amd.source import rotate_vectors
    rng = np.random.default_rng(4)
    n = np.concatenate([rng.normal(size=(200, 3)), [[0, 0, 1], [0, 1, 0], [1, 0, 0], [0, 2, -1], [3, 0, 1],
                                                    [1, -1, 0], [-1e-3, 1, 1e3]]])
    v = rng.normal(size=(len(n), 3))
    got = rotate_vectors(v, n)
    from oracle import ptref
    L = ptref.load("fp64")
    for k in range(len(n)):
        w = np.ascontiguousarray(v[k].copy())
        L.ptref_rotate_vector(ptref._dp(w), ptref._dp(np.ascontiguousarray(n[k])), None)
        assert np.array_equal(got[k], w), k


def test_source_modes_and_parameters(tmp_path):
    from pentrack_amd.source import SurfaceSource, VolumeSource, create_source
    base = CONFIG.split("[SOURCE]")[0]
    def cfg_with(src):
        return C.TConfig(text=base + "[SOURCE]\nparticle neutron\nEmin 1e-7\nEmax 2e-7\n" + src)
    assert isinstance(create_source(cfg_with("sourcemode cylvolume\nparameters 0 1 0 360 0 1\n")), VolumeSource)
    with pytest.raises(C.ConfigError, match="Could not load source"):
        create_source(cfg_with("sourcemode pointsource\n"))
    with pytest.raises(C.ConfigError, match="needs 6 parameters"):
        create_source(cfg_with("sourcemode boxvolume\nparameters 0 1 0 1\n"))
    with pytest.raises(C.ConfigError, match="needs an STLfile"):
        create_source(cfg_with("sourcemode STLvolume\n"))
    s = create_source(cfg_with("sourcemode cylsurface\nparameters 0.1 0.5 -90 90 0 1\nEnormal 5e-8\n"))
    assert isinstance(s, SurfaceSource) and s.Enormal == 5e-8 and np.allclose(s.cyl[2:4], [-np.pi / 2, np.pi / 2])
    write_stl(str(tmp_path / "src.STL"), box_mesh((0, 0, 0), (1, 2, 3)), "src")
    s = create_source(cfg_with("sourcemode STLvolume\nSTLfile src.STL\n"), str(tmp_path))
    assert np.array_equal(s.bbox[0], [0, 0, 0]) and np.array_equal(s.bbox[1], [1, 2, 3])


def test_spin_tracking_settings():
    cfg = C.TConfig(text=CONFIG)
    cfg["neutron"]["spintimes"] = " 500 700 800 900"
    assert C.species_config(cfg, "neutron").spintimes == (500.0, 700.0, 800.0, 900.0)
    cfg["neutron"]["spintimes"] = " 500"
    with pytest.raises(C.ConfigError, match="pairs"):
        C.species_config(cfg, "neutron")
    cfg["neutron"]["spintimes"] = " 0 1"
    cfg["neutron"]["interpolatefields"] = " 1"
    assert C.species_config(cfg, "neutron").interpolatefields == 1
    cfg["neutron"]["spinlog"] = " 1"
    assert C.species_config(cfg, "neutron").spintimes == (0.0, 1.0)
