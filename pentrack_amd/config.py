"""config.in front end: the reference's configuration file and what PENTrack builds from it.

TConfig (src/config.cpp:14-53): sections `[NAME]`, one `key value...` per line, `#`
starts a comment, a repeated key in one section is an error; sections and keys are
held in std::map, so they are iterated in lexicographic (string) order -- the order
in which the reference loads solids (GEOMETRY) and sums fields (FIELDS).  From it:

  * materials (operator>>(istream, material), src/geometry.cpp:10-52) and solids with
    their STL meshes and ignore times (src/geometry.cpp:60-137);
  * field containers of the FIELDS section (TFieldManager, src/fields.cpp:19-136) with
    scale formulas resolved through FORMULAS (ResolveFormula), table files (OPERA2D,
    2Dtable, OPERA3D, 3Dtable, COMSOL) read by pentrack_amd.tables relative to the config
    file;
  * per-species settings: the PARTICLES defaults merged into each species' section
    (src/main.cpp:242-249);
  * the GLOBAL settings (simtype, simcount, simtime, secondaries).
"""
import math
import os
import re
from dataclasses import dataclass
from typing import Dict

from . import _native
from .fields import FieldContainer, TFieldManager, analytic_field, conductor, uniform_efield
from .geometry import Material, Solid, TGeometry, read_stl
from .particle import SpeciesConfig
from .constants import ELECTRON, NEUTRON, PROTON

SPECIES_BY_NAME = {"neutron": NEUTRON, "proton": PROTON, "electron": ELECTRON}


class ConfigError(RuntimeError):
    pass


class TConfig:
    """The sections of a config file: {section: {key: rest of the line}}."""

    def __init__(self, path=None, text=None):
        self.path = path
        self.sections: Dict[str, Dict[str, str]] = {}
        if path is not None:
            self.read(path)
        elif text is not None:
            self.parse(text)

    def read(self, path):
        if not os.path.exists(path):
            raise ConfigError(f"Could not open {path}!")
        with open(path, encoding="latin-1") as f:
            self.parse(f.read())

    def parse(self, text):
        """TConfig::ReadFromFile: a '[' line opens a section (`[/...]` closes it), '#'
        lines are comments, other lines in a section are `key rest`, the rest cut at '#'."""
        section = ""
        for raw in text.splitlines():
            line = raw.lstrip()
            if not line:
                continue
            if line[0] == "[":
                if line[1:2] == "/":
                    section = ""
                else:
                    end = line.find("]", 1)
                    section = line[1:end] if end >= 0 else line[1:]
                    self.sections.setdefault(section, {})
                continue
            if line[0] == "#" or section == "":
                continue
            parts = line.split(None, 1)
            key = parts[0]
            rest = parts[1] if len(parts) > 1 else ""
            rest = " " + rest if rest else rest
            h = rest.find("#")
            if h >= 0:
                rest = rest[:h]
            sec = self.sections.setdefault(section, {})
            if key in sec:
                raise ConfigError(f"Variable {key} in config file is ambiguous!")
            sec[key] = rest

    def __getitem__(self, section):
        return self.sections.setdefault(section, {})

    def items(self, section):
        """Key-value pairs of a section in std::map (lexicographic) order."""
        return sorted(self[section].items(), key=lambda kv: kv[0])


def _float(s):
    """istream >> double on the start of s."""
    m = re.match(r"\s*([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?|[-+]?inf|[-+]?nan)", s)
    if not m:
        raise ConfigError(f"could not read a number from '{s}'")
    return float(m.group(1)), s[m.end():]


def _floats(s):
    out = []
    while s.strip():
        try:
            v, s = _float(s)
        except ConfigError:
            break
        out.append(v)
    return out


def resolve_formula(name, formulas):
    """ResolveFormula (src/fields.cpp): a name defined in FORMULAS stands for its formula."""
    name = name.strip()
    return formulas[name].strip() if name in formulas else name


def load_materials(cfg):
    mats = []
    for name, rest in cfg.items("MATERIALS"):
        vals = _floats(rest)
        if len(vals) < 6:
            raise ConfigError(f"Could not read material {name}!")
        vals = vals + [0.0] * (9 - len(vals))
        mats.append(Material(name, *vals[:9]))
    return mats


def load_geometry(cfg, base_dir):
    """Materials (from the GLOBAL materials_file, relative to the config file, when given),
    the default solid (ID 1) and the meshed solids (src/geometry.cpp:93-137)."""
    matfile = cfg["GLOBAL"].get("materials_file", "").split()
    if matfile:
        cfg.read(os.path.join(base_dir, matfile[0]))
    mats = load_materials(cfg)
    default_mat = None
    solids = []
    for key, rest in cfg.items("GEOMETRY"):
        ID = int(key)
        parts = rest.split()
        if len(parts) < 2:
            raise ConfigError(f"Could not load solid with ID {ID}! Did you define invalid parameters?")
        fname, mat = parts[0], parts[1]
        ign = []
        for tok in parts[2:]:
            m = re.fullmatch(r"([-+0-9.eE]+)-([-+0-9.eE]+)", tok)
            if not m:
                raise ConfigError(f"Invalid ignoretimes for solid {ID}")
            ign.append((float(m.group(1)), float(m.group(2))))
        if ID == 1:
            default_mat = mat
            continue
        name, tris = read_stl(os.path.join(base_dir, fname))
        solids.append(Solid(ID, mat, tris, ign, name))
    if default_mat is None:
        raise ConfigError("no default solid (ID 1) defined in GEOMETRY")
    return TGeometry(mats, default_mat, solids)


def _need(vals, n, what):
    if len(vals) < n:
        raise ConfigError(f"Could not load field {what}! Check config file for invalid field type or parameters.")
    return vals


def load_fields(cfg, base_dir=None):
    """The FIELDS section in key order (TFieldManager::TFieldManager, src/fields.cpp:19-136).
    Table files are found relative to base_dir (default: the config file's folder)."""
    from . import tables
    if base_dir is None:
        base_dir = os.path.dirname(os.path.abspath(cfg.path)) if cfg.path else os.getcwd()
    formulas = cfg["FORMULAS"]
    out = []
    for key, rest in cfg.items("FIELDS"):
        parts = rest.split()
        if not parts:
            continue
        ftype, args = parts[0], parts[1:]
        if ftype == "Conductor":
            v = _need(args, 8, ftype)
            I, p = float(v[0]), [float(x) for x in v[1:7]]
            out.append(conductor(I, p[0:3], p[3:6], bscale=resolve_formula(v[7], formulas)))
        elif ftype == "EDMStaticEField":
            v = _need(args, 4, ftype)
            out.append(uniform_efield(float(v[0]), float(v[1]), float(v[2]), escale=resolve_formula(v[3], formulas)))
        elif ftype in ("ExponentialFieldX", "LinearFieldZ", "B0GradZ", "B0GradX2", "B0GradXY", "B0_XY"):
            npar = {"ExponentialFieldX": 5, "LinearFieldZ": 2, "B0GradZ": 3, "B0GradX2": 4, "B0GradXY": 3,
                    "B0_XY": 2}[ftype]
            v = _need(args, npar + 7, ftype)
            par = [float(x) for x in v[:npar]]
            box = tuple(float(x) for x in v[npar:npar + 6])
            out.append(analytic_field(ftype, par, box=box, bscale=resolve_formula(v[npar + 6], formulas)))
        elif ftype in ("OPERA3D", "3Dtable"):
            out.append(_table(tables.read_opera3d, ftype, args, formulas, base_dir))
        elif ftype == "COMSOL":
            out.append(_table(tables.read_comsol, ftype, args, formulas, base_dir))
        elif ftype in ("OPERA2D", "2Dtable"):
            out.append(_table(tables.read_opera2d, ftype, args, formulas, base_dir))
        else:
            raise ConfigError(f"Could not load field {ftype}: type not supported by this engine")
    return TFieldManager(out)


def _table(reader, ftype, args, formulas, base_dir):
    from .tables import TableError
    try:
        return reader(ftype, args, formulas, base_dir)
    except TableError as e:
        raise ConfigError(str(e)) from None


def _num(section, key, default, conv=float):
    if key not in section or not section[key].strip():
        return default
    return conv(section[key].split()[0])


def particle_settings(cfg, name):
    """The species' section merged with the PARTICLES defaults (src/main.cpp:242-249)."""
    merged = dict(cfg["PARTICLES"])
    merged.update(cfg[name])
    return merged


def species_config(cfg, name):
    """The species' tracking settings, with the BMT spin-tracking windows `spintimes`
    (read as the reference reads them, pairs of times, src/tracking.cpp:51-60) and the
    interpolation of the precession axis; the spin log is read by the command line."""
    s = particle_settings(cfg, name)
    st = _floats(s.get("spintimes", ""))
    if len(st) % 2 or len(st) > 8:
        raise ConfigError(f"{name}: spintimes needs pairs of times (at most four windows)")
    return SpeciesConfig(kind=SPECIES_BY_NAME[name], tau=_num(s, "tau", 0.0), tmax=_num(s, "tmax", 9e99),
                         lmax=_num(s, "lmax", 9e99), flipspin=_num(s, "flipspin", 0, lambda x: int(float(x))),
                         Bmax=_num(s, "Bmax", 0.0), spintimes=tuple(st),
                         interpolatefields=_num(s, "interpolatefields", 0, lambda x: int(float(x))))


@dataclass
class GlobalSettings:
    simtype: int = 1
    simcount: int = 1
    simtime: float = 9e99
    secondaries: int = 1


def global_settings(cfg):
    g = cfg["GLOBAL"]
    return GlobalSettings(simtype=_num(g, "simtype", 1, lambda x: int(float(x))),
                          simcount=_num(g, "simcount", 1, lambda x: int(float(x))),
                          simtime=_num(g, "simtime", 9e99), secondaries=_num(g, "secondaries", 1, lambda x: int(float(x))))


def deg(v):
    """conv = pi / 180 (src/globals.cpp:19)."""
    return v * (math.pi / 180.0)
