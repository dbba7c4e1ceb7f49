"""Readers of tabulated field maps: the OPERA3D (and older 3Dtable) and COMSOL text
formats of 3D maps (src/field_3d.cpp:124-310) and the OPERA2D (and older 2Dtable) format
of axisymmetric r-z maps (src/field_2d.cpp:38-196), plain or compressed with gzip (`.gz`)
or bzip2 (`.bz2`), chosen by the file name's extension as the reference does.

The 3D readers build the TabField3 container the reference builds (src/field_3d.cpp:
417-470): the grid is the sorted list of distinct x, y and z coordinates, each sample goes
to the grid point its coordinates name (lower_bound), grid points without a sample keep
zero field, and the container's bounding box is the extent of the samples with the
boundary width of the config line.  Field components missing from an OPERA3D file are
tabulated as zeros: the reference leaves such a component at the zero the field sum
starts from, which gives the same values and derivatives.  The 2D reader builds a
TabField container (pentrack_amd.fields.tabfield2) without bounds box.
"""
import bz2
import gzip
import os
import re

import numpy as np

from .fields import tabfield2, tabfield3

_NUM = re.compile(r"\s*([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)")
_INT = re.compile(r"\s*([-+]?\d+)")


class TableError(RuntimeError):
    pass


def open_table(path):
    """The file as text, through a gzip or bzip2 decompressor for `.gz` / `.bz2`."""
    if not os.path.exists(path):
        raise TableError("Could not open " + path)
    ext = os.path.splitext(path)[1]
    opener = {".bz2": bz2.open, ".gz": gzip.open}.get(ext, open)
    with opener(path, "rt", encoding="latin-1", newline="") as f:
        return f.read()


def _stod(tok):
    """std::stod: the longest leading decimal number of tok (leading blanks skipped)."""
    try:
        if "_" not in tok:
            return float(tok)
    except ValueError:
        pass
    m = _NUM.match(tok)
    if not m:
        raise ValueError(tok)
    return float(m.group(1))


def _to_doubles(tokens):
    """Whitespace-separated numbers as istream >> double reads them, up to the first token
    that is not a number."""
    try:
        vals = np.array(tokens, dtype=np.float64)
        if np.isfinite(vals).all():
            return vals
    except ValueError:
        pass
    out = []
    for tok in tokens:
        m = _NUM.fullmatch(tok)
        if not m:
            break
        out.append(float(m.group(1)))
    return np.array(out, dtype=np.float64)


def _params(args, n, fieldtype):
    """`ss >> ...` over the config line's tokens after the field type."""
    if len(args) < n:
        raise TableError(f"Could not read all required parameters for field {fieldtype}!")
    return args


def _number(tok, fieldtype):
    m = _NUM.match(tok)
    if not m or m.end() != len(tok):
        raise TableError(f"Could not read all required parameters for field {fieldtype}!")
    return float(m.group(1))


def _container(x, y, z, comps, V, bscale, escale, bw):
    """TabField3 of scattered samples and its TFieldContainer (bounds = sample extent)."""
    grid = [np.unique(c) for c in (x, y, z)]
    shape = tuple(len(g) for g in grid)
    idx = tuple(np.searchsorted(g, c, side="left") for g, c in zip(grid, (x, y, z)))
    B = None
    if any(c is not None for c in comps):
        B = np.zeros((3,) + shape)
        for i, c in enumerate(comps):
            if c is not None:
                B[i][idx] = c
    Vt = None
    if V is not None:
        Vt = np.zeros(shape)
        Vt[idx] = V
    if B is None and Vt is None:
        raise TableError("The table holds no field components")
    if min(shape) < 2:
        raise TableError("A field table needs at least two distinct coordinates in x, y and z")
    box = (x.max(), x.min(), y.max(), y.min(), z.max(), z.min())
    return tabfield3(grid[0], grid[1], grid[2], B=B, V=Vt, bscale=bscale, escale=escale, box=box, bw=bw)


def read_comsol(fieldtype, args, formulas, base_dir):
    """ReadComsolField (src/field_3d.cpp:124-183): `COMSOL file Bscale BoundaryWidth
    lengthconv`; lines `x y z Bx By Bz` split at tabs, commas and blanks, `%` and `#`
    lines skipped; E scale "0"."""
    from .config import resolve_formula
    args = _params(args, 4, fieldtype)
    fname, bscale = args[0], resolve_formula(args[1], formulas)
    bw, lconv = _number(args[2], fieldtype), _number(args[3], fieldtype)
    path = os.path.join(base_dir, fname)
    text = open_table(path)
    rows = []
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    for num, line in enumerate(lines, 1):
        if line[:1] in ("%", "#"):
            continue
        parts = re.split(r"[\t, ]+", line)
        if len(parts) != 6:
            raise TableError(f"Error reading line {num} of file {path}")
        try:
            rows.append([_stod(p) for p in parts])
        except ValueError:
            raise TableError(f"Error reading line {num} of file {path}") from None
    if not rows:
        raise TableError("No data read from " + path)
    a = np.array(rows, dtype=np.float64)
    x, y, z = a[:, 0] * lconv, a[:, 1] * lconv, a[:, 2] * lconv
    return _container(x, y, z, [a[:, 3], a[:, 4], a[:, 5]], None, bscale, "0", bw)


def read_opera3d(fieldtype, args, formulas, base_dir):
    """ReadOperaField3 (src/field_3d.cpp:185-310).

    `OPERA3D file Bscale Escale BoundaryWidth lengthconv`, or the older `3Dtable file
    Bscale Escale BoundaryWidth` whose file is in cm, Gauss and V (scales multiplied by
    1e-4 and 100, lengths by 0.01).  The file starts with the three grid sizes, then
    column descriptions: a line naming BX, BY, BZ each opens that component's column and a
    line with V the potential's; the description ends with a line beginning " 0".  Then
    x y z and the present columns, as many rows as the sizes multiply to.
    """
    from .config import resolve_formula
    if fieldtype == "3Dtable":
        args = _params(args, 4, fieldtype)
        bscale = "(" + resolve_formula(args[1], formulas) + ")*0.0001"
        escale = "(" + resolve_formula(args[2], formulas) + ")*100"
        lconv = 0.01
    elif fieldtype == "OPERA3D":
        args = _params(args, 5, fieldtype)
        bscale, escale = resolve_formula(args[1], formulas), resolve_formula(args[2], formulas)
        lconv = _number(args[4], fieldtype)
    else:
        raise TableError(f"Tried to load 3D table file for unknown field type {fieldtype}!")
    bw = _number(args[3], fieldtype)
    path = os.path.join(base_dir, args[0])
    text = open_table(path)
    pos = 0
    sizes = []
    for _ in range(3):
        m = _INT.match(text, pos)
        if not m:
            raise TableError(f"{path} not found or corrupt!")
        sizes.append(int(m.group(1)))
        pos = m.end()

    def getline():
        nonlocal pos
        e = text.find("\n", pos)
        if e < 0:
            line, pos = text[pos:], len(text)
        else:
            line, pos = text[pos:e], e + 1
        return line

    for _ in range(5):
        line = getline()
    present = []
    for key in ("BX", "BY", "BZ"):
        if key in line:
            present.append(key)
            line = getline()
    has_v = "V" in line
    if has_v:
        line = getline()
    if line[:2] != " 0":
        raise TableError(f"{path} not found or corrupt!")
    ncol = 3 + len(present) + int(has_v)
    vals = _to_doubles(text[pos:].split())
    n_expected = sizes[0] * sizes[1] * sizes[2]
    n_read = len(vals) // ncol + (1 if len(vals) % ncol >= 3 else 0)
    if n_read != n_expected or len(vals) % ncol:
        raise TableError(f"The header says the size is {n_expected}, actually it is {n_read}! Exiting...")
    a = vals.reshape(n_expected, ncol)
    x, y, z = a[:, 0] * lconv, a[:, 1] * lconv, a[:, 2] * lconv
    comps = [None, None, None]
    col = 3
    for k, key in enumerate(("BX", "BY", "BZ")):
        if key in present:
            comps[k] = a[:, col]
            col += 1
    V = a[:, col] if has_v else None
    return _container(x, y, z, comps, V, bscale, escale, bw)


def read_opera2d(fieldtype, args, formulas, base_dir):
    """ReadOperaField2 and TabField::ReadTabFile (src/field_2d.cpp:38-196).

    `OPERA2D file Bscale Escale lengthconv`, or `2Dtable file Bscale Escale` in cm, Gauss
    and V.  The file starts with `m 1 n` (radial and axial sizes) and column descriptions:
    X, then Z (or Y and Z: a third line " 2 Y [LENGU]" adds a column that is skipped), then
    lines naming RBX, RBY, RBZ, EX, EY, EZ, RV open the Br, Bphi, Bz, Er, Ephi, Ez and V
    columns, closed by a line beginning " 0".  Rows run over z for each r: a z smaller than
    the previous one starts the next r.
    """
    from .config import resolve_formula
    if fieldtype == "2Dtable":
        args = _params(args, 3, fieldtype)
        bscale = "(" + resolve_formula(args[1], formulas) + ")*0.0001"
        escale = "(" + resolve_formula(args[2], formulas) + ")*100"
        lconv = 0.01
    elif fieldtype == "OPERA2D":
        args = _params(args, 4, fieldtype)
        bscale, escale = resolve_formula(args[1], formulas), resolve_formula(args[2], formulas)
        lconv = _number(args[3], fieldtype)
    else:
        raise TableError(f"Tried to load 2D table file for unknown field type {fieldtype}!")
    path = os.path.join(base_dir, args[0])
    text = open_table(path)
    pos = 0
    sizes = []
    for _ in range(3):
        mm = _INT.match(text, pos)
        if not mm:
            raise TableError(f"{path} not found or corrupt!")
        sizes.append(int(mm.group(1)))
        pos = mm.end()
    m, n = sizes[0], sizes[2]

    def getline():
        nonlocal pos
        e = text.find("\n", pos)
        if e < 0:
            line, pos = text[pos:], len(text)
        else:
            line, pos = text[pos:e], e + 1
        return line

    for _ in range(3):
        line = getline()
    with_y = line[:12] == " 2 Y [LENGU]"
    line = getline()
    if with_y:
        line = getline()
    present = []
    for key, name in (("RBX", "Br"), ("RBY", "Bphi"), ("RBZ", "Bz"), ("EX", "Er"), ("EY", "Ephi"), ("EZ", "Ez"),
                      ("RV", "V")):
        if key in line:
            present.append(name)
            line = getline()
    if line[:2] != " 0":
        raise TableError(f"{path} not found or corrupt!")
    ncoord = 3 if with_y else 2
    ncol = ncoord + len(present)
    vals = _to_doubles(text[pos:].split())
    if len(vals) % ncol or not len(vals):
        raise TableError(f"{path}: incomplete rows")
    a = vals.reshape(-1, ncol)
    r = a[:, 0] * lconv
    z = a[:, ncoord - 1] * lconv
    new_r = np.concatenate([[False], z[1:] < z[:-1]])
    ri = np.cumsum(new_r)
    starts = np.flatnonzero(np.concatenate([[True], new_r[1:]]))
    zi = np.arange(len(z)) - starts[ri]
    if ri[-1] + 1 != m or zi[-1] + 1 != n or zi.max() >= n or ri.max() >= m or len(z) != m * n:
        raise TableError(f"The header says the size is {m} by {n}, actually it is {ri[-1] + 1} by {zi[-1] + 1}! "
                         "Exiting...")
    rind = np.empty(m)
    zind = np.empty(n)
    rind[ri] = r                       # the last row of each r block, as the reference assigns
    zind[zi] = z
    cols = {}
    for k, name in enumerate(present):
        tab = np.zeros((n, m))
        tab[zi, ri] = a[:, ncoord + k]
        cols[name] = tab
    if not (np.diff(rind) > 0).all() or not (np.diff(zind) > 0).all():
        raise TableError(f"{path}: r and z must increase through the table")
    return tabfield2(rind, zind, bscale=bscale, escale=escale, **cols)
