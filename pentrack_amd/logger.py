"""Text end, snapshot, track and hit logs (TTextLogger, src/logger.cpp:42-232, 314-334).

One file per species and log type, named `%012d<particle><suffix>.out` after the job
number, a header line of the variable names, then one line per particle (end log) or
per wall hit (hit log, particles in order, hits in time order) with the values printed
at std::numeric_limits<double>::digits10 (15) significant digits, each followed by a
space.  The variables are those of TLogger::Print and TLogger::PrintHit; the start and
end fields, potentials and solids are evaluated on the GPU (pt_fields_eval,
pt_geometry_solid_at) at the particles' start and end states, the hits come from the
tracker's hit records.
"""
import os

import numpy as np

from .constants import ELE_E, GRAVCONST, SPECIES_CONSTANTS, NEUTRON
from .particle import kinetic_energy

DEFAULT_ENDLOGVARS = ["jobnumber", "particle", "tstart", "xstart", "ystart", "zstart", "vxstart", "vystart",
                      "vzstart", "polstart", "Sxstart", "Systart", "Szstart", "Hstart", "Estart", "Bstart", "Ustart",
                      "solidstart", "tend", "xend", "yend", "zend", "vxend", "vyend", "vzend", "polend", "Sxend",
                      "Syend", "Szend", "Hend", "Eend", "Bend", "Uend", "solidend", "stopID", "Nspinflip",
                      "spinflipprob", "Nhit", "Nstep", "propert", "trajlength", "Hmax", "wL"]


DEFAULT_TRACKLOGVARS = ["jobnumber", "particle", "polarisation", "t", "x", "y", "z", "vx", "vy", "vz", "H", "E",
                        "Bx", "dBxdx", "dBxdy", "dBxdz", "By", "dBydx", "dBydy", "dBydz", "Bz", "dBzdx", "dBzdy",
                        "dBzdz", "Ex", "Ey", "Ez", "V"]
DEFAULT_SPINLOGVARS = ["jobnumber", "particle", "t", "x", "y", "z", "Sx", "Sy", "Sz", "Wx", "Wy", "Wz", "Bx", "By", "Bz"]
DEFAULT_HITLOGVARS = ["jobnumber", "particle", "t", "x", "y", "z", "v1x", "v1y", "v1z", "pol1", "v2x", "v2y", "v2z",
                      "pol2", "nx", "ny", "nz", "solid1", "solid2"]


def _fmt(v):
    """ostream << double with setprecision(15) (the %g style, 15 significant digits)."""
    s = "%.15g" % v
    if s in ("inf", "-inf", "nan", "-nan"):
        return s
    return s


def potential_energy(kind, y, B, V, mats, solid_mat):
    """TParticle::GetPotentialEnergy (src/particle.cpp:254-269), plus the material potential
    of neutrons (TNeutron::GetPotentialEnergy, src/neutron.cpp:357-359, include/neutron.h:53-55)."""
    q, m, mu, _ = SPECIES_CONSTANTS[kind]
    res = np.zeros(len(y))
    if mu != 0:
        res += -y[:, 7] * mu / ELE_E * np.sqrt((B ** 2).sum(1))
    if q != 0:
        res += q / ELE_E * V
    res += m * GRAVCONST * y[:, 2]
    if kind == NEUTRON:
        mat = mats[solid_mat]
        res += mat[:, 0] * 1e-9 - mat[:, 6] * mu * y[:, 7] / ELE_E
    return res


def print_values(kind, t0, y0, spin0, t, y, spin, rec, geometry, fields, jobnumber=0, first_number=1):
    """{variable: [n] values} of TLogger::Print (src/logger.cpp:42-126) for particles that
    started at (t0, y0, spin0) and are now at (t, y, spin) with the counters of `rec`
    (stopID, Nspinflip, noflipprob, Nhit, Nstep, Hmax).  The solid at the end point is
    TGeometry::GetSolid there, as Print evaluates it."""
    from .queries import fields_eval, solid_at
    n = len(t)
    Bs, _, Vs, _ = fields_eval(fields, y0[:, :3], t0, grad=False)      # BField(x, y, z, t, B) (src/logger.cpp:55,63)
    Be, _, Ve, _ = fields_eval(fields, y[:, :3], t, grad=False)
    sid_s = solid_at(geometry, y0[:, :3], t0)
    sid_e = solid_at(geometry, y[:, :3], t)
    mats = np.stack([m.as_array() for m in geometry.materials.values()])
    names = list(geometry.materials)
    id_to_mat = {1: names.index(geometry.default_material)}
    for s in geometry.solids:
        id_to_mat[s.ID] = names.index(s.material)
    ms = np.array([id_to_mat.get(int(i), id_to_mat[1]) for i in sid_s], dtype=int)
    me = np.array([id_to_mat.get(int(i), id_to_mat[1]) for i in sid_e], dtype=int)
    Es = kinetic_energy(kind, y0[:, 3:6])
    Ee = kinetic_energy(kind, y[:, 3:6])
    Hs = Es + potential_energy(kind, y0, Bs, Vs, mats, ms)
    He = Ee + potential_energy(kind, y, Be, Ve, mats, me)
    wL = np.where(spin[:, 3] > 0, spin[:, 4] / np.where(spin[:, 3] > 0, spin[:, 3], 1.0), 0.0)
    v = {
        "jobnumber": np.full(n, float(jobnumber)), "particle": np.arange(first_number, first_number + n, dtype=float),
        "tstart": t0, "xstart": y0[:, 0], "ystart": y0[:, 1], "zstart": y0[:, 2], "vxstart": y0[:, 3],
        "vystart": y0[:, 4], "vzstart": y0[:, 5], "polstart": y0[:, 7], "Sxstart": spin0[:, 0],
        "Systart": spin0[:, 1], "Szstart": spin0[:, 2], "Hstart": Hs, "Estart": Es,
        "Bstart": np.sqrt((Bs ** 2).sum(1)), "Ustart": Vs, "solidstart": sid_s.astype(float),
        "tend": t, "xend": y[:, 0], "yend": y[:, 1], "zend": y[:, 2], "vxend": y[:, 3],
        "vyend": y[:, 4], "vzend": y[:, 5], "polend": y[:, 7], "Sxend": spin[:, 0], "Syend": spin[:, 1],
        "Szend": spin[:, 2], "Hend": He, "Eend": Ee, "Bend": np.sqrt((Be ** 2).sum(1)), "Uend": Ve,
        "solidend": sid_e.astype(float), "stopID": rec["stopID"].astype(float),
        "Nspinflip": rec["Nspinflip"].astype(float), "spinflipprob": 1 - rec["noflipprob"],
        "Nhit": rec["Nhit"].astype(float), "Nstep": rec["Nstep"].astype(float), "propert": y[:, 6],
        "trajlength": y[:, 8], "Hmax": rec["Hmax"], "wL": wL,
    }
    return v


def end_values(kind, t0, y0, spin0, out, geometry, fields, jobnumber=0, first_number=1):
    """{variable: [n] values} of TLogger::Print for the particles' end states."""
    return print_values(kind, t0, y0, spin0, out["t"], out["y"], out["spin"], out, geometry, fields, jobnumber,
                        first_number)


def snapshot_values(kind, t0, y0, spin0, snaps, counts, numbers, geometry, fields, jobnumber=0):
    """TLogger::Print of the snapshot records snaps [n][cap] (counts[k] taken by particle k,
    at most cap kept): one row per record, particles in order, times ascending."""
    cap = snaps.shape[1]
    k = np.minimum(np.asarray(counts), cap)
    idx = np.repeat(np.arange(len(k)), k)
    rows = np.concatenate([snaps[i, :k[i]] for i in range(len(k))]) if len(idx) else snaps.reshape(-1)[:0]
    v = print_values(kind, t0[idx], y0[idx], spin0[idx], rows["t"], rows["y"], rows["spin"], rows, geometry, fields,
                     jobnumber)
    v["particle"] = np.asarray(numbers, dtype=float)[idx]
    return v


# log files this process has opened: TTextLogger::DoLog (src/logger.cpp:316-329) truncates
# a log file when it first opens it in a run and appends to it afterwards
_opened = set()


def write_log(outdir, jobnumber, particle, suffix, titles, values):
    """Write lines to `<outdir>/%012d<particle><suffix>.out`: the first write of this run
    truncates the file and writes the header, later writes append."""
    os.makedirs(outdir, exist_ok=True)
    path = os.path.join(outdir, "%012d%s%s.out" % (jobnumber, particle, suffix))
    key = os.path.abspath(path)
    new = key not in _opened
    _opened.add(key)
    cols = [np.asarray(values[t], dtype=np.float64) for t in titles]
    with open(path, "w" if new else "a") as f:
        if new:
            f.write(" ".join(titles) + " \n")
        for row in zip(*cols):
            f.write(" ".join(_fmt(v) for v in row) + " \n")
    return path


def hit_values(hits, n_hits, numbers, jobnumber=0):
    """{variable: values} of TLogger::PrintHit (src/logger.cpp:198-232) for the logged hits
    hits [n][cap] (n_hits[k] recorded for particle k, at most cap kept)."""
    cap = hits.shape[1] if hits.ndim == 2 else 0
    k = np.minimum(np.asarray(n_hits), cap)
    rows = np.concatenate([hits[i, :k[i]] for i in range(len(k))]) if cap else hits[:0]
    part = np.repeat(np.asarray(numbers, dtype=float), k)
    v = {"jobnumber": np.full(len(rows), float(jobnumber)), "particle": part, "t": rows["t"],
         "x": rows["pos"][:, 0], "y": rows["pos"][:, 1], "z": rows["pos"][:, 2],
         "v1x": rows["v1"][:, 0], "v1y": rows["v1"][:, 1], "v1z": rows["v1"][:, 2], "pol1": rows["pol1"],
         "v2x": rows["v2"][:, 0], "v2y": rows["v2"][:, 1], "v2z": rows["v2"][:, 2], "pol2": rows["pol2"],
         "nx": rows["normal"][:, 0], "ny": rows["normal"][:, 1], "nz": rows["normal"][:, 2],
         "solid1": rows["leaving"].astype(float), "solid2": rows["entering"].astype(float)}
    return v


def spin_values(recs, counts, numbers, fields, jobnumber=0):
    """TLogger::PrintSpin variables (src/logger.cpp:250-281), particles in order: the recorded
    time, position, spin and precession axis, and B there (value-only, as field.BField)."""
    from .queries import fields_eval
    cap = recs.shape[1]
    k = np.minimum(np.asarray(counts), cap)
    r = np.concatenate([recs[i, :k[i]] for i in range(len(k))]) if len(k) else recs.reshape(-1)[:0]
    part = np.repeat(np.asarray(numbers, dtype=float), k)
    B, _, _, _ = fields_eval(fields, np.ascontiguousarray(r["pos"]), np.ascontiguousarray(r["t"]), grad=False)
    v = {"jobnumber": np.full(len(r), float(jobnumber)), "particle": part, "t": r["t"]}
    for a, ax in enumerate("xyz"):
        v[ax] = r["pos"][:, a]
        v["S" + ax] = r["spin"][:, a]
        v["W" + ax] = r["W"][:, a]
        v["B" + ax] = B[:, a]
    return v


def track_values(kind, t0, y0, sid0, recs, counts, numbers, geometry, fields, jobnumber=0):
    """TLogger::PrintTrack variables (src/logger.cpp:142-196): the start point (logged when
    its trajectory length is 0, src/tracking.cpp:43) and the recorded step ends, particles in
    order.  H uses the potential energy in the particle's current solid."""
    from .queries import fields_eval
    cap = recs.shape[1]
    k = np.minimum(np.asarray(counts), cap)
    rows_t, rows_y, rows_s, part = [], [], [], []
    numbers = np.asarray(numbers, dtype=float)
    for i in range(len(k)):
        if not y0[i, 8] > 0:
            rows_t.append([t0[i]]); rows_y.append(y0[i:i + 1]); rows_s.append([sid0[i]]); part.append([numbers[i]])
        r = recs[i, :k[i]]
        rows_t.append(r["t"]); rows_y.append(r["y"]); rows_s.append(r["solid"]); part.append(np.full(k[i], numbers[i]))
    t = np.concatenate(rows_t) if rows_t else np.zeros(0)
    y = np.concatenate(rows_y) if rows_y else np.zeros((0, 9))
    sid = np.concatenate(rows_s).astype(int) if rows_s else np.zeros(0, int)
    B, dB, V, E = fields_eval(fields, y[:, :3], t, grad=True)
    Bv, _, _, _ = fields_eval(fields, y[:, :3], t, grad=False)
    mats = np.stack([m.as_array() for m in geometry.materials.values()])
    names = list(geometry.materials)
    id_to_mat = {1: names.index(geometry.default_material)}
    for s in geometry.solids:
        id_to_mat[s.ID] = names.index(s.material)
    ms = np.array([id_to_mat.get(int(i), id_to_mat[1]) for i in sid], dtype=int)
    Ek = kinetic_energy(kind, y[:, 3:6])
    H = Ek + potential_energy(kind, y, Bv, V, mats, ms)
    v = {"jobnumber": np.full(len(t), float(jobnumber)), "particle": np.concatenate(part) if part else np.zeros(0),
         "polarisation": y[:, 7], "t": t, "x": y[:, 0], "y": y[:, 1], "z": y[:, 2], "vx": y[:, 3], "vy": y[:, 4],
         "vz": y[:, 5], "H": H, "E": Ek, "Ex": E[:, 0], "Ey": E[:, 1], "Ez": E[:, 2], "V": V}
    for a, ax in enumerate("xyz"):
        v["B" + ax] = B[:, a]
        for b, bx in enumerate("xyz"):
            v[f"dB{ax}d{bx}"] = dB[:, a, b]
    return v
