"""PENTrack-style command line: python -m pentrack_amd [jobnumber [configpath [outpath [seed]]]]

Mirrors the reference's `PENTrack jobnumber configpath outpath seed` for simtype 1
(particle simulation, src/main.cpp:52-200): reads config.in, builds geometry, fields and
the source, tracks `simcount` particles on the GPU (and their decay products when
`secondaries 1`), prints the stopID histogram per species and writes the text end logs
of species whose `endlog` is 1.
"""
import os
import sys
import time

import numpy as np


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    jobnumber = int(argv[0]) if len(argv) > 0 else 0
    configpath = os.path.abspath(argv[1] if len(argv) > 1 else os.path.join("in", "config.in"))
    outpath = os.path.abspath(argv[2] if len(argv) > 2 else "out")
    seed = int(argv[3]) if len(argv) > 3 else 0
    from . import config as cfgmod
    from .logger import DEFAULT_ENDLOGVARS, end_values, write_log
    from .source import VolumeSource
    from .tracker import LaunchConfig, TTracker, secondary_states

    cfg = cfgmod.TConfig(configpath)
    g = cfgmod.global_settings(cfg)
    if g.simtype != 1:
        raise SystemExit(f"simtype {g.simtype} is not supported (particle simulation, simtype 1, only)")
    if seed == 0:
        seed = time.time_ns() & ((1 << 63) - 1)
    print("Random Seed:", seed)
    base = os.path.dirname(configpath)
    geometry = cfgmod.load_geometry(cfg, base)
    fields = cfgmod.load_fields(cfg, base)
    source = VolumeSource(cfg)
    species = [cfgmod.species_config(cfg, name) for name in ("neutron", "proton", "electron")]
    rng = np.random.default_rng(seed)
    n = g.simcount
    t0, y0, spin0 = source.sample(n, rng)
    kind_idx = np.full(n, source.kind, np.int32)
    pid = np.arange(1, n + 1, dtype=np.uint64)
    print(f"Simulating {n} {source.particle}s...")
    tr = TTracker(geometry, fields, LaunchConfig())
    if g.secondaries == 1:
        (out, _), (sout, _), rec = tr.IntegrateWithSecondaries(species, g.simtime, kind_idx, t0, y0, spin0, seed, pid)
    else:
        out, _ = tr.IntegrateBatch(species, g.simtime, kind_idx, t0, y0, spin0, seed, pid)
        sout, rec = np.zeros(0, out.dtype), None
    names = {0: "neutron", 1: "proton", 2: "electron"}
    groups = [(source.kind, out, t0, y0, spin0, pid.astype(float))]
    if rec is not None and len(rec):
        _, sidx, st0, sy0, sspin0, _ = secondary_states(species, rec, pid)
        for k, kind in ((1, 1), (2, 2)):
            sel = rec["kind"] == k
            groups.append((kind, sout[sel], st0[sel], sy0[sel], sspin0[sel], (rec["parent"][sel] + 1).astype(float)))
    for kind, o, a, b, c, numbers in groups:
        if len(o) == 0:
            continue
        ids, counts = np.unique(o["stopID"], return_counts=True)
        print(f"{names[kind]}: " + ", ".join(f"{i}: {c_}" for i, c_ in zip(ids, counts)))
        settings = cfgmod.particle_settings(cfg, names[kind])
        if settings.get("endlog", "").split()[:1] == ["1"]:
            titles = settings.get("endlogvars", "").split() or DEFAULT_ENDLOGVARS
            vals = end_values(kind, a, b, c, o, geometry, fields, jobnumber)
            vals["particle"] = numbers
            missing = [t for t in titles if t not in vals]
            if missing:
                raise SystemExit(f"endlogvars not supported: {missing}")
            path = write_log(outpath, jobnumber, names[kind], "end", titles, vals)
            print("wrote", path)
    steps = int(out["Nstep"].sum() + (sout["Nstep"].sum() if len(sout) else 0))
    print(f"{steps} steps")
    return 0


if __name__ == "__main__":
    sys.exit(main())
