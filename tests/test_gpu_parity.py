"""GPU parity: the HIP tracker against the FP64 oracle on the same inputs.

Integer/index results (stop reasons, hit counts, hit-triangle sequences, step
counts) and the FP64 state must agree bit for bit; the mixed-precision variant
is checked against the FP64 oracle within its stated tolerance.
"""
import numpy as np
import pytest

from pentrack_amd import _native, workloads
from pentrack_amd.constants import NEUTRON, PROTON, ELECTRON
from pentrack_amd.fields import TFieldManager, uniform_efield
from pentrack_amd.particle import SpeciesConfig, make_states
from pentrack_amd.tracker import Batch, LaunchConfig
from pentrack_amd.workloads import Workload

pytestmark = pytest.mark.gpu

FIELDS = ("t", "y", "spin", "stopID", "Nhit", "Nstep", "Nspinflip", "n_acc", "n_rej", "Hmax", "noflipprob", "tau",
          "solidend", "n_hits")


HITS = 512      # hit-log capacity: every test's particles hit fewer walls (checked)


def run_gpu(w, hit_cap=HITS, cfg=None, with_mixed=False):
    b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid,
              hit_cap=hit_cap, with_mixed=with_mixed)
    assert b.run_to_completion(cfg or LaunchConfig(), max_launches=100000)
    return b.download()


def assert_bit_exact(g, gh, o, oh):
    """Every output field bit for bit, and the whole hit log of every particle: the logs
    must hold every hit (no particle hit the walls more often than the log capacity)."""
    msgs = []
    for f in FIELDS:
        if not np.array_equal(g[f], o[f]):
            bad = np.where(~np.all((g[f] == o[f]).reshape(len(g), -1), axis=1))[0]
            k = bad[0]
            msgs.append(f"{f} differs for {len(bad)} particles, first {bad[:5]}: gpu {g[f][k]} oracle {o[f][k]}")
    if msgs:
        raise AssertionError("\n".join(msgs))
    if gh is not None:
        cap = gh.shape[1]
        assert oh.shape[1] == cap and (g["n_hits"] <= cap).all(), "hit log shorter than a particle's hits"
        logged = np.arange(cap)[None, :] < g["n_hits"][:, None]
        for f in ("t", "tri", "leaving", "entering", "normal", "v1", "v2", "pol1", "pol2", "pos"):
            a, b = gh[f][logged], oh[f][logged]
            assert np.array_equal(a, b), f"hit log field {f} differs in {int((a != b).reshape(len(a), -1).any(1).sum())} hits"


def test_cube_specular_bit_exact(gpu, oracle_lib):
    w = workloads.cube(n=3000, seed=1, tmax=10.0)
    print("case: cube n=3000 tmax=10 seed=1", flush=True)
    g, gh = run_gpu(w)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS, threads=8)
    assert_bit_exact(g, gh, o, oh)
    assert (g["stopID"] == -1).mean() > 0.99 and g["Nhit"].mean() > 50


def test_sliced_launches_equal_single_launch(gpu):
    w = workloads.cube(n=1000, seed=3, tmax=5.0)
    print("case: cube sliced n=1000", flush=True)
    a, ah = run_gpu(w)
    b, bh = run_gpu(w, cfg=LaunchConfig(steps_per_launch=7))
    assert_bit_exact(a, ah, b, bh)


def test_cylinder_field_map_lambert_bit_exact(gpu, oracle_lib):
    w = workloads.cylinder(n=2000, seed=2, tmax=3.0, n_grid=64)      # the 64^3 FP64 map of the benchmark
    print("case: cylinder n=2000 tmax=3 grid=64", flush=True)
    g, gh = run_gpu(w)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS, threads=8)
    assert_bit_exact(g, gh, o, oh)
    assert g["Nhit"].sum() > 1000


def test_charged_species_uniform_E_and_map(gpu, oracle_lib):
    """Protons and electrons (Lorentz force, relativistic EoM, absorption inside solids)
    in a uniform E plus the coil map, mixed with neutrons with a finite lifetime."""
    base = workloads.cylinder(n=600, seed=4, tmax=0.2, n_grid=24, Bmax=0.5)
    rng = np.random.default_rng(4)
    n = base.n
    kinds = np.where(np.arange(n) % 3 == 0, 1, np.where(np.arange(n) % 3 == 1, 2, 0)).astype(np.int32)
    E = np.where(kinds == 1, rng.uniform(1, 800, n), np.where(kinds == 2, rng.uniform(1e3, 7e5, n), rng.uniform(0, 250e-9, n)))
    phi = rng.uniform(0, 2 * np.pi, n)
    theta = np.arccos(rng.uniform(-1, 1, n))
    y0 = base.y0.copy()
    for k in (0, 1, 2):
        sel = kinds == k
        _, yk, _ = make_states(k, base.y0[sel, :3], E[sel], phi[sel], theta[sel], 1.0)
        y0[sel] = yk
    species = [SpeciesConfig(NEUTRON, tau=0.5, tmax=9e99), SpeciesConfig(PROTON, tmax=3e-3),
               SpeciesConfig(ELECTRON, tmax=1e-5)]
    fm = TFieldManager(base.fields.containers + [uniform_efield(0.0, 0.0, 1e5)])
    w = Workload("mixed-species", base.geometry, fm, species, base.tmax, kinds, base.t0, y0, base.spin0, base.pid, 4)
    print("case: charged species n=600", flush=True)
    g, gh = run_gpu(w)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS, threads=8)
    assert_bit_exact(g, gh, o, oh)
    assert set(np.unique(g["stopID"])) >= {-4, 1}


def test_analytic_fields_and_time_dependent_scales_bit_exact(gpu, oracle_lib):
    """A finite wire, the six analytic B fields, a tapered coil map and a uniform E, with
    scale formulas of t (ramps, steps, a cosine), on neutrons, protons and electrons."""
    w = workloads.analytic(n=600, seed=6, tmax=0.1)
    print("case: analytic fields n=600", flush=True)
    g, gh = run_gpu(w)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS, threads=8)
    assert_bit_exact(g, gh, o, oh)
    print("case: analytic fields n=600 persistent", flush=True)
    p, ph = run_gpu(w, cfg=LaunchConfig(persistent=1, steps_per_launch=16))
    assert_bit_exact(p, ph, o, oh)
    assert (g["stopID"][w.species_idx == 0] == -1).all() and g["n_acc"].sum() > 10000


def test_neutron_decay_products_and_their_tracks_bit_exact(gpu, oracle_lib):
    """Neutrons with a 20 ms lifetime in the analytic-field box: the decay products of the
    GPU batch against the oracle's TNeutron::Decay at the oracle's end states, then the
    products tracked on the GPU (TTracker.IntegrateWithSecondaries) against the oracle."""
    from pentrack_amd.tracker import TTracker, secondary_states
    base = workloads.analytic(n=600, seed=13, tmax=0.1)
    species = [SpeciesConfig(NEUTRON, tau=0.02, tmax=9e99), SpeciesConfig(PROTON, tmax=3e-3),
               SpeciesConfig(ELECTRON, tmax=1e-5)]
    w = Workload("decay", base.geometry, base.fields, species, base.tmax, base.species_idx, base.t0, base.y0,
                 base.spin0, base.pid, 13)
    print("case: decay products n=600", flush=True)
    tr = TTracker(w.geometry, w.fields, LaunchConfig())
    (g, gh), (sg, sgh), rec = tr.IntegrateWithSecondaries(w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0,
                                                           w.seed, w.pid, hit_cap=HITS)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS, threads=8)
    assert_bit_exact(g, gh, o, oh)
    dec = np.where((o["stopID"] == -4) & (w.species_idx == 0))[0]
    assert len(dec) > 50 and len(rec) == 2 * len(dec)
    out6, pol = oracle_lib.neutron_decay(w.seed, w.pid[dec], o["ctr_end"][dec], o["y"][dec, 3:6], "fp64")
    assert np.array_equal(rec["parent"], np.repeat(dec, 2)) and np.array_equal(rec["kind"], np.tile([1, 2], len(dec)))
    p, e = rec[0::2], rec[1::2]
    for r in (p, e):
        assert np.array_equal(r["t"], o["t"][dec]) and np.array_equal(r["pos"], o["y"][dec, :3])
    assert np.array_equal(np.stack([p["E"], p["phi"], p["theta"], e["E"], e["phi"], e["theta"]], 1), out6)
    assert np.array_equal(np.stack([p["pol"], e["pol"]], 1), pol)
    print("case: decay products tracked", flush=True)
    sspecies, sidx, st0, sy0, sspin0, spid = secondary_states(w.species, rec, w.pid)
    sw = Workload("secondaries", w.geometry, w.fields, sspecies, w.tmax, sidx, st0, sy0, sspin0, spid, w.seed)
    so, soh = oracle_lib.track(sw, "fp64", hit_cap=HITS, threads=8)
    assert_bit_exact(sg, sgh, so, soh)
    assert (sg["n_acc"] > 0).mean() > 0.9


def test_edge_cases(gpu, oracle_lib):
    w = workloads.cube(n=8, seed=5, tmax=1.0)
    w.t0[0] = 2.0                  # starts past tmax: the reference throws (invariant stop)
    w.y0[1, :3] = [3.0, 3.0, 3.0]  # outside every mesh bounding box: ID_HIT_BOUNDARIES
    w.y0[2, 2] = 1.0 - 1e-9        # just below the ceiling
    print("case: edge cases", flush=True)
    g, gh = run_gpu(w)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS)
    assert_bit_exact(g, gh, o, oh)
    assert g["stopID"][0] == _native.ID_INVARIANT and g["stopID"][1] == _native.ID_HIT_BOUNDARIES
    # an empty batch is valid
    e = workloads.cube(n=1, seed=5, tmax=1.0)
    b = Batch(e.geometry, e.fields, e.species, e.tmax, e.species_idx[:0], e.t0[:0], e.y0[:0], e.spin0[:0], e.seed,
              e.pid[:0])
    b.run()
    assert b.active() == 0


def test_mixed_precision_config3(gpu, oracle_lib):
    """The mixed-precision workload as specified: 1 m cube, 128^3 coil map (|B| <= 1 T),
    10,000 neutrons from the config-1 source (the benchmark's error leg).  The FP64 kernel
    is bit-exact with the FP64 oracle on it; the mixed-precision kernel (FP64 constant and
    linear terms, FP32 remainder, FP64 blocks where the remainder's bound exceeds the default
    tolerance) keeps every particle's stop reason and full hit-triangle sequence, and its
    trajectory within 1e-6 m of FP64 (end state at 0.25 s and the snapshots both runs
    recorded on a 10 ms grid), for every particle (those that leave the cube through the
    floor on their path up to the exit).
    The dynamics in this field are chaotic: deviations of any origin grow about e-fold every
    0.8 s, so the bound holds over a horizon, not over 50 s."""
    from pentrack_amd import compare
    w = workloads.cube(n=10000, seed=3, tmax=0.25, with_field=True, n_grid=128, Bmax=1.0)
    times = compare.snapshot_times(0.25)
    print(f"case: mixed precision, cube, 128^3 map, n={w.n}, 0.25 s", flush=True)
    cap = 256
    o, oh = oracle_lib.track(w, "fp64", hit_cap=cap, threads=8)
    assert o["n_hits"].max() <= cap and o["Nhit"].sum() > w.n
    runs = {}
    for name, cfg in (("fp64", LaunchConfig()), ("mixed", LaunchConfig(variant=_native.PT_VARIANT_MIXED))):
        b = Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid,
                  hit_cap=cap, with_mixed=(name == "mixed"))
        b.set_snapshots(times, len(times))
        assert b.run_to_completion(cfg, max_launches=100000)
        out, hits = b.download()
        runs[name] = (out, hits) + b.snapshots()
    assert_bit_exact(runs["fp64"][0], runs["fp64"][1], o, oh)
    blocks, fp64 = w.fields.mixed_info()
    assert 0 < fp64 < 0.5 * blocks
    r = compare.deviation(runs["fp64"], runs["mixed"], times, 0.25)
    print(f"mixed vs FP64: {r}, FP64 blocks {fp64}/{blocks}", flush=True)
    assert r["stopped_before"] > 0                    # the escapes through the floor are covered
    assert r["same_stop_reasons"] == 1.0 and r["same_hit_sequences"] == 1.0
    assert r["trajectory_position_max_m"] <= 1e-6 and r["condition_holds"]
    assert not np.array_equal(runs["mixed"][0]["y"], o["y"])      # the mixed arithmetic is in use


def test_mixed_variant_on_tables_kept_in_fp64(gpu, oracle_lib):
    """A map whose mixed-precision tables keep every block in FP64 (the cylinder's coil map at
    1 T): the mixed variant then computes the FP64 arithmetic and is bit-exact with the FP64
    oracle, in the per-particle and the persistent kernel."""
    w = workloads.cylinder(n=1500, seed=8, tmax=2.0, n_grid=24)
    print("case: mixed variant, cylinder n=1500, every block in FP64", flush=True)
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS, threads=8)
    for persistent in (0, 1):
        g, gh = run_gpu(w, cfg=LaunchConfig(variant=_native.PT_VARIANT_MIXED, persistent=persistent, steps_per_launch=9,
                                            grid=64), with_mixed=True)
        blocks, fp64 = w.fields.mixed_info()
        assert blocks > 0 and fp64 == blocks
        assert_bit_exact(g, gh, o, oh)


def test_persistent_work_queue_bit_exact(gpu, oracle_lib):
    """The persistent kernel (slots refilled from the device work list) gives the same
    particles the same arithmetic: bit-exact with the oracle, in sliced calls."""
    w = workloads.cylinder(n=1500, seed=8, tmax=2.0, n_grid=24)
    print("case: persistent cylinder n=1500", flush=True)
    g, gh = run_gpu(w, cfg=LaunchConfig(persistent=1, steps_per_launch=9, grid=64))
    o, oh = oracle_lib.track(w, "fp64", hit_cap=HITS, threads=8)
    assert_bit_exact(g, gh, o, oh)
    print("case: persistent cylinder n=1500 spatial order", flush=True)
    s, sh = run_gpu(w, cfg=LaunchConfig(persistent=1, steps_per_launch=9, grid=64, spatial_order=1))
    assert_bit_exact(s, sh, o, oh)
    print("case: persistent cylinder n=1500 spatial order split by XCD", flush=True)
    x, xh = run_gpu(w, cfg=LaunchConfig(persistent=1, steps_per_launch=9, grid=64, spatial_order=2))
    assert_bit_exact(x, xh, o, oh)
    print("case: persistent cylinder n=1500 collision queue in candidate-count order", flush=True)
    q, qh = run_gpu(w, cfg=LaunchConfig(persistent=1, steps_per_launch=9, grid=64, collision_order=1))
    assert_bit_exact(q, qh, o, oh)
    print("case: cylinder n=1500 collision queue in candidate-count order", flush=True)
    q, qh = run_gpu(w, cfg=LaunchConfig(steps_per_launch=9, collision_order=1))
    assert_bit_exact(q, qh, o, oh)


def test_graph_replays_match_eager_calls(gpu):
    """hipGraph replays of the launch sequence against eager calls on an identical batch,
    compared after every call, with eager work and synchronisation in between."""
    import torch
    w = workloads.cube(n=2000, seed=12, tmax=3.0)
    print("case: graph replay cube n=2000", flush=True)
    mk = lambda: Batch(w.geometry, w.fields, w.species, w.tmax, w.species_idx, w.t0, w.y0, w.spin0, w.seed, w.pid,
                       hit_cap=HITS)
    a, b = mk(), mk()
    s = torch.cuda.Stream()
    sp = s.cuda_stream
    for cfg_kw in ({}, {"persistent": 1, "grid": 32}, {"persistent": 1, "grid": 32, "spatial_order": 1},
                   {"persistent": 1, "grid": 32, "spatial_order": 2}, {"persistent": 1, "grid": 32, "collision_order": 1}):
        for call in range(6):
            a.run(LaunchConfig(steps_per_launch=13, **cfg_kw), sp)
            b.run(LaunchConfig(steps_per_launch=13, graph=1, **cfg_kw), sp)
            torch.cuda.synchronize()
            x = torch.ones(1000, device="cuda").sum().item()   # eager work between replays
            assert x == 1000.0
            ra, rah = a.download(sp)
            rb, rbh = b.download(sp)
            assert_bit_exact(ra, rah, rb, rbh)
