"""TTracker for the MI355X engine.

`TTracker.IntegrateParticle` keeps the reference signature's meaning
(include/tracking.h:46-47): integrate particles from their current state to a
stop condition, at most until global time tmax.  `IntegrateBatch` is the batch
form the GPU runs; `Batch` exposes the resumable device state for callers that
advance a batch in slices (the benchmark, the persistent variant).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native
from .particle import SpeciesConfig, make_states


@dataclass
class LaunchConfig:
    variant: int = _native.PT_VARIANT_FP64
    block: int = 256
    steps_per_launch: int = 0
    persistent: int = 0
    grid: int = 0
    rounds: int = 0
    graph: int = 0
    spatial_order: int = 0
    collision_order: int = 0
    leaf_size: int = 4

    def native(self):
        c = _native.LaunchCfg()
        c.variant, c.block, c.steps_per_launch, c.persistent, c.grid, c.rounds, c.graph, c.spatial_order = (
            self.variant, self.block, self.steps_per_launch, self.persistent, self.grid, self.rounds, self.graph,
            self.spatial_order)
        c.collision_order = self.collision_order
        return c


class Batch:
    """Particles resident in device memory with their resumable tracker state."""

    def __init__(self, geometry, fields, species, tmax, species_idx, t0, y0, spin0, seed, pid,
                 hit_cap=0, with_mixed=False, leaf_size=4):
        L = _native.lib()
        self.n = len(t0)
        self.hit_cap = hit_cap
        self.snap_cap = 0
        self.trk_cap = 0
        self.slog_cap = 0
        sp = (_native.Species * len(species))(*[s.native() for s in species])
        self._keep = [np.ascontiguousarray(species_idx, dtype=np.int32), np.ascontiguousarray(t0, dtype=np.float64),
                      np.ascontiguousarray(y0, dtype=np.float64).reshape(-1, 9),
                      np.ascontiguousarray(spin0, dtype=np.float64).reshape(-1, 5),
                      np.ascontiguousarray(pid, dtype=np.uint64), sp]
        si, t, y, s, p, _ = self._keep
        self.geometry, self.fields = geometry, fields
        g = geometry.device(leaf_size)
        f = fields.device(with_mixed)
        h = C.c_void_p()
        rc = L.pt_batch_create(g, f, len(species), sp, float(tmax), self.n, si.ctypes.data_as(C.POINTER(C.c_int32)),
                               _native.dptr(t), _native.dptr(y), _native.dptr(s), int(seed) & (2 ** 64 - 1),
                               p.ctypes.data_as(C.POINTER(C.c_uint64)), hit_cap, C.byref(h))
        _native.check(rc, "pt_batch_create")
        self.h = h

    def run(self, cfg=None, stream=None):
        cfg = cfg or LaunchConfig()
        c = cfg.native()
        _native.check(_native.lib().pt_batch_run(self.h, stream, C.byref(c)), "pt_batch_run")

    def active(self, stream=None):
        a = C.c_int()
        _native.check(_native.lib().pt_batch_active(self.h, stream, C.byref(a)), "pt_batch_active")
        return a.value

    def run_to_completion(self, cfg=None, stream=None, max_launches=1 << 30):
        for _ in range(max_launches):
            self.run(cfg, stream)
            if self.active(stream) == 0:
                return True
        return False

    def step_counts(self, stream=None, with_fields=False):
        """Accepted and rejected dopri5 steps summed over the batch (and, with
        with_fields, the step kernel's field-map cell evaluations)."""
        a, r, f = C.c_int64(), C.c_int64(), C.c_int64()
        _native.check(_native.lib().pt_batch_step_counts(self.h, stream, C.byref(a), C.byref(r), C.byref(f)),
                      "pt_batch_step_counts")
        return (a.value, r.value, f.value) if with_fields else (a.value, r.value)

    def set_profiling(self, on=True):
        _native.check(_native.lib().pt_batch_set_profiling(self.h, int(bool(on))), "pt_batch_set_profiling")

    def kernel_times(self, stream=None):
        """Accumulated HIP-event milliseconds and launch counts of the step and
        collision kernels since profiling was switched on."""
        ms = (C.c_double * 2)()
        n = (C.c_int64 * 2)()
        _native.check(_native.lib().pt_batch_kernel_times(self.h, stream, ms, n), "pt_batch_kernel_times")
        return {"step": (ms[0], n[0]), "collision": (ms[1], n[1])}

    def set_loss(self, loss):
        """Per-particle loss per bounce replacing the materials' LossPerBounce (None: off);
        call before the first run."""
        if loss is None:
            _native.check(_native.lib().pt_batch_set_loss(self.h, None), "pt_batch_set_loss")
            return
        v = np.ascontiguousarray(np.asarray(loss, dtype=np.float64))
        if v.shape != (self.n,):
            raise ValueError("one loss per particle")
        self._loss = v
        _native.check(_native.lib().pt_batch_set_loss(self.h, _native.dptr(v)), "pt_batch_set_loss")

    def set_snapshots(self, times, cap):
        """TLogger::PrintSnapshot times (sorted, de-duplicated) and the records kept per
        particle; call before the first run."""
        t = np.ascontiguousarray(np.asarray(times, dtype=np.float64))
        self.snap_cap = int(cap) if len(t) else 0
        _native.check(_native.lib().pt_batch_set_snapshots(self.h, len(t), _native.dptr(t), self.snap_cap),
                      "pt_batch_set_snapshots")

    def snapshots(self, stream=None):
        """(records [n][cap] SNAP_DTYPE, counts [n]): the particles' snapshot states."""
        rec = np.zeros(self.n * self.snap_cap, dtype=_native.SNAP_DTYPE)
        cnt = np.zeros(self.n, dtype=np.int32)
        _native.check(_native.lib().pt_batch_snapshots(self.h, stream, rec.ctypes.data_as(C.c_void_p),
                                                       cnt.ctypes.data_as(C.POINTER(C.c_int32))), "pt_batch_snapshots")
        return rec.reshape(self.n, self.snap_cap), cnt

    def set_track(self, interval, cap):
        """TLogger::PrintTrack records: step ends where the trajectory length crossed a
        multiple of `interval` (m), at most cap per particle; call before the first run."""
        self.trk_cap = int(cap) if interval > 0 else 0
        _native.check(_native.lib().pt_batch_set_track(self.h, float(interval), self.trk_cap), "pt_batch_set_track")

    def track(self, stream=None):
        """(records [n][cap] TRACK_DTYPE, counts [n]): the particles' track-log points."""
        rec = np.zeros(self.n * self.trk_cap, dtype=_native.TRACK_DTYPE)
        cnt = np.zeros(self.n, dtype=np.int32)
        _native.check(_native.lib().pt_batch_track(self.h, stream, rec.ctypes.data_as(C.c_void_p),
                                                   cnt.ctypes.data_as(C.POINTER(C.c_int32))), "pt_batch_track")
        return rec.reshape(self.n, self.trk_cap), cnt

    def set_spinlog(self, interval, cap):
        """TLogger::PrintSpin records inside the BMT spin-tracking windows: the start of each
        spin integration and spin-step ends that crossed a multiple of `interval` (s), at
        most cap per particle; call before the first run."""
        self.slog_cap = int(cap) if interval > 0 else 0
        _native.check(_native.lib().pt_batch_set_spinlog(self.h, float(interval), self.slog_cap),
                      "pt_batch_set_spinlog")

    def spinlog(self, stream=None):
        """(records [n][cap] SPIN_DTYPE, counts [n]): the particles' spin-log rows."""
        rec = np.zeros(self.n * self.slog_cap, dtype=_native.SPIN_DTYPE)
        cnt = np.zeros(self.n, dtype=np.int32)
        _native.check(_native.lib().pt_batch_spinlog(self.h, stream, rec.ctypes.data_as(C.c_void_p),
                                                     cnt.ctypes.data_as(C.POINTER(C.c_int32))), "pt_batch_spinlog")
        return rec.reshape(self.n, self.slog_cap), cnt

    def download(self, stream=None):
        out = np.zeros(self.n, dtype=_native.OUT_DTYPE)
        hits = np.zeros(self.n * self.hit_cap, dtype=_native.HIT_DTYPE) if self.hit_cap else None
        rc = _native.lib().pt_batch_download(self.h, stream, out.ctypes.data_as(C.c_void_p),
                                             hits.ctypes.data_as(C.c_void_p) if hits is not None else None)
        _native.check(rc, "pt_batch_download")
        if hits is not None:
            hits = hits.reshape(self.n, self.hit_cap)
        return out, hits

    def secondaries(self, stream=None):
        """Decay products of the batch's decayed neutrons (TNeutron::Decay), ordered by parent
        index, proton first: a SEC_DTYPE record array."""
        L = _native.lib()
        cnt = C.c_int()
        _native.check(L.pt_batch_secondaries(self.h, stream, 0, None, C.byref(cnt)), "pt_batch_secondaries")
        out = np.zeros(cnt.value, dtype=_native.SEC_DTYPE)
        if cnt.value:
            _native.check(L.pt_batch_secondaries(self.h, stream, cnt.value, out.ctypes.data_as(C.c_void_p),
                                                 C.byref(cnt)), "pt_batch_secondaries")
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and _native._lib is not None:
                _native._lib.pt_batch_destroy(self.h)
        except Exception:
            pass


SECONDARY_STREAM = {1: np.uint64(1) << np.uint64(62), 2: np.uint64(2) << np.uint64(62)}


def secondary_states(species, rec, pid):
    """Species list, indices and initial states of decay products (TProton / TElectron
    constructors, src/particle.cpp:36-90; velocity from E, phi, theta as there).  The
    species' settings come from `species` when it holds a proton or an electron entry,
    else the defaults of the PARTICLES section (in/config.in: no decay, tmax, lmax 9e99).
    A product's random stream id is its parent's id with bit 62 (proton) or 63 (electron) set."""
    from .constants import ELECTRON, PROTON
    by_kind = {s.kind: s for s in species}
    sspecies = [by_kind.get(PROTON, SpeciesConfig(PROTON)), by_kind.get(ELECTRON, SpeciesConfig(ELECTRON))]
    n = len(rec)
    sidx = np.where(rec["kind"] == 1, 0, 1).astype(np.int32)
    st0 = np.asarray(rec["t"], np.float64).copy()
    sy0 = np.zeros((n, 9))
    for k, kind in ((1, PROTON), (2, ELECTRON)):
        sel = rec["kind"] == k
        if sel.any():
            _, yk, _ = make_states(kind, rec["pos"][sel], rec["E"][sel], rec["phi"][sel], rec["theta"][sel],
                                   rec["pol"][sel].astype(np.float64))
            sy0[sel] = yk
    spid = pid[rec["parent"]] | np.array([SECONDARY_STREAM[int(k)] for k in rec["kind"]], np.uint64)
    return sspecies, sidx, st0, sy0, np.zeros((n, 5)), spid


class TTracker:
    """Drop-in for TTracker::IntegrateParticle over batches of particles."""

    def __init__(self, geometry, fields, launch=None):
        self.geometry = geometry
        self.fields = fields
        self.launch = launch or LaunchConfig()

    def IntegrateBatch(self, species, tmax, species_idx, t0, y0, spin0, seed, pid, hit_cap=0, snapshots=None,
                       track=None, spinlog=None):
        """(out, hits); with snapshots = (times, cap) also (records, counts) of the snapshots,
        with track = (interval, cap) also (records, counts) of the track log, with
        spinlog = (interval, cap) also (records, counts) of the spin log, in that order."""
        cfg = self.launch
        b = Batch(self.geometry, self.fields, species, tmax, species_idx, t0, y0, spin0, seed, pid, hit_cap,
                  with_mixed=cfg.variant == _native.PT_VARIANT_MIXED, leaf_size=cfg.leaf_size)
        if snapshots is not None:
            b.set_snapshots(*snapshots)
        if track is not None:
            b.set_track(*track)
        if spinlog is not None:
            b.set_spinlog(*spinlog)
        b.run_to_completion(cfg)
        res = b.download()
        if snapshots is not None:
            res = res + b.snapshots()
        if track is not None:
            res = res + b.track()
        if spinlog is not None:
            res = res + b.spinlog()
        return res

    def IntegrateWithSecondaries(self, species, tmax, species_idx, t0, y0, spin0, seed, pid, hit_cap=0):
        """The reference's main loop for one batch (src/main.cpp:153-172): the primaries,
        then the protons and electrons of their neutron decays.  Returns
        ((out, hits), (sec_out, sec_hits), records): the products' results are in the
        order of `records` (pentrack_amd.tracker.Batch.secondaries)."""
        cfg = self.launch
        b = Batch(self.geometry, self.fields, species, tmax, species_idx, t0, y0, spin0, seed, pid, hit_cap,
                  with_mixed=cfg.variant == _native.PT_VARIANT_MIXED, leaf_size=cfg.leaf_size)
        b.run_to_completion(cfg)
        prim = b.download()
        rec = b.secondaries()
        if len(rec) == 0:
            return prim, (np.zeros(0, _native.OUT_DTYPE), None), rec
        sspecies, sidx, st0, sy0, sspin0, spid = secondary_states(species, rec, np.asarray(pid, np.uint64))
        s = Batch(self.geometry, self.fields, sspecies, tmax, sidx, st0, sy0, sspin0, seed, spid, hit_cap,
                  with_mixed=cfg.variant == _native.PT_VARIANT_MIXED, leaf_size=cfg.leaf_size)
        s.run_to_completion(cfg)
        return prim, s.download(), rec

    def IntegrateParticle(self, species: SpeciesConfig, tmax, t0, y0, spin0=None, seed=0, pid=0, hit_cap=0):
        out, hits = self.IntegrateBatch([species], tmax, np.zeros(1, np.int32), np.array([t0], np.float64),
                                        np.asarray(y0, np.float64).reshape(1, 9),
                                        np.zeros((1, 5)) if spin0 is None else np.asarray(spin0).reshape(1, 5),
                                        seed, np.array([pid], np.uint64), hit_cap)
        return out[0], (hits[0] if hits is not None else None)
