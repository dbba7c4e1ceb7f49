"""TParticle-side helpers: species records and initial states.

`initial_velocity` restates the velocity set-up of the TParticle constructor
(src/particle.cpp:44-68); the state vector layout is
y = (x, y, z, vx, vy, vz, proper time, polarisation, path length)
(include/particle.h:21,25).
"""
from dataclasses import dataclass

import numpy as np

from . import _native
from .constants import C_0, SPECIES_CONSTANTS, NEUTRON, PROTON, ELECTRON

DBL_EPS = np.finfo(np.float64).eps


@dataclass
class SpeciesConfig:
    """Per-species particleconf entries used by TTracker::IntegrateParticle
    (src/tracking.cpp:19-58, in/config.in:296-314)."""
    kind: int = NEUTRON
    tau: float = 0.0
    tmax: float = 9e99
    lmax: float = 9e99
    flipspin: int = 0
    Bmax: float = 0.0
    spintimes: tuple = ()     # BMT spin-tracking windows t0 t1 [t2 t3 ...] (at most four)
    interpolatefields: int = 0  # BMT precession axis from splines over 11 points of each step

    def native(self):
        q, m, mu, gamma = SPECIES_CONSTANTS[self.kind]
        s = _native.Species()
        s.kind = self.kind
        s.q, s.m, s.mu, s.gamma = q, m, mu, gamma
        st = tuple(float(v) for v in self.spintimes)
        if len(st) % 2 or len(st) > 8:
            raise ValueError("spintimes: pairs of times, at most four windows")
        s.n_spintimes = len(st)
        s.interpolatefields = int(self.interpolatefields)
        for k, v in enumerate(st):
            s.spintimes[k] = v
        s.tau, s.tmax, s.lmax = self.tau, self.tmax, self.lmax
        s.flipspin = self.flipspin
        s.Bmax = self.Bmax
        return s


def initial_velocity(kind, E, phi, theta):
    """Velocity vector for kinetic energy E [eV] and direction (phi, theta)."""
    m = SPECIES_CONSTANTS[kind][1]
    E = np.asarray(E, dtype=np.float64)
    Eoverm = E / m / C_0 / C_0
    series = ((((6.0 * Eoverm - 5.0) * Eoverm + 4.0) * Eoverm - 3.0) * Eoverm + 2.0) * Eoverm
    exact = 1.0 - 1.0 / (Eoverm + 1) / (Eoverm + 1)
    beta2 = np.where(Eoverm ** 6 < 2 * DBL_EPS, series, exact)
    v = C_0 * np.sqrt(beta2)
    return np.stack([v * np.cos(phi) * np.sin(theta), v * np.sin(phi) * np.sin(theta), v * np.cos(theta)], axis=-1)


def make_states(kind, pos, E, phi, theta, pol, t0=0.0):
    """Initial (t0, y0[9], spin0[5]) arrays for a batch of particles."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    n = len(pos)
    y = np.zeros((n, 9), dtype=np.float64)
    y[:, 0:3] = pos
    y[:, 3:6] = initial_velocity(kind, E, phi, theta)
    y[:, 7] = pol
    t = np.full(n, t0, dtype=np.float64)
    spin = np.zeros((n, 5), dtype=np.float64)
    return t, y, spin


def kinetic_energy(kind, v):
    """TParticle::GetKineticEnergy (src/particle.cpp:237-251) in eV."""
    m = SPECIES_CONSTANTS[kind][1]
    v = np.asarray(v, dtype=np.float64)
    beta2 = (v[..., 0] ** 2 + v[..., 1] ** 2 + v[..., 2] ** 2) / C_0 / C_0
    series = (0.5 + (3.0 / 8.0 + (5.0 / 16.0 + 35.0 / 128.0 * beta2) * beta2) * beta2) * beta2
    with np.errstate(invalid="ignore", divide="ignore"):
        exact = 1.0 / np.sqrt(1.0 - beta2) - 1.0
    return C_0 * C_0 * m * np.where(beta2 ** 5 < DBL_EPS, series, exact)


__all__ = ["SpeciesConfig", "initial_velocity", "make_states", "kinetic_energy", "NEUTRON", "PROTON", "ELECTRON"]
