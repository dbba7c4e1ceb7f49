"""Particle sources of the SOURCE section (TParticleSource and the volume sources,
src/source.cpp:21-230, include/source.h:138-260).

Supported: `boxvolume` and `cylvolume` with a continuous or pulsed active time, an
energy spectrum and angular distributions given as formulas of x (parse_distribution,
include/mc.h:310-335: the same piecewise linear distribution the reference builds,
pentrack_amd's pt_distribution_table), and a polarisation weight.  Draws come from a
numpy generator seeded with the run's seed, not from the reference's mt19937_64, so a
run reproduces the reference's distributions, not its individual particles.
"""
import ctypes as C
import math

import numpy as np

from . import _native
from .config import ConfigError, SPECIES_BY_NAME, _floats, deg
from .particle import make_states


class PiecewiseLinear:
    """std::piecewise_linear_distribution<double> of a formula of x on [lo, hi]."""

    def __init__(self, formula, lo, hi):
        L = _native.lib()
        n = C.c_int()
        f = formula.strip().encode()
        _native.check(L.pt_distribution_table(f, lo, hi, 0, None, None, None, None, None, C.byref(n)),
                      "pt_distribution_table")
        ni = n.value
        self.x, raw, self.den = np.zeros(ni + 1), np.zeros(ni + 1), np.zeros(ni + 1)
        self.cp, self.m = np.zeros(ni), np.zeros(ni)
        d = _native.dptr
        _native.check(L.pt_distribution_table(f, lo, hi, ni + 1, d(self.x), d(raw), d(self.den), d(self.cp), d(self.m),
                                              C.byref(n)), "pt_distribution_table")

    def sample_from(self, p):
        """libstdc++ operator() for variates p in [0, 1)."""
        i = np.searchsorted(self.cp, p, side="left")
        pref = np.where(i > 0, self.cp[np.maximum(i - 1, 0)], 0.0)
        a = 0.5 * self.m[i]
        b = self.den[i]
        cm = p - pref
        with np.errstate(divide="ignore", invalid="ignore"):
            quad = 0.5 * (np.sqrt(b * b + 4.0 * a * cm) - b) / a
        return self.x[i] + np.where(a == 0, cm / b, quad)

    def __call__(self, rng, n):
        return self.sample_from(rng.random(n))


class VolumeSource:
    """TCuboidVolumeSource / TCylindricalVolumeSource."""

    def __init__(self, cfg):
        sc = cfg["SOURCE"]
        get = lambda k, d="": sc.get(k, d).strip()
        self.mode = get("sourcemode").split()[0] if get("sourcemode") else ""
        if self.mode not in ("boxvolume", "cylvolume"):
            raise ConfigError(f"sourcemode {self.mode!r} not supported (boxvolume, cylvolume)")
        self.particle = get("particle").split()[0]
        if self.particle not in SPECIES_BY_NAME:
            raise ConfigError("Could not create particle " + self.particle)
        self.kind = SPECIES_BY_NAME[self.particle]
        num = lambda k, d=0.0: _floats(get(k))[0] if _floats(get(k)) else d
        if int(num("PhaseSpaceWeighting")):
            raise ConfigError("PhaseSpaceWeighting 1 is not supported by this engine")
        self.active = num("ActiveTime")
        self.pulse_gap, self.pulse_width = num("pulseGap"), num("pulseWidth")
        self.polarization = num("polarization")
        self.spectrum = PiecewiseLinear(get("spectrum") or "1", num("Emin"), num("Emax"))
        self.phi_v = PiecewiseLinear(get("phi_v") or "1", deg(num("phi_v_min")), deg(num("phi_v_max", 360.0)))
        self.theta_v = PiecewiseLinear(get("theta_v") or "sin(x)", deg(num("theta_v_min")), deg(num("theta_v_max", 180.0)))
        p = _floats(get("parameters"))
        if len(p) < 6:
            raise ConfigError("source parameters need 6 values")
        self.par = p[:6]
        # start-time distribution (src/source.cpp:49-78)
        if self.active == 0:
            self.t_int, self.t_w = [0.0, math.nextafter(0.0, math.inf)], [1.0]
        elif self.pulse_gap > 0:
            iv, w = [], []
            i = 0.0
            while i < self.active:
                iv.append(i)
                if self.pulse_width == 0:
                    iv.append(math.nextafter(i, math.inf))
                    w.append(1.0)
                elif i + self.pulse_width > self.active:
                    iv.append(self.active)
                    w.append((self.active - i) / self.pulse_width)
                else:
                    iv.append(i + self.pulse_width)
                    w.append(1.0)
                w.append(0.0)
                i += self.pulse_gap
            w.pop()
            self.t_int, self.t_w = iv, w
        else:
            self.t_int, self.t_w = [0.0, self.active], [1.0]

    def _times(self, rng, n):
        """std::piecewise_constant_distribution over the intervals."""
        x = np.asarray(self.t_int)
        w = np.asarray(self.t_w) * np.diff(x)
        cp = np.cumsum(w) / w.sum()
        p = rng.random(n)
        i = np.minimum(np.searchsorted(cp, p, side="right"), len(cp) - 1)
        pref = np.where(i > 0, cp[np.maximum(i - 1, 0)], 0.0)
        frac = np.where(cp[i] > pref, (p - pref) / (cp[i] - pref), 0.0)
        return x[i] + frac * (x[i + 1] - x[i])

    def sample(self, n, rng):
        """n particles: (t0 [n], y0 [n][9], spin0 [n][5]) as the constructors set them up."""
        t0 = self._times(rng, n)
        if self.mode == "boxvolume":
            xmin, xmax, ymin, ymax, zmin, zmax = self.par
            pos = np.stack([xmin + rng.random(n) * (xmax - xmin), ymin + rng.random(n) * (ymax - ymin),
                            zmin + rng.random(n) * (zmax - zmin)], axis=1)
        else:
            rmin, rmax, phimin, phimax, zmin, zmax = self.par
            r = np.sqrt(rng.random(n) * (rmax * rmax - rmin * rmin) + rmin * rmin)   # linear_distribution(rmin, rmax)
            phr = deg(phimin) + rng.random(n) * (deg(phimax) - deg(phimin))
            pos = np.stack([r * np.cos(phr), r * np.sin(phr), zmin + rng.random(n) * (zmax - zmin)], axis=1)
        pol = np.where(rng.random(n) < 0.5 * (self.polarization + 1.0), 1.0, -1.0)   # polarization_distribution
        E = self.spectrum(rng, n)
        phi = self.phi_v(rng, n)
        theta = self.theta_v(rng, n)
        _, y0, spin0 = make_states(self.kind, pos, E, phi, theta, pol)
        return t0, y0, spin0
