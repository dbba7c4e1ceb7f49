"""Free-neutron decay into a proton and an electron (TNeutron::Decay,
src/neutron.cpp:285-362): the proton recoil spectrum table and its sampling as
std::piecewise_linear_distribution does them, the decay kinematics of the engine
(host build of the device code) against the oracle, and four-momentum conservation.
"""
import ctypes as C
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from pentrack_amd import _native

# constants as the device and the oracle's fp64 mode hold them
ELE_E, C0 = 1.602176487E-19, 299792458.0
M_N, M_P, M_E = 1.674927211E-27 / ELE_E, 1.672621637E-27 / ELE_E, 9.10938215e-31 / ELE_E


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def engine_table():
    L = _native.lib()
    n = C.c_int()
    assert L.pt_proton_beta_table(0, None, None, None, None, None, C.byref(n)) == 0
    ni = n.value
    x, raw, den, cp, m = np.zeros(ni + 1), np.zeros(ni + 1), np.zeros(ni + 1), np.zeros(ni), np.zeros(ni)
    assert L.pt_proton_beta_table(ni + 1, dp(x), dp(raw), dp(den), dp(cp), dp(m), C.byref(n)) == 0
    return x, raw, den, cp, m


def engine_decay(key, ctr, v):
    L = _native.lib()
    out = np.zeros(6)
    pol = np.zeros(2, np.int32)
    cout = C.c_uint64()
    v = np.ascontiguousarray(v, np.float64)
    assert L.pt_neutron_decay_host(key, ctr, dp(v), dp(out), pol.ctypes.data_as(C.POINTER(C.c_int32)),
                                   C.byref(cout)) == 0
    return out, pol, cout.value


def test_spectrum_table_matches_oracle(oracle_lib):
    x, raw, den, cp, m = engine_table()
    ox, oden, ocp, om = oracle_lib.beta_table("fp64")
    for a, b in ((x, ox), (den, oden), (cp, ocp), (m, om)):
        assert np.array_equal(a, b)
    assert x[0] == 0.0 and x[-1] == 750.0 and np.all(np.diff(x) > 0)
    assert cp[-1] == 1.0 and np.all(np.diff(cp) >= 0) and np.all(raw >= 0)
    # the spectrum vanishes at E = 0 and at the recoil endpoint ((m_n - m_p)^2 - m_e^2) c^2 / (2 m_n)
    e_end = ((M_N - M_P) ** 2 - M_E ** 2) * C0 ** 2 / (2 * M_N)
    assert 751.0 < e_end < 752.0 and raw[0] == 0.0
    slope = (raw[-1] - raw[-2]) / (x[-1] - x[-2])
    assert abs(x[-1] - raw[-1] / slope - e_end) < 2.0        # linear extrapolation to the zero


SAMPLER_CPP = r"""
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>
int main(int argc, char **argv) {
    int n = 0, ns = 0;
    if (fread(&n, sizeof n, 1, stdin) != 1 || fread(&ns, sizeof ns, 1, stdin) != 1) return 1;
    std::vector<double> x(n), w(n);
    if (fread(x.data(), sizeof(double), n, stdin) != (size_t)n) return 1;
    if (fread(w.data(), sizeof(double), n, stdin) != (size_t)n) return 1;
    std::piecewise_linear_distribution<double> dist(x.begin(), x.end(), w.begin());
    std::mt19937_64 a(12345), b(12345);
    std::vector<double> den = dist.densities();
    fwrite(den.data(), sizeof(double), den.size(), stdout);
    for (int i = 0; i < ns; ++i) {
        uint64_t r = b();
        double s = dist(a);
        fwrite(&r, sizeof r, 1, stdout);
        fwrite(&s, sizeof s, 1, stdout);
    }
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_sampler_matches_libstdcxx_piecewise_linear_distribution():
    """std::piecewise_linear_distribution built by libstdc++ from the same abscissas and
    spectrum values: identical normalised densities, and identical samples for the same
    64-bit engine outputs."""
    x, raw, den, cp, m = engine_table()
    ns = 20000
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "s.cpp"), os.path.join(d, "s")
        open(src, "w").write(SAMPLER_CPP)
        subprocess.run(["g++", "-O1", "-std=c++11", src, "-o", exe], check=True)
        inp = np.array([len(x), ns], np.int32).tobytes() + x.tobytes() + raw.tobytes()
        res = subprocess.run([exe], input=inp, capture_output=True, check=True).stdout
    lden = np.frombuffer(res[:8 * len(x)], np.float64)
    assert np.array_equal(lden, den)
    rec = np.frombuffer(res[8 * len(x):], dtype=[("r", "<u8"), ("s", "<f8")])
    # generate_canonical<double, 64> of one 64-bit draw: r / 2^64, kept below 1
    p = rec["r"].astype(np.float64) * 2.0 ** -64
    p = np.where(p >= 1.0, np.nextafter(1.0, 0.0), p)
    got = np.zeros(ns)
    assert _native.lib().pt_beta_sample(ns, dp(np.ascontiguousarray(p)), dp(got)) == 0
    assert np.array_equal(got, rec["s"])


def momentum(E, phi, theta, m):
    pabs = np.sqrt(E * E + 2 * E * m * C0 * C0) / C0
    return pabs[:, None] * np.stack([np.cos(phi) * np.sin(theta), np.sin(phi) * np.sin(theta), np.cos(theta)], 1)


@pytest.mark.parametrize("speed", [0.0, 8.0])
def test_four_momentum_conservation(oracle_lib, speed):
    """p_n = p_p + p_e + p_nu with a neutrino of 1 eV/c^2: the neutrino's energy from the
    missing momentum closes the energy balance."""
    rng = np.random.default_rng(3)
    n = 4000
    v = rng.normal(size=(n, 3))
    v = speed * v / np.linalg.norm(v, axis=1)[:, None]
    out, pol = oracle_lib.neutron_decay(5, np.arange(n, dtype=np.uint64), np.zeros(n, np.uint64), v, "refsem")
    Ep, Ee = out[:, 0], out[:, 3]
    assert np.all((Ep >= 0) & (Ep <= 752)) and np.all((Ee > 0) & (Ee < 782400))
    gam = 1 / np.sqrt(1 - (v ** 2).sum(1) / C0 ** 2)
    pn = M_N * gam[:, None] * v
    pp, pe = momentum(Ep, out[:, 1], out[:, 2], M_P), momentum(Ee, out[:, 4], out[:, 5], M_E)
    pnu = pn - pp - pe
    m_nu = 1 / C0 ** 2
    Enu = np.sqrt((pnu ** 2).sum(1) * C0 ** 2 + (m_nu * C0 ** 2) ** 2)
    balance = (M_N * gam * C0 ** 2) - (Ep + M_P * C0 ** 2) - (Ee + M_E * C0 ** 2) - Enu
    assert np.abs(balance).max() < 0.05                        # eV, of 939.6 MeV
    assert set(np.unique(pol)) == {-1, 1}
    # at rest: azimuths uniform; the reference draws the proton's polar angle with
    # sin_distribution(0, pi/2), so protons leave into the upper hemisphere, <cos theta> = 1/2
    if speed == 0.0:
        assert np.abs(np.cos(out[:, 1]).mean()) < 0.05 and np.abs(np.cos(out[:, 4]).mean()) < 0.05
        assert np.all(out[:, 2] <= np.pi / 2) and abs(np.cos(out[:, 2]).mean() - 0.5) < 0.03


def test_engine_decay_matches_oracle(oracle_lib):
    """The device code built for the host against the oracle: bit for bit in the fp64
    semantics; against the reference's long double arithmetic within 1e-6 (energies,
    limited by the reference's own cancellation for slow protons) and 1e-8 rad."""
    rng = np.random.default_rng(9)
    n = 500
    v = rng.normal(size=(n, 3)) * 5
    pid = rng.integers(0, 2 ** 40, n).astype(np.uint64)
    ctr = rng.integers(0, 5000, n).astype(np.uint64)
    o, opol = oracle_lib.neutron_decay(11, pid, ctr, v, "fp64")
    r, rpol = oracle_lib.neutron_decay(11, pid, ctr, v, "refsem")
    for i in range(n):
        key = oracle_lib.stream_key(11, int(pid[i]))
        e, pol, cend = engine_decay(key, int(ctr[i]), v[i])
        assert np.array_equal(e, o[i]) and np.array_equal(pol, opol[i]), i
        assert cend == int(ctr[i]) + 7                             # seven draws per decay
    assert np.array_equal(opol, rpol)
    assert np.allclose(o[:, [0, 3]], r[:, [0, 3]], rtol=1e-6)
    assert np.abs(o[:, [1, 2, 4, 5]] - r[:, [1, 2, 4, 5]]).max() < 1e-8
