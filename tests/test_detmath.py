"""Deterministic elementary functions and the per-particle random stream (pt_detmath.h).

These replace glibc libm on the tracking path so host and device agree bit for
bit; here they are pinned against libm (numpy) and against closed forms.
"""
import numpy as np
import pytest


def ulps(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    sp = np.spacing(np.abs(b))
    return np.where(a == b, 0.0, np.abs(a - b) / sp)


@pytest.fixture(scope="module")
def P(oracle_lib):
    return oracle_lib


def test_log_exp_pow_within_few_ulp(P):
    rng = np.random.default_rng(0)
    x = np.exp(rng.uniform(-700, 700, 200000))
    assert ulps(P.detmath("log", x), np.log(x)).max() <= 2
    e = rng.uniform(-700, 700, 200000)
    assert ulps(P.detmath("exp", e), np.exp(e)).max() <= 2
    # the step adjuster's exponents: err^(-1/3) on rejection, err^(-1/5) on acceptance
    err = np.exp(rng.uniform(np.log(3.2e-4), np.log(1e12), 200000))
    # exp(y log x): the log's rounding is amplified by |y log x| <= 9.2 here
    for y in (-1.0 / 3, -1.0 / 5):
        assert ulps(P.detmath("pow", err, np.full_like(err, y)), err ** y).max() <= 8


def test_special_values(P):
    assert P.detmath("log", np.array([1.0]))[0] == 0.0
    assert np.isneginf(P.detmath("log", np.array([0.0]))[0])
    assert np.isnan(P.detmath("log", np.array([-1.0]))[0])
    assert P.detmath("pow", np.array([1.0]), np.array([-0.2]))[0] == 1.0
    assert P.detmath("exp", np.array([0.0]))[0] == 1.0
    assert P.detmath("acos", np.array([1.0]))[0] == 0.0
    # cos(pi/2) as glibc returns it: the Lambert sampler's cos(amax) term
    assert P.detmath("cos", np.array([np.pi / 2]))[0] == np.cos(np.pi / 2)


def test_trig_on_sampling_range(P):
    rng = np.random.default_rng(1)
    phi = rng.uniform(0, 2 * np.pi, 200000)
    assert np.abs(P.detmath("sin", phi) - np.sin(phi)).max() < 4e-16
    assert np.abs(P.detmath("cos", phi) - np.cos(phi)).max() < 4e-16
    u = rng.uniform(0, 1, 200000)
    assert ulps(P.detmath("acos", np.sqrt(u)), np.arccos(np.sqrt(u))).max() <= 2
    assert ulps(P.detmath("acos", -u), np.arccos(-u)).max() <= 2


def test_atan_and_atan2(P):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-2, 2, 100000), 10 ** rng.uniform(-12, 12, 50000),
                        -(10 ** rng.uniform(-12, 12, 50000)), [0.0, -0.0, 0.5773502691896257, 1.0, -1.0, 1e308]])
    assert ulps(P.detmath("atan", x), np.arctan(x)).max() <= 3
    y = rng.normal(size=200000) * 10 ** rng.uniform(-5, 5, 200000)
    z = rng.normal(size=200000) * 10 ** rng.uniform(-5, 5, 200000)
    assert ulps(P.detmath("atan2", y, z), np.arctan2(y, z)).max() <= 3
    # quadrants, axes and signed zeros follow C atan2
    sy = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 0.0, -0.0, 1.0, -1.0, 2.0, -2.0])
    sx = np.array([0.0, 0.0, -0.0, -0.0, 0.0, 0.0, 1.0, 1.0, -1.0, -1.0, -3.0, -3.0])
    got, ref = P.detmath("atan2", sy, sx), np.arctan2(sy, sx)
    assert np.array_equal(np.signbit(got), np.signbit(ref)) and ulps(got, ref).max() <= 1


def test_complex_sqrt_and_hypot(P):
    rng = np.random.default_rng(2)
    re = rng.normal(size=50000) * 1e-7
    im = -np.abs(rng.normal(size=50000)) * 1e-9
    im[:100] = -0.0  # FermiImag = 0: the branch is chosen by the sign of zero
    z = np.empty(len(re), dtype=np.complex128)
    z.real, z.imag = re, im  # keeps the sign of -0.0
    ref = np.sqrt(z)
    assert np.allclose(P.detmath("csqrt_re", re, im), ref.real, rtol=4e-16, atol=0)
    assert np.allclose(P.detmath("csqrt_im", re, im), ref.imag, rtol=4e-16, atol=0)
    # sqrt(x - 0i) for x < 0 is -i sqrt(|x|) (lower branch), as C99 csqrt
    assert P.detmath("csqrt_im", np.array([-4.0]), np.array([-0.0]))[0] == -2.0
    a, b = rng.normal(size=(2, 50000))
    assert ulps(P.detmath("hypot", a, b), np.hypot(a, b)).max() <= 2


def test_random_stream_is_deterministic_and_uniform(P):
    raw1, u1 = P.rng_stream(1, 7, 100000)
    raw2, u2 = P.rng_stream(1, 7, 100000)
    assert np.array_equal(raw1, raw2)
    raw3, _ = P.rng_stream(1, 8, 1000)
    raw4, _ = P.rng_stream(2, 7, 1000)
    assert not np.array_equal(raw1[:1000], raw3) and not np.array_equal(raw1[:1000], raw4)
    # generate_canonical: r / 2^64, strictly below 1
    assert np.all(u1 < 1.0) and np.all(u1 >= 0.0)
    assert np.array_equal(u1, np.minimum(raw1.astype(np.float64) * 2.0 ** -64, np.nextafter(1.0, 0.0)))
    # uniformity (chi-square over 100 bins) and lag-1 independence
    h, _ = np.histogram(u1, bins=100, range=(0, 1))
    chi2 = ((h - 1000) ** 2 / 1000).sum()
    assert chi2 < 170  # 99 dof, p ~ 1e-5
    assert abs(np.corrcoef(u1[:-1], u1[1:])[0, 1]) < 0.02
