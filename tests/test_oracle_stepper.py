"""The oracle's Boost.Odeint dopri5 restatement (dense_output_runge_kutta<controlled<
runge_kutta_dopri5>>, called at src/tracking.cpp:63-64,79) pinned by mathematics rather
than by agreement with the GPU:

* the Dormand-Prince tableau below is checked against its own order conditions in exact
  rational arithmetic (so a mistyped coefficient here fails), then used to predict the
  oracle's error estimate on dx/dt = lambda x exactly;
* the error norm: max_i |err_i| / (atol + rtol (|x_old,i| + dt |dxdt_old,i|)), atol =
  rtol = 1e-9 (make_dense_output(1e-9, 1e-9), src/tracking.cpp:63), with cases where
  normalising by x_new, dropping the dt |dxdt| term or taking another norm gives another
  value;
* the controller (default_step_adjuster): accept iff err <= 1; on rejection
  dt *= max(0.9 err^(-1/3), 0.2); on acceptance with err < 0.5, dt *= 0.9 max(err,
  5^-5)^(-1/5); the retried and the next step use the proposed size;
* quadrature exactness for polynomial right-hand sides of degree <= 4 (the nodes c_i and
  weights b_i), fifth-order convergence on the Kepler problem, a fourth-order error
  estimate and a fourth-order dense output."""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import ptref

# Dormand & Prince (1980), RK5(4)7FM: Hairer, Norsett, Wanner I, Table II.5.2
A = [[],
     [Fr(1, 5)],
     [Fr(3, 40), Fr(9, 40)],
     [Fr(44, 45), Fr(-56, 15), Fr(32, 9)],
     [Fr(19372, 6561), Fr(-25360, 2187), Fr(64448, 6561), Fr(-212, 729)],
     [Fr(9017, 3168), Fr(-355, 33), Fr(46732, 5247), Fr(49, 176), Fr(-5103, 18656)],
     [Fr(35, 384), Fr(0), Fr(500, 1113), Fr(125, 192), Fr(-2187, 6784), Fr(11, 84)]]
B5 = [Fr(35, 384), Fr(0), Fr(500, 1113), Fr(125, 192), Fr(-2187, 6784), Fr(11, 84), Fr(0)]
B4 = [Fr(5179, 57600), Fr(0), Fr(7571, 16695), Fr(393, 640), Fr(-92097, 339200), Fr(187, 2100), Fr(1, 40)]
C = [sum(r, Fr(0)) for r in A]
FACT = [1, 1, 2, 6, 24, 120, 720, 5040]


def stability_poly(b):
    """Coefficients of R(z) = 1 + z b^T (I - z A)^-1 1 (degree 7 for the 7 stages)."""
    # (I - zA)^-1 1 = sum_k z^k A^k 1; A is strictly lower triangular (A^7 = 0)
    n = 7
    Am = [[A[i][j] if j < len(A[i]) else Fr(0) for j in range(n)] for i in range(n)]
    v = [Fr(1)] * n
    coeffs = [Fr(1)]
    for k in range(n):
        coeffs.append(sum((b[i] * v[i] for i in range(n)), Fr(0)))
        v = [sum((Am[i][j] * v[j] for j in range(n)), Fr(0)) for i in range(n)]
    return coeffs


def test_tableau_order_conditions():
    """The tableau of this test (not the oracle's) against the conditions that fix it."""
    r5, r4 = stability_poly(B5), stability_poly(B4)
    assert r5[:6] == [Fr(1, FACT[k]) for k in range(6)] and r5[6] != Fr(1, 720)
    assert r4[:5] == [Fr(1, FACT[k]) for k in range(5)] and r4[5] != Fr(1, 120)
    for k in range(5):   # quadrature conditions b^T c^k = 1/(k+1)
        assert sum(B5[i] * C[i] ** k for i in range(7)) == Fr(1, k + 1)
        if k < 4:
            assert sum(B4[i] * C[i] ** k for i in range(7)) == Fr(1, k + 1)
    assert A[6] == B5[:6]                                   # FSAL
    # two of the fifth-order tree conditions beyond the linear and quadrature ones
    Ac = [sum((A[i][j] * C[j] for j in range(len(A[i]))), Fr(0)) for i in range(7)]
    assert sum(B5[i] * C[i] * Ac[i] for i in range(7)) == Fr(1, 8)
    assert sum(B5[i] * C[i] ** 2 * Ac[i] for i in range(7)) == Fr(1, 10)


def error_poly():
    """E(z) with err = h sum_i (b5_i - b4_i) k_i = E(z) x0 for dx/dt = lambda x, z = lambda h."""
    r5, r4 = stability_poly(B5), stability_poly(B4)
    return [a - b for a, b in zip(r5, r4)]


def poly_at(c, z):
    z = Fr(z)
    return float(sum(ck * z ** k for k, ck in enumerate(c)))


LAM = np.array([-40.0, 25.0, -3.0, 0.5, -60.0, 10.0, -1.0, 2.0, -7.0])


def expected_err(lam, x0, h, norm="old"):
    """The error norm predicted from the exact error polynomial of the tableau above."""
    E, R = error_poly(), stability_poly(B5)
    worst = 0.0
    for l, x in zip(lam, x0):
        z = Fr(float(l)) * Fr(float(h))
        e = abs(poly_at(E, z) * float(x))
        xr = float(x) if norm != "new" else poly_at(R, z) * float(x)
        dxr = abs(l * xr) if norm != "no_dxdt" else 0.0
        worst = max(worst, e / (1e-9 + 1e-9 * (abs(xr) + h * dxr)))
    return worst


@pytest.mark.parametrize("h", [2e-3, 1e-2, 3e-2])
def test_error_norm_uses_old_state_and_derivative(oracle_lib, h):
    x0 = np.array([1.0, -2.0, 0.5, 3.0, 1e-12, -1.0, 4.0, 0.25, 2.0])
    r = ptref.probe_steps(ptref.PROBE_LINEAR, LAM, x0, 0.0, h, 1, fixed=True)
    want = expected_err(LAM, x0, h)
    assert r["err"][0] == pytest.approx(want, rel=1e-6)
    for other in ("new", "no_dxdt"):
        alt = expected_err(LAM, x0, h, other)
        assert abs(alt / want - 1) > 1e-3, other                    # the case tells them apart
    # a single component: its own ratio, so a norm other than the maximum would differ
    for i in (0, 4):
        lam1 = np.zeros(9); lam1[i] = LAM[i]
        x1 = np.zeros(9); x1[i] = x0[i]
        r1 = ptref.probe_steps(ptref.PROBE_LINEAR, lam1, x1, 0.0, h, 1, fixed=True)
        assert r1["err"][0] == pytest.approx(expected_err(lam1, x1, h), rel=1e-6)
    assert r["err"][0] == pytest.approx(max(expected_err(np.where(np.arange(9) == i, LAM, 0.0),
                                                         np.where(np.arange(9) == i, x0, 0.0), h)
                                            for i in range(9)), rel=1e-6)


def proposed(err, h):
    """default_step_adjuster of Boost.Odeint for dopri5 (stepper order 5, error order 4)."""
    if err > 1.0:
        return h * max(0.9 * err ** (-1.0 / 3), 0.2)
    if err < 0.5:
        return h * (0.9 * max(err, 5.0 ** -5) ** (-1.0 / 5))
    return h


@pytest.mark.parametrize("mode", ["fp64", "refsem"])
def test_controller_branches(oracle_lib, mode):
    """One forced trial per error regime: tiny error (the 5^-5 floor, factor 4.5), growth,
    the dead band 0.5 <= err <= 1 (size kept), mild and strong rejection (the 0.2 clamp)."""
    x0 = np.zeros(9); x0[0] = 1.0
    seen = set()
    for h in [1e-5, 1e-4, 1e-3, 1.7e-3, 1.8e-3, 2e-3, 3e-3, 6e-3, 1e-2]:
        lam = np.zeros(9); lam[0] = -40.0
        r = ptref.probe_steps(ptref.PROBE_LINEAR, lam, x0, 0.0, h, 1, fixed=True, mode=mode)
        err, dt = r["err"][0], r["dt"][0]
        want = proposed(err, h)
        if mode == "refsem":        # glibc pow, as Python's float power
            assert dt == want, (h, err, dt, want)
        else:                       # the deterministic pow of the fp64 mode (<= a few ulp)
            assert dt == pytest.approx(want, rel=4e-16 * 8), (h, err)
        seen.add("floor" if err < 5.0 ** -5 else "grow" if err < 0.5 else "keep" if err <= 1 else
                 "clamp" if 0.9 * err ** (-1 / 3) < 0.2 else "shrink")
    assert seen == {"floor", "grow", "keep", "shrink", "clamp"}, seen
    # the floor gives the largest growth, 0.9 * 5 = 4.5
    r = ptref.probe_steps(ptref.PROBE_LINEAR, np.zeros(9), x0, 0.0, 1e-3, 1, fixed=True, mode=mode)
    assert r["err"][0] == 0 and r["dt"][0] == pytest.approx(4.5e-3, rel=1e-15)


def test_adaptive_steps_use_the_proposed_size(oracle_lib):
    """Adaptive stepping: every accepted step has error <= 1 and the size the previous
    step proposed, rejected trials retry with the shrunk size, and the proposal after an
    accepted step follows the controller."""
    x0 = np.ones(9)
    lam = np.array([-50.0, 3.0, 0, 0, 0, 0, 0, 0, 0])
    h0 = 0.2                                   # far too large: the first trials are rejected
    r = ptref.probe_steps(ptref.PROBE_LINEAR, lam, x0, 0.0, h0, 40)
    assert r["rej"][0] >= 1 and (r["err"] <= 1).all()
    # replay the rejections of the first step with the controller above
    h = h0
    for _ in range(r["rej"][0]):
        e = ptref.probe_steps(ptref.PROBE_LINEAR, lam, x0, 0.0, h, 1, fixed=True)["err"][0]
        assert e > 1
        h = proposed(e, h)
    assert r["t"][0] == pytest.approx(h, rel=1e-15)
    steps = np.diff(np.concatenate([[0.0], r["t"]]))
    for k in range(1, 40):
        if r["rej"][k] == 0:
            assert steps[k] == pytest.approx(r["dt"][k - 1], rel=1e-12), k
    for k in range(40):
        assert r["dt"][k] == pytest.approx(proposed(r["err"][k], steps[k]), rel=1e-12), k
    # the exact solution is followed to the tolerance
    assert np.allclose(r["x"][-1, :2], np.exp(lam[:2] * r["t"][-1]), rtol=2e-7, atol=1e-9)


def test_quadrature_exact_to_degree_four(oracle_lib):
    """dx/dt = p(t): a step is the quadrature sum_i b_i p(t + c_i h); exact for degree <= 4
    (5th-order weights and the nodes c_i = sum_j a_ij), not for degree 5."""
    rng = np.random.default_rng(4)
    par = rng.uniform(-1, 1, (9, 5))
    t0, h = 0.3, 0.7
    r = ptref.probe_steps(ptref.PROBE_POLY_T, par, np.zeros(9), t0, h, 1, fixed=True)
    k = np.arange(5)
    exact = (par * ((t0 + h) ** (k + 1) - t0 ** (k + 1)) / (k + 1)).sum(axis=1)
    assert np.allclose(r["x"][0], exact, rtol=1e-14, atol=1e-15)
    # t^4 is integrated exactly, and the dense output is exact for degree <= 3 integrands
    # (fourth-order interpolant) at interior points
    par3 = par.copy(); par3[:, 4] = 0
    th = np.array([0.0, 0.25, 0.5, 0.9, 1.0])
    r3 = ptref.probe_steps(ptref.PROBE_POLY_T, par3, np.zeros(9), t0, h, 1, fixed=True, theta=th)
    for q, tq in enumerate(th):
        tt = t0 + tq * h
        ex = (par3 * (tt ** (k + 1) - t0 ** (k + 1)) / (k + 1)).sum(axis=1)
        assert np.allclose(r3["dense"][q], ex, rtol=1e-13, atol=1e-14), tq


def kepler_x0():
    x0 = np.zeros(9)
    x0[0], x0[3] = 1.0, 1.0          # circular orbit of radius 1, period 2 pi
    return x0


def test_fifth_order_convergence_kepler(oracle_lib):
    T = 1.0
    errs, ests = [], []
    for n in (24, 48, 96):
        r = ptref.probe_steps(ptref.PROBE_KEPLER, [], kepler_x0(), 0.0, T / n, n, fixed=True)
        exact = np.array([np.cos(T), np.sin(T), -np.sin(T), np.cos(T)])
        errs.append(np.abs(r["x"][-1, :4] - exact).max())
        ests.append(r["err"][0])
    for a, b in zip(errs, errs[1:]):
        assert 30 < a / b < 37, errs                      # 2^5 = 32 (4th order: 16, 6th: 64)
    for a, b in zip(ests, ests[1:]):
        assert 29 < a / b < 35, ests                      # local error estimate ~ h^5


def test_dense_output_fourth_order_kepler(oracle_lib):
    """The dense output at theta = 0 is the old state exactly, at theta = 1 the new state,
    and in between its error is O(h^5) locally (fourth-order interpolant)."""
    th = np.array([0.0, 0.37, 1.0])
    errs = []
    for h in (0.2, 0.1, 0.05):
        r = ptref.probe_steps(ptref.PROBE_KEPLER, [], kepler_x0(), 0.0, h, 1, fixed=True, theta=th)
        assert np.array_equal(r["dense"][0], kepler_x0())
        assert np.allclose(r["dense"][2], r["x"][0], rtol=0, atol=4e-16)
        t = 0.37 * h
        errs.append(np.abs(r["dense"][1, :4] - [np.cos(t), np.sin(t), -np.sin(t), np.cos(t)]).max())
    for a, b in zip(errs, errs[1:]):
        assert 24 < a / b < 40, errs
