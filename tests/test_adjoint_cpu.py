"""Differentiable RK4 path on the CPU restatement (oracle/ptadj.c): the discrete
adjoint against central finite differences of the forward map, for the map scales
s_k and the initial state; the second-derivative tricubic evaluation against finite
differences of the first derivatives is covered through the state gradient."""
import numpy as np
import pytest

from oracle.ptref import AdjModel as oracle_model  # the RK4 model on the CPU (oracle/ptadj.c)
from pentrack_amd import workloads


@pytest.fixture(scope="module")
def case():
    model, y0, s0 = workloads.adjoint_cube(n=6, seed=3, n_grid=12, nmaps=3, dt=2e-4)
    return model, y0, s0


def objective(yT, m_kg):
    """J = mean kinetic energy + mean height (the two objectives summed)."""
    return 0.5 * m_kg * (yT[:, 3:] ** 2).sum(1).mean() + yT[:, 2].mean()


def dJ(yT, m_kg):
    g = np.zeros_like(yT)
    g[:, 3:] = m_kg * yT[:, 3:] / len(yT)
    g[:, 2] = 1.0 / len(yT)
    return g


def test_adjoint_matches_central_differences(case):
    model, y0, s0 = case
    nsteps = 600          # 0.12 s: several wall reflections for the fast particles
    M = oracle_model(model, s0)
    yT = M.forward(y0, nsteps)
    lam, sbar = M.adjoint(y0, dJ(yT, model.mass), nsteps)
    gs = sbar.sum(0)
    for k in range(model.nmaps):
        h = 1e-6
        sp, sm = s0.copy(), s0.copy()
        sp[k] += h
        sm[k] -= h
        fd = (objective(oracle_model(model, sp).forward(y0, nsteps), model.mass) -
              objective(oracle_model(model, sm).forward(y0, nsteps), model.mass)) / (2 * h)
        assert abs(gs[k] - fd) <= 1e-6 * abs(fd) + 1e-12, (k, gs[k], fd)
    for i in range(len(y0)):
        for q in range(6):
            h = 1e-6 * (1.0 if q >= 3 else 0.1)
            yp, ym = y0.copy(), y0.copy()
            yp[i, q] += h
            ym[i, q] -= h
            fd = (objective(M.forward(yp, nsteps), model.mass) - objective(M.forward(ym, nsteps), model.mass)) / (2 * h)
            assert abs(lam[i, q] - fd) <= 1e-5 * abs(fd) + 1e-9, (i, q, lam[i, q], fd)


def test_reflection_map_is_exact_for_free_flight():
    """Without field and gravity the motion is a straight line folded by the walls; the
    wall map applied after each step reproduces the fold exactly."""
    model, y0, s0 = workloads.adjoint_cube(n=4, seed=9, n_grid=8, nmaps=1, dt=1e-3)
    model.grav = 0.0
    M = oracle_model(model, [0.0])
    nsteps = 700
    yT = M.forward(y0, nsteps)
    T = nsteps * model.dt
    for i in range(len(y0)):
        # unfolded free flight, then fold into [0, 1]
        x = y0[i, :3] + y0[i, 3:] * T
        v = y0[i, 3:]
        for a in range(3):
            p = np.mod(x[a], 2.0)
            xa = p if p <= 1.0 else 2.0 - p
            assert abs(yT[i, a] - xa) < 1e-9
            sgn = 1.0 if np.mod(np.floor(x[a]), 2) == 0 else -1.0
            assert abs(yT[i, 3 + a] - sgn * v[a]) < 1e-9
