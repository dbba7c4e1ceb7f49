"""GPU differentiable path: the checkpointed RK4 forward and adjoint kernels against
the CPU restatement (oracle/ptadj.c) on the same model, the torch.autograd wrapper,
and the adjoint gradient against central finite differences of the GPU forward."""
import numpy as np
import pytest

from oracle.ptref import AdjModel as oracle_model  # the RK4 model on the CPU (oracle/ptadj.c)
from pentrack_amd import workloads
from pentrack_amd.adjoint import rk4_adjoint, rk4_forward, rk4_tracker

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(gpu):
    model, y0, s0 = workloads.adjoint_cube(n=500, seed=3, n_grid=16, nmaps=4, dt=2e-4)
    return model, y0, s0


def test_forward_and_adjoint_match_cpu(case, gpu):
    import torch
    model, y0, s0 = case
    print("case: adjoint n=500 maps=4 grid=16 steps=600 K=64", flush=True)
    nsteps, K = 600, 64
    yT, ck = rk4_forward(model, s0, torch.tensor(y0, device=gpu), nsteps, K)
    M = oracle_model(model, s0)
    yc = M.forward(y0, nsteps)
    assert np.allclose(yT.cpu().numpy(), yc, rtol=1e-12, atol=1e-12)
    lamT = np.zeros_like(yc)
    lamT[:, 3:] = model.mass * yc[:, 3:] / len(yc)
    lamT[:, 2] = 1.0 / len(yc)
    gy0, gs = rk4_adjoint(model, s0, ck, torch.tensor(lamT, device=gpu), nsteps, K)
    lc, sc = M.adjoint(y0, lamT, nsteps)
    assert np.allclose(gy0.cpu().numpy(), lc, rtol=1e-10, atol=1e-10 * np.abs(lc).max())
    assert np.allclose(gs.cpu().numpy(), sc, rtol=1e-10, atol=1e-10 * np.abs(sc).max())


def test_autograd_gradient_against_finite_differences(case, gpu):
    import torch
    model, y0, s0 = case
    print("case: autograd vs FD", flush=True)
    nsteps = 400
    f = rk4_tracker(model, nsteps, K=128)
    s = torch.tensor(s0, dtype=torch.float64, device=gpu, requires_grad=True)
    y = torch.tensor(y0, device=gpu, requires_grad=True)
    yT = f(s, y)
    J = 0.5 * model.mass * (yT[:, 3:] ** 2).sum(1).mean() + yT[:, 2].mean()
    J.backward()
    gs = s.grad.cpu().numpy()

    def J_of(sv):
        yv, _ = rk4_forward(model, sv, torch.tensor(y0, device=gpu), nsteps, keep_ckpt=False)
        yv = yv.cpu().numpy()
        return 0.5 * model.mass * (yv[:, 3:] ** 2).sum(1).mean() + yv[:, 2].mean()

    for k in range(model.nmaps):
        h = 1e-6
        sp, sm = s0.copy(), s0.copy()
        sp[k] += h
        sm[k] -= h
        fd = (J_of(sp) - J_of(sm)) / (2 * h)
        assert abs(gs[k] - fd) <= 1e-5 * abs(fd) + 1e-12, (k, gs[k], fd)
    assert y.grad is not None and torch.isfinite(y.grad).all()


def test_chunked_horizon_equals_one_pass(case, gpu):
    """A horizon run in chunks (config 5's 2e5 steps are run that way): the same final
    states and dJ/dy0 bit for bit, dJ/ds to rounding (the per-chunk sums are added)."""
    import torch
    from pentrack_amd.adjoint import rk4_adjoint_chunks, rk4_forward_chunks
    model, y0, s0 = case
    nsteps, K, chunk = 1100, 64, 256
    print(f"case: adjoint chunks n=500 steps={nsteps} chunk={chunk} K={K}", flush=True)
    y0t = torch.tensor(y0, device=gpu)
    yT, ck = rk4_forward(model, s0, y0t, nsteps, K)
    yc, cks, lens = rk4_forward_chunks(model, s0, y0t, nsteps, chunk, K)
    assert lens == [256, 256, 256, 256, 76]
    assert torch.equal(yT, yc)
    lamT = torch.zeros_like(yT)
    lamT[:, 3:] = model.mass * yT[:, 3:] / len(yT)
    lamT[:, 2] = 1.0 / len(yT)
    gy0, gs = rk4_adjoint(model, s0, ck, lamT, nsteps, K)
    cy0, cgs = rk4_adjoint_chunks(model, s0, cks, lens, lamT, K)
    assert torch.equal(gy0, cy0)
    assert np.allclose(cgs.cpu().numpy(), gs.cpu().numpy(), rtol=1e-12, atol=1e-14 * np.abs(gs.cpu().numpy()).max())
    with pytest.raises(ValueError):
        rk4_forward_chunks(model, s0, y0t, nsteps, 100, K)
