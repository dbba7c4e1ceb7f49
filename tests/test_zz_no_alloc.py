"""pt_rk4_adjoint_noalloc (the adjoint with its workspace in the library's device array)
against pt_rk4_adjoint (the same kernel with a caller-allocated workspace), through the
C ABI: bit-identical gradients on two calls in a row, the second with fewer particles
and a shorter horizon, and the size limit reported as PT_EINVAL before any launch."""
import ctypes as C

import numpy as np
import pytest

from pentrack_amd import _native, workloads
from pentrack_amd.adjoint import _lib, rk4_forward

pytestmark = pytest.mark.gpu

PT_ADJ_NOALLOC_MAX_BYTES = 64 << 20


def _adjoint_both(model, s0, y0, lamT, nsteps, K, gpu):
    """(lam, sbar) from pt_rk4_adjoint and from pt_rk4_adjoint_noalloc on the same inputs."""
    import torch
    n = y0.shape[0]
    _, ck = rk4_forward(model, s0, torch.tensor(y0, device=gpu), nsteps, K)
    P = model.params([float(v) for v in s0], gpu)
    stream = torch.cuda.current_stream(gpu).cuda_stream
    L = _lib()
    res = []
    for noalloc in (False, True):
        lam = torch.tensor(lamT.T.copy(), device=gpu).contiguous()
        sbar = torch.zeros((max(model.nmaps, 1), n), dtype=torch.float64, device=gpu)
        if noalloc:
            rc = L.pt_rk4_adjoint_noalloc(stream, C.byref(P), n, ck.data_ptr(), lam.data_ptr(), sbar.data_ptr(),
                                          nsteps, K)
        else:
            seg = torch.empty((min(K, nsteps), 24, n), dtype=torch.float64, device=gpu)
            flags = torch.empty((min(K, nsteps), n), dtype=torch.int32, device=gpu)
            rc = L.pt_rk4_adjoint(stream, C.byref(P), n, ck.data_ptr(), lam.data_ptr(), sbar.data_ptr(),
                                  seg.data_ptr(), flags.data_ptr(), nsteps, K)
        _native.check(rc, "pt_rk4_adjoint_noalloc" if noalloc else "pt_rk4_adjoint")
        torch.cuda.synchronize(gpu)
        res.append((lam.cpu().numpy(), sbar.cpu().numpy()))
    return res


def test_noalloc_adjoint_matches_allocating_adjoint(gpu):
    model, y0, s0 = workloads.adjoint_cube(n=600, seed=8, n_grid=16, nmaps=4, dt=2e-4)
    # two calls in a row: the second with fewer particles and steps reuses the array
    for n, nsteps, K in ((600, 500, 64), (250, 300, 128)):
        print(f"case: noalloc adjoint n={n} steps={nsteps} K={K}", flush=True)
        yn = np.ascontiguousarray(y0[:n])
        lamT = np.zeros((n, 6))
        lamT[:, 3:] = model.mass * yn[:, 3:] / n
        lamT[:, 2] = 1.0 / n
        (la, sa), (lb, sb) = _adjoint_both(model, s0, yn, lamT, nsteps, K, gpu)
        assert np.isfinite(la).all() and np.abs(sa).max() > 0
        assert np.array_equal(la, lb), "dJ/dy0 differs between the two entry points"
        assert np.array_equal(sa, sb), "dJ/ds differs between the two entry points"


def test_noalloc_adjoint_rejects_oversized_workspace(gpu):
    import torch
    model, y0, s0 = workloads.adjoint_cube(n=8, seed=8, n_grid=8, nmaps=1, dt=2e-4)
    P = model.params([float(v) for v in s0], gpu)
    L = _lib()
    K = 256
    n = PT_ADJ_NOALLOC_MAX_BYTES // (K * 196) + 1          # one particle beyond the limit
    dummy = torch.zeros(1, dtype=torch.float64, device=gpu)  # never read: the call returns first
    rc = L.pt_rk4_adjoint_noalloc(None, C.byref(P), n, dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(),
                                  1000, K)
    assert rc == _native.PT_EINVAL
    assert b"PT_ADJ_NOALLOC_MAX_BYTES" in _native.lib().pt_last_error()
