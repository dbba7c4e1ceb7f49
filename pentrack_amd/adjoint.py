"""Differentiable tracking: fixed-step RK4 with a checkpointed discrete adjoint.

`RK4Tracker` is a torch.autograd.Function over the C ABI entry points
`pt_rk4_forward_ckpt` / `pt_rk4_adjoint` (pentrack_amd/csrc/pt_adjoint.hip): the
forward pass integrates neutrons (non-relativistic TParticle::EquationOfMotion,
src/particle.cpp:105-135) through B = sum_k s_k B_k and gravity, with specular box
walls applied as fixed maps after each step; the backward pass returns dJ/ds and
dJ/dy0 for any scalar J of the final states.  The reference has no derivatives:
the adjoint is an addition of this engine.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _native
from .constants import GRAVCONST, MASS_N_KG, MU_N_SI

MAX_MAPS = 8


class AdjMaps(C.Structure):
    _fields_ = [
        ("nmaps", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("x0", C.c_double * 3), ("inv_h", C.c_double * 3),
        ("g", C.c_void_p * 3), ("coef", C.c_void_p * MAX_MAPS), ("s", C.c_double * MAX_MAPS),
    ]


class AdjParams(C.Structure):
    _fields_ = [
        ("maps", AdjMaps), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("dt", C.c_double), ("coupling", C.c_double), ("grav", C.c_double),
    ]


def device_layout(coef):
    """libtricubic order a[i + 4j + 16k] -> device order a[16i + 4j + k] (coefficient of x^i y^j z^k)."""
    c = np.asarray(coef, dtype=np.float64)
    shp = c.shape
    c = c.reshape(-1, 4, 4, 4)            # [.., k, j, i]
    return np.ascontiguousarray(c.transpose(0, 3, 2, 1)).reshape(shp)


@dataclass
class AdjointModel:
    """Maps on one grid (each [cells][3][64], libtricubic order), box walls, RK4 step."""
    grid: tuple
    coefs: List[np.ndarray]
    lo: tuple = (0.0, 0.0, 0.0)
    hi: tuple = (1.0, 1.0, 1.0)
    dt: float = 1e-4
    pol: float = 1.0
    mass: float = MASS_N_KG
    mu: float = MU_N_SI
    grav: float = GRAVCONST
    _dev: dict = field(default_factory=dict, repr=False)

    @property
    def nmaps(self):
        return len(self.coefs)

    @property
    def coupling(self):
        return self.pol * self.mu / self.mass

    def device(self, torch_device):
        """Upload grid and coefficient tables once (device layout) and keep them."""
        import torch
        key = str(torch_device)
        if key not in self._dev:
            g = [torch.tensor(np.asarray(a, np.float64), device=torch_device) for a in self.grid]
            c = [torch.tensor(device_layout(k), device=torch_device) for k in self.coefs]
            self._dev[key] = (g, c)
        return self._dev[key]

    def params(self, s, torch_device):
        if self.nmaps > MAX_MAPS:
            raise ValueError(f"at most {MAX_MAPS} maps")
        for k in self.coefs:
            if k.shape[1] != 3:
                raise ValueError("maps must hold the three B components")
        g, c = self.device(torch_device)
        P = AdjParams()
        m = P.maps
        m.nmaps = self.nmaps
        m.nx, m.ny, m.nz = (len(a) for a in self.grid)
        for a in range(3):
            ga = np.asarray(self.grid[a])
            m.x0[a] = float(ga[0])
            m.inv_h[a] = (len(ga) - 1) / float(ga[-1] - ga[0])
            m.g[a] = g[a].data_ptr()
            P.lo[a], P.hi[a] = self.lo[a], self.hi[a]
        for k in range(self.nmaps):
            m.coef[k] = c[k].data_ptr()
            m.s[k] = float(s[k])
        P.dt, P.coupling, P.grav = self.dt, self.coupling, self.grav
        return P


def _lib():
    L = _native.lib()
    vp = C.c_void_p
    L.pt_rk4_forward_ckpt.argtypes = [vp, C.POINTER(AdjParams), C.c_int, vp, vp, vp, C.c_int, C.c_int]
    L.pt_rk4_adjoint_noalloc.argtypes = [vp, C.POINTER(AdjParams), C.c_int, vp, vp, vp, C.c_int, C.c_int]
    L.pt_rk4_adjoint.argtypes = [vp, C.POINTER(AdjParams), C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_int]
    return L


def rk4_forward(model, s, y0, nsteps, K=256, keep_ckpt=True):
    """y0: torch [n][6] on the GPU.  Returns (yT [n][6], checkpoints [nc][6][n] or None)."""
    import torch
    n = y0.shape[0]
    dev = y0.device
    y0s = y0.detach().t().contiguous()
    yT = torch.empty_like(y0s)
    nc = (nsteps + K - 1) // K
    ck = torch.empty((max(nc, 1), 6, n), dtype=torch.float64, device=dev) if keep_ckpt else None
    P = model.params([float(v) for v in s], dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = _lib().pt_rk4_forward_ckpt(stream, C.byref(P), n, y0s.data_ptr(), yT.data_ptr(),
                                    ck.data_ptr() if ck is not None else None, nsteps, K)
    _native.check(rc, "pt_rk4_forward_ckpt")
    return yT.t(), ck


def rk4_adjoint(model, s, ckpt, lamT, nsteps, K=256):
    """lamT: torch [n][6] = dJ/dyT.  Returns (dJ/dy0 [n][6], dJ/ds per particle [n][nmaps])."""
    import torch
    n = lamT.shape[0]
    dev = lamT.device
    lam = lamT.detach().t().contiguous().clone()
    sbar = torch.zeros((max(model.nmaps, 1), n), dtype=torch.float64, device=dev)
    seg = torch.empty((K, 24, n), dtype=torch.float64, device=dev)
    flags = torch.empty((K, n), dtype=torch.int32, device=dev)
    P = model.params([float(v) for v in s], dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = _lib().pt_rk4_adjoint(stream, C.byref(P), n, ckpt.data_ptr(), lam.data_ptr(), sbar.data_ptr(),
                               seg.data_ptr(), flags.data_ptr(), nsteps, K)
    _native.check(rc, "pt_rk4_adjoint")
    return lam.t(), sbar[:model.nmaps].t()


def rk4_forward_chunks(model, s, y0, nsteps, chunk, K=256, progress=None):
    """The forward pass of a long horizon in chunks of `chunk` steps (a multiple of K, so
    the checkpoints fall where one pass puts them): (yT, [checkpoints of each chunk],
    [steps of each chunk]).  progress(j, nchunks, steps) is called after each chunk."""
    if chunk < nsteps and chunk % K:
        raise ValueError("the chunk length must be a multiple of the checkpoint interval K")
    lens = [chunk] * (nsteps // chunk) + ([nsteps % chunk] if nsteps % chunk else [])
    cks, y = [], y0
    for j, m in enumerate(lens):
        y, ck = rk4_forward(model, s, y, m, K)
        cks.append(ck)
        if progress:
            progress(j, len(lens), m)
    return y, cks, lens


def rk4_adjoint_chunks(model, s, cks, lens, lamT, K=256, progress=None):
    """The adjoint of rk4_forward_chunks' pass: the chunks backwards, dJ/dy at a chunk's end
    the adjoint state at its start, dJ/ds summed over the chunks.  Returns (dJ/dy0 [n][6],
    dJ/ds per particle [n][nmaps]); each chunk's checkpoints are released once used."""
    lam, gs = lamT, None
    for j in reversed(range(len(lens))):
        lam, g = rk4_adjoint(model, s, cks[j], lam, lens[j], K)
        cks[j] = None
        gs = g if gs is None else gs + g
        if progress:
            progress(j, len(lens), lens[j])
    return lam, gs


def rk4_tracker(model, nsteps, K=256):
    """torch.autograd.Function applying the tracker: f(s, y0) -> yT."""
    import torch

    class RK4Tracker(torch.autograd.Function):
        @staticmethod
        def forward(ctx, s, y0):
            yT, ck = rk4_forward(model, s.detach().cpu().numpy(), y0, nsteps, K)
            ctx.save_for_backward(s, ck)
            return yT

        @staticmethod
        def backward(ctx, gyT):
            s, ck = ctx.saved_tensors
            gy0, gs_part = rk4_adjoint(model, s.detach().cpu().numpy(), ck, gyT.contiguous(), nsteps, K)
            return gs_part.sum(0).to(s.dtype).to(s.device), gy0

    return RK4Tracker.apply
