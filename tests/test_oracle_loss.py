"""The per-particle loss per bounce (a build extension beside the per-material LossPerBounce,
include/geometry.h:30) in the oracle's TNeutron::OnHit (src/neutron.cpp:241): on total
reflection the absorption probability is 1 - |R|^2 + loss = loss (|R| = 1 below the Fermi
potential), so a loss of 1 absorbs at the first bounce and 0 never does."""
import numpy as np

from oracle import ptref
from pentrack_amd.constants import NEUTRON
from pentrack_amd.fields import TFieldManager
from pentrack_amd.geometry import Material, Solid, TGeometry, hollow_box_mesh
from pentrack_amd.particle import SpeciesConfig
from pentrack_amd.workloads import Workload


def test_loss_per_particle(oracle_lib):
    n = 40
    geom = TGeometry([Material("vacuum"), Material("wall", FermiReal=335.0, LossPerBounce=0.0)], "vacuum",
                     [Solid(2, "wall", hollow_box_mesh((0, 0, 0), (1, 1, 1), 0.001))])
    y0 = np.zeros((n, 9))
    y0[:, :3] = [0.5, 0.5, 0.5]
    y0[:, 5] = -2.0                                   # down onto the floor, E_perp << 335 neV
    y0[:, 7] = 1.0
    sp = [SpeciesConfig(kind=NEUTRON, tau=0.0, tmax=9e99, lmax=9e99)]
    loss = np.where(np.arange(n) % 2 == 0, 1.0, 0.0)
    w = Workload("loss", geom, TFieldManager([]), sp, 2.0, np.zeros(n, np.int32), np.zeros(n), y0, np.zeros((n, 5)),
                 np.arange(n, dtype=np.uint64), 1, loss=loss)
    o, _ = ptref.track(w, "fp64")
    assert (o["stopID"][0::2] == 2).all() and (o["Nhit"][0::2] == 1).all()
    assert (o["stopID"][1::2] == -1).all() and (o["Nhit"][1::2] > 2).all()
    # without the per-particle values, the material's loss (0) applies to every neutron
    w.loss = None
    o2, _ = ptref.track(w, "fp64")
    assert (o2["stopID"] == -1).all()
    # a loss of 1 in the material itself absorbs everyone
    geom2 = TGeometry([Material("vacuum"), Material("wall", FermiReal=335.0, LossPerBounce=1.0)], "vacuum",
                      [Solid(2, "wall", hollow_box_mesh((0, 0, 0), (1, 1, 1), 0.001))])
    w2 = Workload("loss", geom2, TFieldManager([]), sp, 2.0, np.zeros(n, np.int32), np.zeros(n), y0, np.zeros((n, 5)),
                  np.arange(n, dtype=np.uint64), 1)
    o3, _ = ptref.track(w2, "fp64")
    assert (o3["stopID"] == 2).all()
