"""ctypes binding of libpentrack_gpu.so (the C ABI in include/pentrack_gpu.h).

The library is built in-tree by pentrack_amd.build (hipcc, gfx950).  Loading
fails loudly when it is missing: there is no CPU fallback for the tracking path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpentrack_gpu.so")

PT_OK = 0
PT_EINVAL = -1
PT_ENOMEM = -2
PT_EDEVICE = -3
PT_FIELD_TRICUBIC = 1
PT_FIELD_UNIFORM_E = 2
PT_FIELD_CONDUCTOR = 3
PT_FIELD_EXPONENTIAL_X = 4
PT_FIELD_LINEAR_Z = 5
PT_FIELD_B0_GRAD_Z = 6
PT_FIELD_B0_GRAD_X2 = 7
PT_FIELD_B0_GRAD_XY = 8
PT_FIELD_B0_XY = 9
PT_FIELD_TAB2D = 10
PT_VARIANT_FP64 = 0
PT_VARIANT_MIXED = 1

# stop IDs, include/globals.h:13-23 (+ extensions)
ID_UNKNOWN = 0
ID_NOT_FINISH = -1
ID_HIT_BOUNDARIES = -2
ID_ODEINT_ERROR = -3
ID_DECAYED = -4
ID_CGAL_ERROR = -6
ID_GEOMETRY_ERROR = -7
ID_INVARIANT = -8
ID_UNSUPPORTED = -9
ID_ABSORBED_IN_MATERIAL = 1
ID_ABSORBED_ON_SURFACE = 2


class FieldDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("bscale", C.c_double), ("escale", C.c_double),
        ("box", C.c_double * 6), ("bw", C.c_double),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xg", C.POINTER(C.c_double)), ("yg", C.POINTER(C.c_double)), ("zg", C.POINTER(C.c_double)),
        ("has_b", C.c_int), ("has_v", C.c_int),
        ("coef", C.POINTER(C.c_double)),
        ("e", C.c_double * 3),
        ("p", C.c_double * 8),
        ("bscale_formula", C.c_char_p), ("escale_formula", C.c_char_p),
        ("values", C.POINTER(C.c_double)),
        ("mixed_tol", C.c_double),
    ]


class Species(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("q", C.c_double), ("m", C.c_double), ("mu", C.c_double), ("gamma", C.c_double),
        ("tau", C.c_double), ("tmax", C.c_double), ("lmax", C.c_double),
        ("flipspin", C.c_int), ("Bmax", C.c_double),
        ("n_spintimes", C.c_int), ("spintimes", C.c_double * 8), ("interpolatefields", C.c_int),
    ]


class ParticleOut(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("y", C.c_double * 9), ("spin", C.c_double * 5),
        ("stopID", C.c_int32), ("Nhit", C.c_int32), ("Nstep", C.c_int32), ("Nspinflip", C.c_int32),
        ("n_acc", C.c_int64), ("n_rej", C.c_int64),
        ("Hmax", C.c_double), ("noflipprob", C.c_double), ("tau", C.c_double),
        ("solidend", C.c_int32), ("n_hits", C.c_int32),
    ]


class HitRec(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("tri", C.c_int32), ("leaving", C.c_int32), ("entering", C.c_int32),
        ("normal", C.c_double * 3), ("v1", C.c_double * 3), ("v2", C.c_double * 3),
        ("pol1", C.c_double), ("pol2", C.c_double), ("pos", C.c_double * 3),
    ]


class SecondaryRec(C.Structure):
    _fields_ = [
        ("parent", C.c_int32), ("kind", C.c_int32), ("t", C.c_double), ("pos", C.c_double * 3),
        ("E", C.c_double), ("phi", C.c_double), ("theta", C.c_double), ("pol", C.c_int32), ("pad", C.c_int32),
    ]


class LaunchCfg(C.Structure):
    _fields_ = [
        ("variant", C.c_int), ("block", C.c_int), ("steps_per_launch", C.c_int),
        ("persistent", C.c_int), ("grid", C.c_int), ("rounds", C.c_int), ("graph", C.c_int),
        ("spatial_order", C.c_int), ("collision_order", C.c_int),
    ]


# numpy views of the output records
OUT_DTYPE = np.dtype([
    ("t", "<f8"), ("y", "<f8", (9,)), ("spin", "<f8", (5,)),
    ("stopID", "<i4"), ("Nhit", "<i4"), ("Nstep", "<i4"), ("Nspinflip", "<i4"),
    ("n_acc", "<i8"), ("n_rej", "<i8"),
    ("Hmax", "<f8"), ("noflipprob", "<f8"), ("tau", "<f8"),
    ("solidend", "<i4"), ("n_hits", "<i4"),
], align=True)
HIT_DTYPE = np.dtype([
    ("t", "<f8"), ("tri", "<i4"), ("leaving", "<i4"), ("entering", "<i4"),
    ("normal", "<f8", (3,)), ("v1", "<f8", (3,)), ("v2", "<f8", (3,)),
    ("pol1", "<f8"), ("pol2", "<f8"), ("pos", "<f8", (3,)),
], align=True)
SEC_DTYPE = np.dtype([
    ("parent", "<i4"), ("kind", "<i4"), ("t", "<f8"), ("pos", "<f8", (3,)),
    ("E", "<f8"), ("phi", "<f8"), ("theta", "<f8"), ("pol", "<i4"), ("pad", "<i4"),
], align=True)
SNAP_DTYPE = np.dtype([
    ("t", "<f8"), ("y", "<f8", (9,)), ("spin", "<f8", (5,)), ("Hmax", "<f8"), ("noflipprob", "<f8"),
    ("stopID", "<i4"), ("Nhit", "<i4"), ("Nstep", "<i4"), ("Nspinflip", "<i4"),
], align=True)
assert SNAP_DTYPE.itemsize == 152
TRACK_DTYPE = np.dtype([("t", "<f8"), ("y", "<f8", (9,)), ("solid", "<i4"), ("pad", "<i4")], align=True)
assert TRACK_DTYPE.itemsize == 88
SPIN_DTYPE = np.dtype([("t", "<f8"), ("pos", "<f8", (3,)), ("spin", "<f8", (3,)), ("W", "<f8", (3,))])
assert SPIN_DTYPE.itemsize == 80
assert OUT_DTYPE.itemsize == C.sizeof(ParticleOut), (OUT_DTYPE.itemsize, C.sizeof(ParticleOut))
assert SEC_DTYPE.itemsize == C.sizeof(SecondaryRec), (SEC_DTYPE.itemsize, C.sizeof(SecondaryRec))
assert HIT_DTYPE.itemsize == C.sizeof(HitRec), (HIT_DTYPE.itemsize, C.sizeof(HitRec))

# every symbol include/pentrack_gpu.h declares (checked by the test suite)
EXPORTS = [
    "pt_geometry_create", "pt_geometry_destroy", "pt_geometry_info",
    "pt_fields_create", "pt_fields_destroy", "pt_tricubic_precompute",
    "pt_batch_create", "pt_batch_destroy", "pt_batch_run", "pt_batch_active",
    "pt_batch_download", "pt_batch_step_counts", "pt_track_batch", "pt_batch_set_profiling",
    "pt_batch_kernel_times",
    "pt_request_quit", "pt_last_error", "pt_clearance_grid", "pt_tricubic_matrix",
    "pt_spline_grid_diff", "pt_rk4_forward_ckpt", "pt_rk4_adjoint", "pt_formula_eval",
    "pt_batch_secondaries", "pt_proton_beta_table", "pt_neutron_decay_host", "pt_beta_sample",
    "pt_distribution_table", "pt_fields_eval", "pt_fields_eval_tables", "pt_geometry_solid_at", "pt_geometry_segment_hits", "pt_bicubic_precompute",
    "pt_tricubic_precompute_gpu", "pt_batch_set_snapshots", "pt_batch_snapshots",
    "pt_batch_set_track", "pt_batch_track", "pt_batch_set_spinlog", "pt_batch_spinlog",
    "pt_rk4_adjoint_noalloc", "pt_fields_mixed_info", "pt_gather_bandwidth", "pt_gather_bandwidth_coop", "pt_batch_set_loss",
    "pt_seg_tri",
]

_lib = None


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def lib():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc, gfx950). The tracking path has no CPU fallback.")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    ip = C.POINTER(C.c_int32)
    L.pt_geometry_create.argtypes = [C.c_int, C.POINTER(C.c_float), ip, C.c_int, C.POINTER(C.c_uint32), ip, ip,
                                     C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int,
                                     C.POINTER(vp)]
    L.pt_geometry_destroy.argtypes = [vp]
    L.pt_geometry_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pt_fields_create.argtypes = [C.c_int, C.POINTER(FieldDesc), C.c_int, C.POINTER(vp)]
    L.pt_fields_destroy.argtypes = [vp]
    L.pt_fields_mixed_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.pt_batch_set_loss.argtypes = [vp, C.POINTER(C.c_double)]
    L.pt_gather_bandwidth.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
    L.pt_gather_bandwidth_coop.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.pt_tricubic_precompute.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.c_int]
    L.pt_tricubic_matrix.argtypes = [ip]
    L.pt_seg_tri.argtypes = [C.POINTER(C.c_double)] * 6
    L.pt_tricubic_precompute_gpu.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double)]
    L.pt_bicubic_precompute.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.pt_batch_create.argtypes = [vp, vp, C.c_int, C.POINTER(Species), C.c_double, C.c_int, ip,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.c_uint64, C.POINTER(C.c_uint64), C.c_int, C.POINTER(vp)]
    L.pt_batch_destroy.argtypes = [vp]
    L.pt_batch_run.argtypes = [vp, vp, C.POINTER(LaunchCfg)]
    L.pt_batch_active.argtypes = [vp, vp, C.POINTER(C.c_int)]
    L.pt_batch_download.argtypes = [vp, vp, vp, vp]
    L.pt_batch_step_counts.argtypes = [vp, vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.pt_batch_set_profiling.argtypes = [vp, C.c_int]
    L.pt_batch_set_snapshots.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.c_int]
    L.pt_batch_snapshots.argtypes = [vp, vp, vp, C.POINTER(C.c_int32)]
    L.pt_batch_set_track.argtypes = [vp, C.c_double, C.c_int]
    L.pt_batch_track.argtypes = [vp, vp, vp, C.POINTER(C.c_int32)]
    L.pt_batch_set_spinlog.argtypes = [vp, C.c_double, C.c_int]
    L.pt_batch_spinlog.argtypes = [vp, vp, vp, C.POINTER(C.c_int32)]
    L.pt_batch_kernel_times.argtypes = [vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.pt_track_batch.argtypes = [vp, vp, vp, C.c_int, C.POINTER(Species), C.c_double, C.c_int, ip,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.c_uint64, C.POINTER(C.c_uint64), vp, vp, C.c_int, C.POINTER(LaunchCfg)]
    L.pt_request_quit.argtypes = [C.c_int]
    L.pt_spline_grid_diff.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.pt_clearance_grid.argtypes = [C.c_int, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.pt_formula_eval.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    dp = C.POINTER(C.c_double)
    L.pt_batch_secondaries.argtypes = [vp, vp, C.c_int, vp, C.POINTER(C.c_int)]
    L.pt_proton_beta_table.argtypes = [C.c_int, dp, dp, dp, dp, dp, C.POINTER(C.c_int)]
    L.pt_beta_sample.argtypes = [C.c_int, dp, dp]
    L.pt_distribution_table.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_int, dp, dp, dp, dp, dp,
                                         C.POINTER(C.c_int)]
    L.pt_fields_eval.argtypes = [vp, vp, C.c_int, dp, dp, dp, dp, dp, dp]
    L.pt_fields_eval_tables.argtypes = [vp, vp, C.c_int, dp, dp, C.c_int, C.c_int, dp, dp, dp, dp]
    L.pt_geometry_solid_at.argtypes = [vp, vp, C.c_int, dp, dp, C.POINTER(C.c_int32)]
    L.pt_geometry_segment_hits.argtypes = [vp, vp, C.c_int, dp, dp, C.c_int, C.POINTER(C.c_int32)]
    L.pt_neutron_decay_host.argtypes = [C.c_uint64, C.c_uint64, dp, dp, C.POINTER(C.c_int32), C.POINTER(C.c_uint64)]
    L.pt_request_quit.restype = None
    L.pt_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(rc, what):
    if rc != PT_OK:
        msg = lib().pt_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed with status {rc}: {msg}")
