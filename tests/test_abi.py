"""The C-ABI library loads and exports every entry point include/pentrack_gpu.h declares."""
import ctypes
import os
import re

from conftest import ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "pentrack_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_expected_api():
    names = declared_functions()
    for n in ["pt_geometry_create", "pt_fields_create", "pt_tricubic_precompute", "pt_batch_create", "pt_batch_run",
              "pt_batch_download", "pt_track_batch", "pt_request_quit"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from pentrack_amd import _native
    lib = _native.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_native.EXPORTS) == declared_functions()


def test_struct_layouts_match_header():
    from pentrack_amd import _native
    # pt_particle_out: 1+9+5 doubles, 4 ints, 2 int64, 3 doubles, 2 ints
    assert ctypes.sizeof(_native.ParticleOut) == 8 * 15 + 16 + 16 + 24 + 8
    assert ctypes.sizeof(_native.HitRec) == 8 + 12 + 4 + 9 * 8 + 16 + 24


def test_host_entry_points_reject_bad_arguments():
    """Status codes instead of exceptions: invalid inputs return PT_EINVAL without touching a GPU."""
    from pentrack_amd import _native
    lib = _native.lib()
    h = ctypes.c_void_p()
    assert lib.pt_geometry_create(-1, None, None, 0, None, None, None, None, 0, 0, None, 4, ctypes.byref(h)) == -1
    assert b"pt_geometry_create" in lib.pt_last_error()
    assert lib.pt_fields_create(-1, None, 0, ctypes.byref(h)) == -1
    assert lib.pt_tricubic_precompute(1, 2, 2, None, None, None, 1, None, None, 1) == -1
    assert lib.pt_batch_run(None, None, None) == -1
