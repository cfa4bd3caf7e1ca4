"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/viss.h declares, and rejects bad arguments before touching a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "viss.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(viss_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_table():
    from viss_optimize_3d_amd import _lib

    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from viss_optimize_3d_amd import _lib

    L = _lib.load()
    for name in _declared():
        assert hasattr(L, name), name
    assert L.viss_version() == 100
    assert L.viss_status_string(-1) == b"invalid argument"


def test_argument_validation_without_gpu():
    """Entry points validate sizes and pointers before any HIP call."""
    from viss_optimize_3d_amd import _lib

    L = _lib.load()
    rc = L.viss_pix2ang(0, 0, None, None, None, None, None)
    assert rc == -1
    assert b"nside" in L.viss_last_error()
    assert L.viss_pix2ang(16, 7, None, None, None, None, None) == -1
    assert L.viss_pixel_dirs(None, None, None, -1, 1.0, None, None, None) == -1
    assert L.viss_pixel_dirs(None, None, None, 0, 1.0, None, None, None) == 0  # empty input is a no-op
    assert L.viss_forward(None, None, 0, None, None, None, 0, None, 0, None, None) == 0
    assert L.viss_forward(None, None, -3, None, None, None, 0, None, 0, None, None) == -1
    assert L.viss_backproject(None, 10, None, None, None, None, None, 5, None, 0, None, None) == -1
    assert L.viss_weights(None, None, 10, 0, 0, None, None, None, None, None, None, None) == -1
    assert b"nr" in L.viss_last_error()
    assert L.viss_weights(None, None, 0, 4, 64, None, None, None, None, None, None, None) == -1
    out = ctypes.c_void_p()
    assert L.viss_sky_create(None, 10, ctypes.byref(out), None) == -1
    assert L.viss_sky_destroy(None) == 0
    assert L.viss_sky_npix(None) == -1
    assert L.viss_sky_forward(None, None, None, None, None, 1, None, 0, None, None) == -1
    # the allocation-free back-projection validates before touching a GPU
    assert L.viss_sky_backproject_noalloc(None, None, None, None, None, None, 5, None, 0, None, None) == -1
    assert b"sky" in L.viss_last_error()
