"""ctypes binding of libviss_hip.so (the C ABI declared in include/viss.h).

The library is built in-tree by ``viss_optimize_3d_amd.build``.  There is no
CPU fallback: if the library is missing, loading raises immediately.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libviss_hip.so")

# flags and constants (include/viss.h)
OK = 0
MASK = 1
PHASE_TRACK = 2
DCF_BLOCKAGE_SHIFT = 4
PREC_FP64 = 8
HERMITIAN = 16
SCHEME_RING = 0
SCHEME_NEST = 1
NOALLOC_SCRATCH_BYTES = 32 << 20

# Every entry point include/viss.h declares (tests check the export table).
EXPORTS = (
    "viss_version", "viss_status_string", "viss_last_error", "viss_pix2ang", "viss_pixel_dirs", "viss_forward",
    "viss_backproject", "viss_weights", "viss_sky_create", "viss_sky_destroy", "viss_sky_npix", "viss_sky_pairs",
    "viss_sky_quads", "viss_sky_forward",
    "viss_sky_backproject", "viss_sky_backproject_noalloc", "viss_sky_count_visible", "viss_sky_count_work", "viss_batch_create", "viss_batch_destroy", "viss_batch_forward",
    "viss_batch_autotune", "viss_batch_tuning_save", "viss_batch_tuning_load", "viss_batch_stats",
    "viss_batch_weights", "viss_batch_backproject",
    "viss_image_accumulate", "viss_reorder", "viss_sky_backproject_snapshots", "viss_sky_forward_snapshots",
)


class VissError(RuntimeError):
    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} failed with status {status}: {msg}")
        self.status = status


class MaskGeometry(ctypes.Structure):
    _fields_ = [("R_moon", ctypes.c_double), ("orbit_radius", ctypes.c_double)]


_lib = None
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_int = ctypes.c_int
_geom_p = ctypes.POINTER(MaskGeometry)


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libviss_hip.so not found at {LIB_PATH}: build it with "
                          "`python -m viss_optimize_3d_amd.build` (there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "viss_version": ([], _int),
        "viss_status_string": ([_int], ctypes.c_char_p),
        "viss_last_error": ([], ctypes.c_char_p),
        "viss_pix2ang": ([_i64, _int, _vp, _vp, _vp, _vp, _vp], _int),
        "viss_pixel_dirs": ([_vp, _vp, _vp, _i64, ctypes.c_double, _vp, _vp, _vp], _int),
        "viss_forward": ([_vp, _vp, _i64, _vp, _vp, _vp, _i64, _geom_p, _int, _vp, _vp], _int),
        "viss_backproject": ([_vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _geom_p, _int, _vp, _vp], _int),
        "viss_weights": ([_vp, _vp, _i64, _i32, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _int),
        "viss_sky_create": ([_vp, _i64, ctypes.POINTER(_vp), _vp], _int),
        "viss_sky_destroy": ([_vp], _int),
        "viss_sky_npix": ([_vp], _i64),
        "viss_sky_pairs": ([_vp, _vp, _vp, _vp], _int),
        "viss_sky_quads": ([_vp, _vp, _vp, _vp], _int),
        "viss_sky_forward": ([_vp, _vp, _vp, _vp, _vp, _i64, _geom_p, _int, _vp, _vp], _int),
        "viss_sky_backproject": ([_vp, _vp, _vp, _vp, _vp, _vp, _i64, _geom_p, _int, _vp, _vp], _int),
        "viss_sky_backproject_noalloc": ([_vp, _vp, _vp, _vp, _vp, _vp, _i64, _geom_p, _int, _vp, _vp], _int),
        "viss_sky_count_visible": ([_vp, _vp, _vp, _i64, _geom_p, _vp, _vp], _int),
        "viss_sky_count_work": ([_vp, _vp, _vp, _i64, _geom_p, _vp, _vp], _int),
        "viss_batch_create": ([_vp, _i64, _vp, _vp, _vp, _vp, _geom_p, _int, ctypes.POINTER(_vp), _vp], _int),
        "viss_batch_destroy": ([_vp], _int),
        "viss_batch_forward": ([_vp, _vp, _vp, _int, _vp], _int),
        "viss_batch_autotune": ([_vp, _vp, _vp, _int, _vp], _int),
        "viss_batch_tuning_save": ([_vp, ctypes.c_char_p], _int),
        "viss_batch_tuning_load": ([_vp, ctypes.c_char_p, _vp], _int),
        "viss_batch_stats": ([_vp, _vp, _vp, _vp, _vp], _int),
        "viss_batch_weights": ([_vp, _vp, _i32, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _int),
        "viss_batch_backproject": ([_vp, _vp, _vp, _int, _vp, _int, _int, _vp], _int),
        "viss_image_accumulate": ([_vp, _i64, _int, _vp, _vp], _int),
        "viss_reorder": ([_vp, _i64, _int, _vp, _vp], _int),
        "viss_sky_backproject_snapshots": ([_vp, _vp, _i64, _i32, ctypes.c_double, _vp, _vp, _geom_p, _int, _vp, _vp],
                                           _int),
        "viss_sky_forward_snapshots": ([_vp, _vp, _vp, _i64, _i32, ctypes.c_double, _geom_p, _int, _vp, _vp], _int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(fn: str, status: int):
    if status != OK:
        msg = load().viss_last_error().decode(errors="replace")
        raise VissError(fn, status, msg)
