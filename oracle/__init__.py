"""FP64 CPU oracle for the lunar-orbit interferometer hot path.

TEST INFRASTRUCTURE ONLY: the checker that the HIP library is compared
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` import this package; the product (``viss_optimize_3d_amd``)
never does and fails loudly when its HIP library is missing.

The compiled part is ``oracle/viss_oracle.c`` (built by ``oracle/Makefile``
into ``oracle/_build/liboracle.so``); each C function cites the reference
file:line it restates.  This module wraps it with numpy and adds the small
pure-Python pieces used as independent cross-checks (``nest2ring`` by bit
twiddling, ``ang2pix_ring``) -- written from the HEALPix definition (Gorski et
al. 2005), not from the C code, so that a mistake in one does not hide in the
other.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

MASK = 1
PHASE_TRACK = 2
DCF_BLOCKAGE_SHIFT = 4
HERMITIAN = 16

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


def build(force: bool = False) -> str:
    """Compile the oracle with make (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_pix2ang_nest.argtypes = [_i64, _d, _d, _vp, _vp]
        L.oracle_pix2ang_ring.argtypes = [_i64, _d, _d, _vp, _vp]
        L.oracle_pix2ang_nest.restype = ctypes.c_int
        L.oracle_pix2ang_ring.restype = ctypes.c_int
        L.oracle_pixel_dirs.argtypes = [_d, _d, _d, _i64, ctypes.c_double, _d, _d, _d, _d]
        L.oracle_pixel_dirs.restype = None
        L.oracle_mask_threshold.argtypes = [ctypes.c_double] * 5
        L.oracle_mask_threshold.restype = ctypes.c_double
        L.oracle_forward.argtypes = [_d, _d, _d, _d, _i64, _d, _d, _d, _vp, _vp, _i64, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, _d, _d]
        L.oracle_forward.restype = None
        L.oracle_forward_ref32.argtypes = [_f, _f, _f, _f, _i64, _f, _f, _f, _f, _f, _i64, ctypes.c_float, _f, _f]
        L.oracle_forward_ref32.restype = None
        L.oracle_shell_index.argtypes = [ctypes.c_float]
        L.oracle_shell_index.restype = ctypes.c_int32
        L.oracle_weights.argtypes = [_f, _d, _i64, ctypes.c_int32, ctypes.c_int, _i32, _u32, _d, _d, _d]
        L.oracle_weights.restype = ctypes.c_int64
        L.oracle_backproject.argtypes = [_d, _d, _d, _i64, _d, _d, _d, _vp, _vp, _d, _d, _vp, _i64, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double, _d]
        L.oracle_backproject.restype = None
        L.oracle_count_visible.argtypes = [_d, _d, _d, _i64, _d, _d, _i64, ctypes.c_double, ctypes.c_double]
        L.oracle_count_visible.restype = ctypes.c_int64
        L.oracle_mask_flips.argtypes = [_d, _d, _d, _i64, _d, _d, _i64, ctypes.c_float, ctypes.c_double,
                                        ctypes.c_double, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
        L.oracle_mask_flips.restype = None
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- pix2ang
def pix2ang(nside: int, nest: bool):
    """(theta, phi, ring, iphi) for every pixel; ring/iphi are 1-based."""
    npix = 12 * nside * nside
    th = np.empty(npix, np.float64)
    ph = np.empty(npix, np.float64)
    ring = np.empty(npix, np.int32)
    iphi = np.empty(npix, np.int32)
    fn = lib().oracle_pix2ang_nest if nest else lib().oracle_pix2ang_ring
    rc = fn(nside, th, ph, _ptr(ring), _ptr(iphi))
    if rc != 0:
        raise ValueError(f"bad nside {nside}")
    return th, ph, ring, iphi


def pixel_dirs(theta, phi, B, scale):
    theta, phi, B = _c(theta, np.float64), _c(phi, np.float64), _c(B, np.float64)
    P = theta.size
    l, m, n, I = (np.empty(P) for _ in range(4))
    lib().oracle_pixel_dirs(theta, phi, B, P, float(scale), l, m, n, I)
    return l, m, n, I


# ------------------------------------------------------- independent checks
def ring_start(nside: int, ring):
    """0-based RING index of the first pixel of 1-based ring ``ring`` (vectorized)."""
    ring = np.asarray(ring, np.int64)
    ns = nside
    ncap = 2 * ns * (ns - 1)
    npix = 12 * ns * ns
    out = np.empty_like(ring)
    north = ring < ns
    south = ring > 3 * ns
    belt = ~(north | south)
    out[north] = 2 * ring[north] * (ring[north] - 1)
    out[belt] = ncap + (ring[belt] - ns) * 4 * ns
    i = 4 * ns - ring[south]
    out[south] = npix - 2 * i * (i + 1)
    return out


def nest2ring(nside: int, ipnest):
    """Independent NEST -> RING index map from the HEALPix face layout.

    Face f has its corner pixel (ix=iy=0, the face's southmost pixel in the
    "jr" sense of the ring numbering) on ring jrll[f]*nside - 1, and in-ring
    position given by jpll[f] (both are the HEALPix base-pixel tables, also
    pix2ang.cu:183-184).  Written with numpy bit operations rather than loops.
    """
    ipnest = np.asarray(ipnest, np.int64)
    ns = nside
    npface = ns * ns
    face = ipnest // npface
    ipf = ipnest % npface
    ix = np.zeros_like(ipf)
    iy = np.zeros_like(ipf)
    for b in range(31):
        ix |= ((ipf >> (2 * b)) & 1) << b
        iy |= ((ipf >> (2 * b + 1)) & 1) << b
    jrll = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4])
    jpll = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7])
    jr = jrll[face] * ns - (ix + iy) - 1
    nr = np.where(jr < ns, jr, np.where(jr > 3 * ns, 4 * ns - jr, ns))
    kshift = np.where((jr >= ns) & (jr <= 3 * ns), (jr - ns) & 1, 0)
    jp = (jpll[face] * nr + (ix - iy) + 1 + kshift) // 2
    jp = np.where(jp > 4 * ns, jp - 4 * ns, jp)
    jp = np.where(jp < 1, jp + 4 * ns, jp)
    return ring_start(ns, jr) + jp - 1


def ang2pix_ring(nside: int, theta, phi):
    """Independent RING ang2pix (HEALPix standard algorithm, Gorski 2005 sec. 4)."""
    z = np.cos(np.asarray(theta, np.float64))
    tt = np.mod(np.asarray(phi, np.float64), 2 * np.pi) / (0.5 * np.pi)  # in [0,4)
    za = np.abs(z)
    ns = nside
    out = np.empty(z.shape, np.int64)
    eq = za <= 2.0 / 3.0
    # equatorial region
    temp1 = ns * (0.5 + tt[eq])
    temp2 = ns * z[eq] * 0.75
    jp = np.floor(temp1 - temp2).astype(np.int64)
    jm = np.floor(temp1 + temp2).astype(np.int64)
    ir = ns + 1 + jp - jm  # in {1, 2n+1}
    kshift = 1 - (ir & 1)
    ip = (jp + jm - ns + kshift + 1) // 2
    ip = np.mod(ip, 4 * ns)
    out[eq] = 2 * ns * (ns - 1) + (ir - 1) * 4 * ns + ip
    # polar caps
    pc = ~eq
    tp = tt[pc] - np.floor(tt[pc])
    tmp = ns * np.sqrt(3 * (1 - za[pc]))
    jp = np.floor(tp * tmp).astype(np.int64)
    jm = np.floor((1.0 - tp) * tmp).astype(np.int64)
    ir = jp + jm + 1
    ip = np.floor(tt[pc] * ir).astype(np.int64)
    ip = np.mod(ip, 4 * ir)
    north = z[pc] > 0
    res = np.where(north, 2 * ir * (ir - 1) + ip, 12 * ns * ns - 2 * ir * (ir + 1) + ip)
    out[pc] = res
    return out


# ------------------------------------------------------------- the operator
def mask_threshold(x, y, z, R_moon, orbit_radius):
    return lib().oracle_mask_threshold(float(x), float(y), float(z), float(R_moon), float(orbit_radius))


def forward(l, m, n, I, uvw, xyz_a=None, xyz_b=None, flags=0, R_moon=0.0, orbit_radius=0.0):
    """V (complex128 [K], or [2K] with HERMITIAN).  uvw, xyz_*: (3, K) arrays."""
    l, m, n, I = (_c(a, np.float64) for a in (l, m, n, I))
    uvw = _c(uvw, np.float64)
    K = uvw.shape[1]
    xa = None if xyz_a is None else _c(xyz_a, np.float64)
    xb = None if xyz_b is None else _c(xyz_b, np.float64)
    if flags & MASK and (xa is None or xb is None):
        raise ValueError("mask needs satellite positions")
    rows = 2 * K if flags & HERMITIAN else K
    vr = np.empty(rows)
    vi = np.empty(rows)
    lib().oracle_forward(l, m, n, I, l.size, _c(uvw[0], np.float64), _c(uvw[1], np.float64), _c(uvw[2], np.float64),
                         _ptr(xa), _ptr(xb), K, int(flags), float(R_moon), float(orbit_radius), vr, vi)
    return vr + 1j * vi


def forward_ref32(l, m, n, B, uvw, xyz_a, xyz_b, phi_cut):
    f = np.float32
    l, m, n, B = (_c(a, f) for a in (l, m, n, B))
    uvw = _c(uvw, f)
    K = uvw.shape[1]
    vr = np.empty(K, f)
    vi = np.empty(K, f)
    lib().oracle_forward_ref32(l, m, n, B, l.size, _c(uvw[0], f), _c(uvw[1], f), _c(uvw[2], f), _c(xyz_a, f),
                               _c(xyz_b, f), K, float(phi_cut), vr, vi)
    return vr.astype(np.float64) + 1j * vi.astype(np.float64)


def shell_index(bll):
    return np.array([lib().oracle_shell_index(float(b)) for b in np.asarray(bll, np.float32).ravel()], np.int32)


def weights(bll, w, nr: int, flags: int = 0):
    """(gs, mb, dcf, dg, wt, n_out_of_range)."""
    bll = _c(bll, np.float32)
    w = _c(w, np.float64)
    K = bll.size
    gs = np.empty(K, np.int32)
    mb = np.empty(max(nr, 1), np.uint32)
    dcf = np.empty(nr + 1)
    dg = np.empty(K)
    wt = np.empty(K)
    bad = lib().oracle_weights(bll, w, K, int(nr), int(flags), gs, mb, dcf, dg, wt)
    return gs, mb[:nr], dcf, dg, wt, int(bad)


def backproject(l, m, n, uvw, y, wt=None, xyz_a=None, xyz_b=None, flags=0, R_moon=0.0, orbit_radius=0.0):
    """x (float64 [P]) = Re sum_k [mask] wt_k y_k exp(+2 pi i b_k.s_p)."""
    l, m, n = (_c(a, np.float64) for a in (l, m, n))
    uvw = _c(uvw, np.float64)
    K = uvw.shape[1]
    y = np.asarray(y, np.complex128)
    yr = _c(y.real, np.float64)
    yi = _c(y.imag, np.float64)
    wtc = None if wt is None else _c(wt, np.float64)
    xa = None if xyz_a is None else _c(xyz_a, np.float64)
    xb = None if xyz_b is None else _c(xyz_b, np.float64)
    if flags & MASK and (xa is None or xb is None):
        raise ValueError("mask needs satellite positions")
    x = np.empty(l.size)
    lib().oracle_backproject(l, m, n, l.size, _c(uvw[0], np.float64), _c(uvw[1], np.float64), _c(uvw[2], np.float64),
                             _ptr(xa), _ptr(xb), yr, yi, _ptr(wtc), K, int(flags), float(R_moon), float(orbit_radius),
                             x)
    return x


def count_visible(l, m, n, xyz_a, xyz_b, R_moon, orbit_radius=0.0):
    l, m, n = (_c(a, np.float64) for a in (l, m, n))
    xa = _c(xyz_a, np.float64)
    xb = _c(xyz_b, np.float64)
    return int(lib().oracle_count_visible(l, m, n, l.size, xa, xb, xa.shape[1], float(R_moon), float(orbit_radius)))


def mask_flips(l, m, n, xyz_a, xyz_b, phi_cut, R_moon, orbit_radius=0.0):
    """(acos_only, dot_only, both): (row, pixel) pairs that the reference's FP32
    acos mask against Phi keeps and the FP64 dot form s.r >= t drops, the
    reverse, and pairs both keep."""
    l, m, n = (_c(a, np.float64) for a in (l, m, n))
    xa = _c(xyz_a, np.float64)
    xb = _c(xyz_b, np.float64)
    out = np.zeros(3, np.int64)
    lib().oracle_mask_flips(l, m, n, l.size, xa, xb, xa.shape[1], float(phi_cut), float(R_moon), float(orbit_radius),
                            out)
    return int(out[0]), int(out[1]), int(out[2])


def reference_cut_angle(R_moon=1737.1e3, h=300e3):
    """Phi = pi - asin(R/(R+h)) (vissgen3d_noblockage.cu:240-243)."""
    return math.pi - math.asin(R_moon / (R_moon + h))
