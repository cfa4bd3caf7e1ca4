"""Torch-tensor front end of the HIP library.

Mirrors the reference's kernels as operators on device tensors (the
reference's names in brackets):

* :func:`pix2ang`      [mk_pix2xy + f_pix2and_nest, pix2ang.cu:61-160]
* :func:`pixel_dirs`   [healpix_moonback_pre, vissgen3d_noblockage.cu:52-73]
* :class:`Sky`.forward [healpix_moonback_viss, vissgen3d_noblockage.cu:76-102]
* :class:`Sky`.backproject [computeC, vissgen3d_noblockage.cu:167-182 and
  vissgen3d_blockage.cu:167-192; the adjoint when ``wt`` is None]
* :func:`weights`      [ceilAndScale ... viss_gamma_trans, vissgen3d_noblockage.cu:105-164]

Torch only supplies device memory and the current stream; every computation is
a call into libviss_hip.so.  Tensors must live on a HIP device; inputs are
made contiguous in the dtype the C ABI declares.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib
from ._lib import MaskGeometry, check

REF_R_MOON = 1737.1e3
REF_ORBIT_RADIUS = 1737.1e3 + 300e3


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _dev(t: torch.Tensor, dtype) -> torch.Tensor:
    if not t.is_cuda:
        raise ValueError("viss ops take device tensors")
    return t.to(dtype).contiguous()


def _flags(mask=False, phase_track=False, fp64=False, hermitian=False):
    return ((_lib.MASK if mask else 0) | (_lib.PHASE_TRACK if phase_track else 0) |
            (_lib.PREC_FP64 if fp64 else 0) | (_lib.HERMITIAN if hermitian else 0))


def _geom(R_moon: float, orbit_radius: float):
    return ctypes.byref(MaskGeometry(float(R_moon), float(orbit_radius)))


def pix2ang(nside: int, nest: bool = True, device="cuda", with_ring: bool = True):
    """theta, phi (float64 [12 nside^2]) and, optionally, 1-based ring / in-ring index (int32)."""
    L = _lib.load()
    P = 12 * nside * nside
    th = torch.empty(P, dtype=torch.float64, device=device)
    ph = torch.empty(P, dtype=torch.float64, device=device)
    ring = torch.empty(P, dtype=torch.int32, device=device) if with_ring else None
    iphi = torch.empty(P, dtype=torch.int32, device=device) if with_ring else None
    check("viss_pix2ang", L.viss_pix2ang(nside, _lib.SCHEME_NEST if nest else _lib.SCHEME_RING, _p(th), _p(ph),
                                         _p(ring), _p(iphi), _stream()))
    return (th, ph, ring, iphi) if with_ring else (th, ph)


def pixel_dirs(theta: torch.Tensor, phi: torch.Tensor, B: Optional[torch.Tensor] = None,
               scale: Optional[float] = None):
    """lmn (float64 [3, P]) and I = B * scale (default scale 4 pi / P)."""
    L = _lib.load()
    theta = _dev(theta, torch.float64)
    phi = _dev(phi, torch.float64)
    P = theta.numel()
    lmn = torch.empty(3, P, dtype=torch.float64, device=theta.device)
    I = None
    if B is not None:
        B = _dev(B, torch.float64)
        I = torch.empty(P, dtype=torch.float64, device=theta.device)
        if scale is None:
            scale = 4.0 * torch.pi / P
    check("viss_pixel_dirs", L.viss_pixel_dirs(_p(theta), _p(phi), _p(B), P, float(scale or 0.0), _p(lmn), _p(I),
                                               _stream()))
    return lmn, I


class Sky:
    """A pixel set sorted into compact tiles on the device (``viss_sky``)."""

    def __init__(self, lmn: torch.Tensor):
        L = _lib.load()
        self.lmn = _dev(lmn, torch.float64)
        if self.lmn.dim() != 2 or self.lmn.shape[0] != 3:
            raise ValueError("lmn must be [3, P]")
        self.P = self.lmn.shape[1]
        self.device = self.lmn.device
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check("viss_sky_create", L.viss_sky_create(_p(self.lmn), self.P, ctypes.byref(h), _stream()))
        self._h = h

    def pairs(self):
        """Antipodal pairs and unpaired pixels found at creation, and whether the
        mixed per-row kernels use the pairs; equatorial-mirror quads, their
        16-quad tiles, whether the mixed per-row forward uses them, and whether
        it forms their phases by the azimuth recurrence (HEALPix rings)."""
        n_pair, n_single, uses = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        check("viss_sky_pairs", _lib.load().viss_sky_pairs(self._h, ctypes.byref(n_pair), ctypes.byref(n_single),
                                                           ctypes.byref(uses)))
        out = {"pairs": n_pair.value, "singles": n_single.value, "uses_pairs": bool(uses.value)}
        n_quad, n_qtile, uses_q = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        check("viss_sky_quads", _lib.load().viss_sky_quads(self._h, ctypes.byref(n_quad), ctypes.byref(n_qtile),
                                                           ctypes.byref(uses_q)))
        out.update({"quads": n_quad.value, "quad_tiles": n_qtile.value, "uses_quads": bool(uses_q.value),
                    "quad_recurrence": uses_q.value == 2})
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            torch.cuda.synchronize(self.device)
            _lib.load().viss_sky_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _rows(self, uvw, xyz_a, xyz_b, mask):
        uvw = _dev(uvw, torch.float64)
        if uvw.dim() != 2 or uvw.shape[0] != 3:
            raise ValueError("uvw must be [3, K]")
        K = uvw.shape[1]
        xa = xb = None
        if mask:
            if xyz_a is None or xyz_b is None:
                raise ValueError("mask=True needs xyz_a and xyz_b")
            xa = _dev(xyz_a, torch.float64)
            xb = _dev(xyz_b, torch.float64)
            if tuple(xa.shape) != (3, K) or tuple(xb.shape) != (3, K):
                raise ValueError("xyz_a / xyz_b must be [3, K]")
        return uvw, xa, xb, K

    def forward(self, I: torch.Tensor, uvw: torch.Tensor, xyz_a=None, xyz_b=None, *, mask: bool = False,
                R_moon: float = REF_R_MOON, orbit_radius: float = 0.0, phase_track: bool = False,
                hermitian: bool = False, fp64: bool = False, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """V (complex128 [K] or [2K]).  Accumulates into ``out`` when given."""
        L = _lib.load()
        I = _dev(I, torch.float64)
        if I.numel() != self.P:
            raise ValueError("I must have P entries")
        uvw, xa, xb, K = self._rows(uvw, xyz_a, xyz_b, mask)
        rows = 2 * K if hermitian else K
        if out is None:
            out = torch.zeros(rows, dtype=torch.complex128, device=self.device)
        elif out.dtype != torch.complex128 or out.numel() != rows or not out.is_contiguous():
            raise ValueError("out must be a contiguous complex128 tensor of K (2K) rows")
        check("viss_sky_forward", L.viss_sky_forward(self._h, _p(I), _p(uvw), _p(xa), _p(xb), K,
                                                     _geom(R_moon, orbit_radius),
                                                     _flags(mask, phase_track, fp64, hermitian), _p(out), _stream()))
        return out

    def backproject(self, uvw: torch.Tensor, y: torch.Tensor, wt: Optional[torch.Tensor] = None, xyz_a=None,
                    xyz_b=None, *, mask: bool = False, R_moon: float = REF_R_MOON, orbit_radius: float = 0.0,
                    phase_track: bool = False, hermitian: bool = False, fp64: bool = False,
                    out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """x (float64 [P]) = herm * Re sum_k [mask] wt_k y'_k exp(+2 pi i b_k . s_p)."""
        L = _lib.load()
        uvw, xa, xb, K = self._rows(uvw, xyz_a, xyz_b, mask)
        y = _dev(y, torch.complex128)
        if y.numel() != K:
            raise ValueError("y must have K entries")
        w = None if wt is None else _dev(wt, torch.float64)
        if out is None:
            out = torch.zeros(self.P, dtype=torch.float64, device=self.device)
        check("viss_sky_backproject", L.viss_sky_backproject(self._h, _p(uvw), _p(xa), _p(xb), _p(y), _p(w), K,
                                                             _geom(R_moon, orbit_radius),
                                                             _flags(mask, phase_track, fp64, hermitian), _p(out),
                                                             _stream()))
        return out

    def forward_snapshots(self, I: torch.Tensor, sat_pos: torch.Tensor, wavelength: float, *, mask: bool = False,
                          R_moon: float = REF_R_MOON, orbit_radius: float = 0.0, phase_track: bool = False,
                          hermitian: bool = False, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Forward of snapshot-structured rows (sat_pos [S, N, 3]; rows = pairs a<b per snapshot)."""
        L = _lib.load()
        I = _dev(I, torch.float64)
        pos = _dev(sat_pos, torch.float64)
        if pos.dim() != 3 or pos.shape[2] != 3:
            raise ValueError("sat_pos must be [S, N, 3]")
        S, N = pos.shape[0], pos.shape[1]
        rows = S * N * (N - 1) // 2 * (2 if hermitian else 1)
        if out is None:
            out = torch.zeros(rows, dtype=torch.complex128, device=self.device)
        elif out.dtype != torch.complex128 or out.numel() != rows or not out.is_contiguous():
            raise ValueError("out must be a contiguous complex128 tensor of the row count")
        check("viss_sky_forward_snapshots",
              L.viss_sky_forward_snapshots(self._h, _p(I), _p(pos), S, N, float(wavelength),
                                           _geom(R_moon, orbit_radius), _flags(mask, phase_track, False, hermitian),
                                           _p(out), _stream()))
        return out

    def backproject_snapshots(self, sat_pos: torch.Tensor, wavelength: float, y: torch.Tensor,
                              wt: Optional[torch.Tensor] = None, *, mask: bool = False, R_moon: float = REF_R_MOON,
                              orbit_radius: float = 0.0, phase_track: bool = False, hermitian: bool = False,
                              out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Back-projection of snapshot-structured rows (sat_pos [S, N, 3]; rows = pairs a<b per snapshot)."""
        L = _lib.load()
        pos = _dev(sat_pos, torch.float64)
        if pos.dim() != 3 or pos.shape[2] != 3:
            raise ValueError("sat_pos must be [S, N, 3]")
        S, N = pos.shape[0], pos.shape[1]
        y = _dev(y, torch.complex128)
        if y.numel() != S * N * (N - 1) // 2:
            raise ValueError("y must have S * N(N-1)/2 entries")
        w = None if wt is None else _dev(wt, torch.float64)
        if out is None:
            out = torch.zeros(self.P, dtype=torch.float64, device=self.device)
        check("viss_sky_backproject_snapshots",
              L.viss_sky_backproject_snapshots(self._h, _p(pos), S, N, float(wavelength), _p(y), _p(w),
                                               _geom(R_moon, orbit_radius), _flags(mask, phase_track, False, hermitian),
                                               _p(out), _stream()))
        return out

    def count_visible(self, xyz_a, xyz_b, *, R_moon: float = REF_R_MOON, orbit_radius: float = 0.0) -> torch.Tensor:
        L = _lib.load()
        xa = _dev(xyz_a, torch.float64)
        xb = _dev(xyz_b, torch.float64)
        K = xa.shape[1]
        cnt = torch.zeros(1, dtype=torch.int64, device=self.device)
        check("viss_sky_count_visible", L.viss_sky_count_visible(self._h, _p(xa), _p(xb), K,
                                                                 _geom(R_moon, orbit_radius), _p(cnt), _stream()))
        return cnt


    def count_work(self, xyz_a, xyz_b, *, R_moon: float = REF_R_MOON, orbit_radius: float = 0.0) -> dict:
        """Work of the mixed per-row quad kernels for these rows (viss_sky_count_work)."""
        L = _lib.load()
        xa = _dev(xyz_a, torch.float64)
        xb = _dev(xyz_b, torch.float64)
        cnt = torch.zeros(6, dtype=torch.int64, device=self.device)
        check("viss_sky_count_work", L.viss_sky_count_work(self._h, _p(xa), _p(xb), xa.shape[1],
                                                           _geom(R_moon, orbit_radius), _p(cnt), _stream()))
        c = cnt.cpu().tolist()
        return {"visible_pairs": c[0], "row_tiles": c[1], "row_quads": c[2], "row_rings": c[3], "row_rest_pairs": c[4],
                "row_runs": c[5]}


def weights(bll: torch.Tensor, w: torch.Tensor, nr: int, blockage_shift: bool = False, want_dg: bool = True):
    """Dict with gs (int32 [K]), mb (the uint32 shell counts in an int32 tensor [nr]), dcf (float64 [nr + 1]),
    dg and wt (float64 [K]; dg None without want_dg) and n_out_of_range (int64 [1])."""
    L = _lib.load()
    bll = _dev(bll, torch.float32)
    w = _dev(w, torch.float64)
    K = bll.numel()
    dev = bll.device
    gs = torch.empty(K, dtype=torch.int32, device=dev)
    mb = torch.zeros(nr, dtype=torch.int32, device=dev)  # uint32 counts (< 2^31 in practice)
    dcf = torch.empty(nr + 1, dtype=torch.float64, device=dev)
    dg = torch.empty(K, dtype=torch.float64, device=dev) if want_dg else None
    wt = torch.empty(K, dtype=torch.float64, device=dev)
    bad = torch.zeros(1, dtype=torch.int64, device=dev)
    check("viss_weights", L.viss_weights(_p(bll), _p(w), K, int(nr), _lib.DCF_BLOCKAGE_SHIFT if blockage_shift else 0,
                                         _p(gs), _p(mb), _p(dcf), _p(dg), _p(wt), _p(bad), _stream()))
    return {"gs": gs, "mb": mb, "dcf": dcf, "dg": dg, "wt": wt, "n_out_of_range": bad}


def forward(lmn, I, uvw, xyz_a=None, xyz_b=None, **kw):
    """One-shot forward (builds a temporary Sky)."""
    return Sky(lmn).forward(I, uvw, xyz_a, xyz_b, **kw)


def backproject(lmn, uvw, y, wt=None, xyz_a=None, xyz_b=None, **kw):
    return Sky(lmn).backproject(uvw, y, wt, xyz_a, xyz_b, **kw)


class Batch:
    """Many visibility groups over one Sky, forward in one launch over a work table
    (``viss_batch``).  ``row_offsets``: host sequence of n_groups + 1 row
    offsets into the concatenated [3, K_total] uvw / xyz arrays."""

    def __init__(self, sky: Sky, row_offsets, uvw, xyz_a=None, xyz_b=None, *, mask: bool = True,
                 R_moon: float = REF_R_MOON, orbit_radius: float = 0.0, phase_track: bool = False,
                 fp64: bool = False):
        import numpy as np

        L = _lib.load()
        self.sky = sky
        self.uvw, self.xa, self.xb, self.K = sky._rows(uvw, xyz_a, xyz_b, mask)
        self.offs = np.ascontiguousarray(np.asarray(row_offsets, dtype=np.int64))
        if self.offs[-1] != self.K:
            raise ValueError("row_offsets[-1] must equal K_total")
        self.G = self.offs.size - 1
        h = ctypes.c_void_p()
        flags = _flags(mask, phase_track, fp64, False)
        with torch.cuda.device(sky.device):
            check("viss_batch_create", L.viss_batch_create(sky._h, self.G, self.offs.ctypes.data_as(ctypes.c_void_p),
                                                           _p(self.uvw), _p(self.xa), _p(self.xb),
                                                           _geom(R_moon, orbit_radius), flags, ctypes.byref(h),
                                                           _stream()))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            torch.cuda.synchronize(self.sky.device)
            _lib.load().viss_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, I: torch.Tensor, out: Optional[torch.Tensor] = None, use_graph: bool = False) -> torch.Tensor:
        L = _lib.load()
        I = _dev(I, torch.float64)
        if out is None:
            out = torch.zeros(self.K, dtype=torch.complex128, device=self.sky.device)
        check("viss_batch_forward", L.viss_batch_forward(self._h, _p(I), _p(out), 1 if use_graph else 0, _stream()))
        return out

    def autotune(self, I: torch.Tensor, reps: int = 2):
        L = _lib.load()
        I = _dev(I, torch.float64)
        scratch = torch.zeros(self.K, dtype=torch.complex128, device=self.sky.device)
        check("viss_batch_autotune", L.viss_batch_autotune(self._h, _p(I), _p(scratch), int(reps), _stream()))

    def weights(self, bll: torch.Tensor, nr: int, blockage_shift: bool = False, want_dg: bool = True):
        """Every group's density-compensation weights (viss_batch_weights): dict with gs, wt, dg
        ([K_total]), mb ([G, nr]), dcf ([G, nr + 1]) and n_out_of_range (total)."""
        L = _lib.load()
        bll = _dev(bll, torch.float32)
        if bll.numel() != self.K:
            raise ValueError("bll must have K_total entries")
        dev = self.sky.device
        gs = torch.empty(self.K, dtype=torch.int32, device=dev)
        mb = torch.zeros(self.G, nr, dtype=torch.int32, device=dev)
        dcf = torch.empty(self.G, nr + 1, dtype=torch.float64, device=dev)
        dg = torch.empty(self.K, dtype=torch.float64, device=dev) if want_dg else None
        wt = torch.empty(self.K, dtype=torch.float64, device=dev)
        bad = torch.zeros(1, dtype=torch.int64, device=dev)
        check("viss_batch_weights", L.viss_batch_weights(self._h, _p(bll), int(nr),
                                                         _lib.DCF_BLOCKAGE_SHIFT if blockage_shift else 0, _p(gs),
                                                         _p(mb), _p(dcf), _p(dg), _p(wt), _p(bad), _stream()))
        return {"gs": gs, "mb": mb, "dcf": dcf, "dg": dg, "wt": wt, "n_out_of_range": bad}

    def backproject(self, y: torch.Tensor, wt: Optional[torch.Tensor] = None, *, per_group: bool = False,
                    phase_track: bool = False, hermitian: bool = False, out: Optional[torch.Tensor] = None,
                    use_graph: bool = False) -> torch.Tensor:
        """Back-projection of every group's rows (viss_batch_backproject): the summed image [P],
        or one image per group [G, P] with per_group."""
        L = _lib.load()
        y = _dev(y, torch.complex128)
        if y.numel() != self.K:
            raise ValueError("y must have K_total entries")
        w = None if wt is None else _dev(wt, torch.float64)
        shape = (self.G, self.sky.P) if per_group else (self.sky.P,)
        if out is None:
            out = torch.zeros(shape, dtype=torch.float64, device=self.sky.device)
        elif out.dtype != torch.float64 or tuple(out.shape) != shape or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous float64 tensor of shape {shape}")
        check("viss_batch_backproject", L.viss_batch_backproject(
            self._h, _p(y), _p(w), _flags(False, phase_track, False, hermitian), _p(out), 1 if per_group else 0,
            1 if use_graph else 0, _stream()))
        return out

    def save_tuning(self, path: str):
        check("viss_batch_tuning_save", _lib.load().viss_batch_tuning_save(self._h, path.encode()))

    def load_tuning(self, path: str):
        check("viss_batch_tuning_load", _lib.load().viss_batch_tuning_load(self._h, path.encode(), _stream()))

    def stats(self):
        vt, tt, ni, gb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check("viss_batch_stats", _lib.load().viss_batch_stats(self._h, ctypes.byref(vt), ctypes.byref(tt),
                                                               ctypes.byref(ni), ctypes.byref(gb)))
        return {"visible_tiles": vt.value, "total_tiles": tt.value, "n_items": ni.value, "grid_blocks": gb.value}


def image_accumulate(img: torch.Tensor, total: torch.Tensor, nan_to_zero: bool = True) -> torch.Tensor:
    """total += nan_to_num(img) (the notebooks' multi-day sum), in place."""
    img = _dev(img, torch.float64)
    if total.dtype != torch.float64 or not total.is_contiguous() or total.numel() != img.numel():
        raise ValueError("total must be a contiguous float64 tensor of the image's size")
    check("viss_image_accumulate", _lib.load().viss_image_accumulate(_p(img), img.numel(), 1 if nan_to_zero else 0,
                                                                     _p(total), _stream()))
    return total


def reorder(m: torch.Tensor, nside: int, to_ring: bool) -> torch.Tensor:
    """NEST -> RING (to_ring) or RING -> NEST reorder of a HEALPix map."""
    m = _dev(m, torch.float64)
    out = torch.empty_like(m)
    check("viss_reorder", _lib.load().viss_reorder(_p(m), int(nside), 1 if to_ring else 0, _p(out), _stream()))
    return out
