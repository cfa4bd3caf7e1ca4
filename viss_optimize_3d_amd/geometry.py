"""Synthetic lunar-orbit array inputs (the producer the reference's data
directories expect but do not ship: ``earth_1Mhz/``, ``earth_10Mhz/`` hold only
``.gitkeep``).

Everything is generated in FP64 on the host with numpy's PCG64 from a recorded
seed, so a workload is named by (config, seed).

Geometry (all lengths in metres, Moon-centred frame, z = lunar spin axis):

* satellites fly a circular orbit of radius ``orbit_radius`` (2037.4 km) around
  a Moon of radius ``R_moon`` (1737.4 km); satellite i sits at the orbit point
  plus a fixed offset delta_i ~ N(0, sigma^2 I3) per axis;
* the orbit plane is polar; its normal either stays fixed (``rotating=False``)
  or turns uniformly through 360 degrees about the spin axis over the run;
* a visibility row is one (snapshot, pair a<b) with baseline
  b = (r_a - r_b) / lambda in wavelengths, ordered snapshot-major;
* pixel direction s is occulted for satellite i when
  s . r_i < -sqrt(|r_i|^2 - R_moon^2)  (equivalently s . r^_i < -sqrt(1-(R/|r_i|)^2)).

The reference's row files carry 2*amount rows, the second half being the
negated baselines with the conjugate visibility (vissgen3d_noblockage.cu:100,
400); ``hermitian_rows`` builds that layout.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional, Sequence

import numpy as np

C_LIGHT = 299792458.0
R_MOON = 1737.4e3
ORBIT_RADIUS = 2037.4e3
# Constants of the reference driver (vissgen3d_noblockage.cu:240-250).
REF_R_MOON = 1737.1e3
REF_ALTITUDE = 300e3
REF_BL_MAX = 100e3


@dataclasses.dataclass(frozen=True)
class Workload:
    name: str
    nside: int
    freq_hz: float
    n_sat: int
    sigma_m: float
    n_snap: int
    blockage: bool
    rotating: bool
    snaps_per_orbit: int
    n_sources: int = 20
    float32_sky: bool = False
    max_rows: Optional[int] = None
    seed: int = 1

    @property
    def npix(self) -> int:
        return 12 * self.nside * self.nside

    @property
    def wavelength(self) -> float:
        return C_LIGHT / self.freq_hz

    @property
    def n_baselines(self) -> int:
        return self.n_sat * (self.n_sat - 1) // 2

    @property
    def n_vis(self) -> int:
        k = self.n_snap * self.n_baselines
        return k if self.max_rows is None else min(k, self.max_rows)


# The five workloads the measurements and parity tests use.
WORKLOADS = {
    "plain-nside16": Workload("plain-nside16", 16, 1e6, 4, 10e3, 500, False, False, 500, seed=1),
    "blockage-nside64": Workload("blockage-nside64", 64, 1e6, 8, 10e3, 10_000, True, True, 100, seed=2),
    "precision-nside256": Workload("precision-nside256", 256, 10e6, 10, 50e3, 20_000, True, True, 100, seed=3),
    "diff-nside128": Workload("diff-nside128", 128, 10e6, 8, 10e3, 7_143, True, True, 100, n_sources=50,
                              float32_sky=True, max_rows=200_000, seed=5),
}


def make_sky(nside: int, rng: np.random.Generator, n_sources: int = 20, amp=(10.0, 100.0),
             float32: bool = False) -> np.ndarray:
    """Brightness B_p ~ lognormal(0,1) plus ``n_sources`` point sources U(amp)."""
    P = 12 * nside * nside
    B = rng.lognormal(0.0, 1.0, P)
    idx = rng.choice(P, size=n_sources, replace=False)
    B[idx] += rng.uniform(amp[0], amp[1], n_sources)
    return B.astype(np.float32).astype(np.float64) if float32 else B


def satellite_positions(n_sat: int, n_snap: int, sigma_m: float, rotating: bool, snaps_per_orbit: int,
                        rng: np.random.Generator, orbit_radius: float = ORBIT_RADIUS) -> np.ndarray:
    """Positions [n_snap, n_sat, 3] in metres."""
    offsets = rng.normal(0.0, sigma_m, (n_sat, 3))
    j = np.arange(n_snap, dtype=np.float64)
    alpha = 2.0 * math.pi * j / snaps_per_orbit  # orbital phase
    psi = 2.0 * math.pi * j / n_snap if rotating else np.zeros(n_snap)  # ascending node
    # polar orbit: normal (cos psi, sin psi, 0); in-plane axes e1 = (-sin psi, cos psi, 0), e2 = z
    e1 = np.stack([-np.sin(psi), np.cos(psi), np.zeros(n_snap)], axis=1)
    e2 = np.tile(np.array([0.0, 0.0, 1.0]), (n_snap, 1))
    centre = orbit_radius * (np.cos(alpha)[:, None] * e1 + np.sin(alpha)[:, None] * e2)
    return centre[:, None, :] + offsets[None, :, :]


def baseline_rows(pos: np.ndarray, wavelength: float, max_rows: Optional[int] = None):
    """uvw (3,K) [wavelengths], xyz_a, xyz_b (3,K) [m], bll (K,) [wavelengths]."""
    n_snap, n_sat, _ = pos.shape
    ia, ib = np.triu_indices(n_sat, k=1)
    ra = pos[:, ia, :].reshape(-1, 3)
    rb = pos[:, ib, :].reshape(-1, 3)
    if max_rows is not None:
        ra, rb = ra[:max_rows], rb[:max_rows]
    uvw = ((ra - rb) / wavelength).T.copy()
    bll = np.sqrt((uvw * uvw).sum(axis=0))
    return uvw, ra.T.copy(), rb.T.copy(), bll


def hermitian_rows(uvw, xyz_a, xyz_b, bll):
    """The reference's 2*amount row layout: second half = negated baselines,
    satellites swapped (the occultation mask is symmetric in a, b)."""
    return (np.concatenate([uvw, -uvw], axis=1), np.concatenate([xyz_a, xyz_b], axis=1),
            np.concatenate([xyz_b, xyz_a], axis=1), np.concatenate([bll, bll]))


@dataclasses.dataclass
class Inputs:
    workload: Workload
    B: np.ndarray
    uvw: np.ndarray
    xyz_a: np.ndarray
    xyz_b: np.ndarray
    bll: np.ndarray
    positions: np.ndarray


def generate(w: Workload, seed: Optional[int] = None) -> Inputs:
    rng = np.random.default_rng(w.seed if seed is None else seed)
    B = make_sky(w.nside, rng, w.n_sources, float32=w.float32_sky)
    pos = satellite_positions(w.n_sat, w.n_snap, w.sigma_m, w.rotating, w.snaps_per_orbit, rng)
    uvw, xa, xb, bll = baseline_rows(pos, w.wavelength, w.max_rows)
    return Inputs(w, B, uvw, xa, xb, bll, pos)


def shell_count(freq_hz: float, bl_max: float = REF_BL_MAX, wavelength: Optional[float] = None) -> int:
    """nr = ceil(2 bl_max / lambda) (vissgen3d_noblockage.cu:248-250)."""
    lam = C_LIGHT / freq_hz if wavelength is None else wavelength
    return int(math.ceil(bl_max / lam * 2.0))


def subset(inp: Inputs, rows: Sequence[int]) -> Inputs:
    rows = np.asarray(rows)
    return Inputs(inp.workload, inp.B, inp.uvw[:, rows].copy(), inp.xyz_a[:, rows].copy(),
                  inp.xyz_b[:, rows].copy(), inp.bll[rows].copy(), inp.positions)


# Batched workload: 5 frequencies x 64 snapshot groups x 500 snapshots, 8
# satellites, nside 128.  Each group flies a short arc (500 snapshots over 2
# degrees of orbital phase) at its own orbit radius, from 1.01 R to 1.2 R, so
# the occultation removes a different share of the sky per group (a single
# occulting sphere can remove at most half of it).
BATCH_FREQS = (1e6, 2.5e6, 5e6, 7.5e6, 10e6)


def batched_workload(nside: int = 128, freqs=BATCH_FREQS, n_groups: int = 64, snaps: int = 500, n_sat: int = 8,
                     sigma_m: float = 10e3, seed: int = 4, arc_deg: float = 2.0):
    """(B, uvw, xyz_a, xyz_b, row_offsets, group_freq) with groups ordered frequency-major."""
    rng = np.random.default_rng(seed)
    B = make_sky(nside, rng, 20)
    offsets = rng.normal(0.0, sigma_m, (n_sat, 3))
    pos_groups = []
    for g in range(n_groups):
        r = R_MOON * (1.01 + 0.19 * g / max(1, n_groups - 1))
        psi = 2 * math.pi * g / n_groups
        alpha0 = rng.uniform(0, 2 * math.pi)
        alpha = alpha0 + np.deg2rad(arc_deg) * np.arange(snaps) / snaps
        e1 = np.array([-math.sin(psi), math.cos(psi), 0.0])
        e2 = np.array([0.0, 0.0, 1.0])
        centre = r * (np.cos(alpha)[:, None] * e1 + np.sin(alpha)[:, None] * e2)
        pos_groups.append(centre[:, None, :] + offsets[None, :, :] * (r / ORBIT_RADIUS))
    uvws, xas, xbs, offs, gf = [], [], [], [0], []
    for f in freqs:
        for pos in pos_groups:
            uvw, xa, xb, _ = baseline_rows(pos, C_LIGHT / f)
            uvws.append(uvw)
            xas.append(xa)
            xbs.append(xb)
            offs.append(offs[-1] + uvw.shape[1])
            gf.append(f)
    return (B, np.concatenate(uvws, axis=1), np.concatenate(xas, axis=1), np.concatenate(xbs, axis=1),
            np.array(offs, np.int64), np.array(gf))
