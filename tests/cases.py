"""Seeded workloads shared by the parity tests (small enough for the oracle)."""
import numpy as np

import oracle
from viss_optimize_3d_amd import geometry as geo


def make_case(nside=16, n_snap=60, n_sat=4, freq=1e6, sigma=10e3, rotating=True, snaps_per_orbit=50, nest=False,
              seed=1, n_sources=20, max_rows=None):
    rng = np.random.default_rng(seed)
    th, ph, _, _ = oracle.pix2ang(nside, nest)
    B = geo.make_sky(nside, rng, n_sources)
    l, m, n, I = oracle.pixel_dirs(th, ph, B, 4 * np.pi / B.size)
    pos = geo.satellite_positions(n_sat, n_snap, sigma, rotating, snaps_per_orbit, rng)
    uvw, xa, xb, bll = geo.baseline_rows(pos, geo.C_LIGHT / freq, max_rows)
    return dict(lmn=np.stack([l, m, n]), I=I, B=B, uvw=uvw, xa=xa, xb=xb, bll=bll, theta=th, phi=ph)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
