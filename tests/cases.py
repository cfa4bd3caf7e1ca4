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


def check_weights(out, bll, w, nr, shift=False):
    """Device weights (ops.weights output) against the oracle on the same rows:
    gs, mb, dcf bit-exact; dg and wt within 1e-14 relative or 1e-13 absolute,
    plus dg's own rounding sensitivity.  dg = 1.5 sqrt|1 - 4 x^2| / cos g with
    x = w / bll = sin g cancels near |x| = 1/2: an ulp-level difference of
    1 - 4 x^2 between two evaluations (<= e = 4.4e-16) moves dg by up to
    1.5 min(sqrt(e), e / sqrt|1 - 4 x^2|) / cos g; near |x| = 1, cos^2 g =
    1 - x^2 cancels, and the same difference in it moves dg by up to
    dg min(1, e / (2 |1 - x^2|)).  wt moves by dcf times that.
    Returns the oracle's outputs."""
    ref = oracle.weights(bll, w, nr, oracle.DCF_BLOCKAGE_SHIFT if shift else 0)
    gs, mb, dcf, dg, wt, bad = ref
    np.testing.assert_array_equal(out["gs"].cpu().numpy(), gs)
    np.testing.assert_array_equal(out["mb"].cpu().numpy().astype(np.uint32), mb)
    np.testing.assert_array_equal(out["dcf"].cpu().numpy(), dcf)
    x = w / bll.astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        e, sa = 4.4e-16, np.abs(1.0 - 4.0 * x * x)
        sq = np.abs((1.0 - x) * (1.0 + x))
        cond = 1.5 / np.sqrt(sq) * np.minimum(np.sqrt(e), e / np.sqrt(sa))
        cond = np.nan_to_num(cond, nan=0.0, posinf=0.0) + np.nan_to_num(np.abs(dg) * np.minimum(1.0, e / (2 * sq)),
                                                                         nan=0.0, posinf=0.0)
    d_dev, w_dev = out["dg"].cpu().numpy(), out["wt"].cpu().numpy()
    fin = np.isfinite(dg)
    assert np.array_equal(np.isnan(d_dev), np.isnan(dg))
    assert np.all(np.abs(d_dev[fin] - dg[fin]) <= 1e-13 + 1e-14 * np.abs(dg[fin]) + cond[fin])
    dcf_k = np.where((gs >= 0) & (gs <= nr), dcf[np.clip(gs, 0, nr)], 0.0)
    with np.errstate(invalid="ignore"):
        tol_w = 1e-13 + 1e-14 * np.abs(wt) + np.nan_to_num(dcf_k * cond, nan=0.0)  # dcf is inf for an empty shell
    assert np.all(np.abs(w_dev - wt) <= tol_w)
    assert int(out["n_out_of_range"].item()) == bad
    return ref
