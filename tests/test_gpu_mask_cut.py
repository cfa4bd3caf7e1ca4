"""Occultation decisions at the cut: every kernel path against the oracle.

Each satellite sits on the cut of a chosen pixel (tests/cut_cases.py), so that
pair's s.r - t is within a few ulps of |r| of 0 and the decision depends on
the evaluation order; test_oracle_cut.py shows that the oracle takes the
canonical order, fma(l, x, fma(m, y, n z)) >= t, and that another order flips
a share of these decisions.  The kernels evaluate the same expression on the
coordinates they hold for a pixel: its own (per-pixel kernels, all-FP64), the
negation of its antipode's (pair layout), or (+-x, +-y, +-z) of its quad
(quad layout).  So on skies whose pairs and quads are exact negations and
mirrors (an exact +-s set, HEALPix made exactly symmetric) every decision must
equal the oracle's, and on random directions too.

Decisions are read as integers (zero baselines, unit inputs): visible pixels per
row (forward) and seeing rows per pixel (back-projection), for the per-row
kernels in mixed precision and FP64, the Hermitian rows, the snapshot kernels
(wavelength 1e30 m, phases below 1e-23 cycles), the batched form (3 groups)
and the allocation-free back-projection; both cuts (the reference's fixed cut
and each satellite's own).

On a HEALPix map computed in FP64 the antipode and mirror of a pixel differ
from its exact negation by rounding, and the pair / quad kernels evaluate the
negation: decisions may then differ from the oracle's only for pixels within
that rounding of the cut.  The test checks that every difference is at a
pixel a satellite was placed on and counts them.
"""
import zlib

import numpy as np
import pytest

import oracle
from tests import cut_cases as cc
from viss_optimize_3d_amd import geometry as geo

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
R = cc.R


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import ops as _ops

    return _ops


def make_sky(kind, rng):
    if kind in ("healpix32-sym", "healpix32"):
        th, ph, ring, iphi = oracle.pix2ang(32, False)
        lmn = np.stack(oracle.pixel_dirs(th, ph, np.ones_like(th), 1.0)[:3])
        if kind == "healpix32":
            return lmn
        return cc.symmetric_healpix(32, ring, iphi, lmn)[:, rng.permutation(th.size)]
    if kind == "pm-exact":
        v = rng.standard_normal((3, 3000))
        v /= np.linalg.norm(v, axis=0)
        return np.concatenate([v, -v], axis=1)[:, rng.permutation(6000)]
    if kind == "random":
        v = rng.standard_normal((3, 6000))
        return v / np.linalg.norm(v, axis=0)
    raise ValueError(kind)


def oracle_counts(lmn, xa, xb, orb):
    K = xa.shape[1]
    zero = np.zeros((3, K))
    per_row = oracle.forward(*lmn, np.ones(lmn.shape[1]), zero, xa, xb, flags=oracle.MASK, R_moon=R,
                             orbit_radius=orb).real
    per_pix = oracle.backproject(*lmn, zero, np.ones(K, dtype=complex), None, xa, xb, flags=oracle.MASK, R_moon=R,
                                 orbit_radius=orb)
    return per_row, per_pix


EXPECT_LAYOUT = {"healpix32-sym": "quads", "healpix32": "quads", "pm-exact": "pairs", "random": "pixels"}


@pytest.mark.parametrize("orb", [cc.FIXED_ORBIT, 0.0])
@pytest.mark.parametrize("kind", ["healpix32-sym", "pm-exact", "random", "healpix32"])
def test_decisions_at_the_cut(ops, kind, orb):
    from viss_optimize_3d_amd import _lib

    rng = np.random.default_rng(zlib.crc32(f"cut/{kind}/{orb}".encode()))
    lmn = make_sky(kind, rng)
    P = lmn.shape[1]
    pos, targets = cc.cut_satellites(lmn, 100, 6, rng, orb)
    _, xa, xb, _ = geo.baseline_rows(pos, 1.0)
    K = xa.shape[1]
    rows, pix = oracle_counts(lmn, xa, xb, orb)
    s = ops.Sky(torch.from_numpy(lmn).cuda())
    lay = s.pairs()
    layout = "quads" if lay["uses_quads"] else ("pairs" if lay["uses_pairs"] else "pixels")
    assert layout == EXPECT_LAYOUT[kind], lay
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    zero, ta, tb = d(np.zeros((3, K))), d(xa), d(xb)
    ones_p = torch.ones(P, dtype=torch.float64, device="cuda")
    ones_k = torch.ones(K, dtype=torch.complex128, device="cuda")
    kw = dict(mask=True, R_moon=R, orbit_radius=orb)
    got_rows, got_pix = {}, {}
    for fp64 in (False, True):
        tag = "fp64" if fp64 else "mixed"
        got_rows[tag] = s.forward(ones_p, zero, ta, tb, fp64=fp64, **kw).real.cpu().numpy()
        got_pix[tag] = s.backproject(zero, ones_k, None, ta, tb, fp64=fp64, **kw).cpu().numpy()
    got_rows["hermitian"] = s.forward(ones_p, zero, ta, tb, hermitian=True, **kw).real.cpu().numpy()[K:]
    tpos = d(pos)
    got_rows["snapshot"] = np.round(s.forward_snapshots(ones_p, tpos, 1e30, **kw).real.cpu().numpy())
    got_pix["snapshot"] = np.round(s.backproject_snapshots(tpos, 1e30, ones_k, **kw).cpu().numpy())
    offs = [0, K // 3, 2 * K // 3, K]
    for fp64 in (False, True):
        b = ops.Batch(s, offs, zero, ta, tb, fp64=fp64, **kw)
        got_rows["batch-fp64" if fp64 else "batch"] = b.forward(ones_p).real.cpu().numpy()
        b.close()
    x_na = torch.zeros(P, dtype=torch.float64, device="cuda")
    L = _lib.load()
    _lib.check("viss_sky_backproject_noalloc",
               L.viss_sky_backproject_noalloc(s._h, ops._p(zero), ops._p(ta), ops._p(tb), ops._p(ones_k), None, K,
                                              ops._geom(R, orb), ops._flags(True, False, False, False),
                                              ops._p(x_na), ops._stream()))
    got_pix["noalloc"] = x_na.cpu().numpy()
    total = int(s.count_visible(ta, tb, R_moon=R, orbit_radius=orb).item())
    assert total == int(rows.sum())
    on_cut = np.zeros(P, bool)
    on_cut[targets.ravel()] = True
    exact = kind != "healpix32"
    flips = {}
    for name, v in got_rows.items():
        bad = np.flatnonzero(v != rows)
        flips[f"rows/{name}"] = int(np.abs(v - rows).sum())
        if exact:
            assert bad.size == 0, f"{name}: {bad.size} rows differ, first {bad[:5]}: {v[bad[:5]]} vs {rows[bad[:5]]}"
        else:  # a row has two satellites, each on the cut of one pixel
            assert np.all(np.abs(v - rows) <= 2), name
    for name, v in got_pix.items():
        bad = np.flatnonzero(v != pix)
        flips[f"pixels/{name}"] = int(np.abs(v - pix).sum())
        if exact:
            assert bad.size == 0, f"{name}: {bad.size} pixels differ, first {bad[:5]}: {v[bad[:5]]} vs {pix[bad[:5]]}"
        else:
            assert np.all(on_cut[bad]), f"{name}: a pixel off the cut changed its decision"
    print(f"{kind} orbit_radius={orb}: K={K} P={P} layout={layout} {targets.size} satellites on the cut; "
          f"decisions differing from the oracle's: {sum(flips.values())} {flips if not exact else ''}", flush=True)
    s.close()
