"""Skies far smaller than one tile: HEALPix nside 1, 2 and 4 (12, 48 and 192
pixels, RING and NEST) and random directions with 1, 2, 3, 5, 63 and 65
pixels, plus a single antipodal pair.  Every layout the sky builder can pick
for them (quads, pairs, per-pixel; partial tiles everywhere) runs the forward
and the back-projection, mixed and all-FP64, masked and unmasked, against the
FP64 oracle: relative L2 <= 1e-5 mixed, <= 1e-11 FP64.

The back-projection of a sky of one or a few pixels is a sum over all rows
whose terms have random phases, so it can cancel far below the size of its
terms; its error is measured against the larger of the oracle image's norm
and the norm such a sum has without cancellation, sqrt(P sum_k (wt_k |y_k|)^2).
"""
import zlib

import numpy as np
import pytest

import oracle
from tests.cases import make_case, rel_l2
from viss_optimize_3d_amd import geometry as geo

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
R = geo.R_MOON


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import ops as _ops

    return _ops


def tiny_sky(kind, rng):
    if kind.startswith("healpix"):
        nside, nest = int(kind.split("-")[1]), kind.endswith("nest")
        th, ph, _, _ = oracle.pix2ang(nside, nest)
        return np.stack(oracle.pixel_dirs(th, ph, np.ones_like(th), 1.0)[:3])
    if kind == "pair":
        v = rng.standard_normal((3, 1))
        v /= np.linalg.norm(v)
        return np.concatenate([v, -v], axis=1)
    n = int(kind.split("-")[1])
    v = rng.standard_normal((3, n))
    return v / np.linalg.norm(v, axis=0, keepdims=True)


KINDS = ["healpix-1-ring", "healpix-1-nest", "healpix-2-ring", "healpix-2-nest", "healpix-4-ring", "healpix-4-nest",
         "random-1", "random-2", "random-3", "random-5", "random-63", "random-65", "pair"]


@pytest.mark.parametrize("kind", KINDS)
def test_tiny_sky_against_oracle(ops, kind):
    rng = np.random.default_rng(zlib.crc32(kind.encode()))
    lmn = tiny_sky(kind, rng)
    P = lmn.shape[1]
    I = rng.lognormal(0.0, 1.0, P)
    c = make_case(nside=4, n_snap=30, n_sat=5, freq=10e6, sigma=30e3, seed=3)  # rows only
    uvw, xa, xb = c["uvw"], c["xa"], c["xb"]
    K = uvw.shape[1]
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    sky = ops.Sky(d(lmn))
    l, m, n = lmn
    y = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    wt = rng.uniform(0.5, 1.5, K)
    incoherent = float(np.sqrt(P * np.sum((wt * np.abs(y)) ** 2)))
    worst = {}
    for mask in (True, False):
        flags = oracle.MASK if mask else 0
        Vo = oracle.forward(l, m, n, I, uvw, xa, xb, flags=flags, R_moon=R)
        xo = oracle.backproject(l, m, n, uvw, y, wt, xa, xb, flags=flags, R_moon=R)
        for fp64, tol in ((False, 1e-5), (True, 1e-11)):
            V = sky.forward(d(I), d(uvw), d(xa), d(xb), mask=mask, R_moon=R, fp64=fp64).cpu().numpy()
            x = sky.backproject(d(uvw), d(y), d(wt), d(xa), d(xb), mask=mask, R_moon=R, fp64=fp64).cpu().numpy()
            rv = rel_l2(V, Vo)
            rx = float(np.linalg.norm(x - xo) / max(np.linalg.norm(xo), incoherent))
            worst[(mask, fp64)] = max(rv, rx)
            assert rv <= tol and rx <= tol, (mask, fp64, rv, rx)
    print(f"{kind}: P={P} layout {sky.pairs()}; worst rel L2 {worst}", flush=True)
