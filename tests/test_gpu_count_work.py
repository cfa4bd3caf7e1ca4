"""viss_sky_count_work against a count on the host.

The quads of a HEALPix map are found independently here: each northern pixel
(x, y, z) with its partners at (-x, -y, z), (x, y, -z), (-x, -y, -z), matched by
a k-d tree; the equator ring's pixels pair with their antipodes.  With each
(row, pixel) decision taken on the host (rows of random satellites: no pixel
lies within rounding of a cut, so the evaluation order does not matter), the
visible pixels, the (row, quad) and (row, quad, ring) units with a visible
pixel and the (row, equator pair) units must equal the device's counts
exactly; a 16-quad tile with a visible pixel is counted at most once per row,
so the tile count lies between the quad count / 16 and the quad count.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle
from tests.cases import make_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
R = 1737.4e3


def host_quads(lmn):
    tree = cKDTree(lmn.T)
    north = np.flatnonzero(lmn[2] > 1e-9)
    x, y, z = lmn[:, north]
    _, b = tree.query(np.stack([-x, -y, z]).T)
    _, c = tree.query(np.stack([x, y, -z]).T)
    _, d = tree.query(np.stack([-x, -y, -z]).T)
    quads = np.stack([north, b, c, d], axis=1)
    quads = quads[north < b]  # each quad once
    eq = np.flatnonzero(np.abs(lmn[2]) <= 1e-9)
    _, anti = tree.query(-lmn[:, eq].T)
    pairs = np.stack([eq, anti], axis=1)
    pairs = pairs[eq < anti]
    assert 4 * quads.shape[0] + 2 * pairs.shape[0] == lmn.shape[1]
    return quads, pairs


@pytest.mark.parametrize("nside,nest", [(16, False), (32, True)])
def test_count_work_matches_host(nside, nest):
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import ops

    c = make_case(nside=nside, n_snap=60, n_sat=6, freq=10e6, sigma=50e3, seed=9, nest=nest)
    lmn = c["lmn"]
    quads, pairs = host_quads(lmn)
    K = c["xa"].shape[1]
    vis = np.empty((K, lmn.shape[1]), bool)
    for k in range(K):
        ta = oracle.mask_threshold(*c["xa"][:, k], R, 0.0)
        tb = oracle.mask_threshold(*c["xb"][:, k], R, 0.0)
        vis[k] = (lmn.T @ c["xa"][:, k] >= ta) & (lmn.T @ c["xb"][:, k] >= tb)
    q = vis[:, quads]  # K x Q x 4
    want = {"visible_pairs": int(vis.sum()), "row_quads": int(q.any(axis=2).sum()),
            "row_rings": int((q[:, :, 0] | q[:, :, 1]).sum() + (q[:, :, 2] | q[:, :, 3]).sum()),
            "row_rest_pairs": int(vis[:, pairs].any(axis=2).sum())}
    sky = ops.Sky(torch.from_numpy(lmn).cuda())
    assert sky.pairs()["quads"] == quads.shape[0]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    got = sky.count_work(t(c["xa"]), t(c["xb"]), R_moon=R)
    print(f"nside {nside} {'NEST' if nest else 'RING'}: {got}", flush=True)
    for key, v in want.items():
        assert got[key] == v, (key, got[key], v)
    assert got["visible_pairs"] == int(sky.count_visible(t(c["xa"]), t(c["xb"]), R_moon=R).item())
    assert want["row_quads"] / 16 <= got["row_tiles"] <= want["row_quads"]


def test_count_work_needs_quads():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import _lib, ops

    v = np.random.default_rng(1).standard_normal((3, 2000))
    sky = ops.Sky(torch.from_numpy(v / np.linalg.norm(v, axis=0)).cuda())
    xa = torch.zeros((3, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.VissError, match="quad layout"):
        sky.count_work(xa, xa, R_moon=R)
