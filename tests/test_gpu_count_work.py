"""viss_sky_count_work against a count on the host.

The quads of a HEALPix map are found independently here: each northern pixel
(x, y, z) with its partners at (-x, -y, z), (x, y, -z), (-x, -y, -z), matched by
a k-d tree; the equator ring's pixels pair with their antipodes.  With each
(row, pixel) decision taken on the host (rows of random satellites: no pixel
lies within rounding of a cut, so the evaluation order does not matter), the
visible pixels, the (row, quad) and (row, quad, ring) units with a visible
pixel and the (row, equator pair) units must equal the device's counts
exactly; a 16-quad tile with a visible pixel is counted at most once per row,
so the tile count lies between the quad count / 16 and the quad count.  The
tiles and the runs that share a ring factor are rebuilt on the host as well
(quads sorted by ring height, then azimuth, 16 to a tile; runs of up to 8
consecutive tiles with a visible pixel on one ring within a chunk of the
forward's), so both counts are checked exactly too.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle
from tests.cases import make_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
R = 1737.4e3


def host_tiles(lmn, quads):
    """Tile index (z, then azimuth order, 16 quads a tile) of every quad, and the
    ring height of every tile."""
    x, y, z = lmn[:, quads[:, 0]]
    bx, by = lmn[0, quads[:, 1]], lmn[1, quads[:, 1]]
    up = np.where(np.abs(y) > 1e-9, y > 0, x > 0)  # the quad's pixel in the upper half-plane
    ux, uy = np.where(up, x, bx), np.where(up, y, by)
    phi = np.maximum(np.arctan2(uy, ux) + 0.0, 0.0)
    order = np.lexsort((phi, z))
    tile_of = np.empty(len(quads), np.int64)
    tile_z = []
    t = 0
    i = 0
    while i < len(order):
        j = i
        while j < len(order) and z[order[j]] == z[order[i]]:
            j += 1
        for s in range(i, j, 16):
            tile_of[order[s:min(j, s + 16)]] = t
            tile_z.append(z[order[i]])
            t += 1
        i = j
    return tile_of, np.array(tile_z)


def host_runs(vis_tiles, tile_z, qpc):
    """Runs of the forward for rows on their own: per chunk of qpc tiles, up to 8
    consecutive tiles with a visible pixel on one ring."""
    runs = 0
    for row in vis_tiles:
        for c0 in range(0, len(tile_z), qpc):
            zrun, n = None, 0
            for t in np.flatnonzero(row[c0:c0 + qpc]) + c0:
                if tile_z[t] != zrun or n == 8:
                    runs += 1
                    zrun, n = tile_z[t], 0
                n += 1
    return runs


def quad_chunk(K, nqtile, nrtile):
    """Quad tiles per chunk of the per-row forward (its launch rule)."""
    bx, qpc = -(-K // 256), 512
    while qpc > 32 and bx * (-(-nqtile // qpc) + -(-nrtile // (qpc // 2))) < 4096:
        qpc //= 2
    return max(qpc, -(-nqtile // 32768))


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
    tile_of, tile_z = host_tiles(lmn, quads)
    vis_tiles = np.zeros((K, len(tile_z)), bool)
    for k in range(K):
        vis_tiles[k, tile_of[q[k].any(axis=1)]] = True
    want["row_tiles"] = int(vis_tiles.sum())
    nrtile = -(-pairs.shape[0] // 64)  # the equator pairs; a HEALPix map has no singles
    want["row_runs"] = host_runs(vis_tiles, tile_z, quad_chunk(K, len(tile_z), nrtile))
    sky = ops.Sky(torch.from_numpy(lmn).cuda())
    assert sky.pairs()["quads"] == quads.shape[0] and sky.pairs()["quad_tiles"] == len(tile_z)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    got = sky.count_work(t(c["xa"]), t(c["xb"]), R_moon=R)
    print(f"nside {nside} {'NEST' if nest else 'RING'}: {got}", flush=True)
    for key, v in want.items():
        assert got[key] == v, (key, got[key], v)
    assert got["visible_pairs"] == int(sky.count_visible(t(c["xa"]), t(c["xb"]), R_moon=R).item())
    assert want["row_quads"] / 16 <= got["row_tiles"] <= want["row_quads"]
    assert got["row_tiles"] / 8 <= got["row_runs"] <= got["row_tiles"]


def test_count_work_needs_quads():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import _lib, ops

    v = np.random.default_rng(1).standard_normal((3, 2000))
    sky = ops.Sky(torch.from_numpy(v / np.linalg.norm(v, axis=0)).cuda())
    xa = torch.zeros((3, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.VissError, match="quad layout"):
        sky.count_work(xa, xa, R_moon=R)
