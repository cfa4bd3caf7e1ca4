"""Occultation decisions of every kernel path, bit for bit.

With all baselines zero every phase is 0 (sin 0, cos 1 exactly), so with unit
brightness the forward returns, per row, the number of pixels both satellites
see, and with unit visibilities the back-projection returns, per pixel, the
number of rows that see it.  Both are integers far below 2^24, summed exactly
in FP32 and FP64, so they must equal the FP64 oracle's counts exactly: one
pixel included or dropped by a tile classification, a ring mode (whole, one
side tested, both tested) or a per-pixel test changes a count by one.  The
tile classification and ring modes depend on the satellites only, so zero
baselines exercise the same decisions as real ones.

Satellite geometries: a tight cluster (10 satellites, sigma 50 km, as the
precision workload: the wavefront's rows share their cap edges, so tiles take
the one-side-tested ring modes) and a scattered one (sigma 500 km: many edges
per wavefront, tiles tested on both sides).  The quad forward first settles
tiles against bounds of the wavefront's 64 rows; with the rows in random order
each wavefront holds unrelated snapshots, the bounds settle few tiles, and the
per-row classification takes the rest.  Skies: HEALPix RING and NEST
(quad layout), rings of random azimuths (quads), the same with 5% of the
pixels removed (quads, pairs and singles), random directions (per-pixel
kernels) and an exact +-s set (pairs only).
"""
import zlib

import numpy as np
import pytest

import oracle
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


def unit(v):
    return v / np.linalg.norm(v, axis=0, keepdims=True)


def rings(rng, nring=24, per_ring=64, drop=0.0):
    """Rings of random azimuths in mirror pairs z, -z, every pixel with its antipode."""
    out = []
    for z in rng.uniform(0.02, 0.98, nring):
        phi = rng.uniform(0, np.pi, per_ring)
        rho = np.sqrt(1 - z * z)
        x, y = rho * np.cos(phi), rho * np.sin(phi)
        for sx, sz in ((1, 1), (-1, 1), (1, -1), (-1, -1)):
            out.append(np.stack([sx * x, sx * y, np.full(per_ring, sz * z)]))
    lmn = np.concatenate(out, axis=1)
    if drop:
        lmn = lmn[:, np.sort(rng.choice(lmn.shape[1], int(lmn.shape[1] * (1 - drop)), replace=False))]
    return lmn[:, rng.permutation(lmn.shape[1])]


def sky(kind, rng):
    if kind in ("healpix32", "healpix16-nest"):
        nside, nest = (32, False) if kind == "healpix32" else (16, True)
        th, ph, _, _ = oracle.pix2ang(nside, nest)
        return np.stack(oracle.pixel_dirs(th, ph, np.ones_like(th), 1.0)[:3])
    if kind == "rings":
        return rings(rng)
    if kind == "rings-minus":
        return rings(rng, drop=0.05)
    if kind == "random":
        return unit(rng.standard_normal((3, 4000)))
    if kind == "exact-pm":
        v = unit(rng.standard_normal((3, 2000)))
        return np.concatenate([v, -v], axis=1)[:, rng.permutation(4000)]
    raise ValueError(kind)


GEOMETRIES = {"cluster": dict(n_sat=10, sigma=50e3, n_snap=60), "scattered": dict(n_sat=8, sigma=500e3, n_snap=60)}


def positions(name, seed):
    g = GEOMETRIES[name]
    return geo.satellite_positions(g["n_sat"], g["n_snap"], g["sigma"], True, 100, np.random.default_rng(seed))


def oracle_counts(lmn, xa, xb):
    """(visible pixels per row, rows that see each pixel), exact integers."""
    K = xa.shape[1]
    zero = np.zeros((3, K))
    l, m, n = lmn
    per_row = oracle.forward(l, m, n, np.ones(lmn.shape[1]), zero, xa, xb, flags=oracle.MASK, R_moon=R)
    per_pix = oracle.backproject(l, m, n, zero, np.ones(K, dtype=complex), None, xa, xb, flags=oracle.MASK,
                                 R_moon=R)
    assert np.array_equal(per_row.imag, np.zeros(K)) and np.array_equal(per_row.real, np.round(per_row.real))
    assert np.array_equal(per_pix, np.round(per_pix))
    return per_row.real, per_pix


@pytest.mark.parametrize("geom", ["cluster", "scattered"])
@pytest.mark.parametrize("kind", ["healpix32", "healpix16-nest", "rings", "rings-minus", "random", "exact-pm"])
def test_mask_counts_per_row_kernels(ops, kind, geom):
    rng = np.random.default_rng(zlib.crc32(f"{kind}/{geom}".encode()))
    lmn = sky(kind, rng)
    pos = positions(geom, 3)
    _, xa, xb, _ = geo.baseline_rows(pos, 1.0)
    K, P = xa.shape[1], lmn.shape[1]
    rows, pix = oracle_counts(lmn, xa, xb)
    s = ops.Sky(torch.from_numpy(lmn).cuda())
    layout = s.pairs()
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    zero, ta, tb = d(np.zeros((3, K))), d(xa), d(xb)
    ones_p = torch.ones(P, dtype=torch.float64, device="cuda")
    ones_k = torch.ones(K, dtype=torch.complex128, device="cuda")
    print(f"{kind} {geom}: K={K} P={P} layout={layout} visible fraction {rows.sum() / (K * P):.3f}", flush=True)
    for fp64 in (False, True):
        V = s.forward(ones_p, zero, ta, tb, mask=True, R_moon=R, fp64=fp64).cpu().numpy()
        bad = np.flatnonzero(V.real != rows)
        assert bad.size == 0, f"fp64={fp64}: {bad.size} rows differ, first {bad[:5]}: {V.real[bad[:5]]} vs {rows[bad[:5]]}"
        assert np.array_equal(V.imag, np.zeros(K))
        x = s.backproject(zero, ones_k, None, ta, tb, mask=True, R_moon=R, fp64=fp64).cpu().numpy()
        bad = np.flatnonzero(x != pix)
        assert bad.size == 0, f"fp64={fp64}: {bad.size} pixels differ, first {bad[:5]}: {x[bad[:5]]} vs {pix[bad[:5]]}"
    Vh = s.forward(ones_p, zero, ta, tb, mask=True, R_moon=R, hermitian=True).cpu().numpy()
    assert np.array_equal(Vh.real, np.concatenate([rows, rows]))


@pytest.mark.parametrize("geom", ["cluster", "scattered"])
@pytest.mark.parametrize("kind", ["healpix32", "rings-minus", "random"])
def test_mask_counts_snapshot_and_batched_kernels(ops, kind, geom):
    """The snapshot kernels phase with positions / wavelength: at 1e30 m the
    phases are below 1e-23 cycles, so the counts come out within 1e-12.  The
    batched form runs the same rows as 3 groups."""
    rng = np.random.default_rng(zlib.crc32(f"{kind}/{geom}/snap".encode()))
    lmn = sky(kind, rng)
    pos = positions(geom, 4)
    _, xa, xb, _ = geo.baseline_rows(pos, 1.0)
    K, P = xa.shape[1], lmn.shape[1]
    rows, pix = oracle_counts(lmn, xa, xb)
    s = ops.Sky(torch.from_numpy(lmn).cuda())
    ones_p = torch.ones(P, dtype=torch.float64, device="cuda")
    ones_k = torch.ones(K, dtype=torch.complex128, device="cuda")
    tpos = torch.from_numpy(pos).cuda()
    V = s.forward_snapshots(ones_p, tpos, 1e30, mask=True, R_moon=R).cpu().numpy()
    assert np.max(np.abs(V.real - rows)) < 1e-12 and np.max(np.abs(V.imag)) < 1e-12
    x = s.backproject_snapshots(tpos, 1e30, ones_k, mask=True, R_moon=R).cpu().numpy()
    assert np.max(np.abs(x - pix)) < 1e-12
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    offs = [0, K // 3, 2 * K // 3, K]
    b = ops.Batch(s, offs, d(np.zeros((3, K))), d(xa), d(xb), mask=True, R_moon=R)
    Vb = b.forward(ones_p).cpu().numpy()
    bad = np.flatnonzero(Vb.real != rows)
    assert bad.size == 0, f"batched: {bad.size} rows differ"
    print(f"{kind} {geom}: snapshot and batched counts exact over K={K} P={P}", flush=True)


@pytest.mark.parametrize("kind", ["healpix32", "rings-minus"])
def test_mask_counts_rows_shuffled(ops, kind):
    """Rows in random order through the per-row and batched forward: exact counts."""
    rng = np.random.default_rng(zlib.crc32(f"{kind}/shuffled".encode()))
    lmn = sky(kind, rng)
    pos = positions("cluster", 5)
    _, xa, xb, _ = geo.baseline_rows(pos, 1.0)
    perm = rng.permutation(xa.shape[1])
    xa, xb = xa[:, perm], xb[:, perm]
    K, P = xa.shape[1], lmn.shape[1]
    rows, _ = oracle_counts(lmn, xa, xb)
    s = ops.Sky(torch.from_numpy(lmn).cuda())
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    zero, ta, tb = d(np.zeros((3, K))), d(xa), d(xb)
    ones_p = torch.ones(P, dtype=torch.float64, device="cuda")
    V = s.forward(ones_p, zero, ta, tb, mask=True, R_moon=R).cpu().numpy()
    bad = np.flatnonzero(V.real != rows)
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}: {V.real[bad[:5]]} vs {rows[bad[:5]]}"
    b = ops.Batch(s, [0, K // 2, K], zero, ta, tb, mask=True, R_moon=R)
    Vb = b.forward(ones_p).cpu().numpy()
    assert np.array_equal(Vb.real, rows), "batched"
    print(f"{kind}: shuffled rows, counts exact over K={K} P={P} (layout {s.pairs()})", flush=True)


def test_mask_counts_with_nan_rows(ops):
    """Rows with a NaN satellite coordinate see no pixel (every comparison with NaN is false,
    as in the oracle); the other rows of their wavefronts keep their exact counts, through the
    per-row and the batched forward."""
    rng = np.random.default_rng(77)
    lmn = sky("healpix32", rng)
    pos = positions("cluster", 6)
    _, xa, xb, _ = geo.baseline_rows(pos, 1.0)
    K, P = xa.shape[1], lmn.shape[1]
    bad = np.array([5, 70, 71, 200, K - 1])
    xa, xb = xa.copy(), xb.copy()
    xa[0, bad[:2]] = np.nan
    xb[2, bad[2]] = np.nan
    xa[1, bad[3]] = np.nan
    xb[:, bad[4]] = np.nan
    rows, _ = oracle_counts(lmn, xa, xb)
    assert np.all(rows[bad] == 0) and np.all(rows[np.setdiff1d(np.arange(K), bad)] > 0)
    s = ops.Sky(torch.from_numpy(lmn).cuda())
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    zero, ta, tb = d(np.zeros((3, K))), d(xa), d(xb)
    ones_p = torch.ones(P, dtype=torch.float64, device="cuda")
    for fp64 in (False, True):
        V = s.forward(ones_p, zero, ta, tb, mask=True, R_moon=R, fp64=fp64).cpu().numpy()
        diff = np.flatnonzero(V.real != rows)
        assert diff.size == 0, f"fp64={fp64}: {diff.size} rows differ, first {diff[:5]}: {V.real[diff[:5]]} vs {rows[diff[:5]]}"
    b = ops.Batch(s, [0, K // 2, K], zero, ta, tb, mask=True, R_moon=R)
    Vb = b.forward(ones_p).cpu().numpy()
    diff = np.flatnonzero(Vb.real != rows)
    assert diff.size == 0, f"batched: {diff.size} rows differ, first {diff[:5]}"
    print(f"NaN rows {bad.tolist()}: no pixel; the other {K - bad.size} rows exact", flush=True)
