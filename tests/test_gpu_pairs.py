"""Antipodal-pair layout of the sky (the mixed-precision per-row forward and
back-projection evaluate one phase and one sin/cos per pair s_b = -s_a).

HEALPix skies pair every pixel.  Skies with pixels removed (pairs and singles
in one layout), skies with no symmetry (the per-pixel fallback) and exact +-s
sets exercise the pairing itself; every case is checked against the FP64
oracle with and without occultation, phase tracking and the Hermitian rows.

Tolerances: mixed precision against the FP64 oracle, relative L2 <= 1e-5 (the
precision target; measured values are printed); pair counts exact.
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


def unit(v):
    return v / np.linalg.norm(v, axis=0, keepdims=True)


def sky_of(kind, rng):
    """lmn [3, P] of one test sky."""
    if kind == "healpix":
        th, ph, _, _ = oracle.pix2ang(16, False)
        return np.stack(oracle.pixel_dirs(th, ph, np.ones_like(th), 1.0)[:3])
    if kind == "healpix-nest":
        th, ph, _, _ = oracle.pix2ang(8, True)
        return np.stack(oracle.pixel_dirs(th, ph, np.ones_like(th), 1.0)[:3])
    if kind == "healpix-minus-10pct":
        lmn = sky_of("healpix", rng)
        keep = np.sort(rng.choice(lmn.shape[1], int(lmn.shape[1] * 0.9), replace=False))
        return lmn[:, keep]
    if kind == "random":
        return unit(rng.standard_normal((3, 3000)))
    if kind == "exact-pm":
        v = unit(rng.standard_normal((3, 1500)))
        lmn = np.concatenate([v, -v], axis=1)
        return lmn[:, rng.permutation(lmn.shape[1])]
    raise ValueError(kind)


def test_pair_counts(ops):
    rng = np.random.default_rng(3)
    expect = {"healpix": (1536, 0, True), "healpix-nest": (384, 0, True), "random": (0, 3000, False),
              "exact-pm": (1500, 0, True)}
    for kind, (npair, nsingle, uses) in expect.items():
        sky = ops.Sky(torch.from_numpy(sky_of(kind, rng)).cuda())
        got = sky.pairs()
        print(f"{kind}: {got}", flush=True)
        assert (got["pairs"], got["singles"], got["uses_pairs"]) == (npair, nsingle, uses)
    sky = ops.Sky(torch.from_numpy(sky_of("healpix-minus-10pct", rng)).cuda())
    got = sky.pairs()
    print(f"healpix-minus-10pct: {got}", flush=True)
    assert got["uses_pairs"] and got["singles"] > 0 and 2 * got["pairs"] + got["singles"] == sky.P


@pytest.mark.parametrize("kind", ["healpix", "healpix-nest", "healpix-minus-10pct", "random", "exact-pm"])
def test_pair_layouts_against_oracle(ops, kind):
    rng = np.random.default_rng(zlib.crc32(kind.encode()))
    lmn = sky_of(kind, rng)
    P = lmn.shape[1]
    I = rng.lognormal(0.0, 1.0, P)
    c = make_case(nside=4, n_snap=40, n_sat=6, freq=10e6, sigma=20e3, seed=5)  # rows only
    uvw, xa, xb = c["uvw"], c["xa"], c["xb"]
    K = uvw.shape[1]
    sky = ops.Sky(torch.from_numpy(lmn).cuda())
    t = {k: torch.from_numpy(v).cuda() for k, v in dict(uvw=uvw, xa=xa, xb=xb, I=I).items()}
    l, m, n = lmn
    for mask, track, herm in ((True, False, False), (True, True, True), (False, False, False), (False, True, False)):
        flags = (oracle.MASK if mask else 0) | (oracle.PHASE_TRACK if track else 0) | (oracle.HERMITIAN if herm else 0)
        Vo = oracle.forward(l, m, n, I, uvw, xa, xb, flags=flags, R_moon=R)
        V = sky.forward(t["I"], t["uvw"], t["xa"], t["xb"], mask=mask, phase_track=track, hermitian=herm, R_moon=R)
        rv = rel_l2(V.cpu().numpy(), Vo)
        y = rng.standard_normal(K) + 1j * rng.standard_normal(K)
        wt = rng.uniform(0.5, 1.5, K)
        bflags = flags & ~oracle.HERMITIAN
        xo = oracle.backproject(l, m, n, uvw, y, wt, xa, xb, flags=bflags, R_moon=R)
        x = sky.backproject(t["uvw"], torch.from_numpy(y).cuda(), torch.from_numpy(wt).cuda(), t["xa"], t["xb"],
                            mask=mask, phase_track=track, R_moon=R)
        rx = rel_l2(x.cpu().numpy(), xo)
        print(f"{kind} mask={mask} track={track} herm={herm}: forward {rv:.2e}, back-projection {rx:.2e}",
              flush=True)
        assert rv <= 1e-5 and rx <= 1e-5


@pytest.mark.parametrize("kind", ["healpix", "healpix-minus-10pct", "random", "exact-pm"])
@pytest.mark.parametrize("n_sat", [5, 14])
def test_pair_layouts_snapshots_against_oracle(ops, kind, n_sat):
    """The snapshot-structured kernels over the same layouts (5 satellites:
    the 10-satellite register tile; 14: the 16-satellite one)."""
    rng = np.random.default_rng(n_sat * 7 + len(kind))
    lmn = sky_of(kind, rng)
    P = lmn.shape[1]
    I = rng.lognormal(0.0, 1.0, P)
    pos = geo.satellite_positions(n_sat, 12, 30e3, True, 15, rng)
    lam = geo.C_LIGHT / 10e6
    uvw, xa, xb, _ = geo.baseline_rows(pos, lam)
    K = uvw.shape[1]
    y = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    wt = rng.uniform(0.5, 1.5, K)
    c = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    sky = ops.Sky(c(lmn))
    l, m, n = lmn
    for mask, track, herm in ((True, False, False), (True, True, True), (False, False, False)):
        flags = (oracle.MASK if mask else 0) | (oracle.PHASE_TRACK if track else 0) | (oracle.HERMITIAN if herm else 0)
        kw = dict(mask=mask, phase_track=track, hermitian=herm, R_moon=R)
        Vo = oracle.forward(l, m, n, I, uvw, xa, xb, flags=flags, R_moon=R)
        V = sky.forward_snapshots(c(I), c(pos), lam, **kw).cpu().numpy()
        xo = oracle.backproject(l, m, n, uvw, y, wt, xa, xb, flags=flags, R_moon=R)
        x = sky.backproject_snapshots(c(pos), lam, c(y), c(wt), **kw).cpu().numpy()
        rv, rx = rel_l2(V, Vo), rel_l2(x, xo)
        print(f"{kind} n_sat={n_sat} mask={mask} track={track} herm={herm}: snapshot forward {rv:.2e}, "
              f"back-projection {rx:.2e}", flush=True)
        assert rv <= 1e-5 and rx <= 1e-5


def ring_sky(rng, nring=20, per_ring=48):
    """Quads on rings: for each of nring heights z, per_ring azimuths in [0, pi)
    and the four mirror images (x, y, +-z), (-x, -y, +-z)."""
    out = []
    for z in rng.uniform(0.05, 0.95, nring):
        phi = rng.uniform(0.0, np.pi, per_ring)
        r = np.sqrt(1.0 - z * z)
        x, y = r * np.cos(phi), r * np.sin(phi)
        zz = np.full(per_ring, z)
        for sx, sz in ((1, 1), (-1, 1), (1, -1), (-1, -1)):
            out.append(np.stack([sx * x, sx * y, sz * zz]))
    lmn = np.concatenate(out, axis=1)
    return lmn[:, rng.permutation(lmn.shape[1])]


def test_quad_counts(ops):
    """HEALPix: every pair off the equator ring is in a quad; the tiles are
    full enough for the quad forward from nside 8 on, and their quads sit at
    equal azimuth steps, so the forward takes the recurrent phase.  Quads at
    random azimuths keep the direct phase."""
    rng = np.random.default_rng(4)
    for nside, nest in ((16, False), (8, True), (64, False)):
        th, ph, _, _ = oracle.pix2ang(nside, nest)
        lmn = np.stack(oracle.pixel_dirs(th, ph, np.ones_like(th), 1.0)[:3])
        got = ops.Sky(torch.from_numpy(lmn).cuda()).pairs()
        print(f"healpix nside={nside} nest={nest}: {got}", flush=True)
        P = lmn.shape[1]
        assert got["pairs"] == P // 2 and got["quads"] == (P // 2 - 2 * nside) // 2 and got["uses_quads"]
        assert got["quad_recurrence"]
    # one quad of a HEALPix sky moved off its ring's azimuth grid by 8e-13 in x
    # (mirror images kept exact): still a quad, but the sky keeps the direct phase
    l, m, n = lmn = lmn.copy()
    i = int(np.argmax(n < 0.6)) + 2  # third pixel of a ring of 256: the third quad of its tile
    for s in (1.0, -1.0):
        j = np.where((np.abs(l - s * l[i]) < 1e-9) & (np.abs(m - s * m[i]) < 1e-9) & (np.abs(np.abs(n) - n[i]) < 1e-9))[0]
        assert len(j) == 2
        l[j] += s * 8e-13
    got = ops.Sky(torch.from_numpy(lmn).cuda()).pairs()
    print(f"healpix nside=64, one quad moved: {got}", flush=True)
    assert got["uses_quads"] and got["quads"] == (P // 2 - 2 * 64) // 2 and not got["quad_recurrence"]
    got = ops.Sky(torch.from_numpy(ring_sky(rng)).cuda()).pairs()
    print(f"rings: {got}", flush=True)
    assert got["quads"] == 20 * 48 and got["quad_tiles"] == 20 * 3 and got["uses_quads"]
    assert not got["quad_recurrence"]
    got = ops.Sky(torch.from_numpy(sky_of("exact-pm", rng)).cuda()).pairs()
    assert got["quads"] == 0 and not got["uses_quads"]
    # quads in tiles under 2/3 full (3 per ring): the pair kernels run instead
    got = ops.Sky(torch.from_numpy(ring_sky(rng, nring=40, per_ring=3)).cuda()).pairs()
    print(f"sparse rings: {got}", flush=True)
    assert got["quads"] == 120 and got["uses_pairs"] and not got["uses_quads"]


@pytest.mark.parametrize("kind", ["rings", "rings-minus", "rings-sparse"])
def test_quad_layouts_against_oracle(ops, kind):
    """Forward and back-projection over quad skies that are not HEALPix (rings
    of random azimuths; with 5% of the pixels removed: quads, leftover pairs
    and singles; sparse rings: quads too few per tile, so the pair kernels)."""
    rng = np.random.default_rng({"rings": 11, "rings-minus": 12, "rings-sparse": 13}[kind])
    lmn = ring_sky(rng, nring=40, per_ring=3) if kind == "rings-sparse" else ring_sky(rng)
    if kind == "rings-minus":
        keep = np.sort(rng.choice(lmn.shape[1], int(lmn.shape[1] * 0.95), replace=False))
        lmn = lmn[:, keep]
    P = lmn.shape[1]
    I = rng.lognormal(0.0, 1.0, P)
    c = make_case(nside=4, n_snap=40, n_sat=6, freq=10e6, sigma=20e3, seed=5)
    uvw, xa, xb = c["uvw"], c["xa"], c["xb"]
    sky = ops.Sky(torch.from_numpy(lmn).cuda())
    info = sky.pairs()
    print(f"{kind}: {info}", flush=True)
    t = {k: torch.from_numpy(v).cuda() for k, v in dict(uvw=uvw, xa=xa, xb=xb, I=I).items()}
    l, m, n = lmn
    for mask, track, herm in ((True, False, False), (True, True, True), (False, False, False)):
        flags = (oracle.MASK if mask else 0) | (oracle.PHASE_TRACK if track else 0) | (oracle.HERMITIAN if herm else 0)
        Vo = oracle.forward(l, m, n, I, uvw, xa, xb, flags=flags, R_moon=R)
        V = sky.forward(t["I"], t["uvw"], t["xa"], t["xb"], mask=mask, phase_track=track, hermitian=herm, R_moon=R)
        rv = rel_l2(V.cpu().numpy(), Vo)
        K = uvw.shape[1]
        y = rng.standard_normal(K) + 1j * rng.standard_normal(K)
        wt = rng.uniform(0.5, 1.5, K)
        bflags = flags & ~oracle.HERMITIAN
        xo = oracle.backproject(l, m, n, uvw, y, wt, xa, xb, flags=bflags, R_moon=R)
        x = sky.backproject(t["uvw"], torch.from_numpy(y).cuda(), torch.from_numpy(wt).cuda(), t["xa"], t["xb"],
                            mask=mask, phase_track=track, R_moon=R)
        rx = rel_l2(x.cpu().numpy(), xo)
        print(f"{kind} mask={mask} track={track} herm={herm}: forward {rv:.2e}, back-projection {rx:.2e}", flush=True)
        assert rv <= 1e-5 and rx <= 1e-5


def test_quad_recurrence_matches_direct_phase(ops, monkeypatch):
    """The quad forward's recurrent phase (A_{j+1} = 2 cos(d) A_j - A_{j-1} along a
    tile's quads) against its direct phase (u x + v y per quad) on one HEALPix
    sky, nside 64 RING, with masked, phase-tracked and unmasked rows.  The phases
    differ by under 2e-9 cycles, so the two V differ only where that moves an FP32
    argument by an ulp: within 1e-7 relative L2 (the mixed path's own error
    against FP64 is about 2e-7), and each within 1e-6 of the all-FP64 kernels."""
    th, ph, _, _ = oracle.pix2ang(64, False)
    lmn = torch.from_numpy(np.stack(oracle.pixel_dirs(th, ph, np.ones_like(th), 1.0)[:3])).cuda()
    rng = np.random.default_rng(21)
    I = torch.from_numpy(rng.lognormal(0.0, 1.0, lmn.shape[1])).cuda()
    c = make_case(nside=4, n_snap=60, n_sat=8, freq=10e6, sigma=50e3, seed=21)
    uvw, xa, xb = (torch.from_numpy(np.ascontiguousarray(c[k])).cuda() for k in ("uvw", "xa", "xb"))
    sky_rec = ops.Sky(lmn)
    monkeypatch.setenv("VISS_QUAD_RECURRENCE", "0")
    sky_dir = ops.Sky(lmn)
    monkeypatch.delenv("VISS_QUAD_RECURRENCE")
    assert sky_rec.pairs()["quad_recurrence"] and sky_dir.pairs()["uses_quads"]
    assert not sky_dir.pairs()["quad_recurrence"]
    for mask, track in ((True, False), (True, True), (False, False)):
        kw = dict(mask=mask, phase_track=track, R_moon=R)
        Va = sky_rec.forward(I, uvw, xa, xb, **kw).cpu().numpy()
        Vb = sky_dir.forward(I, uvw, xa, xb, **kw).cpu().numpy()
        V64 = sky_dir.forward(I, uvw, xa, xb, fp64=True, **kw).cpu().numpy()
        d, ea, eb = rel_l2(Va, Vb), rel_l2(Va, V64), rel_l2(Vb, V64)
        print(f"mask={mask} track={track}: recurrent vs direct {d:.2e}; vs FP64 {ea:.2e} / {eb:.2e}", flush=True)
        assert d < 1e-7 and ea < 1e-6 and eb < 1e-6
