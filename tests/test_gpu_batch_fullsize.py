"""The batched workload at its full size: 5 frequencies (1, 2.5, 5, 7.5, 10 MHz)
x 64 snapshot groups x 500 snapshots x 28 baselines = 4,480,000 rows over a
HEALPix RING nside 128 sky (196,608 pixels), occultation on, each group at its
own orbit radius so that it sees its own share of the sky.

  * the one-launch batch (eager, then hipGraph-replayed after the autotuner
    has rebuilt the work table) against one viss_sky_forward per group, on
    every row: mixed within 1e-6 relative L2 (both run the quad tile loop;
    the batch cuts a group's tile list into items and joins their partial
    sums in a different FP64 order), all-FP64 within 1e-12;
  * rows of the first and the last group of every frequency against the FP64
    oracle: mixed within 1e-5, all-FP64 within 1e-11;
  * every occultation decision, exactly: with zero baselines every phase is
    0, so unit brightness makes each row's V the number of pixels its two
    satellites both see (integers, summed exactly); the mixed batch and the
    all-FP64 batch must give the all-FP64 per-group kernels' counts row by
    row, and their total must equal viss_sky_count_visible's.
"""
import numpy as np
import pytest

import oracle
from tests.cases import rel_l2
from viss_optimize_3d_amd import geometry as geo

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
R = geo.R_MOON


@pytest.fixture(scope="module")
def wl():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import ops

    B, uvw, xa, xb, offs, gf = geo.batched_workload()
    assert uvw.shape[1] == 4_480_000 and len(offs) == 321
    th, ph = ops.pix2ang(128, nest=False, with_ring=False)
    lmn, I = ops.pixel_dirs(th, ph, torch.from_numpy(B).cuda())
    sky = ops.Sky(lmn)
    lay = sky.pairs()
    assert lay["uses_quads"]
    c = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    print(f"batched workload: {uvw.shape[1]} rows, {len(offs) - 1} groups, {sky.P} pixels, "
          f"{lay['quads']} quads", flush=True)
    return ops, sky, I, lmn, c(uvw), c(xa), c(xb), offs, gf, (uvw, xa, xb)


def _per_group(sky, I, U, A, Bx, offs, fp64, out):
    for g in range(len(offs) - 1):
        s = slice(offs[g], offs[g + 1])
        sky.forward(I, U[:, s].contiguous(), A[:, s].contiguous(), Bx[:, s].contiguous(), mask=True, R_moon=R,
                    fp64=fp64, out=out[s])
    return out


def _oracle_rows(offs, gf):
    """100 rows of the first and of the last group of each frequency."""
    rows = []
    freqs = sorted(set(gf.tolist()))
    for f in freqs:
        gs = np.flatnonzero(gf == f)
        for g in (gs[0], gs[-1]):
            rows.append(np.arange(offs[g], offs[g] + 100))
    return np.concatenate(rows)


def test_batch_mixed_full_size(wl):
    ops, sky, I, lmn, U, A, Bx, offs, gf, host = wl
    K = U.shape[1]
    batch = ops.Batch(sky, offs, U, A, Bx, mask=True, R_moon=R)
    st = batch.stats()
    print(f"mixed batch: {st}", flush=True)
    V_eager = batch.forward(I)
    V_group = _per_group(sky, I, U, A, Bx, offs, False, torch.zeros(K, dtype=torch.complex128, device="cuda"))
    e = rel_l2(V_eager.cpu().numpy(), V_group.cpu().numpy())
    print(f"  eager batch vs per-group launches: rel L2 {e:.2e}", flush=True)
    assert e < 1e-6
    batch.autotune(I, reps=1)
    st2 = batch.stats()
    print(f"  after autotune: {st2['n_items']} items (untuned {st['n_items']})", flush=True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    out = torch.zeros(K, dtype=torch.complex128, device="cuda")
    with torch.cuda.stream(side):
        for rep in range(3):
            out.zero_()
            batch.forward(I, out=out, use_graph=True)
            side.synchronize()
            eg = rel_l2(out.cpu().numpy(), V_group.cpu().numpy())
            print(f"  graph replay {rep}: rel L2 vs per-group {eg:.2e}", flush=True)
            assert eg < 1e-6
            chk = batch.forward(I)  # eager work between replays
            side.synchronize()
            assert rel_l2(chk.cpu().numpy(), out.cpu().numpy()) < 1e-12
    torch.cuda.synchronize()
    rows = _oracle_rows(offs, gf)
    u, a, b = (h[:, rows] for h in host)
    ref = oracle.forward(*lmn.cpu().numpy(), I.cpu().numpy(), u, a, b, flags=oracle.MASK, R_moon=R)
    eo = rel_l2(out.cpu().numpy()[rows], ref)
    print(f"  graph-replayed batch vs the FP64 oracle on {rows.size} rows: rel L2 {eo:.2e}", flush=True)
    assert eo < 1e-5
    for f in sorted(set(gf.tolist())):  # each frequency's own error
        sel = np.flatnonzero(np.isin(rows, np.concatenate([np.arange(offs[g], offs[g + 1])
                                                           for g in np.flatnonzero(gf == f)])))
        assert rel_l2(out.cpu().numpy()[rows[sel]], ref[sel]) < 1e-5, f


def test_batch_fp64_full_size(wl):
    ops, sky, I, lmn, U, A, Bx, offs, gf, host = wl
    K = U.shape[1]
    batch = ops.Batch(sky, offs, U, A, Bx, mask=True, R_moon=R, fp64=True)
    st = batch.stats()
    assert 0 < st["visible_tiles"] < st["total_tiles"]  # occultation compacted the per-group tile lists
    print(f"FP64 batch: {st}, kept tile fraction {st['visible_tiles'] / st['total_tiles']:.3f}", flush=True)
    V = batch.forward(I)
    Vg = _per_group(sky, I, U, A, Bx, offs, True, torch.zeros(K, dtype=torch.complex128, device="cuda"))
    e = rel_l2(V.cpu().numpy(), Vg.cpu().numpy())
    print(f"  FP64 batch vs per-group launches: rel L2 {e:.2e}", flush=True)
    assert e < 1e-12
    rows = _oracle_rows(offs, gf)
    u, a, b = (h[:, rows] for h in host)
    ref = oracle.forward(*lmn.cpu().numpy(), I.cpu().numpy(), u, a, b, flags=oracle.MASK, R_moon=R)
    eo = rel_l2(V.cpu().numpy()[rows], ref)
    print(f"  FP64 batch vs the FP64 oracle on {rows.size} rows: rel L2 {eo:.2e}", flush=True)
    assert eo < 1e-11


def test_batch_visible_counts_exact(wl):
    ops, sky, I, lmn, U, A, Bx, offs, gf, host = wl
    K = U.shape[1]
    zero = torch.zeros_like(U)
    ones = torch.ones(sky.P, dtype=torch.float64, device="cuda")
    ref = _per_group(sky, ones, zero, A, Bx, offs, True, torch.zeros(K, dtype=torch.complex128, device="cuda"))
    cnt = ref.real.cpu().numpy()
    assert np.array_equal(cnt, np.round(cnt)) and np.all(ref.imag.cpu().numpy() == 0)
    total = int(sky.count_visible(A, Bx, R_moon=R).item())
    assert int(cnt.sum()) == total
    frac = cnt.reshape(-1, 14_000).sum(axis=1) / (14_000 * sky.P)  # each group's visible share
    print(f"visible pairs {total}; per-group visible share {frac.min():.3f} .. {frac.max():.3f}", flush=True)
    assert frac.min() >= 0.4 and frac.max() <= 0.8  # blockage removes 20-60% of the pixels per group
    for fp64 in (False, True):
        batch = ops.Batch(sky, offs, zero, A, Bx, mask=True, R_moon=R, fp64=fp64)
        V = batch.forward(ones)
        got = V.real.cpu().numpy()
        bad = np.flatnonzero(got != cnt)
        print(f"  {'FP64' if fp64 else 'mixed'} batch: {bad.size} rows with a different visible count", flush=True)
        assert bad.size == 0, (bad[:10], got[bad[:10]], cnt[bad[:10]])
