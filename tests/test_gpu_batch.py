"""Batched one-launch forward (viss_batch_*): many snapshot groups x frequencies
with uneven compacted tile lists in one launch, eager and hipGraph-replayed,
against per-group launches and the FP64 oracle.

A HEALPix sky turned off its axes keeps its antipodal pairs but loses the
equatorial-mirror quads: there the batch runs its pair-tile kernel, checked
the same way, with the batched back-projection beside it.

Tolerances: batched vs per-group launches within 1e-12 relative L2 for the
all-FP64 kernels (the same kernel on the same tiles; only the FP64 combine
order differs) and 1e-6 for mixed precision (the same FP32 tile formulations;
FP32 partial sums combine in a different order); eager vs graph-replayed
batches within 1e-12; against the FP64 oracle within 1e-5 (mixed) / 1e-11
(FP64).
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
def setup():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import ops

    B, uvw, xa, xb, offs, gf = geo.batched_workload(nside=32, freqs=(1e6, 10e6), n_groups=6, snaps=40, n_sat=6)
    th, ph = ops.pix2ang(32, nest=False, with_ring=False)
    lmn, I = ops.pixel_dirs(th, ph, torch.from_numpy(B).cuda())
    sky = ops.Sky(lmn)
    c = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    return ops, sky, I, c(uvw), c(xa), c(xb), offs, lmn, (uvw, xa, xb)


@pytest.mark.parametrize("fp64", [False, True])
def test_batch_matches_per_group_and_oracle(setup, fp64):
    ops, sky, I, uvw, xa, xb, offs, lmn, host = setup
    batch = ops.Batch(sky, offs, uvw, xa, xb, mask=True, R_moon=R, fp64=fp64)
    st = batch.stats()
    if fp64:
        assert 0 < st["visible_tiles"] < st["total_tiles"]  # occultation compacted the tile lists
    else:
        # tiles of antipodal pairs: one side of a pair is always in view (the occulted cap is
        # smaller than a hemisphere), so the lists keep every tile; the kernel drops the hidden side
        assert 0 < st["visible_tiles"] <= st["total_tiles"]
    V = batch.forward(I)
    parts = [sky.forward(I, uvw[:, offs[g]:offs[g + 1]], xa[:, offs[g]:offs[g + 1]], xb[:, offs[g]:offs[g + 1]],
                         mask=True, R_moon=R, fp64=fp64) for g in range(len(offs) - 1)]
    Vg = torch.cat(parts)
    assert rel_l2(V.cpu().numpy(), Vg.cpu().numpy()) < (1e-12 if fp64 else 1e-6)
    ln = lmn.cpu().numpy()
    rows = np.r_[0:200, offs[-2]:offs[-1]]
    ref = oracle.forward(*ln, I.cpu().numpy(), host[0][:, rows], host[1][:, rows], host[2][:, rows],
                         flags=oracle.MASK, R_moon=R)
    assert rel_l2(V.cpu().numpy()[rows], ref) < (1e-11 if fp64 else 1e-5)


def test_batch_graph_replays_and_autotune(setup, tmp_path):
    ops, sky, I, uvw, xa, xb, offs, lmn, host = setup
    batch = ops.Batch(sky, offs, uvw, xa, xb, mask=True, R_moon=R)
    eager = batch.forward(I)
    side = torch.cuda.Stream()
    out = torch.zeros_like(eager)
    with torch.cuda.stream(side):
        for rep in range(3):
            out.zero_()
            batch.forward(I, out=out, use_graph=True)
            side.synchronize()
            # eager work between replays
            chk = batch.forward(I)
            side.synchronize()
            assert rel_l2(out.cpu().numpy(), eager.cpu().numpy()) < 1e-12, rep
            assert rel_l2(chk.cpu().numpy(), eager.cpu().numpy()) < 1e-12, rep
    torch.cuda.synchronize()
    batch.autotune(I, reps=1)
    V2 = batch.forward(I)
    assert rel_l2(V2.cpu().numpy(), eager.cpu().numpy()) < 1e-12
    path = str(tmp_path / "tune.txt")
    batch.save_tuning(path)
    lines = [ln for ln in open(path) if not ln.startswith("#")]
    assert lines, "autotuner wrote no bucket"
    b2 = ops.Batch(sky, offs, uvw, xa, xb, mask=True, R_moon=R)
    b2.load_tuning(path)
    V3 = b2.forward(I)
    assert rel_l2(V3.cpu().numpy(), eager.cpu().numpy()) < 1e-12


def test_batch_over_pairs_without_quads(setup):
    """Rotated nside-32 HEALPix sky: pairs, no quads; batch forward (eager and graph) and
    back-projection against per-group launches, forward rows against the oracle."""
    ops, _, _, uvw, xa, xb, offs, lmn0, host = setup
    a, b = 0.3, 0.7
    rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    lmn_h = (rx @ rz) @ lmn0.cpu().numpy()
    lmn = torch.from_numpy(np.ascontiguousarray(lmn_h)).cuda()
    sky = ops.Sky(lmn)
    lay = sky.pairs()
    print(f"rotated nside 32: {lay}", flush=True)
    assert lay["uses_pairs"] and not lay["uses_quads"] and lay["singles"] == 0
    rng = np.random.default_rng(21)
    I = torch.from_numpy(rng.lognormal(0.0, 1.0, sky.P)).cuda()
    batch = ops.Batch(sky, offs, uvw, xa, xb, mask=True, R_moon=R)
    V = batch.forward(I)
    G = len(offs) - 1
    Vg = torch.cat([sky.forward(I, uvw[:, offs[g]:offs[g + 1]], xa[:, offs[g]:offs[g + 1]],
                                xb[:, offs[g]:offs[g + 1]], mask=True, R_moon=R) for g in range(G)])
    e = rel_l2(V.cpu().numpy(), Vg.cpu().numpy())
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    out = torch.zeros_like(V)
    with torch.cuda.stream(side):
        for rep in range(2):
            out.zero_()
            batch.forward(I, out=out, use_graph=True)
            side.synchronize()
            assert rel_l2(out.cpu().numpy(), V.cpu().numpy()) < 1e-12, rep
    torch.cuda.synchronize()
    rows = np.r_[0:200, offs[-2]:offs[-1]]
    ref = oracle.forward(*lmn_h, I.cpu().numpy(), host[0][:, rows], host[1][:, rows], host[2][:, rows],
                         flags=oracle.MASK, R_moon=R)
    eo = rel_l2(V.cpu().numpy()[rows], ref)
    y = torch.from_numpy(rng.standard_normal(V.numel()) + 1j * rng.standard_normal(V.numel())).cuda()
    wt = torch.from_numpy(rng.uniform(0.5, 1.5, V.numel())).cuda()
    x = batch.backproject(y, wt)
    xg = torch.zeros(sky.P, dtype=torch.float64, device="cuda")
    for g in range(G):
        s = slice(offs[g], offs[g + 1])
        sky.backproject(uvw[:, s].contiguous(), y[s].contiguous(), wt[s].contiguous(), xa[:, s].contiguous(),
                        xb[:, s].contiguous(), mask=True, R_moon=R, out=xg)
    ex = rel_l2(x.cpu().numpy(), xg.cpu().numpy())
    print(f"pairs batch: forward vs per-group {e:.2e}, vs oracle {eo:.2e}; back-projection vs per-group {ex:.2e}",
          flush=True)
    assert e < 1e-6 and eo < 1e-5 and ex < 1e-6


def test_batch_with_empty_groups(setup):
    """Groups without rows (first, middle, last) in a batch: forward (eager and graph), weights
    and per-group images against per-group calls; an empty group's weights are those of
    viss_weights on no rows and its image stays 0."""
    ops, sky, I, uvw, xa, xb, offs, lmn, host = setup
    K = 350
    offs2 = [0, 0, 100, 100, 350, 350]
    U, A, Bx = (t[:, :K].contiguous() for t in (uvw, xa, xb))
    batch = ops.Batch(sky, offs2, U, A, Bx, mask=True, R_moon=R)
    V = batch.forward(I)
    groups = [(offs2[g], offs2[g + 1]) for g in range(len(offs2) - 1)]
    Vg = torch.zeros_like(V)
    for k0, k1 in groups:
        if k1 > k0:
            sky.forward(I, U[:, k0:k1].contiguous(), A[:, k0:k1].contiguous(), Bx[:, k0:k1].contiguous(), mask=True,
                        R_moon=R, out=Vg[k0:k1])
    assert rel_l2(V.cpu().numpy(), Vg.cpu().numpy()) < 1e-6
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    out = torch.zeros_like(V)
    with torch.cuda.stream(side):
        batch.forward(I, out=out, use_graph=True)
        side.synchronize()
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), V.cpu().numpy()) < 1e-12
    bll = torch.sqrt((U * U).sum(0)).float()
    nr = int(torch.ceil((bll.max() - 0.25) / 0.5).item()) + 1
    w = batch.weights(bll, nr)
    for g, (k0, k1) in enumerate(groups):
        ref = ops.weights(bll[k0:k1].contiguous(), U[2, k0:k1].contiguous(), nr)
        assert torch.equal(w["mb"][g], ref["mb"]), g
        assert torch.equal(w["dcf"][g], ref["dcf"]), g
        for key in ("gs", "wt", "dg"):
            assert torch.equal(w[key][k0:k1], ref[key]), (g, key)
    rng = np.random.default_rng(9)
    y = torch.from_numpy(rng.standard_normal(K) + 1j * rng.standard_normal(K)).cuda()
    x = batch.backproject(y, w["wt"], per_group=True)
    for g, (k0, k1) in enumerate(groups):
        if k1 == k0:
            assert float(x[g].abs().max()) == 0.0, g
        else:
            xr = sky.backproject(U[:, k0:k1].contiguous(), y[k0:k1].contiguous(), w["wt"][k0:k1].contiguous(),
                                 A[:, k0:k1].contiguous(), Bx[:, k0:k1].contiguous(), mask=True, R_moon=R)
            assert rel_l2(x[g].cpu().numpy(), xr.cpu().numpy()) < 1e-6, g
    print(f"batch with empty groups {[g for g, (a, b) in enumerate(groups) if a == b]}: forward, weights and images "
          f"as per-group calls", flush=True)


def test_batch_without_rows(setup):
    """A batch whose groups are all empty: the forward returns nothing, the weights are those of
    viss_weights on no rows, the images stay 0."""
    ops, sky, I, uvw, xa, xb, offs, lmn, host = setup
    e = torch.zeros(3, 0, dtype=torch.float64, device="cuda")
    batch = ops.Batch(sky, [0, 0, 0], e, e, e, mask=True, R_moon=R)
    V = batch.forward(I)
    assert V.numel() == 0
    V = batch.forward(I, out=V, use_graph=False)
    bll = torch.zeros(0, dtype=torch.float32, device="cuda")
    w = batch.weights(bll, 8)
    ref = ops.weights(bll, torch.zeros(0, dtype=torch.float64, device="cuda"), 8)
    for g in range(2):
        assert torch.equal(w["mb"][g], ref["mb"]) and torch.equal(w["dcf"][g], ref["dcf"]), g
    assert int(w["n_out_of_range"].item()) == 0
    x = batch.backproject(torch.zeros(0, dtype=torch.complex128, device="cuda"), w["wt"], per_group=True)
    assert tuple(x.shape) == (2, sky.P) and float(x.abs().max()) == 0.0
