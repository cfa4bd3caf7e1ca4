"""Batched one-launch forward (viss_batch_*): many snapshot groups x frequencies
with uneven compacted tile lists in one launch, eager and hipGraph-replayed,
against per-group launches and the FP64 oracle.

Tolerances: batched vs per-group launches within 1e-12 relative L2 for the
all-FP64 kernels (the same kernel on the same tiles; only the FP64 combine
order differs) and 1e-6 for mixed precision (per-group mixed launches on a
paired sky use the antipodal-pair kernel, the batch the per-pixel one: two
FP32 formulations); eager vs graph-replayed batches within 1e-12; against the
FP64 oracle within 1e-5 (mixed) / 1e-11 (FP64).
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
