"""Density-compensated imaging of the batched workload at its full size (5
frequencies x 64 groups x 500 snapshots x 28 baselines = 4,480,000 rows, nside
128), every group's weights and images from one batch:

  * viss_batch_weights against viss_weights on each group's rows alone, with
    and without the blockage indexing of dcf: gs, mb, dcf, dg and wt bit-exact
    (the same expressions; only the launch layout differs), and the first and
    last group against the oracle (check_weights);
  * viss_batch_backproject against viss_sky_backproject of each group's rows
    with those weights, summed over groups and per group: within 1e-6 relative
    L2 in mixed precision (the same tile kernels; row chunks and the order of
    the FP64 atomic additions differ), 1e-12 in FP64; graph replays against
    eager (1e-12);
  * per-group images of 100 rows of the first and the last group of every
    frequency (weights of the other rows set to 0) against the oracle's
    back-projection of those rows: 1e-5 mixed, 1e-11 FP64.
"""
import numpy as np
import pytest

import oracle
from tests.cases import check_weights, rel_l2
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
    bll = np.sqrt((uvw * uvw).sum(axis=0)).astype(np.float32)
    shells = np.ceil((bll - np.float32(0.25)) / np.float32(0.5)).astype(np.int64)
    nr = int(shells.max()) + 1
    th, ph = ops.pix2ang(128, nest=False, with_ring=False)
    lmn, I = ops.pixel_dirs(th, ph, torch.from_numpy(B).cuda())
    sky = ops.Sky(lmn)
    c = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    g = np.random.default_rng(44)
    y = c(g.normal(size=uvw.shape[1]) + 1j * g.normal(size=uvw.shape[1]))
    print(f"batched workload: {uvw.shape[1]} rows, {len(offs) - 1} groups, nr {nr}", flush=True)
    return ops, sky, lmn, c(uvw), c(xa), c(xb), c(bll), offs, gf, nr, y, (uvw, xa, xb, bll)


def _sample_rows(offs, gf):
    rows, groups = [], []
    for f in sorted(set(gf.tolist())):
        gs_ = np.flatnonzero(gf == f)
        for g in (gs_[0], gs_[-1]):
            rows.append(np.arange(offs[g], offs[g] + 100))
            groups.append(g)
    return rows, groups


@pytest.mark.parametrize("shift", [False, True])
def test_batch_weights_match_per_group(wl, shift):
    ops, sky, lmn, U, A, Bx, bll, offs, gf, nr, y, host = wl
    batch = ops.Batch(sky, offs, U, A, Bx, mask=True, R_moon=R)
    out = batch.weights(bll, nr, blockage_shift=shift)
    G = len(offs) - 1
    for g in range(G):
        s = slice(offs[g], offs[g + 1])
        ref = ops.weights(bll[s].contiguous(), U[2, s].contiguous(), nr, blockage_shift=shift)
        for key in ("gs", "wt", "dg"):
            assert torch.equal(out[key][s], ref[key]), (g, key)
        assert torch.equal(out["mb"][g], ref["mb"]), g
        assert torch.equal(out["dcf"][g], ref["dcf"]), g
    assert int(out["n_out_of_range"].item()) == 0
    print(f"batch weights of {G} groups bit-exact with per-group viss_weights (blockage shift {shift})", flush=True)
    uvw, xa, xb, bllh = host
    for g in (0, G - 1):
        s = slice(offs[g], offs[g + 1])
        sub = {k: (out[k][s] if k in ("gs", "wt", "dg") else out[k][g]) for k in ("gs", "wt", "dg", "mb", "dcf")}
        sub["n_out_of_range"] = torch.zeros(1, dtype=torch.int64)
        check_weights(sub, bllh[s], uvw[2, s], nr, shift=shift)


@pytest.mark.parametrize("fp64", [False, True])
def test_batch_backproject_matches_per_group(wl, fp64):
    ops, sky, lmn, U, A, Bx, bll, offs, gf, nr, y, host = wl
    G, P = len(offs) - 1, sky.P
    batch = ops.Batch(sky, offs, U, A, Bx, mask=True, R_moon=R, fp64=fp64)
    wt = batch.weights(bll, nr, want_dg=False)["wt"]
    x_sum = batch.backproject(y, wt)
    x_grp = batch.backproject(y, wt, per_group=True)
    ref = torch.zeros(G, P, dtype=torch.float64, device="cuda")
    for g in range(G):
        s = slice(offs[g], offs[g + 1])
        sky.backproject(U[:, s].contiguous(), y[s].contiguous(), wt[s].contiguous(), A[:, s].contiguous(),
                        Bx[:, s].contiguous(), mask=True, R_moon=R, fp64=fp64, out=ref[g])
    tol = 1e-12 if fp64 else 1e-6
    e_sum = float(torch.linalg.vector_norm(x_sum - ref.sum(0)) / torch.linalg.vector_norm(ref.sum(0)))
    e_grp = float(torch.linalg.vector_norm(x_grp - ref) / torch.linalg.vector_norm(ref))
    e_max = float(((x_grp - ref).norm(dim=1) / ref.norm(dim=1)).max())
    print(f"{'FP64' if fp64 else 'mixed'}: summed image rel L2 {e_sum:.2e}, per-group {e_grp:.2e} "
          f"(worst group {e_max:.2e})", flush=True)
    assert e_sum < tol and e_grp < tol and e_max < 10 * tol
    # graph replays against eager
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    xg = torch.zeros(P, dtype=torch.float64, device="cuda")
    with torch.cuda.stream(side):
        for rep in range(3):
            xg.zero_()
            batch.backproject(y, wt, out=xg, use_graph=True)
            side.synchronize()
            assert rel_l2(xg.cpu().numpy(), x_sum.cpu().numpy()) < 1e-12, rep
            batch.backproject(y, wt, per_group=True)  # eager work between replays
            side.synchronize()
    torch.cuda.synchronize()
    # oracle: per-group images of sampled rows (weights of all other rows 0)
    rows, groups = _sample_rows(offs, gf)
    mask = torch.zeros_like(wt)
    allr = torch.from_numpy(np.concatenate(rows)).cuda()
    mask[allr] = 1.0
    xs = batch.backproject(y, wt * mask, per_group=True)
    uvw, xa, xb, _ = host
    l, m, n = (lmn[i].cpu().numpy() for i in range(3))
    yh, wh = y.cpu().numpy(), wt.cpu().numpy()
    for r, g in zip(rows, groups):
        xo = oracle.backproject(l, m, n, uvw[:, r], yh[r], wh[r], xa[:, r], xb[:, r], flags=oracle.MASK, R_moon=R)
        e = rel_l2(xs[g].cpu().numpy(), xo)
        assert e < (1e-11 if fp64 else 1e-5), (g, e)
    print(f"  per-group images of {len(groups)} groups' sampled rows against the oracle within "
          f"{1e-11 if fp64 else 1e-5:g}", flush=True)
