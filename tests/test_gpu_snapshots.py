"""Snapshot-structured forward and back-projection (viss_sky_forward_snapshots,
viss_sky_backproject_snapshots) against the per-row kernels and the FP64
oracle on the same rows.

Tolerances: against the per-row mixed kernel 1e-6 relative L2 (both mixed
precision; different phase association d_a - d_b vs b.s); against the FP64
oracle 1e-5."""
import numpy as np
import pytest

import oracle
from tests.cases import rel_l2
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


@pytest.mark.parametrize("n_sat,nside,freq,sigma", [(2, 16, 1e6, 10e3), (4, 16, 1e6, 10e3), (10, 32, 10e6, 50e3),
                                                    (16, 16, 10e6, 50e3)])
def test_snapshot_backprojection(ops, n_sat, nside, freq, sigma):
    rng = np.random.default_rng(n_sat)
    th, ph, _, _ = oracle.pix2ang(nside, False)
    l, m, n, _ = oracle.pixel_dirs(th, ph, np.ones(th.size), 1.0)
    pos = geo.satellite_positions(n_sat, 30, sigma, True, 15, rng)
    lam = geo.C_LIGHT / freq
    uvw, xa, xb, bll = geo.baseline_rows(pos, lam)
    K = uvw.shape[1]
    y = rng.normal(size=K) + 1j * rng.normal(size=K)
    wt = rng.uniform(0, 0.125, K)
    c = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    sky = ops.Sky(c(np.stack([l, m, n])))
    for flags in ({}, {"mask": True}, {"mask": True, "phase_track": True, "hermitian": True}):
        f = (oracle.MASK if flags.get("mask") else 0) | (oracle.PHASE_TRACK if flags.get("phase_track") else 0) | \
            (oracle.HERMITIAN if flags.get("hermitian") else 0)
        ref = oracle.backproject(l, m, n, uvw, y, wt, xa, xb, flags=f, R_moon=R)
        per_row = sky.backproject(c(uvw), c(y), c(wt), c(xa), c(xb), R_moon=R, **flags).cpu().numpy()
        snap = sky.backproject_snapshots(c(pos), lam, c(y), c(wt), R_moon=R, **flags).cpu().numpy()
        assert rel_l2(snap, per_row) < 1e-6, flags
        assert rel_l2(snap, ref) < 1e-5, flags


@pytest.mark.parametrize("n_sat,nside,freq,sigma", [(2, 16, 1e6, 10e3), (3, 16, 1e6, 10e3), (8, 32, 10e6, 50e3),
                                                    (10, 32, 10e6, 50e3),
                                                    (12, 16, 1e6, 10e3), (16, 16, 10e6, 50e3)])
def test_snapshot_forward(ops, n_sat, nside, freq, sigma):
    rng = np.random.default_rng(100 + n_sat)
    th, ph, _, _ = oracle.pix2ang(nside, False)
    B = geo.make_sky(nside, rng, 5)
    l, m, n, I = oracle.pixel_dirs(th, ph, B, 4 * np.pi / B.size)
    pos = geo.satellite_positions(n_sat, 20, sigma, True, 15, rng)
    lam = geo.C_LIGHT / freq
    uvw, xa, xb, _ = geo.baseline_rows(pos, lam)
    c = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    sky = ops.Sky(c(np.stack([l, m, n])))
    for flags in ({}, {"mask": True}, {"mask": True, "phase_track": True, "hermitian": True}):
        f = (oracle.MASK if flags.get("mask") else 0) | (oracle.PHASE_TRACK if flags.get("phase_track") else 0) | \
            (oracle.HERMITIAN if flags.get("hermitian") else 0)
        ref = oracle.forward(l, m, n, I, uvw, xa, xb, flags=f, R_moon=R)
        per_row = sky.forward(c(I), c(uvw), c(xa), c(xb), R_moon=R, **flags).cpu().numpy()
        snap = sky.forward_snapshots(c(I), c(pos), lam, R_moon=R, **flags).cpu().numpy()
        assert rel_l2(snap, per_row) < 1e-6, flags
        assert rel_l2(snap, ref) < 1e-5, flags


def test_snapshot_backprojection_rejects_bad_input(ops):
    from viss_optimize_3d_amd._lib import VissError

    sky = ops.Sky(torch.tensor([[1.0], [0.0], [0.0]], dtype=torch.float64).cuda())
    pos = torch.zeros(2, 17, 3, dtype=torch.float64).cuda()
    with pytest.raises(VissError):
        sky.backproject_snapshots(pos, 1.0, torch.zeros(2 * 136, dtype=torch.complex128).cuda())
    with pytest.raises(VissError):
        sky.backproject_snapshots(pos[:, :4], -1.0, torch.zeros(12, dtype=torch.complex128).cuda())


def test_snapshot_counts_at_the_limits(ops):
    """No snapshot: nothing to do (the arrays may be empty); 1 or 17 satellites: a status, not a launch."""
    from viss_optimize_3d_amd._lib import VissError

    sky = ops.Sky(torch.tensor([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0]], dtype=torch.float64).cuda())
    I = torch.ones(2, dtype=torch.float64).cuda()
    V = sky.forward_snapshots(I, torch.zeros(0, 5, 3, dtype=torch.float64).cuda(), 1.0)
    assert V.numel() == 0
    x = sky.backproject_snapshots(torch.zeros(0, 5, 3, dtype=torch.float64).cuda(), 1.0,
                                  torch.zeros(0, dtype=torch.complex128).cuda())
    assert float(x.abs().max()) == 0.0
    for n in (1, 17):
        pos = torch.ones(3, n, 3, dtype=torch.float64).cuda()
        with pytest.raises((VissError, ValueError)):
            sky.forward_snapshots(I, pos, 1.0)
