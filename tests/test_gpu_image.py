"""Image utilities after the hot path: the multi-day sum (the notebooks'
total += nan_to_num(C_i)) and the NEST <-> RING reorder, against numpy and the
oracle's independently written nest2ring.  Bit-exact (sums of exact doubles in
the same order; permutations)."""
import numpy as np
import pytest

import oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import ops as _ops

    return _ops


def test_image_accumulate(ops):
    rng = np.random.default_rng(1)
    imgs = [rng.normal(size=3000) for _ in range(5)]
    imgs[2][[3, 17]] = np.nan
    total = torch.zeros(3000, dtype=torch.float64, device="cuda")
    for im in imgs:
        ops.image_accumulate(torch.from_numpy(im).cuda(), total, nan_to_zero=True)
    ref = np.zeros(3000)
    for im in imgs:
        ref += np.nan_to_num(im)
    np.testing.assert_array_equal(total.cpu().numpy(), ref)
    t2 = torch.zeros(3000, dtype=torch.float64, device="cuda")
    ops.image_accumulate(torch.from_numpy(imgs[2]).cuda(), t2, nan_to_zero=False)
    assert np.isnan(t2.cpu().numpy()[3])


@pytest.mark.parametrize("nside", [1, 2, 16, 128])
def test_reorder_matches_nest2ring(ops, nside):
    P = 12 * nside * nside
    x = np.random.default_rng(nside).normal(size=P)
    perm = oracle.nest2ring(nside, np.arange(P))
    ring = ops.reorder(torch.from_numpy(x).cuda(), nside, to_ring=True).cpu().numpy()
    expect = np.empty(P)
    expect[perm] = x
    np.testing.assert_array_equal(ring, expect)
    back = ops.reorder(torch.from_numpy(ring).cuda(), nside, to_ring=False).cpu().numpy()
    np.testing.assert_array_equal(back, x)
