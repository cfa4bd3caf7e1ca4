"""The allocation-free back-projection (viss_sky_backproject_noalloc) through
the C ABI, against viss_sky_backproject on the same rows.

Both run the same kernels on the same row records; only the order of the FP64
atomic additions that join row chunks differs, so the results agree to 1e-12
relative L2.  Two calls in a row, the second with fewer rows, show that the
shared static scratch keeps no state between calls.  Device free memory is
unchanged across a call on a fresh sky (the sky's own entry point allocates
its row-record buffer on the same first call), and row counts past the static
array return VISS_ERR_SIZE.
"""
import gc

import numpy as np
import pytest

from tests.cases import make_case, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
R = 1737.4e3


@pytest.fixture(scope="module")
def env():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import _lib, ops

    c = make_case(nside=32, n_snap=7000, n_sat=6, freq=10e6, sigma=30e3, seed=21)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    rows = dict(uvw=t(c["uvw"]), xa=t(c["xa"]), xb=t(c["xb"]))
    K = c["uvw"].shape[1]
    g = np.random.default_rng(5)
    y = t(g.normal(size=K) + 1j * g.normal(size=K))
    wt = t(g.uniform(0.0, 0.125, K))
    return _lib, ops, c, rows, y, wt, K


def _noalloc(L, sky, uvw, xa, xb, y, wt, K, flags, x, stream):
    from viss_optimize_3d_amd import ops

    geom = ops._geom(R, 0.0)
    return L.viss_sky_backproject_noalloc(sky._h, ops._p(uvw), ops._p(xa), ops._p(xb), ops._p(y), ops._p(wt), K,
                                          geom, flags, ops._p(x), stream)


def _sub(rows, y, wt, n):
    return (rows["uvw"][:, :n].contiguous(), rows["xa"][:, :n].contiguous(), rows["xb"][:, :n].contiguous(),
            y[:n].contiguous(), wt[:n].contiguous())


@pytest.mark.parametrize("fp64,mask,track,herm", [(False, True, False, False), (True, True, False, False),
                                                 (False, False, True, True), (False, True, True, False)])
def test_noalloc_matches_backproject(env, fp64, mask, track, herm):
    _lib, ops, c, rows, y, wt, K = env
    L = _lib.load()
    sky = ops.Sky(torch.from_numpy(c["lmn"]).cuda())
    assert sky.pairs()["uses_quads"]
    flags = ops._flags(mask, track, fp64, herm)
    stream = ops._stream()
    for n in (K, 37_123):  # the second call has fewer rows
        print(f"noalloc fp64={fp64} mask={mask} track={track} herm={herm} rows={n}", flush=True)
        uvw, xa, xb, yy, ww = _sub(rows, y, wt, n)
        x_na = torch.zeros(sky.P, dtype=torch.float64, device="cuda")
        _lib.check("viss_sky_backproject_noalloc", _noalloc(L, sky, uvw, xa, xb, yy, ww, n, flags, x_na, stream))
        x_ref = sky.backproject(uvw, yy, ww, xa, xb, mask=mask, R_moon=R, phase_track=track, hermitian=herm,
                                fp64=fp64)
        e = rel_l2(x_na.cpu().numpy(), x_ref.cpu().numpy())
        print(f"  rel L2 vs viss_sky_backproject {e:.2e}", flush=True)
        assert e < 1e-12


def test_noalloc_allocates_nothing_and_rejects_oversize(env):
    _lib, ops, c, rows, y, wt, K = env
    L = _lib.load()
    lmn = torch.from_numpy(c["lmn"]).cuda()
    flags = ops._flags(True, False, False, False)
    stream = ops._stream()
    uvw, xa, xb, yy, ww = _sub(rows, y, wt, K)
    x = torch.zeros(lmn.shape[1], dtype=torch.float64, device="cuda")
    warm = ops.Sky(lmn)  # the first call on the device looks up the scratch address
    _lib.check("viss_sky_backproject_noalloc", _noalloc(L, warm, uvw, xa, xb, yy, ww, K, flags, x, stream))
    torch.cuda.synchronize()
    fresh_a, fresh_b = ops.Sky(lmn), ops.Sky(lmn)  # no back-projection has run on these skies yet
    # objects of earlier tests freed by the collector during the measurement would raise free memory
    gc.collect()
    gc.disable()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    _lib.check("viss_sky_backproject_noalloc", _noalloc(L, fresh_a, uvw, xa, xb, yy, ww, K, flags, x, stream))
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    print(f"free device memory across the allocation-free call: {free0} -> {free1}", flush=True)
    if free1 != free0:
        gc.enable()
    assert free1 == free0
    # the sky's own entry point allocates its row records (K * 128 B) on its first call
    fresh_b.backproject(uvw, yy, ww, xa, xb, mask=True, R_moon=R, out=x)
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info()[0]
    print(f"free device memory across viss_sky_backproject's first call: {free1} -> {free2}", flush=True)
    gc.enable()
    assert free1 - free2 >= K * 128
    # row counts past the static array: rejected on the host, nothing launched
    cap = (_lib.NOALLOC_SCRATCH_BYTES - 256) // 128
    big = cap + 1
    rc = L.viss_sky_backproject_noalloc(fresh_a._h, ops._p(uvw), ops._p(xa), ops._p(xb), ops._p(yy), ops._p(ww),
                                        big, ops._geom(R, 0.0), flags, ops._p(x), stream)
    assert rc == -3
    assert b"static" in L.viss_last_error()
    assert cap == 262_142
    assert (_lib.NOALLOC_SCRATCH_BYTES - 256) // 160 == 209_713
    fresh_a.close()
    fresh_b.close()
    warm.close()
