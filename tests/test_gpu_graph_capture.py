"""The steady state of the per-row entry points neither synchronises nor
allocates, so it can be captured into a HIP graph.

After one warm-up call (which sizes the sky's row-record buffer), a torch CUDA
graph captures viss_sky_forward followed by viss_sky_backproject of its output
(mixed precision, occultation on), with the outputs cleared by fill kernels
inside the graph.  Capture fails if any captured call synchronises or
allocates.  Replays, with eager work of the program between them, must equal
the eager results (to 1e-12: only the order of the FP64 atomic additions that
join partial sums may differ).  viss_sky_backproject_noalloc is captured the
same way on a sky that never ran a back-projection.
"""
import numpy as np
import pytest

from tests.cases import make_case, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
R = 1737.4e3


def test_forward_backproject_graph_replays():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import _lib, ops

    c = make_case(nside=32, n_snap=400, n_sat=8, freq=10e6, sigma=50e3, seed=31)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    lmn, I, uvw, xa, xb = t(c["lmn"]), t(c["I"]), t(c["uvw"]), t(c["xa"]), t(c["xb"])
    K = uvw.shape[1]
    sky = ops.Sky(lmn)
    assert sky.pairs()["uses_quads"]
    V_e = sky.forward(I, uvw, xa, xb, mask=True, R_moon=R)
    x_e = sky.backproject(uvw, V_e, None, xa, xb, mask=True, R_moon=R)
    torch.cuda.synchronize()
    V = torch.zeros(K, dtype=torch.complex128, device="cuda")
    y = torch.zeros_like(V)
    x = torch.zeros(sky.P, dtype=torch.float64, device="cuda")

    def step():
        V.zero_()
        sky.forward(I, uvw, xa, xb, mask=True, R_moon=R, out=V)
        y.copy_(V)
        x.zero_()
        sky.backproject(uvw, y, None, xa, xb, mask=True, R_moon=R, out=x)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()  # warm-up on the capture stream
    side.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        step()
    for rep in range(4):
        V.fill_(7.0)
        x.fill_(7.0)
        g.replay()
        torch.cuda.synchronize()
        ev = rel_l2(V.cpu().numpy(), V_e.cpu().numpy())
        ex = rel_l2(x.cpu().numpy(), x_e.cpu().numpy())
        print(f"replay {rep}: V rel L2 {ev:.2e}, x rel L2 {ex:.2e}", flush=True)
        assert ev < 1e-12 and ex < 1e-12
        # eager work of the program between replays
        chk = sky.forward(I, uvw, xa, xb, mask=True, R_moon=R)
        torch.cuda.synchronize()
        assert rel_l2(chk.cpu().numpy(), V_e.cpu().numpy()) < 1e-12

    # the allocation-free back-projection, captured on a sky whose own buffer was never sized
    L = _lib.load()
    fresh = ops.Sky(lmn)
    flags = ops._flags(True, False, False, False)
    x2 = torch.zeros(sky.P, dtype=torch.float64, device="cuda")
    with torch.cuda.stream(side):
        _lib.check("viss_sky_backproject_noalloc", L.viss_sky_backproject_noalloc(
            sky._h, ops._p(uvw), ops._p(xa), ops._p(xb), ops._p(V_e), None, K, ops._geom(R, 0.0), flags,
            ops._p(x2), ops._stream()))  # first call on the device: the scratch address lookup
    side.synchronize()
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=side):
        x2.zero_()
        _lib.check("viss_sky_backproject_noalloc", L.viss_sky_backproject_noalloc(
            fresh._h, ops._p(uvw), ops._p(xa), ops._p(xb), ops._p(V_e), None, K, ops._geom(R, 0.0), flags,
            ops._p(x2), ops._stream()))
    for rep in range(3):
        x2.fill_(-1.0)
        g2.replay()
        torch.cuda.synchronize()
        e2 = rel_l2(x2.cpu().numpy(), x_e.cpu().numpy())
        print(f"no-alloc replay {rep}: x rel L2 {e2:.2e}", flush=True)
        assert e2 < 1e-12
    del g, g2
    fresh.close()
    sky.close()
