"""The torch.autograd.Function around the forward kernel: its backward (the
hand-written adjoint kernel) against central finite differences and the FP64
CPU adjoint, and a short CG reconstruction.

Tolerances: finite differences (FP64 kernels, h = 1e-3, quadratic loss so the
central difference has no truncation error) within 1e-7 relative; mixed
gradient vs the FP64 oracle adjoint within 1e-5 relative L2.
"""
import numpy as np
import pytest

import oracle
from tests.cases import make_case, rel_l2
from viss_optimize_3d_amd import geometry as geo

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
R_MOON = 1737.4e3


@pytest.fixture(scope="module")
def problem():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import autograd, ops

    c = make_case(nside=32, n_snap=120, n_sat=8, freq=10e6, sigma=10e3, seed=31, n_sources=10)
    sky = ops.Sky(torch.from_numpy(c["lmn"]).cuda())
    ops_ = {}
    for fp64 in (True, False):
        ops_[fp64] = autograd.VisibilityOperator(sky, torch.from_numpy(c["uvw"]), torch.from_numpy(c["xa"]),
                                                 torch.from_numpy(c["xb"]), mask=True, R_moon=R_MOON, fp64=fp64)
    x_true = torch.from_numpy(c["B"]).cuda()
    rng = np.random.default_rng(4)
    V = ops_[True].forward(x_true)
    noise = (rng.normal(size=V.numel()) + 1j * rng.normal(size=V.numel())) * float(V.abs().mean()) / 20 / np.sqrt(2)
    Vn = V + torch.from_numpy(noise).cuda()
    return c, ops_, x_true, Vn, autograd


def loss_fn(op, x, V, lam=1e-3):
    r = op(x) - V
    return (r.real ** 2 + r.imag ** 2).sum() + lam * (x ** 2).sum()


def test_gradient_vs_finite_differences(problem):
    c, ops_, x_true, V, _ = problem
    op = ops_[True]
    x = (x_true * 0.7).clone().requires_grad_(True)
    L = loss_fn(op, x, V)
    L.backward()
    g = x.grad.detach().cpu().numpy()
    rng = np.random.default_rng(5)
    h = 1e-3
    worst = 0.0
    for p in rng.choice(x.numel(), 32, replace=False):
        e = torch.zeros_like(x_true)
        e[p] = h
        with torch.no_grad():
            fd = (float(loss_fn(op, x + e, V)) - float(loss_fn(op, x - e, V))) / (2 * h)
        worst = max(worst, abs(fd - g[p]) / (np.abs(g).max()))
    assert worst < 1e-7, worst


def test_mixed_gradient_vs_oracle_adjoint(problem):
    c, ops_, x_true, V, _ = problem
    x = (x_true * 0.7).clone().requires_grad_(True)
    L = loss_fn(ops_[False], x, V)
    L.backward()
    g = x.grad.detach().cpu().numpy()
    xn = x.detach().cpu().numpy()
    r = oracle.forward(*c["lmn"], xn, c["uvw"], c["xa"], c["xb"], flags=oracle.MASK, R_moon=R_MOON) - V.cpu().numpy()
    g_ref = 2 * oracle.backproject(*c["lmn"], c["uvw"], r, None, c["xa"], c["xb"], flags=oracle.MASK,
                                   R_moon=R_MOON) + 2e-3 * xn
    assert rel_l2(g, g_ref) < 1e-5


def test_cg_reduces_residual(problem):
    c, ops_, x_true, V, autograd = problem
    hist = []
    x = autograd.lsq_reconstruct(ops_[False], V, lam=1e-3, iters=15, callback=lambda i, x, r: hist.append(r))
    assert hist[-1] < 0.1 * hist[0]
    L0 = float(loss_fn(ops_[True], torch.zeros_like(x_true), V))
    L1 = float(loss_fn(ops_[True], x, V))
    assert L1 < 0.05 * L0


def test_reconstruction_workload_full_size():
    """The differentiable-reconstruction workload at full size: nside 128
    (196,608 pixels), 200,000 visibilities at 10 MHz from 8 satellites, an
    FP32-valued sky (lognormal + 50 sources), complex Gaussian noise at SNR 20,
    loss ||Ax - V||^2 + 1e-3 ||x||^2.
    - The FP64 gradient (autograd through the adjoint kernel) against central
      finite differences, h = 1e-3, on 32 pixels, within 1e-7 of the largest
      gradient entry.  The loss is ~1e12 here, so its own FP64 rounding is far
      above the change a step of 1e-3 makes; the difference
      L(x + h e_p) - L(x - h e_p) is therefore summed row by row as
      sum_k |r+_k|^2 - |r-_k|^2 (the same quantity, without subtracting two
      large totals).
    - The mixed gradient against the FP64 gradient (relative L2 <= 1e-5).
    - 50 CG iterations with the mixed kernels: the normal-equation residual
      falls 10x or more and the loss below 5% of its value at x = 0."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import autograd, ops

    wl = geo.WORKLOADS["diff-nside128"]
    inp = geo.generate(wl)
    K, P = inp.uvw.shape[1], wl.npix
    assert K == 200_000 and P == 196_608
    dev = torch.device("cuda:0")
    th, ph = ops.pix2ang(wl.nside, nest=False, with_ring=False)
    lmn, _ = ops.pixel_dirs(th, ph, torch.from_numpy(inp.B).to(dev))
    sky = ops.Sky(lmn)
    rows = [torch.from_numpy(a) for a in (inp.uvw, inp.xyz_a, inp.xyz_b)]
    op = {f: autograd.VisibilityOperator(sky, *rows, mask=True, R_moon=R_MOON, fp64=f) for f in (True, False)}
    x_true = torch.from_numpy(inp.B).to(dev)
    V = op[True].forward(x_true)
    rng = np.random.default_rng(55)
    sigma = float(V.abs().pow(2).mean().sqrt()) / 20.0  # SNR 20 in RMS amplitude
    noise = (rng.normal(size=K) + 1j * rng.normal(size=K)) * sigma / np.sqrt(2.0)
    Vn = V + torch.from_numpy(noise).to(dev)
    print(f"diff-nside128: K={K} P={P}, |V| rms {sigma * 20:.4g}, noise rms {sigma:.4g}", flush=True)

    grads = {}
    for f in (True, False):
        x = (x_true * 0.7).clone().requires_grad_(True)
        loss_fn(op[f], x, Vn).backward()
        grads[f] = x.grad.detach()
    g = grads[True]
    gmax = float(g.abs().max())
    x0 = (x_true * 0.7).detach()
    h, worst = 1e-3, 0.0
    for p in rng.choice(P, 32, replace=False):
        e = torch.zeros_like(x0)
        e[int(p)] = h
        with torch.no_grad():
            rp = op[True].forward(x0 + e) - Vn
            rm = op[True].forward(x0 - e) - Vn
            fd = float(((rp.real ** 2 + rp.imag ** 2) - (rm.real ** 2 + rm.imag ** 2)).sum()) / (2 * h)
        xp = float(x0[int(p)])
        fd += 1e-3 * ((xp + h) ** 2 - (xp - h) ** 2) / (2 * h)
        worst = max(worst, abs(fd - float(g[int(p)])) / gmax)
    rg = float(torch.linalg.vector_norm(grads[False] - g) / torch.linalg.vector_norm(g))
    print(f"  finite differences on 32 pixels: worst {worst:.2e} of max |grad| {gmax:.4g}; "
          f"mixed vs FP64 gradient {rg:.2e}", flush=True)
    assert worst < 1e-7 and rg <= 1e-5

    hist = []
    x = autograd.lsq_reconstruct(op[False], Vn, lam=1e-3, iters=50, callback=lambda i, x, r: hist.append(r))
    L0 = float(loss_fn(op[True], torch.zeros_like(x_true), Vn))
    L1 = float(loss_fn(op[True], x, Vn))
    print(f"  CG 50 iterations: residual {hist[0]:.3e} -> {hist[-1]:.3e}; loss {L0:.4g} -> {L1:.4g} "
          f"({L1 / L0:.2e})", flush=True)
    assert len(hist) == 50 and hist[-1] < 0.1 * hist[0] and L1 < 0.05 * L0
