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
