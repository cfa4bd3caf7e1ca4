"""Least-squares sky reconstruction beside the density-compensated dirty image
(the differentiation workload: nside=128 sky, 200,000 visibilities at 10 MHz
with occultation, complex Gaussian noise at SNR 20, 50 CG iterations on
||A x - V||^2 + 1e-3 ||x||^2).

    python -m viss_optimize_3d_amd.reconstruct [--iters 50] [--workload diff-nside128]

Prints one JSON line: relative image errors against the true sky for the CG
solution and for the dirty image (raw and after the best scalar fit), the
gradient finite-difference check and timings.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from . import autograd, geometry as geo, ops


def rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def zm_scale(a, b):
    a0, b0 = a - a.mean(), b - b.mean()
    return float(torch.dot(a0, b0) / torch.dot(a0, a0))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="diff-nside128")
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--lam", type=float, default=1e-3)
    ap.add_argument("--snr", type=float, default=20.0)
    ap.add_argument("--fd-pixels", type=int, default=32)
    args = ap.parse_args(argv)
    wl = geo.WORKLOADS[args.workload]
    inp = geo.generate(wl)
    dev = torch.device("cuda:0")
    th, ph = ops.pix2ang(wl.nside, nest=False, with_ring=False)
    lmn, _ = ops.pixel_dirs(th, ph)
    sky = ops.Sky(lmn)
    op = autograd.VisibilityOperator(sky, torch.from_numpy(inp.uvw), torch.from_numpy(inp.xyz_a),
                                     torch.from_numpy(inp.xyz_b), mask=True, R_moon=geo.R_MOON)
    op64 = autograd.VisibilityOperator(sky, op.uvw, op.xyz_a, op.xyz_b, mask=True, R_moon=geo.R_MOON, fp64=True)
    x_true = torch.from_numpy(inp.B).to(dev)
    V0 = op64.forward(x_true)
    rng = np.random.default_rng(wl.seed + 100)
    sigma = float(torch.sqrt((V0.abs() ** 2).mean())) / args.snr / math.sqrt(2.0)
    noise = (rng.normal(size=V0.numel()) + 1j * rng.normal(size=V0.numel())) * sigma
    V = V0 + torch.from_numpy(noise).to(dev)

    # gradient check: central differences (FP64 kernels) on random pixels
    def loss(o, x):
        r = o(x) - V
        return (r.real ** 2 + r.imag ** 2).sum() + args.lam * (x ** 2).sum()

    x0 = (x_true * 0.5).clone().requires_grad_(True)
    loss(op64, x0).backward()
    g = x0.grad.detach()
    h = 1e-3
    worst = 0.0
    for p in rng.choice(x_true.numel(), args.fd_pixels, replace=False):
        e = torch.zeros_like(x_true)
        e[p] = h
        with torch.no_grad():
            fd = (float(loss(op64, x0 + e)) - float(loss(op64, x0 - e))) / (2 * h)
        worst = max(worst, abs(fd - float(g[p])) / float(g.abs().max()))
    x1 = (x_true * 0.5).clone().requires_grad_(True)
    loss(op, x1).backward()
    grad_mixed_vs_fp64 = rel(x1.grad, g)

    # CG reconstruction (mixed precision kernels)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = []
    x = autograd.lsq_reconstruct(op, V, lam=args.lam, iters=args.iters, callback=lambda i, x, r: hist.append(r))
    torch.cuda.synchronize()
    t_cg = time.perf_counter() - t0

    # density-compensated dirty image (weights of the reference's DCF scheme)
    bll = torch.from_numpy(inp.bll.astype(np.float32)).to(dev)
    nr = max(geo.shell_count(wl.freq_hz), int(math.ceil(float(bll.max()) * 2)))
    wts = ops.weights(bll, op.uvw[2], nr)
    dirty = sky.backproject(op.uvw, V, wts["wt"], op.xyz_a, op.xyz_b, mask=True, R_moon=geo.R_MOON, hermitian=True)
    alpha = float(torch.dot(dirty, x_true) / torch.dot(dirty, dirty))
    beta = float(torch.dot(x, x_true) / torch.dot(x, x))
    out = {
        "workload": wl.name, "npix": sky.P, "n_vis": op.shape[0], "iters": args.iters, "lam": args.lam,
        "snr": args.snr, "fd_max_rel_err": worst, "grad_mixed_vs_fp64_rel_l2": grad_mixed_vs_fp64,
        "cg_residual_first_last": [hist[0], hist[-1]], "cg_seconds": t_cg,
        "cg_rel_err": rel(x, x_true), "cg_rel_err_scaled": rel(beta * x, x_true),
        "dirty_rel_err": rel(dirty, x_true), "dirty_rel_err_scaled": rel(alpha * dirty, x_true),
        # the monopole (sky mean) is barely constrained by non-zero baselines: errors without it
        "cg_rel_err_zero_mean": rel(x - x.mean(), x_true - x_true.mean()),
        "dirty_rel_err_zero_mean_scaled": rel(zm_scale(dirty, x_true) * (dirty - dirty.mean()), x_true - x_true.mean()),
    }
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
