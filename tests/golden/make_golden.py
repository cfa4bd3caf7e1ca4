"""Writes the committed golden fixtures (tests/golden/*.npz) with the FP64 oracle.

The reference ships no golden vectors (its data directories hold only
.gitkeep), so these fixtures are produced by the oracle -- itself pinned by the
closed-form / brute-force tests in tests/test_oracle_*.py -- and freeze its
outputs so that a later change to the oracle or the HIP path is caught
against fixed numbers:

  pix2ang_nside{1,2,4}.npz : NEST and RING theta, phi, ring, iphi
  plain_nside16.npz        : the plain-method workload (nside 16 RING, 4 satellites,
                             500 snapshots over one orbit, 1 MHz, 3,000 rows): inputs,
                             forward V (unmasked), weights (gs, mb, dcf, wt) and the
                             density-compensated image

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from viss_optimize_3d_amd import geometry as geo  # noqa: E402


def main():
    oracle.build()
    for nside in (1, 2, 4):
        d = {}
        for nest in (False, True):
            th, ph, ring, iphi = oracle.pix2ang(nside, nest)
            k = "nest" if nest else "ring"
            d.update({f"{k}_theta": th, f"{k}_phi": ph, f"{k}_ring": ring, f"{k}_iphi": iphi})
        np.savez(os.path.join(HERE, f"pix2ang_nside{nside}.npz"), **d)
    wl = geo.WORKLOADS["plain-nside16"]
    inp = geo.generate(wl)
    th, ph, _, _ = oracle.pix2ang(wl.nside, False)
    l, m, n, I = oracle.pixel_dirs(th, ph, inp.B, 4 * np.pi / inp.B.size)
    V = oracle.forward(l, m, n, I, inp.uvw)
    nr = geo.shell_count(wl.freq_hz)
    bll = inp.bll.astype(np.float32)
    gs, mb, dcf, dg, wt, bad = oracle.weights(bll, inp.uvw[2], nr)
    image = oracle.backproject(l, m, n, inp.uvw, V, wt, flags=oracle.HERMITIAN)
    np.savez(os.path.join(HERE, "plain_nside16.npz"), B=inp.B, uvw=inp.uvw, bll=bll, lmn=np.stack([l, m, n]), I=I,
             V=V, nr=nr, gs=gs, mb=mb, dcf=dcf, wt=wt, n_out_of_range=bad, image=image)


if __name__ == "__main__":
    main()
