"""The oracle against the committed golden fixtures (tests/golden/), and the
HIP path against the same fixtures (gpu-marked).  Fixtures: see
tests/golden/make_golden.py."""
import os

import numpy as np
import pytest

import oracle
from tests.cases import rel_l2
from viss_optimize_3d_amd import geometry as geo

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


@pytest.mark.parametrize("nside", [1, 2, 4])
def test_oracle_pix2ang_matches_golden(oracle_lib, nside):
    g = load(f"pix2ang_nside{nside}.npz")
    for nest in (False, True):
        th, ph, ring, iphi = oracle.pix2ang(nside, nest)
        k = "nest" if nest else "ring"
        np.testing.assert_array_equal(th, g[f"{k}_theta"])
        np.testing.assert_array_equal(ph, g[f"{k}_phi"])
        np.testing.assert_array_equal(ring, g[f"{k}_ring"])
        np.testing.assert_array_equal(iphi, g[f"{k}_iphi"])


def test_oracle_plain_workload_matches_golden(oracle_lib):
    g = load("plain_nside16.npz")
    inp = geo.generate(geo.WORKLOADS["plain-nside16"])
    np.testing.assert_array_equal(inp.B, g["B"])
    np.testing.assert_array_equal(inp.uvw, g["uvw"])
    l, m, n = g["lmn"]
    V = oracle.forward(l, m, n, g["I"], g["uvw"])
    np.testing.assert_allclose(V, g["V"], rtol=0, atol=1e-13 * np.abs(g["V"]).max())
    gs, mb, dcf, dg, wt, bad = oracle.weights(g["bll"], g["uvw"][2], int(g["nr"]))
    np.testing.assert_array_equal(gs, g["gs"])
    np.testing.assert_array_equal(mb, g["mb"])
    np.testing.assert_array_equal(dcf, g["dcf"])
    np.testing.assert_array_equal(wt, g["wt"])
    img = oracle.backproject(l, m, n, g["uvw"], g["V"], g["wt"], flags=oracle.HERMITIAN)
    np.testing.assert_allclose(img, g["image"], rtol=0, atol=1e-13 * np.abs(g["image"]).max())


@pytest.mark.gpu
def test_gpu_matches_golden():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import ops

    for nside in (1, 2, 4):
        g = load(f"pix2ang_nside{nside}.npz")
        for nest in (False, True):
            th, ph, ring, iphi = ops.pix2ang(nside, nest=nest)
            k = "nest" if nest else "ring"
            np.testing.assert_array_equal(ring.cpu().numpy(), g[f"{k}_ring"])
            np.testing.assert_array_equal(iphi.cpu().numpy(), g[f"{k}_iphi"])
            np.testing.assert_array_equal(ph.cpu().numpy(), g[f"{k}_phi"])
            assert np.all(np.abs(th.cpu().numpy() - g[f"{k}_theta"]) <= np.spacing(g[f"{k}_theta"]))
    g = load("plain_nside16.npz")
    c = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    sky = ops.Sky(c(g["lmn"]))
    for fp64, tol in ((True, 1e-12), (False, 1e-5)):
        V = sky.forward(c(g["I"]), c(g["uvw"]), fp64=fp64)
        assert rel_l2(V.cpu().numpy(), g["V"]) < tol
        img = sky.backproject(c(g["uvw"]), c(g["V"]), c(g["wt"]), hermitian=True, fp64=fp64)
        assert rel_l2(img.cpu().numpy(), g["image"]) < tol
    w = ops.weights(c(g["bll"]), c(g["uvw"][2]), int(g["nr"]))
    np.testing.assert_array_equal(w["gs"].cpu().numpy(), g["gs"])
    np.testing.assert_array_equal(w["mb"].cpu().numpy().astype(np.uint32), g["mb"])
    np.testing.assert_array_equal(w["dcf"].cpu().numpy(), g["dcf"])
    np.testing.assert_allclose(w["wt"].cpu().numpy(), g["wt"], rtol=1e-14, atol=1e-13)
