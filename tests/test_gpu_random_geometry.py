"""Seeded random geometries through the per-row kernels: satellites from just
above the surface (the visible cap is then barely larger than a hemisphere,
so a pixel and its antipode can both lie near its edge and rings take the
both-sides-tested mode) to ten lunar radii, clustered or scattered, HEALPix
RING / NEST skies with and without pixels removed, and row counts that are
not multiples of the 64-row wavefront.

Per case: the occultation decisions exactly (zero baselines and unit inputs
give visible-pixel and seeing-row counts, against the FP64 oracle), and the
mixed kernels against the all-FP64 kernels on random baselines (relative L2
<= 1e-5, the precision target).
"""
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


def random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    nside = int(rng.choice([8, 16, 32]))
    nest = bool(rng.integers(2))
    th, ph, _, _ = oracle.pix2ang(nside, nest)
    lmn = np.stack(oracle.pixel_dirs(th, ph, np.ones_like(th), 1.0)[:3])
    if rng.integers(3) == 0:  # some pixels removed: quads, leftover pairs and singles
        keep = np.sort(rng.choice(lmn.shape[1], int(lmn.shape[1] * rng.uniform(0.7, 0.98)), replace=False))
        lmn = lmn[:, keep]
    n_sat = int(rng.integers(3, 9))
    if seed % 2 == 0:  # low orbits: caps barely larger than a hemisphere
        radius = R * float(rng.uniform(1.001, 1.1))
    else:
        radius = R * float(np.exp(rng.uniform(np.log(1.1), np.log(10.0))))
    spread = float(rng.choice([1e3, 50e3, 0.3 * radius]))
    pos = geo.satellite_positions(n_sat, int(rng.integers(5, 60)), spread, True, int(rng.integers(20, 200)), rng,
                                  orbit_radius=radius)
    norm = np.linalg.norm(pos, axis=2, keepdims=True)  # keep every satellite above the surface
    pos = pos * np.maximum(1.0, 1.0005 * R / norm)
    wavelength = geo.C_LIGHT / float(rng.choice([1e6, 10e6]))
    uvw, xa, xb, _ = geo.baseline_rows(pos, wavelength)
    K = uvw.shape[1]
    cut = int(rng.integers(max(1, K // 2), K + 1))  # ragged row count
    return dict(nside=nside, nest=nest, lmn=lmn, uvw=uvw[:, :cut], xa=xa[:, :cut], xb=xb[:, :cut], radius=radius,
                spread=spread, rng=rng)


@pytest.mark.parametrize("seed", range(16))
def test_random_geometry(ops, seed):
    c = random_case(seed)
    lmn, uvw, xa, xb = c["lmn"], c["uvw"], c["xa"], c["xb"]
    K, P = uvw.shape[1], lmn.shape[1]
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    sky = ops.Sky(d(lmn))
    lay = sky.pairs()
    print(f"seed {seed}: nside {c['nside']} nest={c['nest']} P={P} K={K} orbit {c['radius'] / R:.3f} R, "
          f"spread {c['spread'] / 1e3:.0f} km, layout {lay}", flush=True)
    l, m, n = lmn
    zero = np.zeros((3, K))
    rows = oracle.forward(l, m, n, np.ones(P), zero, xa, xb, flags=oracle.MASK, R_moon=R).real
    pix = oracle.backproject(l, m, n, zero, np.ones(K, dtype=complex), None, xa, xb, flags=oracle.MASK, R_moon=R)
    ta, tb = d(xa), d(xb)
    for fp64 in (False, True):
        V = sky.forward(torch.ones(P, dtype=torch.float64, device="cuda"), d(zero), ta, tb, mask=True, R_moon=R,
                        fp64=fp64).cpu().numpy()
        x = sky.backproject(d(zero), torch.ones(K, dtype=torch.complex128, device="cuda"), None, ta, tb, mask=True,
                            R_moon=R, fp64=fp64).cpu().numpy()
        assert np.array_equal(V.real, rows), f"fp64={fp64}: {np.count_nonzero(V.real != rows)} row counts differ"
        assert np.array_equal(x, pix), f"fp64={fp64}: {np.count_nonzero(x != pix)} pixel counts differ"
    I = d(c["rng"].lognormal(0.0, 1.0, P))
    y = d(c["rng"].standard_normal(K) + 1j * c["rng"].standard_normal(K))
    tu = d(uvw)
    V32 = sky.forward(I, tu, ta, tb, mask=True, R_moon=R)
    V64 = sky.forward(I, tu, ta, tb, mask=True, R_moon=R, fp64=True)
    x32 = sky.backproject(tu, y, None, ta, tb, mask=True, R_moon=R)
    x64 = sky.backproject(tu, y, None, ta, tb, mask=True, R_moon=R, fp64=True)
    rv, rx = rel_l2(V32.cpu().numpy(), V64.cpu().numpy()), rel_l2(x32.cpu().numpy(), x64.cpu().numpy())
    print(f"  visible fraction {rows.sum() / (K * P):.3f}; mixed vs FP64: forward {rv:.2e}, adjoint {rx:.2e}",
          flush=True)
    assert rv <= 1e-5 and rx <= 1e-5
