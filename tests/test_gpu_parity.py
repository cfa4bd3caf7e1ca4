"""GPU parity: the HIP library through its C ABI against the FP64 oracle on the
same seeded inputs.

Tolerances (stated here, per quantity):
  * pix2ang ring / in-ring indices and phi: bit-exact; theta within 1 ulp;
  * pixel directions: 4e-16 absolute;
  * weights: gs, mb, dcf bit-exact; dg, wt within 1e-14 relative or 1e-13
    absolute, plus dg's own rounding sensitivity (dg = 1.5 sqrt|1 - 4 sin^2 g|
    / |cos g| loses digits by cancellation where sin g ~ 1/2: an ulp of
    1 - 4 sin^2 g in either evaluation moves dg by up to 1.5 sqrt(4.4e-16) /
    |cos g|);
  * all-FP64 forward / back-projection: relative L2 <= 1e-11 (FP64 phase
    rounding alone is ~1e-12 relative at |b.s| ~ 10^4 cycles);
  * mixed-precision forward / back-projection: relative L2 <= 1e-5 (the
    precision target) -- the measured value is printed;
  * visible-pair counts: exact.
"""
import numpy as np
import pytest

import oracle
from tests.cases import check_weights, make_case, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

R_MOON = 1737.4e3
FP64_TOL = 1e-11
MIXED_TOL = 1e-5


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from viss_optimize_3d_amd import ops as _ops

    return _ops


def cuda(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.to("cuda", dtype) if dtype is not None else t.to("cuda")


@pytest.mark.parametrize("nest", [False, True])
@pytest.mark.parametrize("nside", [1, 2, 3, 4, 16, 64, 100, 256, 600, 1024])
def test_pix2ang_parity(ops, nside, nest):
    """Small maps compute acos per pixel, maps of 2^22 pixels and more once per
    ring (nside 600 and 1024 here); NEST needs a power of two."""
    if nest and nside & (nside - 1):
        pytest.skip("NEST needs a power-of-two nside")
    th, ph, ring, iphi = ops.pix2ang(nside, nest=nest)
    oth, oph, oring, oiphi = oracle.pix2ang(nside, nest)
    np.testing.assert_array_equal(ring.cpu().numpy(), oring)
    np.testing.assert_array_equal(iphi.cpu().numpy(), oiphi)
    np.testing.assert_array_equal(ph.cpu().numpy(), oph)
    d = np.abs(th.cpu().numpy() - oth)
    assert np.all(d <= np.spacing(oth)), f"theta off by {np.max(d / np.spacing(oth))} ulp"


@pytest.mark.parametrize("nest", [False, True])
@pytest.mark.parametrize("nside", [4096, 8192])
def test_pix2ang_largest_maps(ops, nside, nest):
    """The reference's largest map (nside 4096, 201,326,592 pixels) and this library's largest
    (nside 8192, 805,306,368 pixels, 19 GB of outputs): every pixel against the oracle, as
    test_pix2ang_parity, compared in chunks of 2^26 pixels."""
    th, ph, ring, iphi = ops.pix2ang(nside, nest=nest)
    oth, oph, oring, oiphi = oracle.pix2ang(nside, nest)
    P = 12 * nside * nside
    step = 1 << 26
    worst = 0.0
    for a in range(0, P, step):
        b = min(P, a + step)
        np.testing.assert_array_equal(ring[a:b].cpu().numpy(), oring[a:b])
        np.testing.assert_array_equal(iphi[a:b].cpu().numpy(), oiphi[a:b])
        np.testing.assert_array_equal(ph[a:b].cpu().numpy(), oph[a:b])
        d = np.abs(th[a:b].cpu().numpy() - oth[a:b]) / np.spacing(oth[a:b])
        worst = max(worst, float(d.max()))
    assert worst <= 1.0, f"theta off by {worst} ulp"
    print(f"nside {nside} {'NEST' if nest else 'RING'}: {P} pixels, ring / iphi / phi exact, theta within "
          f"{worst:g} ulp", flush=True)


def test_pix2ang_rejects_bad_nside(ops):
    from viss_optimize_3d_amd._lib import VissError

    with pytest.raises(VissError):
        ops.pix2ang(0)
    with pytest.raises(VissError):
        ops.pix2ang(12, nest=True)  # NEST needs a power of two


def test_pixel_dirs_parity(ops):
    th, ph, _, _ = oracle.pix2ang(64, True)
    B = np.random.default_rng(0).lognormal(size=th.size)
    lmn, I = ops.pixel_dirs(cuda(th), cuda(ph), cuda(B))
    l, m, n, oI = oracle.pixel_dirs(th, ph, B, 4 * np.pi / th.size)
    np.testing.assert_allclose(lmn.cpu().numpy(), np.stack([l, m, n]), rtol=0, atol=4e-16)
    np.testing.assert_allclose(I.cpu().numpy(), oI, rtol=1e-15)


CASES = {
    # plain method: nside 16, 4 satellites (6 baselines) x 500 snapshots over one orbit, 1 MHz, no mask
    "plain-nside16": dict(nside=16, n_snap=500, n_sat=4, freq=1e6, sigma=10e3, rotating=False, snaps_per_orbit=500,
                          seed=1),
    # blockage: nside 64, 8 satellites, rotating orbit normal (a 2,800-row slice of the 280,000)
    "blockage-nside64": dict(nside=64, n_snap=100, n_sat=8, freq=1e6, sigma=10e3, rotating=True,
                             snaps_per_orbit=100, seed=2),
    # precision: nside 256, 10 MHz, sigma 50 km (|b| to thousands of wavelengths), a 225-row slice
    "precision-nside256": dict(nside=256, n_snap=5, n_sat=10, freq=10e6, sigma=50e3, rotating=True,
                               snaps_per_orbit=100, seed=3),
}
MASKED = {"plain-nside16": False, "blockage-nside64": True, "precision-nside256": True}


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    c = make_case(**CASES[request.param])
    c["name"] = request.param
    c["mask"] = MASKED[request.param]
    return c


@pytest.fixture(scope="module")
def oracle_fwd(case):
    flags = oracle.MASK if case["mask"] else 0
    return oracle.forward(*case["lmn"], case["I"], case["uvw"], case["xa"], case["xb"], flags=flags, R_moon=R_MOON)


@pytest.mark.parametrize("fp64", [True, False])
def test_forward_parity(ops, case, oracle_fwd, fp64):
    sky = ops.Sky(cuda(case["lmn"]))
    V = sky.forward(cuda(case["I"]), cuda(case["uvw"]), cuda(case["xa"]), cuda(case["xb"]), mask=case["mask"],
                    R_moon=R_MOON, fp64=fp64)
    err = rel_l2(V.cpu().numpy(), oracle_fwd)
    print(f"\n{case['name']} forward fp64={fp64}: rel L2 {err:.3e}")
    assert err <= (FP64_TOL if fp64 else MIXED_TOL)


@pytest.mark.parametrize("fp64", [True, False])
def test_backproject_parity(ops, case, fp64):
    rng = np.random.default_rng(9)
    K = case["uvw"].shape[1]
    y = rng.normal(size=K) + 1j * rng.normal(size=K)
    wt = rng.uniform(0.0, 0.125, K)
    flags = oracle.MASK if case["mask"] else 0
    ref = oracle.backproject(*case["lmn"], case["uvw"], y, wt, case["xa"], case["xb"], flags=flags, R_moon=R_MOON)
    sky = ops.Sky(cuda(case["lmn"]))
    x = sky.backproject(cuda(case["uvw"]), cuda(y), cuda(wt), cuda(case["xa"]), cuda(case["xb"]), mask=case["mask"],
                        R_moon=R_MOON, fp64=fp64)
    err = rel_l2(x.cpu().numpy(), ref)
    print(f"\n{case['name']} backproject fp64={fp64}: rel L2 {err:.3e}")
    assert err <= (FP64_TOL if fp64 else MIXED_TOL)


def test_forward_flags(ops):
    """Phase tracking and Hermitian rows (vissgen3d_noblockage.cu:95,100) and the
    fixed reference cut (orbit_radius > 0)."""
    c = make_case(nside=16, n_snap=40, n_sat=5, freq=1e6, seed=4)
    args = (*c["lmn"], c["I"], c["uvw"], c["xa"], c["xb"])
    orbit = 1737.1e3 + 300e3
    ref = oracle.forward(*args, flags=oracle.MASK | oracle.PHASE_TRACK | oracle.HERMITIAN, R_moon=1737.1e3,
                         orbit_radius=orbit)
    sky = ops.Sky(cuda(c["lmn"]))
    for fp64 in (True, False):
        V = sky.forward(cuda(c["I"]), cuda(c["uvw"]), cuda(c["xa"]), cuda(c["xb"]), mask=True, R_moon=1737.1e3,
                        orbit_radius=orbit, phase_track=True, hermitian=True, fp64=fp64)
        assert rel_l2(V.cpu().numpy(), ref) <= (FP64_TOL if fp64 else MIXED_TOL)
    # accumulate semantics: a second call adds
    V2 = sky.forward(cuda(c["I"]), cuda(c["uvw"]), cuda(c["xa"]), cuda(c["xb"]), mask=True, R_moon=1737.1e3,
                     orbit_radius=orbit, phase_track=True, hermitian=True, fp64=True, out=V.clone())
    assert rel_l2(V2.cpu().numpy(), 2 * ref) <= 1e-5


def test_backproject_flags(ops):
    c = make_case(nside=16, n_snap=40, n_sat=5, freq=1e6, seed=5)
    K = c["uvw"].shape[1]
    rng = np.random.default_rng(2)
    y = rng.normal(size=K) + 1j * rng.normal(size=K)
    f = oracle.MASK | oracle.PHASE_TRACK | oracle.HERMITIAN
    ref = oracle.backproject(*c["lmn"], c["uvw"], y, None, c["xa"], c["xb"], flags=f, R_moon=R_MOON)
    sky = ops.Sky(cuda(c["lmn"]))
    for fp64 in (True, False):
        x = sky.backproject(cuda(c["uvw"]), cuda(y), None, cuda(c["xa"]), cuda(c["xb"]), mask=True, R_moon=R_MOON,
                            phase_track=True, hermitian=True, fp64=fp64)
        assert rel_l2(x.cpu().numpy(), ref) <= (FP64_TOL if fp64 else MIXED_TOL)


def test_adjoint_dot_product_gpu(ops):
    c = make_case(nside=32, n_snap=50, n_sat=6, freq=10e6, sigma=30e3, seed=6)
    K = c["uvw"].shape[1]
    rng = np.random.default_rng(3)
    x = rng.normal(size=c["I"].size)
    y = rng.normal(size=K) + 1j * rng.normal(size=K)
    sky = ops.Sky(cuda(c["lmn"]))
    for fp64, tol in ((True, 1e-11), (False, 1e-6)):
        Ax = sky.forward(cuda(x), cuda(c["uvw"]), cuda(c["xa"]), cuda(c["xb"]), mask=True, R_moon=R_MOON, fp64=fp64)
        Ahy = sky.backproject(cuda(c["uvw"]), cuda(y), None, cuda(c["xa"]), cuda(c["xb"]), mask=True, R_moon=R_MOON,
                              fp64=fp64)
        lhs = np.real(np.vdot(Ax.cpu().numpy(), y))
        rhs = np.dot(x, Ahy.cpu().numpy())
        scale = np.linalg.norm(Ax.cpu().numpy()) * np.linalg.norm(y)
        assert abs(lhs - rhs) <= tol * scale


def test_count_visible_exact(ops):
    c = make_case(nside=64, n_snap=30, n_sat=8, freq=1e6, seed=7)
    sky = ops.Sky(cuda(c["lmn"]))
    cnt = int(sky.count_visible(cuda(c["xa"]), cuda(c["xb"]), R_moon=R_MOON).item())
    assert cnt == oracle.count_visible(*c["lmn"], c["xa"], c["xb"], R_MOON)


def test_weights_parity(ops):
    rng = np.random.default_rng(8)
    nr = 667
    K = 200_000
    bll = rng.uniform(0.0, 360.0, K).astype(np.float32)
    bll[:5] = [0.25, 0.75, 1.25, 333.75, 2e6]
    w = rng.uniform(-1, 1, K) * bll * 0.95
    w[7] = 5 * bll[7]  # NaN dg -> weight 1/8
    w[3], w[9], w[10] = bll[3], bll[9], -bll[10]  # |w / bll| = 1 (row 3 in shell nr: dcf 0 with the shift)
    w[11:20] = 0.5 * bll[11:20] * (1.0 + np.array([-1e-9, -1e-12, -1e-15, 0.0, 1e-15, 1e-12, 1e-9, 0.2, -0.2]))
    # |w / bll| just below 1 (1 - x^2 cancels), and bll of 0, infinity and NaN
    w[20:25] = bll[20:25] * np.array([1 - 1e-15, 1 - 1e-12, 1 - 1e-9, -(1 - 1e-12), 1 - 1e-6])
    bll[25:29] = [0.0, 0.0, np.inf, np.nan]
    w[25:29] = [0.0, 1.0, 7.0, 1.0]
    for shift in (False, True):
        check_weights(ops.weights(cuda(bll), cuda(w), nr, blockage_shift=shift), bll, w, nr, shift)


def test_weights_shell_count_past_2_24(ops):
    """A shell holding 20,000,000 rows: its count passes 2^24, where the
    reference's float atomicAdd counts (vissgen3d_noblockage.cu:114-123) stop
    changing; the uint32 counts here stay exact, and so do dcf and wt."""
    K, nr = 20_000_000, 40
    bll = np.full(K, 10.3, np.float32)
    bll[::1000] = 3.3  # a second shell with 20,000 rows
    w = np.zeros(K)
    out = ops.weights(cuda(bll), cuda(w), nr)
    mb = out["mb"].cpu().numpy()
    assert mb[20] == K - K // 1000 and mb[6] == K // 1000 and mb.sum() == K
    check_weights(out, bll, w, nr)


def test_weights_shell_table_beyond_lds(ops):
    """50,000 shells: the counts no longer fit one workgroup's LDS (160 KB on
    gfx950), so the histogram takes its global-atomic path."""
    rng = np.random.default_rng(21)
    K, nr = 300_000, 50_000
    bll = rng.uniform(0.0, nr / 2, K).astype(np.float32)
    w = rng.uniform(-1, 1, K) * bll * 0.9
    check_weights(ops.weights(cuda(bll), cuda(w), nr), bll, w, nr)


def test_weights_mostly_out_of_range(ops):
    """Most rows outside the shell table (shell 0, past nr, NaN bll): the
    histogram adds them into spare counters past its LDS table, all on a few
    addresses; counts, the out-of-range tally, dcf and wt as the oracle's.
    2e6 rows over 2,000 shells take the LDS paths of both passes."""
    rng = np.random.default_rng(23)
    K, nr = 2_000_000, 2_000
    kind = rng.integers(0, 10, K)
    bll = rng.uniform(0.3, nr / 2.0, K).astype(np.float32)  # in range
    bll[kind < 4] = 0.0                                       # shell 0: not counted, not out of range
    bll[(kind >= 4) & (kind < 7)] = np.float32(nr)            # past shell nr
    bll[kind == 7] = np.nan
    w = rng.uniform(-1, 1, K) * np.nan_to_num(bll) * 0.9
    out = ops.weights(cuda(bll), cuda(w), nr)
    check_weights(out, bll, w, nr)
    assert int(out["n_out_of_range"].item()) == int(np.sum((kind >= 4) & (kind < 8)))


def test_weights_unaligned_and_ragged(ops):
    """Row counts that are not multiples of the 4-row / 2-row vector steps
    (the tail loops), and views one element into their buffers (the scalar
    kernels, taken when bll / w are not 16-B / 8-B aligned)."""
    rng = np.random.default_rng(22)
    nr = 667
    for K in (1, 3, 5, 100_003):
        bll = rng.uniform(0.0, 340.0, K + 1).astype(np.float32)
        w = rng.uniform(-1, 1, K + 1) * bll * 0.9
        check_weights(ops.weights(cuda(bll[:K]), cuda(w[:K]), nr), bll[:K], w[:K], nr)
        tb, tw = cuda(bll), cuda(w)
        assert tb[1:].data_ptr() % 16 != 0 and tw[1:].data_ptr() % 16 != 0
        check_weights(ops.weights(tb[1:], tw[1:], nr), bll[1:], w[1:], nr)


def test_one_shot_entry_points(ops):
    """viss_forward / viss_backproject (a temporary sky per call) through the C
    ABI: the same sums as the reusable sky's entry points, masked, with phase
    tracking and Hermitian rows, mixed and FP64."""
    import ctypes

    from viss_optimize_3d_amd import _lib

    L = _lib.load()
    c = make_case(nside=16, n_snap=40, n_sat=6, freq=1e6, seed=11)
    lmn, I, uvw, xa, xb = (cuda(c[k]) for k in ("lmn", "I", "uvw", "xa", "xb"))
    P, K = lmn.shape[1], uvw.shape[1]
    sky = ops.Sky(lmn)
    geom = ctypes.byref(_lib.MaskGeometry(R_MOON, 0.0))
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    y = cuda(np.random.default_rng(12).standard_normal(2 * K)).view(torch.complex128)
    for fp64 in (False, True):
        flags = _lib.MASK | _lib.PHASE_TRACK | _lib.HERMITIAN | (_lib.PREC_FP64 if fp64 else 0)
        V = torch.zeros(2 * K, dtype=torch.complex128, device="cuda")
        assert L.viss_forward(p(lmn), p(I), P, p(uvw), p(xa), p(xb), K, geom, flags, p(V), stream) == 0
        Vs = sky.forward(I, uvw, xa, xb, mask=True, R_moon=R_MOON, phase_track=True, hermitian=True, fp64=fp64)
        x = torch.zeros(P, dtype=torch.float64, device="cuda")
        assert L.viss_backproject(p(lmn), P, p(uvw), p(xa), p(xb), p(y), None, K, geom, flags, p(x), stream) == 0
        xs = sky.backproject(uvw, y, None, xa, xb, mask=True, R_moon=R_MOON, phase_track=True, hermitian=True,
                             fp64=fp64)
        torch.cuda.synchronize()
        assert rel_l2(V.cpu().numpy(), Vs.cpu().numpy()) < 1e-13
        assert rel_l2(x.cpu().numpy(), xs.cpu().numpy()) < 1e-13


def test_empty_inputs(ops):
    c = make_case(nside=4, n_snap=2, n_sat=3, seed=1)
    sky = ops.Sky(cuda(c["lmn"]))
    z = torch.zeros(3, 0, dtype=torch.float64, device="cuda")
    V = sky.forward(cuda(c["I"]), z, z, z, mask=True)
    assert V.numel() == 0
    x = sky.backproject(z, torch.zeros(0, dtype=torch.complex128, device="cuda"), None, z, z, mask=True)
    assert float(x.abs().sum()) == 0.0
