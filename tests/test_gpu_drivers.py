"""End-to-end GPU test of the drop-in executables (pix2ang, vissgen3d) on
synthetic day files written in the reference's layouts, against the FP64
oracle run on the same parsed text inputs.

Tolerances: pix2ang text output identical, line by line, to the oracle's
angles printed the way the reference prints them (float, 6 significant
digits); images with --fp64 --precision 17 within 1e-10 relative L2, default
(mixed precision, reference float text) within 1e-5.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.cases import rel_l2
from viss_optimize_3d_amd import geometry as geo
from viss_optimize_3d_amd import textio

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "viss_optimize_3d_amd", "bin")


def run(args, cwd):
    r = subprocess.run([os.path.join(BIN, args[0])] + args[1:], cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return r.stdout


def ref_float_text(x):
    """How the reference writes a value: `file << (float)v` (6 significant digits)."""
    return [f"{float(np.float32(v)):g}" for v in x]


def oracle_day(in_dir, day, P, theta, phi, B, blockage, freq=1e6):
    """The reference pipeline for one day, restated with the oracle."""
    tag = textio.day_tag(day)
    uvw = textio.read_rows(os.path.join(in_dir, "uvw" + tag), 3, 3).T
    xa = textio.read_rows(os.path.join(in_dir, "xyza" + tag), 3, 3).T
    xb = textio.read_rows(os.path.join(in_dir, "xyzb" + tag), 3, 3).T
    bll = textio.read_rows(os.path.join(in_dir, "bll" + tag), 1, 1)[:, 0].astype(np.float32)
    rows = uvw.shape[1]
    amount = rows // 2
    l, m, n, I = oracle.pixel_dirs(theta, phi, B, 4 * np.pi / P)
    R, orbit = 1737.1e3, 1737.1e3 + 300e3
    V = np.zeros(rows, np.complex128)
    V[:2 * amount] = oracle.forward(l, m, n, I, uvw[:, :amount], xa[:, :amount], xb[:, :amount],
                                    flags=oracle.MASK | oracle.PHASE_TRACK | oracle.HERMITIAN, R_moon=R,
                                    orbit_radius=orbit)
    lam = np.float32(3e8 / freq)
    nr = int(np.ceil(np.float32(100e3) / lam * np.float32(2.0)))
    nr = max(nr, int(oracle.shell_index(bll).max()))
    _, _, _, _, wt, _ = oracle.weights(bll, uvw[2], nr)
    flags = oracle.PHASE_TRACK | (oracle.MASK if blockage else 0)
    return oracle.backproject(l, m, n, uvw, V, wt, xa, xb, flags=flags, R_moon=R, orbit_radius=orbit)


@pytest.mark.parametrize("variant", ["noblockage", "blockage"])
def test_pix2ang_and_vissgen3d_end_to_end(tmp_path, variant):
    nside = 8
    P = 12 * nside * nside
    rng = np.random.default_rng(21)
    B = geo.make_sky(nside, rng, 5)
    pos = geo.satellite_positions(4, 40, 10e3, True, 20, rng, orbit_radius=1737.1e3 + 300e3)
    uvw, xa, xb, bll = geo.baseline_rows(pos, 3e8 / 1e6)
    inputs = geo.Inputs(None, B, uvw, xa, xb, bll, pos)
    in_dir = tmp_path / "earth_1Mhz"
    textio.write_dataset(str(in_dir), inputs, days=2)
    # pix2ang drop-in: reads ./B.txt, writes theta_heal.txt / phi_heal.txt (NEST, float text)
    run(["pix2ang"], cwd=str(in_dir))
    th_txt = open(in_dir / "theta_heal.txt").read().split()
    ph_txt = open(in_dir / "phi_heal.txt").read().split()
    oth, oph, _, _ = oracle.pix2ang(nside, True)
    assert th_txt == ref_float_text(oth)
    assert ph_txt == ref_float_text(oph)
    theta = np.array(th_txt, np.float64)
    phi = np.array(ph_txt, np.float64)
    # vissgen3d drop-in, all-FP64 with full-precision text, then the default
    run(["vissgen3d", "--variant", variant, "--days", "2", "--fp64", "--precision", "17", "--out", "exact"],
        cwd=str(tmp_path))
    run(["vissgen3d", "--variant", variant, "--days", "2"], cwd=str(tmp_path))
    outdir = tmp_path / ("3d" + variant)
    for day in (1, 2):
        ref = oracle_day(str(in_dir), day, P, theta, phi, B, blockage=(variant == "blockage"))
        exact = np.loadtxt(tmp_path / "exact" / f"C{day}day1M.txt")
        assert rel_l2(exact, ref) < 1e-10
        dflt = np.loadtxt(outdir / f"C{day}day1M.txt")
        assert rel_l2(dflt, ref) < 1e-5
        assert np.isfinite(dflt).all()
    # multi-day aggregation (the notebooks' sum), from the text files and from the binary sidecars
    run(["vissgen3d", "--variant", variant, "--days", "2", "--binary", "--out", "withbin"], cwd=str(tmp_path))
    run(["vissimage_sum", "--dir", str(outdir), "--days", "2", "--nan-to-zero", "--precision", "17", "--out",
         "sum_txt.txt"], cwd=str(tmp_path))
    run(["vissimage_sum", "--dir", "withbin", "--days", "2", "--binary", "--precision", "17", "--out", "sum_bin.txt",
         "--to-ring", str(nside)], cwd=str(tmp_path))
    s_txt = np.loadtxt(tmp_path / "sum_txt.txt")
    np.testing.assert_array_equal(s_txt, np.loadtxt(outdir / "C1day1M.txt") + np.loadtxt(outdir / "C2day1M.txt"))
    b1 = np.fromfile(tmp_path / "withbin" / "C1day1M.bin", dtype=np.float64)
    b2 = np.fromfile(tmp_path / "withbin" / "C2day1M.bin", dtype=np.float64)
    ring = np.empty(P)
    ring[oracle.nest2ring(nside, np.arange(P))] = b1 + b2
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "sum_bin.txt"), ring)


def test_vissgen3d_missing_day_files_give_zero_image(tmp_path):
    """A missing per-day file reads as 0 rows (vissgen3d_noblockage.cu:307), so the image is 0."""
    nside = 4
    in_dir = tmp_path / "earth_1Mhz"
    in_dir.mkdir()
    textio.write_sky(str(in_dir / "B.txt"), np.ones(12 * nside * nside))
    run(["pix2ang"], cwd=str(in_dir))
    run(["vissgen3d", "--days", "1"], cwd=str(tmp_path))
    C = np.loadtxt(tmp_path / "3dnoblockage" / "C1day1M.txt")
    assert C.shape == (192,) and np.all(C == 0)


def test_vissgen3d_exact_healpix(tmp_path):
    """--exact-healpix: the reference's 6-digit angle files give, through the
    exact HEALPix angles, the image of full-precision angle files (the same
    directions, bit for bit), with the quad kernels; a file that is not a
    HEALPix map is refused."""
    nside = 8
    P = 12 * nside * nside
    rng = np.random.default_rng(22)
    B = geo.make_sky(nside, rng, 5)
    pos = geo.satellite_positions(4, 40, 10e3, True, 20, rng, orbit_radius=1737.1e3 + 300e3)
    uvw, xa, xb, bll = geo.baseline_rows(pos, 3e8 / 1e6)
    in_dir = tmp_path / "earth_1Mhz"
    textio.write_dataset(str(in_dir), geo.Inputs(None, B, uvw, xa, xb, bll, pos), days=1)
    run(["pix2ang"], cwd=str(in_dir))  # NEST, 6 significant digits
    run(["pix2ang", "--precision", "17", "--theta", "theta17.txt", "--phi", "phi17.txt"], cwd=str(in_dir))
    common = ["vissgen3d", "--variant", "blockage", "--days", "1", "--fp64", "--precision", "17"]
    out6 = run(common + ["--out", "six"], cwd=str(tmp_path))
    outx = run(common + ["--exact-healpix", "--out", "exact"], cwd=str(tmp_path))
    out17 = run(common + ["--theta", "theta17.txt", "--phi", "phi17.txt", "--out", "full"], cwd=str(tmp_path))
    assert "per-pixel kernels" in out6 and "quad kernels" in outx and "quad kernels" in out17
    assert "exact HEALPix angles: NEST, nside 8" in outx
    # same directions bit for bit; the FP64 atomics may sum in another order
    exact = np.loadtxt(tmp_path / "exact" / "C1day1M.txt")
    assert rel_l2(exact, np.loadtxt(tmp_path / "full" / "C1day1M.txt")) < 1e-13
    oth, oph, _, _ = oracle.pix2ang(nside, True)
    ref = oracle_day(str(in_dir), 1, P, oth, oph, B, blockage=True)
    assert rel_l2(exact, ref) < 1e-10
    # one angle moved by 1e-3 rad: not this map, refused
    th = np.array(open(in_dir / "theta_heal.txt").read().split(), np.float64)
    th[7] += 1e-3
    textio.write_column(str(in_dir / "theta_bad.txt"), th)
    r = subprocess.run([os.path.join(BIN, "vissgen3d"), "--days", "1", "--exact-healpix", "--theta", "theta_bad.txt",
                        "--out", "bad"], cwd=str(tmp_path), stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode != 0 and "not HEALPix RING or NEST angles" in r.stdout
