"""Pins the oracle's pix2ang (NEST restated from pix2ang.cu:61-160 with the
0-based indexing fixed; RING from Gorski et al. 2005) against closed forms and
independently written HEALPix routines."""
import math

import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("nest", [False, True])
def test_nside1_closed_form(oracle_lib, nest):
    th, ph, ring, iphi = oracle.pix2ang(1, nest)
    z = np.cos(th)
    # pixels 0-3: z = 2/3, phi = pi/4 + k pi/2; 4-7: z = 0, phi = k pi/2; 8-11: z = -2/3
    np.testing.assert_allclose(z[:4], 2 / 3, rtol=0, atol=1e-15)
    np.testing.assert_allclose(z[4:8], 0, atol=1e-15)
    np.testing.assert_allclose(z[8:], -2 / 3, atol=1e-15)
    k = np.arange(4)
    if not nest:
        np.testing.assert_allclose(ph[:4], np.pi / 4 + k * np.pi / 2, atol=1e-15)
        np.testing.assert_allclose(ph[4:8], k * np.pi / 2, atol=1e-15)
        np.testing.assert_allclose(ph[8:], np.pi / 4 + k * np.pi / 2, atol=1e-15)
    else:
        # NEST at nside=1: face f = pixel f; equatorial faces 4..7 are centred at phi = 0, pi/2, pi, 3pi/2
        np.testing.assert_allclose(np.sort(ph[4:8]), k * np.pi / 2, atol=1e-15)
        np.testing.assert_allclose(ph[:4], np.pi / 4 + k * np.pi / 2, atol=1e-15)
    assert list(ring) == [1] * 4 + [2] * 4 + [3] * 4


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 64])
def test_ring_structure(oracle_lib, nside):
    th, ph, ring, iphi = oracle.pix2ang(nside, False)
    P = 12 * nside * nside
    assert th.size == P
    counts = np.bincount(ring, minlength=4 * nside)[1:]
    i = np.arange(1, 4 * nside)
    expect = np.where(i < nside, 4 * i, np.where(i > 3 * nside, 4 * (4 * nside - i), 4 * nside))
    np.testing.assert_array_equal(counts, expect)
    # z of each ring: 1 - i^2/(3n^2) in the north cap, (4/3 - 2i/(3n)) in the belt
    zr = np.where(i < nside, 1 - i**2 / (3 * nside**2),
                  np.where(i > 3 * nside, -1 + (4 * nside - i) ** 2 / (3 * nside**2), 4 / 3 - 2 * i / (3 * nside)))
    np.testing.assert_allclose(np.cos(th), zr[ring - 1], atol=2e-15)
    # ring pixels are consecutive in RING order and iphi runs 1..4nr
    assert np.all(np.diff(ring) >= 0)
    assert np.all(iphi >= 1) and np.all(iphi <= 4 * nside)
    assert np.all((ph >= 0) & (ph < 2 * np.pi))


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64, 256])
def test_ring_ang2pix_roundtrip(oracle_lib, nside):
    th, ph, _, _ = oracle.pix2ang(nside, False)
    np.testing.assert_array_equal(oracle.ang2pix_ring(nside, th, ph), np.arange(12 * nside * nside))


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64, 256])
def test_nest_matches_ring(oracle_lib, nside):
    """NEST pixel p and RING pixel nest2ring(p) are the same point, bit for bit,
    and the geometric ang2pix agrees with the bit-twiddling nest2ring."""
    thn, phn, rn, jn = oracle.pix2ang(nside, True)
    thr, phr, rr, jr = oracle.pix2ang(nside, False)
    perm = oracle.nest2ring(nside, np.arange(12 * nside * nside))
    assert np.array_equal(np.sort(perm), np.arange(perm.size))
    np.testing.assert_array_equal(thn, thr[perm])
    np.testing.assert_array_equal(phn, phr[perm])
    np.testing.assert_array_equal(rn, rr[perm])
    np.testing.assert_array_equal(jn, jr[perm])
    np.testing.assert_array_equal(oracle.ang2pix_ring(nside, thn, phn), perm)


@pytest.mark.parametrize("nside", [2, 16, 64])
def test_antipodal_and_centroid(oracle_lib, nside):
    th, ph, _, _ = oracle.pix2ang(nside, False)
    P = th.size
    z = np.cos(th)
    np.testing.assert_allclose(z[::-1], -z, atol=2e-15)
    s = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), z])
    np.testing.assert_allclose(s.sum(axis=1), 0, atol=1e-10 * P)
    # equal-area pixels: the pixel centres' z is uniform on [-1,1] in distribution
    assert abs(np.mean(z)) < 1e-12


def test_pixel_dirs_closed_form(oracle_lib):
    th = np.array([0.0, np.pi / 2, np.pi / 2, np.pi, np.pi / 3])
    ph = np.array([0.3, 0.0, np.pi / 2, 1.0, 3 * np.pi / 2])
    B = np.arange(1.0, 6.0)
    l, m, n, I = oracle.pixel_dirs(th, ph, B, 4 * np.pi / 12)
    # l = sin t cos p, m = -sin t sin p, n = cos t  (vissgen3d_noblockage.cu:58-70)
    np.testing.assert_allclose(l, [0, 1, 0, 0, 0], atol=1e-15)
    np.testing.assert_allclose(m, [0, 0, -1, 0, math.sin(np.pi / 3)], atol=1e-15)
    np.testing.assert_allclose(n, [1, 0, 0, -1, 0.5], atol=1e-15)
    np.testing.assert_allclose(I, B * np.pi / 3, rtol=1e-15)
    np.testing.assert_allclose(l * l + m * m + n * n, 1, atol=1e-15)


def test_reference_nest_bug_is_not_reproduced(oracle_lib):
    """pix2ang.cu:109-110 reads pix2x[ip_low+1]; with 0-based tables that adds 1
    at every empty level and sends z far below -1 (theta = NaN).  The oracle's
    angles must all be finite."""
    th, ph, _, _ = oracle.pix2ang(512, True)
    assert np.isfinite(th).all() and np.isfinite(ph).all()
