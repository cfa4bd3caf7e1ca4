"""Readers and writers for the reference's on-disk text layouts.

Layouts (vissgen3d_noblockage.cu:207-230, 302-378, 536-556; pix2ang.cu:42-58,
166-176):

* ``B.txt``            npix, then npix brightness values (whitespace separated)
* ``theta_heal.txt``   npix values, one per line, no header (``phi_heal.txt`` alike)
* ``uvw{p}day1M.txt``  3 header tokens, then rows ``u v w`` [wavelengths]
* ``xyza{p}day1M.txt`` 3 header tokens, then rows ``x y z`` [m] (``xyzb`` alike)
* ``bll{p}day1M.txt``  1 header token, then one baseline length per row [wavelengths]
* ``C{p}day1M.txt``    npix lines of Re(C)

The reference reads every header with ``operator>>`` into a float, so headers
must be numeric; the writers put the row count there.  Values are written with
17 significant digits by default so FP64 inputs survive the round trip.
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from . import geometry as geo


def _fmt(precision: int) -> str:
    return f"%.{precision}g"


def write_sky(path: str, B: np.ndarray, precision: int = 17) -> None:
    B = np.asarray(B, np.float64)
    with open(path, "w") as f:
        f.write(f"{B.size}\n")
        np.savetxt(f, B, fmt=_fmt(precision))


def write_column(path: str, x: np.ndarray, precision: int = 17) -> None:
    np.savetxt(path, np.asarray(x, np.float64), fmt=_fmt(precision))


def write_rows(path: str, rows: np.ndarray, header_tokens: int, precision: int = 17) -> None:
    """rows: (n, width) array; header = row count followed by zeros."""
    rows = np.asarray(rows, np.float64)
    if rows.ndim == 1:
        rows = rows[:, None]
    with open(path, "w") as f:
        f.write(" ".join([str(rows.shape[0])] + ["0"] * (header_tokens - 1)) + "\n")
        np.savetxt(f, rows, fmt=_fmt(precision))


def read_numbers(path: str) -> np.ndarray:
    with open(path) as f:
        return np.array(f.read().split(), dtype=np.float64)


def read_rows(path: str, header_tokens: int, width: int) -> np.ndarray:
    """Rows of a per-day file (missing file -> 0 rows, as the reference)."""
    if not os.path.exists(path):
        return np.zeros((0, width))
    v = read_numbers(path)[header_tokens:]
    n = v.size // width
    return v[: n * width].reshape(n, width)


def read_sky(path: str) -> np.ndarray:
    v = read_numbers(path)
    npix = int(v[0])
    return v[1:1 + npix]


def day_tag(day: int) -> str:
    return f"{day}day1M.txt"


def write_day(in_dir: str, day: int, uvw: np.ndarray, xyz_a: np.ndarray, xyz_b: np.ndarray, bll: np.ndarray,
              hermitian: bool = True, precision: int = 17) -> int:
    """Write one day's four row files; uvw/xyz_* are (3, K).  With ``hermitian``
    the files hold the reference's 2K-row layout (second half negated)."""
    if hermitian:
        uvw, xyz_a, xyz_b, bll = geo.hermitian_rows(uvw, xyz_a, xyz_b, bll)
    tag = day_tag(day)
    write_rows(os.path.join(in_dir, "uvw" + tag), uvw.T, 3, precision)
    write_rows(os.path.join(in_dir, "xyza" + tag), xyz_a.T, 3, precision)
    write_rows(os.path.join(in_dir, "xyzb" + tag), xyz_b.T, 3, precision)
    write_rows(os.path.join(in_dir, "bll" + tag), bll, 1, precision)
    return uvw.shape[1]


def write_dataset(in_dir: str, inputs: geo.Inputs, days: int = 1, theta: Optional[np.ndarray] = None,
                  phi: Optional[np.ndarray] = None, sky_name: str = "B.txt", theta_name: str = "theta_heal.txt",
                  phi_name: str = "phi_heal.txt") -> None:
    """Sky files plus ``days`` per-day row files (the snapshots split evenly)."""
    os.makedirs(in_dir, exist_ok=True)
    write_sky(os.path.join(in_dir, sky_name), inputs.B)
    if theta is not None:
        write_column(os.path.join(in_dir, theta_name), theta)
    if phi is not None:
        write_column(os.path.join(in_dir, phi_name), phi)
    K = inputs.uvw.shape[1]
    edges = np.linspace(0, K, days + 1).astype(int)
    for d in range(days):
        s = slice(edges[d], edges[d + 1])
        write_day(in_dir, d + 1, inputs.uvw[:, s], inputs.xyz_a[:, s], inputs.xyz_b[:, s], inputs.bll[s])
