"""Reference file layouts (textio) and the synthetic generators, on the CPU."""
import numpy as np

from viss_optimize_3d_amd import geometry as geo
from viss_optimize_3d_amd import textio


def test_day_files_round_trip(tmp_path):
    inp = geo.generate(geo.WORKLOADS["plain-nside16"])
    textio.write_dataset(str(tmp_path), inp, days=3)
    assert textio.read_sky(str(tmp_path / "B.txt")).tolist() == inp.B.tolist()
    rows = 0
    for d in (1, 2, 3):
        tag = textio.day_tag(d)
        uvw = textio.read_rows(str(tmp_path / ("uvw" + tag)), 3, 3)
        xa = textio.read_rows(str(tmp_path / ("xyza" + tag)), 3, 3)
        bll = textio.read_rows(str(tmp_path / ("bll" + tag)), 1, 1)
        n = uvw.shape[0] // 2
        # 2*amount layout: the second half is the negated first half, satellites swapped
        np.testing.assert_array_equal(uvw[n:], -uvw[:n])
        assert xa.shape == uvw.shape and bll.shape == (uvw.shape[0], 1)
        np.testing.assert_array_equal(uvw[:n].T, inp.uvw[:, rows:rows + n])
        rows += n
    assert rows == inp.uvw.shape[1]
    assert textio.read_rows(str(tmp_path / "uvw9day1M.txt"), 3, 3).shape == (0, 3)  # missing file: 0 rows


def test_batched_workload_shape():
    B, uvw, xa, xb, offs, gf = geo.batched_workload(nside=8, freqs=(1e6, 1e7), n_groups=3, snaps=5, n_sat=4)
    assert B.size == 768 and offs.tolist() == [0, 30, 60, 90, 120, 150, 180]
    np.testing.assert_allclose(uvw[:, :30] * 10, uvw[:, 90:120], rtol=1e-12)  # same geometry, 10x frequency
    r = np.linalg.norm(xa, axis=0)
    assert r.min() > geo.R_MOON
