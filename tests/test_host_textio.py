"""The drivers' text I/O (csrc/host/textio.h: threaded parse and write)
against strtod and snprintf, compiled with g++ on the host.  Parsed values
must be bit-identical and written files byte-identical (the reference's text
layouts, Appendix-B formats)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_textio_matches_strtod_and_printf(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    exe = tmp_path / "textio_check"
    src = os.path.join(ROOT, "tests", "host", "textio_check.cpp")
    inc = os.path.join(ROOT, "viss_optimize_3d_amd", "csrc", "host")
    subprocess.run([cxx, "-O2", "-std=c++17", "-pthread", "-I", inc, src, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr


def test_pow5_table_is_generated():
    """The committed 128-bit power-of-five table is what its generator writes
    (exact big-integer arithmetic in Python)."""
    d = os.path.join(ROOT, "viss_optimize_3d_amd", "csrc", "host")
    out = subprocess.run(["python3", os.path.join(d, "gen_pow5_table.py")], capture_output=True, text=True,
                         check=True).stdout
    with open(os.path.join(d, "pow5_table.h")) as f:
        assert f.read() == out
