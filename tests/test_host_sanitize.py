"""CPU sanitizer run: the oracle (oracle/viss_oracle.c) and the drivers' text
I/O (csrc/host/textio.h) built with -fsanitize=address,undefined
(`make -C oracle sanitize`) and driven by tests/host/oracle_check.c and
tests/host/textio_check.cpp: every oracle function on small and edge-case
inputs (all flag combinations, empty inputs, NaN / infinite / out-of-range
shell lengths, a satellite at the origin, non-power-of-two NEST), and the
threaded number parser on malformed, truncated, huge-exponent, long-digit and
fuzzed text against strtod.  Any out-of-bounds access, leak or undefined
behaviour aborts the program."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def built():
    if shutil.which("gcc") is None or shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("no gcc / g++ / make")
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True, capture_output=True, text=True)
    return os.path.join(ORACLE, "_build")


@pytest.mark.parametrize("prog,threads", [("oracle_check_san", "4"), ("textio_check_san", None)])
def test_sanitized_host_code(built, prog, threads, tmp_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    if threads:
        env["OMP_NUM_THREADS"] = threads
    r = subprocess.run([os.path.join(built, prog)], cwd=tmp_path, env=env, capture_output=True, text=True,
                       timeout=600)
    tail = (r.stdout + r.stderr)[-3000:]
    print(tail)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), tail
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, tail
