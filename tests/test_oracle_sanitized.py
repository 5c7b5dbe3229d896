"""The FP64 oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(`make -C oracle san`) runs the oracle's CPU tests in a child process with
libasan preloaded: any out-of-bounds access, use after free or undefined
behaviour in the oracle aborts that run and fails this test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SAN_LIB = os.path.join(ORACLE, "liboracle_dts_san.so")

SUITES = ["tests/test_oracle.py", "tests/test_oracle_pins.py", "tests/test_golden.py",
          "tests/test_oracle_rf.py", "tests/test_fixed_point_bounds.py"]


def _libasan():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_suites_under_sanitizers():
    asan = _libasan()
    if asan is None:
        pytest.skip("gcc has no libasan")
    subprocess.check_call(["make", "-s", "san"], cwd=ORACLE)
    env = dict(os.environ, DTS_ORACLE_LIB=SAN_LIB, LD_PRELOAD=asan, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="detect_leaks=0:exitcode=86",
               UBSAN_OPTIONS="halt_on_error=1:exitcode=87:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        *SUITES], cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    tail = (r.stdout[-2000:] + r.stderr[-4000:])
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "failed" not in r.stdout, tail
