"""Host-side sanitizer runs (`make -C oracle SAN=1 check`): the CPU oracle and the product's
host-only translation units (scene generator, camera basis, host BVH4 builder) under
AddressSanitizer + UndefinedBehaviorSanitizer through oracle/san_check.cpp, and the oracle's own
CPU tests against the UndefinedBehaviorSanitizer build of the oracle library."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

pytestmark = pytest.mark.skipif(shutil.which("make") is None or shutil.which("g++") is None,
                                reason="needs make and g++")


def _san_build():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "SAN=1", "all"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_asan_ubsan_harness_is_clean():
    _san_build()
    r = subprocess.run(["make", "-s", "-C", ORACLE, "SAN=1", "check"], capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "san_check: ok (0 failures)" in out
    assert "runtime error" not in out and "AddressSanitizer" not in out


def test_oracle_tests_under_ubsan():
    _san_build()
    env = dict(os.environ, MRT_ORACLE_LIB=os.path.join(ORACLE, "_build_san", "liboracle.so"),
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    tests = [os.path.join(ROOT, "tests", f) for f in
             ("test_oracle.py", "test_oracle_cache.py", "test_oracle_sky.py", "test_oracle_post.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *tests],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out
