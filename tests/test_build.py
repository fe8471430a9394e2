"""The HIP library's Makefile reports compile failures: a translation unit that does not compile
makes `make` exit non-zero and leaves no object behind (so no stale object from an earlier
build can be linked into the library)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "minotert_amd", "csrc")

pytestmark = pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                                reason="needs make and hipcc")


def _copy_tree(tmp_path):
    # the Makefile reads ../../include/minotert.h and writes ../_build, ../_lib
    pkg = tmp_path / "pkg"
    shutil.copytree(CSRC, pkg / "csrc")
    shutil.copytree(os.path.join(ROOT, "include"), tmp_path / "include")
    return pkg


def _make(pkg, target):
    return subprocess.run(["make", "-s", "-C", str(pkg / "csrc"), target], capture_output=True, text=True)


def test_compile_error_fails_the_build_and_leaves_no_object(tmp_path):
    pkg = _copy_tree(tmp_path)
    obj = pkg / "_build" / "util.hip.o"
    r = _make(pkg, "../_build/util.hip.o")
    assert r.returncode == 0, r.stderr
    assert obj.exists()
    src = pkg / "csrc" / "util.hip"
    src.write_text("this is not C++;\n" + src.read_text())
    r = _make(pkg, "../_build/util.hip.o")
    assert r.returncode != 0
    assert "error" in (r.stdout + r.stderr)
    assert not obj.exists(), "a failed compile must not leave the previous object in place"


def test_warning_filter_keeps_a_clean_compile_quiet_and_successful(tmp_path):
    pkg = _copy_tree(tmp_path)
    r = _make(pkg, "../_build/util.hip.o")
    assert r.returncode == 0
    assert "packed-fp32-ops' is not a recognized feature" not in r.stdout + r.stderr
