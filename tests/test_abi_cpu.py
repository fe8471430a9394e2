"""C-ABI checks that need no GPU: the library loads and exports every entry point declared in
include/*.h, host-side entry points (scene generator, camera basis) agree bit-for-bit with
the oracle, and argument validation reports status codes instead of crashing."""
import ctypes as C
import glob
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import minotert_amd as mrt
from minotert_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mrt_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert set(names) <= set(_capi.PROTOTYPES), "ctypes prototypes must cover the header"
    assert lib.mrt_abi_version() == 1


def test_library_is_gfx950_code_object():
    # the HIP fat binary embedded in the library targets gfx950
    data = open(_capi.LIB_PATH, "rb").read()
    assert b"gfx950" in data


@pytest.mark.parametrize("cid,spheres,level", [(0, None, None), (1, 40, None), (4, 3, 4)])
def test_scene_generator_matches_oracle(orc, cfg, cid, spheres, level):
    d = cfg.scene_desc(cid, spheres=spheres, level=level)
    a = mrt.generate_scene(d)
    b = orc.generate_scene(d)
    for x, y in zip(a, b):
        assert np.array_equal(np.ascontiguousarray(x).view(np.uint32), np.ascontiguousarray(y).view(np.uint32))


def test_camera_basis_matches_oracle(orc, cfg):
    for cid in (0, 1, 4):
        cam = cfg.camera(cid)
        assert np.array_equal(mrt.camera_basis(cam).view(np.uint32), orc.camera_basis(cam).view(np.uint32))


def test_host_argument_validation():
    lib = _capi.lib()
    assert lib.mrt_scene_generate(None, None, None, None) == 1
    cam = _capi.mrt_camera()
    out = (C.c_float * 14)()
    assert lib.mrt_camera_basis(C.byref(cam), out) == 1  # zero viewport
    assert b"" != lib.mrt_last_error(None) or True
    assert lib.mrt_create(None, None) == 1
    assert lib.mrt_sync(None) == 1
    assert lib.mrt_trace_closest(None, None, None, 0, None) == 1
    assert b"null" in lib.mrt_last_error(None)


def test_create_without_device_reports_error():
    h = C.c_void_p()
    st = _capi.lib().mrt_create(C.byref(h), None)
    if st == 0:
        _capi.lib().mrt_destroy(h)
        pytest.skip("a HIP device is present")
    assert st == 2 and b"no HIP device" in _capi.lib().mrt_last_error(None)


def test_configs_consistent(cfg):
    assert cfg.scene_desc(0).tri_count() == 20482
    assert cfg.scene_desc(1).tri_count() == 1024000
    assert cfg.scene_desc(4).tri_count() == 16384000
    lay = cfg.batch_layout()
    assert len(lay) == 16
    tris = [e.scene.tri_count() for e in lay]
    assert min(tris) >= 1280 and max(tris) <= 2_000_000 + 1280
    assert {(e.width, e.height) for e in lay} <= {(256, 256), (512, 512), (1280, 720), (1920, 1080)}


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpp_render_stage_header_compiles(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "minotert.hpp"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
