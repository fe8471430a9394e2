"""The C++ render-stage API (include/minotert.hpp) on the GPU: the example program renders the
config-0 scene through Pathtracer::primaryRays + Lambert and through path tracing, then
Denoiser::bilateral + the AMD Tonemapper; its images must equal the oracle's bit for bit."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "minotert_amd", "_lib", "pathtracer_demo")


def _run(tmp_path, mode, *extra):
    out = tmp_path / f"{mode}.f32"
    print(f"[gpu-cpp] pathtracer_demo {mode}", flush=True)
    r = subprocess.run([DEMO, str(out), "64", "2", "10", "256", "256", mode, *extra], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return np.fromfile(out, np.float32).reshape(-1, 3)


def test_cpp_lambert_matches_oracle(tmp_path, orc, cfg, scene0):
    img = _run(tmp_path, "lambert")
    _, ref = scene0["bvh"].render_lambert(scene0["mat"], scene0["alb"], cfg.camera(0), cfg.CONFIG["sun_dir"],
                                          [1.0, 1.0, 1.0])
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


def test_cpp_path_tracing_matches_oracle(tmp_path, orc, cfg, scene0):
    img = _run(tmp_path, "path", "0x1234", "3")
    p = cfg.path_params(1)
    p.frame_seed, p.bounces, p.rr_start = 0x1234, 3, 3
    p.sky, p.sun_dir, p.sun_e = (1.0, 1.0, 1.0), tuple(cfg.CONFIG["sun_dir"]), (2.0, 2.0, 2.0)
    ref = scene0["bvh"].render_path(scene0["mat"], scene0["alb"], cfg.camera(0), p)
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


def test_cpp_denoise_tonemap_matches_oracle(tmp_path, ctx, orc, cfg, scene0):
    """RGBA8 of the C++ Denoiser + Tonemapper over the demo's path-traced image, against the
    oracle's denoiser and AMD tonemapper applied to the same image and G-buffer planes."""
    import minotert_amd as mrt
    from oracle import oracle_post as op
    q_path = tmp_path / "post.rgba8"
    img = _run(tmp_path, "path", "0x1234", "3", str(q_path))
    q = np.fromfile(q_path, np.uint8).reshape(-1, 4)
    cam = cfg.camera(0)
    s = mrt.Scene(ctx, scene0["verts"], scene0["mat"], scene0["alb"])
    try:
        gb = mrt.Pathtracer(ctx, s).primary_rays((cam.width, cam.height), cam)
        depth = gb.depth.download().reshape(cam.height, cam.width)
        nrm = gb.normal.download().reshape(cam.height, cam.width, 4)
    finally:
        s.close()
    den = op.denoise_bilateral(img.reshape(cam.height, cam.width, 3), depth, nrm, 0.001, frame=0)
    _, want = op.tonemap(den, op.tonemap_params())
    assert np.array_equal(q, want)
