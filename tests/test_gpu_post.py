"""Post-processing on the GPU (post.hip) against the oracle (oracle/oracle_post.cpp).

Both sides take exp and pow from the same basic-operation methods (csrc/transc.h,
oracle/oracle_transc.hpp), the same bilinear rule and the same operation order, so the bar is
bit-exact: every tonemap operator over HDR values from 0 to 1e4 (float and RGBA8 outputs), and
the denoiser on a rendered G-buffer (1-spp path-traced colour, its depth and normals) and on
random ragged images, through both the LDS-tiled kernel (radius <= 12) and the global-memory
kernel (radius 13 and above).
"""
import numpy as np
import pytest

import minotert_amd as mrt
from minotert_amd import post
from oracle import oracle_post as op

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-post] {msg}", flush=True)


def _hdr(n, seed):
    rng = np.random.default_rng(seed)
    x = np.exp(rng.uniform(np.log(1e-5), np.log(1e4), (n, 3))).astype(np.float32)
    x[rng.random((n, 3)) < 0.05] = 0.0
    x[:4] = [[0, 0, 0], [1, 1, 1], [0.18, 0.18, 0.18], [16, 16, 16]]
    return x


@pytest.mark.parametrize("name", list(post.OPERATORS))
def test_tonemap_bit_exact(ctx, name):
    _say(f"tonemap {name}")
    x = _hdr(70001, 11)
    tp = post.TonemapParams(op=post.OPERATORS[name], exposure=1.5)
    d_in = ctx.to_device(x)
    d_out = ctx.empty((x.shape[0], 3), np.float32)
    d_q = ctx.empty((x.shape[0], 4), np.uint8)
    post.Tonemapper(ctx).apply(d_in, tp, d_out, d_q)
    got, q = d_out.download(), d_q.download()
    want, wq = op.tonemap(x, op.tonemap_params(op=tp.op, exposure=1.5))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(q, wq)
    # RGBA8 alone
    d_q2 = ctx.empty((x.shape[0], 4), np.uint8)
    post.Tonemapper(ctx).apply(d_in, tp, None, d_q2)
    np.testing.assert_array_equal(d_q2.download(), wq)


def test_tonemap_defaults_and_errors(ctx):
    d, r = post.tonemap_defaults(), post.TonemapParams()
    assert (d.op, d.exposure, d.reinhard_max) == (r.op, r.exposure, r.reinhard_max)
    np.testing.assert_array_equal(np.float32(d.uchimura), np.float32(r.uchimura))
    np.testing.assert_array_equal(np.float32(d.amd), np.float32(r.amd))
    x = ctx.to_device(np.ones((4, 3), np.float32))
    with pytest.raises(mrt.MrtError):
        post.Tonemapper(ctx).apply(x, post.TonemapParams(op=9), ctx.empty((4, 3), np.float32))
    with pytest.raises(mrt.MrtError):
        post.Tonemapper(ctx).apply(x, post.TonemapParams())  # no output


def _denoise(ctx, rgb, depth, normal, near, params, frame):
    h, w = depth.shape
    d_rgb, d_d, d_n = ctx.to_device(rgb.reshape(-1, 3)), ctx.to_device(depth.reshape(-1)), ctx.to_device(normal.reshape(-1, 4))
    d_out = ctx.empty((w * h, 3), np.float32)
    post.Denoiser(ctx).bilateral(d_rgb, d_d, d_n, w, h, near, d_out, params, frame)
    return d_out.download().reshape(h, w, 3)


@pytest.mark.parametrize("w,h,sigma,ks", [(37, 23, 5.0, 2.0), (64, 48, 6.0, 2.0), (45, 19, 6.5, 2.0), (45, 19, 4.0, 4.0),
                                          (1, 1, 5.0, 2.0), (130, 3, 1.0, 3.0)])
def test_denoise_random_bit_exact(ctx, w, h, sigma, ks):
    """Random colour, depth (with sky pixels) and normals; radius 10 and 12 (LDS tile), 13 and 16
    (global memory), a single pixel and a three-row strip."""
    _say(f"denoise random {w}x{h} radius {round(sigma * ks)}")
    rng = np.random.default_rng(w * 1000 + h)
    rgb = rng.uniform(0, 3, (h, w, 3)).astype(np.float32)
    depth = rng.uniform(0.02, 0.06, (h, w)).astype(np.float32)
    depth[rng.random((h, w)) < 0.1] = 0.0
    n = rng.normal(size=(h, w, 3))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    normal = np.zeros((h, w, 4), np.float32)
    normal[..., :3] = n
    # piecewise-flat regions so that the range weights are not all at their minimum
    normal[: h // 2, : w // 2, :3] = [0, 0, 1]
    depth[: h // 2, : w // 2] = 0.05
    bp = post.BilateralParams(sigma=sigma, k_sigma=ks)
    got = _denoise(ctx, rgb, depth, normal, 0.1, bp, 3)
    want = op.denoise_bilateral(rgb, depth, normal, 0.1, sigma=sigma, k_sigma=ks, threshold=bp.threshold, frame=3)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_denoise_rendered_frame_bit_exact(ctx, cfg, scene0):
    """The reference's use: the 1-spp path-traced image filtered with its own G-buffer."""
    cam = cfg.camera(0, width=96, height=64)
    s = mrt.Scene(ctx, scene0["verts"], scene0["mat"], scene0["alb"])
    try:
        pt = mrt.Pathtracer(ctx, s)
        env = mrt.Environment(sky=tuple(cfg.CONFIG["sky"]), sun_dir=tuple(cfg.CONFIG["sun_dir"]),
                              sun_e=tuple(cfg.CONFIG["sun_e"]))
        gb = pt.primary_rays((cam.width, cam.height), cam)
        rgb = pt.secondary_rays(gb, cam, env, mrt.RenderOptions(frame_seed=4))
        d_out = ctx.empty((cam.width * cam.height, 3), np.float32)
        post.Denoiser(ctx).bilateral(rgb, gb.depth, gb.normal, cam.width, cam.height, cam.near_plane, d_out, frame=4)
        got = d_out.download().reshape(cam.height, cam.width, 3)
        c = rgb.download().reshape(cam.height, cam.width, 3)
        d = gb.depth.download().reshape(cam.height, cam.width)
        nrm = gb.normal.download().reshape(cam.height, cam.width, 4)
    finally:
        s.close()
    assert np.any(d == 0.0) and np.any(d > 0.0)
    want = op.denoise_bilateral(c, d, nrm, cam.near_plane, frame=4)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    # the filter reduces the noise of the 1-spp image on hit pixels
    assert np.std(np.diff(got[d > 0], axis=0)) < np.std(np.diff(c[d > 0], axis=0))


def test_denoise_errors(ctx):
    x = ctx.empty((16, 3), np.float32)
    dd = ctx.empty(16, np.float32)
    nn = ctx.empty((16, 4), np.float32)
    with pytest.raises(mrt.MrtError):
        post.Denoiser(ctx).bilateral(x, dd, nn, 4, 4, 0.1, x)  # in place
    with pytest.raises(mrt.MrtError):
        post.Denoiser(ctx).bilateral(x, dd, nn, 4, 4, 0.1, ctx.empty((16, 3), np.float32),
                                     post.BilateralParams(sigma=40.0, k_sigma=2.0))  # radius 80 > 64
