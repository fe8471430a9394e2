"""The allocation-free ray order (mrt_ray_order_noalloc) against mrt_ray_order, both called
through the C ABI: two calls in a row, the second with a smaller input (the shared scratch array
still holds the first call's keys and histograms), each bit-identical to the allocating entry
point; a size past the limit is refused with a status code before any device work."""
import numpy as np
import pytest

import minotert_amd as mrt
from minotert_amd import _capi

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)


def _order(ctx, scene, d_rays, n, noalloc):
    lib = _capi.lib()
    out = ctx.empty(n, np.uint32)
    fn = lib.mrt_ray_order_noalloc if noalloc else lib.mrt_ray_order
    ctx.check(fn(ctx.handle, scene.handle, d_rays.ptr, n, out.ptr))
    ctx.sync()
    got = out.download()
    out.free()
    return got


def test_noalloc_ray_order_equals_allocating_order(ctx, orc, cfg):
    verts, mat, alb = mrt.generate_scene(cfg.scene_desc(1, spheres=200))
    scene = mrt.Scene(ctx, verts, mat, alb)
    limit = _capi.RAY_ORDER_NOALLOC_MAX
    for n, seed in ((limit, 5), (300001, 6)):  # the largest size, then a smaller one
        rays = orc.random_rays(seed, 44.0, n)
        d_rays = ctx.to_device(rays)
        _say(f"ray order of {n} rays: allocation-free, then allocating")
        got = _order(ctx, scene, d_rays, n, noalloc=True)
        want = _order(ctx, scene, d_rays, n, noalloc=False)
        assert np.array_equal(got, want)
        assert np.array_equal(np.sort(got), np.arange(n, dtype=np.uint32))
        d_rays.free()
    lib = _capi.lib()
    d_rays = ctx.empty(1, mrt.RAY_DTYPE)
    d_out = ctx.empty(1, np.uint32)
    _say("size past the limit")
    st = lib.mrt_ray_order_noalloc(ctx.handle, scene.handle, d_rays.ptr, limit + 1, d_out.ptr)
    assert st == 1 and b"exceed" in lib.mrt_last_error(ctx.handle)
    assert lib.mrt_ray_order_noalloc(ctx.handle, scene.handle, d_rays.ptr, 0, d_out.ptr) == 0
    scene.close()
