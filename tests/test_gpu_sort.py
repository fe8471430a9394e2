"""GPU radix sort and ray reordering (sort.hip): stable sort equals numpy's stable argsort;
the ray order is a permutation grouped by direction octant; ordered traversal and sorted
wavefront rendering give results identical to the unordered ones."""
import numpy as np
import pytest

import minotert_amd as mrt
from minotert_amd import runtime as rtm

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-sort] {msg}", flush=True)


@pytest.mark.parametrize("n,bits", [(1, 32), (2047, 8), (2048, 30), (300001, 30), (1 << 20, 32), (5000, 13)])
def test_sort_pairs_matches_numpy_stable(ctx, n, bits):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    keys[::7] = keys[0]  # many duplicates: stability matters
    vals = np.arange(n, dtype=np.uint32)
    _say(f"sort_pairs n={n} bits={bits}")
    dk, dv = ctx.to_device(keys), ctx.to_device(vals)
    rtm.sort_pairs(ctx, dk, dv, n, bits)
    mask = np.uint64((1 << bits) - 1)
    low = (keys.astype(np.uint64) & mask)
    order = np.argsort(low, kind="stable")
    assert np.array_equal(dv.download(), vals[order])
    assert np.array_equal(dk.download(), keys[order])


def test_ray_order_and_ordered_trace(ctx, orc, cfg, scene1):
    s = mrt.Scene(ctx, scene1["verts"], scene1["mat"], scene1["alb"])
    rays = orc.random_rays(77, 40.0, 300000)
    _say("ray_order n=300000")
    d_rays = ctx.to_device(rays)
    d_order = ctx.empty(rays.shape[0], np.uint32)
    rtm.ray_order(s, d_rays, d_order)
    order = d_order.download()
    assert np.array_equal(np.sort(order), np.arange(rays.shape[0], dtype=np.uint32))
    d = rays[order, 4:7]
    oct_ = (d[:, 0] < 0) * 4 + (d[:, 1] < 0) * 2 + (d[:, 2] < 0) * 1
    assert np.all(np.diff(oct_) >= 0)
    h1 = ctx.empty(rays.shape[0], mrt.HIT_DTYPE)
    h2 = ctx.empty(rays.shape[0], mrt.HIT_DTYPE)
    s.trace_closest(d_rays, h1)
    rtm.trace_closest_ordered(s, d_rays, d_order, h2)
    a, b = h1.download(), h2.download()
    assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))
    s.close()


def test_sorted_wavefront_equals_unsorted(ctx, orc, cfg, scene1):
    s = mrt.Scene(ctx, scene1["verts"], scene1["mat"], scene1["alb"])
    pt = mrt.Pathtracer(ctx, s)
    cam = cfg.camera(1, width=320, height=180)
    p = cfg.path_params(1, frame=4)
    env = mrt.Environment(sky=tuple(cfg.CONFIG["sky"]), sun_dir=tuple(cfg.CONFIG["sun_dir"]), sun_e=tuple(cfg.CONFIG["sun_e"]))
    _say("sorted vs unsorted wavefront 320x180 spp=2")
    a = pt.render(cam, env, mrt.RenderOptions(frame_seed=p.frame_seed, spp=2, bounces=4, rr_start=3)).download()
    b = pt.render(cam, env, mrt.RenderOptions(frame_seed=p.frame_seed, spp=2, bounces=4, rr_start=3,
                                              sort_rays=True)).download()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    s.close()
