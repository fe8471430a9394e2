"""GPU radix sort and ray reordering (sort.hip): stable sort equals numpy's stable argsort;
the ray order is the stable sort of the closed-form octant + Morton keys; ordered traversal and
sorted wavefront rendering give results identical to the unordered ones."""
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


def test_ray_order_keys_closed_form(ctx, orc, cfg, scene1):
    """The order is the stable sort of the closed-form keys (octant above a 27-bit Morton code of
    the origin in the scene bounds): on hand-made rays whose keys were worked out by hand, and on
    random rays against the numpy restatement."""
    from sortkey import expected_order, hand_made_rays, scene_bounds
    s = mrt.Scene(ctx, scene1["verts"], scene1["mat"], scene1["alb"])
    lo, hi = scene_bounds(scene1["verts"])
    rows, keys = hand_made_rays(lo, hi)
    _say(f"ray_order hand-made n={rows.shape[0]}")
    d_rays = ctx.to_device(rows)
    d_order = ctx.empty(rows.shape[0], np.uint32)
    rtm.ray_order(s, d_rays, d_order)
    assert np.array_equal(d_order.download(), np.argsort(keys, kind="stable").astype(np.uint32))
    rays = orc.random_rays(91, 44.0, 200000)  # a few origins outside the bounds: clamped cells
    _say("ray_order random n=200000")
    d_rays = ctx.to_device(rays)
    d_order = ctx.empty(rays.shape[0], np.uint32)
    rtm.ray_order(s, d_rays, d_order)
    assert np.array_equal(d_order.download(), expected_order(rays, lo, hi))
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
