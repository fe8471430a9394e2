"""The autotuner (minotert_amd/autotune.py): a measured choice is recorded in the cache file and
reused on the next call without measuring, and results never depend on the choice — every
traversal variant and every frame variant (shading block, megakernel shape) gives the same hits
and the same images bit for bit."""
import json

import numpy as np
import pytest

import minotert_amd as mrt
from minotert_amd import autotune
from minotert_amd.pathtracer import Batch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-autotune] {msg}", flush=True)


@pytest.fixture
def restore_variants(ctx):
    before = (ctx.trace_variant(False), ctx.trace_variant(True), ctx.frame_variant())
    yield
    ctx.set_trace_variant(before[0], False)
    ctx.set_trace_variant(before[1], True)
    ctx.set_frame_variant(before[2])


def _env(cfg):
    return mrt.Environment(sky=tuple(cfg.CONFIG["sky"]), sun_dir=tuple(cfg.CONFIG["sun_dir"]),
                           sun_e=tuple(cfg.CONFIG["sun_e"]))


def test_trace_tune_records_and_reuses(ctx, orc, scene1, tmp_path, restore_variants):
    s = mrt.Scene(ctx, scene1["verts"], scene1["mat"], scene1["alb"])
    rays = orc.random_rays(5, 40.0, 100000)
    d_rays = ctx.to_device(rays)
    d_hits = ctx.empty(rays.shape[0], mrt.HIT_DTYPE)
    ctx.set_trace_variant(0)
    s.trace_closest(d_rays, d_hits)
    ref = d_hits.download()
    path = str(tmp_path / "tuning.json")
    _say("traversal tune, measured")
    a = autotune.tune(ctx, s, d_rays, d_hits, "test_random_rays", cache_path=path)
    assert a["source"] == "measured" and set(a["ms"]) == set(mrt.runtime.trace_variants())
    rec = json.load(open(path))
    key = next(k for k in rec if k.endswith("|trace_closest|test_random_rays"))
    assert rec[key]["variant"] == a["variant"] == mrt.runtime.trace_variants()[ctx.trace_variant()]
    ctx.set_trace_variant(0)
    b = autotune.tune(ctx, s, d_rays, d_hits, "test_random_rays", cache_path=path)
    assert b["source"] == "cache" and b["variant"] == a["variant"]
    assert ctx.trace_variant() == b["index"]
    s.trace_closest(d_rays, d_hits)
    assert np.array_equal(np.ascontiguousarray(d_hits.download()).view(np.uint32), np.ascontiguousarray(ref).view(np.uint32))
    s.close()


def test_frame_tune_records_reuses_and_keeps_images(ctx, cfg, tmp_path, restore_variants):
    layout = cfg.batch_layout(2)[:3]
    scenes, views = [], []
    for j, e in enumerate(layout):
        d = cfg.scene_desc(1, spheres=20 + 10 * j, box=12.0, seed=e.scene.seed, level=3)
        v, m, a = mrt.generate_scene(d)
        scenes.append(mrt.Scene(ctx, v, m, a))
        cam = cfg.camera(1, width=96 + 32 * j, height=64, box=12.0)
        views.append((cam, ctx.empty((cam.width * cam.height, 3), np.float32)))
    env = _env(cfg)
    opts = mrt.RenderOptions(frame_seed=cfg.frame_seed(3), spp=2, bounces=4, rr_start=3)
    b = Batch(ctx, scenes)
    for any_hit in (False, True):
        ctx.set_trace_variant(0, any_hit)
    ctx.set_frame_variant(0)
    b.render(views, env, opts)
    ref = [o.download() for _, o in views]
    single = mrt.Pathtracer(ctx, scenes[0]).render(views[0][0], env, opts).download()
    for fv, name in enumerate(mrt.runtime.frame_variants()):
        _say(f"frame variant {name}")
        ctx.set_frame_variant(fv)
        b.render(views, env, opts)
        for (_, o), r in zip(views, ref):
            assert np.array_equal(o.download().view(np.uint32), r.view(np.uint32)), name
        img = mrt.Pathtracer(ctx, scenes[0]).render(views[0][0], env, opts).download()
        assert np.array_equal(img.view(np.uint32), single.view(np.uint32)), name
    path = str(tmp_path / "tuning.json")
    _say("frame tune, measured")
    t1 = autotune.tune_frame(ctx, lambda: b.render(views, env, opts), "test_batch3", cache_path=path, reps=3)
    assert t1["source"] == "measured"
    assert len(t1["ms"]) == len(mrt.runtime.trace_variants()) + len(mrt.runtime.frame_variants())
    rec = json.load(open(path))
    key = next(k for k in rec if k.endswith("|frame|test_batch3"))
    assert (rec[key]["trace_variant"], rec[key]["frame_variant"]) == (t1["trace_variant"], t1["frame_variant"])
    ctx.set_frame_variant(0)
    t2 = autotune.tune_frame(ctx, lambda: b.render(views, env, opts), "test_batch3", cache_path=path)
    assert t2["source"] == "cache"
    assert (t2["trace_variant"], t2["frame_variant"]) == (t1["trace_variant"], t1["frame_variant"])
    assert mrt.runtime.frame_variants()[ctx.frame_variant()] == t1["frame_variant"]
    b.render(views, env, opts)
    for (_, o), r in zip(views, ref):
        assert np.array_equal(o.download().view(np.uint32), r.view(np.uint32))
    b.close()
    for sc in scenes:
        sc.close()


def test_every_traversal_variant_renders_the_same_frames(ctx, cfg, restore_variants):
    """Every traversal variant renders bit-identical frames: single-scene wavefront frames (mixed
    shadow + extension launches, device-side counts) and a two-scene batch (per-ray roots)."""
    layout = cfg.batch_layout(2)[:2]
    scenes, views = [], []
    for j, e in enumerate(layout):
        d = cfg.scene_desc(1, spheres=30 + 10 * j, box=12.0, seed=e.scene.seed, level=3)
        v, m, a = mrt.generate_scene(d)
        scenes.append(mrt.Scene(ctx, v, m, a))
        cam = cfg.camera(1, width=80 + 16 * j, height=48, box=12.0)
        views.append((cam, ctx.empty((cam.width * cam.height, 3), np.float32)))
    env = _env(cfg)
    opts = mrt.RenderOptions(frame_seed=cfg.frame_seed(5), spp=2, bounces=4, rr_start=3)
    b = Batch(ctx, scenes)
    pt = mrt.Pathtracer(ctx, scenes[0])
    for any_hit in (False, True):
        ctx.set_trace_variant(0, any_hit)
    b.render(views, env, opts)
    ref = [o.download() for _, o in views]
    single = pt.render(views[0][0], env, opts).download()
    for i, name in enumerate(mrt.runtime.trace_variants()):
        _say(f"trace variant {i} {name}")
        for any_hit in (False, True):
            ctx.set_trace_variant(i, any_hit)
        b.render(views, env, opts)
        for (_, o), r in zip(views, ref):
            assert np.array_equal(o.download().view(np.uint32), r.view(np.uint32)), name
        img = pt.render(views[0][0], env, opts).download()
        assert np.array_equal(img.view(np.uint32), single.view(np.uint32)), name
    b.close()
    for sc in scenes:
        sc.close()
