"""GPU batches (paths.hip): several scenes in one frame with one traversal launch per bounce
over a shared queue, and the same frame captured in a HIP graph. Every scene's image must equal
its single-scene render bit for bit; graph replays (with eager work and synchronizations in
between) must equal eager batch renders of the same frame index; one scene is also checked
against the CPU oracle. Edge cases: an empty scene, one scene twice, 1x1 and odd view sizes, and
the 1..24 scene-count limit."""
import numpy as np
import pytest

import minotert_amd as mrt
from minotert_amd.pathtracer import Batch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-batch] {msg}", flush=True)


@pytest.fixture(scope="module")
def small_batch(ctx, cfg):
    descs = [cfg.scene_desc(1, spheres=40, box=12.0, seed=11), cfg.scene_desc(1, spheres=120, box=20.0, seed=12),
             cfg.scene_desc(0)]
    cams = [cfg.camera(1, width=64, height=48, box=12.0), cfg.camera(1, width=96, height=64, box=20.0),
            cfg.camera(0, width=80, height=80)]
    data = [mrt.generate_scene(d) for d in descs]
    scenes = [mrt.Scene(ctx, *x) for x in data]
    yield {"scenes": scenes, "cams": cams, "data": data}
    for s in scenes:
        s.close()


def _env(cfg):
    return mrt.Environment(sky=tuple(cfg.CONFIG["sky"]), sun_dir=tuple(cfg.CONFIG["sun_dir"]), sun_e=tuple(cfg.CONFIG["sun_e"]))


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_batch_equals_single_scene_renders(ctx, cfg, orc, small_batch):
    b = Batch(ctx, small_batch["scenes"])
    info = b.info()
    assert info["scenes"] == 3 and info["triangles"] == sum(s.n_tris for s in small_batch["scenes"])
    outs = [ctx.empty((c.width * c.height, 3), np.float32) for c in small_batch["cams"]]
    opts = mrt.RenderOptions(frame_seed=0x51, spp=2, bounces=3, rr_start=3)
    _say("batch render of 3 scenes")
    b.render(list(zip(small_batch["cams"], outs)), _env(cfg), opts)
    for j, (s, cam, out) in enumerate(zip(small_batch["scenes"], small_batch["cams"], outs)):
        _say(f"single render of scene {j}")
        ref = mrt.Pathtracer(ctx, s).render(cam, _env(cfg), opts).download()
        got = out.download()
        assert np.array_equal(_u32(got), _u32(ref)), f"scene {j} differs inside the batch"
        assert got.mean() > 0.01
    # scene 2 (the config-0 soup) against the oracle, sample 0 and 1 averaged as the kernel does
    v, m, a = small_batch["data"][2]
    bvh = orc.BVH(v)
    acc = None
    for smp in range(2):
        p = cfg.path_params(1)
        p.frame_seed, p.sample, p.bounces, p.rr_start = 0x51, smp, 3, 3
        r = bvh.render_path(m, a, small_batch["cams"][2], p)
        acc = r if acc is None else acc + r
    assert np.array_equal(_u32(outs[2].download()), _u32(acc / np.float32(2.0)))
    # the megakernel finishes the batch's paths with the same results
    mk = [ctx.empty((c.width * c.height, 3), np.float32) for c in small_batch["cams"]]
    _say("batch render with the megakernel")
    b.render(list(zip(small_batch["cams"], mk)), _env(cfg),
             mrt.RenderOptions(frame_seed=0x51, spp=2, bounces=3, rr_start=3, megakernel=True))
    for j in range(3):
        assert np.array_equal(_u32(mk[j].download()), _u32(outs[j].download())), f"megakernel scene {j}"
    b.close()


def test_graph_replays_match_eager_frames(ctx, cfg, small_batch):
    b = Batch(ctx, small_batch["scenes"])
    cams = small_batch["cams"]
    g_out = [ctx.empty((c.width * c.height, 3), np.float32) for c in cams]
    e_out = [ctx.empty((c.width * c.height, 3), np.float32) for c in cams]
    base = 0x1234
    _say("capture graph")
    g = b.graph(list(zip(cams, g_out)), _env(cfg), mrt.RenderOptions(frame_seed=base, spp=1, bounces=2, rr_start=3))
    g.set_frame(7)
    for rep in range(3):
        _say(f"graph replay {rep} (frame {7 + rep})")
        g.launch()
        ctx.sync()
        got = [o.download() for o in g_out]
        # eager work between replays: the same frame rendered without the graph
        b.render(list(zip(cams, e_out)), _env(cfg), mrt.RenderOptions(frame_seed=base ^ (7 + rep), spp=1, bounces=2,
                                                                      rr_start=3))
        ctx.sync()
        for j in range(len(cams)):
            assert np.array_equal(_u32(got[j]), _u32(e_out[j].download())), f"replay {rep} scene {j}"
    g.close()
    b.close()


def test_batch_rejects_bad_views(ctx, cfg, small_batch):
    b = Batch(ctx, small_batch["scenes"])
    out = ctx.empty((64 * 48, 3), np.float32)
    with pytest.raises(mrt.MrtError):
        b.render([(small_batch["cams"][0], out)], _env(cfg), mrt.RenderOptions())
    b.close()


@pytest.mark.parametrize("megakernel", [False, True])
def test_batch_edge_scenes_equal_single_renders(ctx, cfg, small_batch, megakernel):
    """A batch holding an empty scene (its BVH is one empty node: every path of its view ends in
    the sky), the same scene twice (two copies of one tree, rebased to different offsets), a 1x1
    view and odd view sizes: every view equals its single-scene render bit for bit."""
    empty = mrt.Scene(ctx, np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), np.zeros((1, 3), np.float32))
    s0, s1 = small_batch["scenes"][0], small_batch["scenes"][1]
    scenes = [s0, empty, s1, s0]
    cams = [cfg.camera(1, width=1, height=1, box=12.0), cfg.camera(0, width=7, height=3),
            cfg.camera(1, width=33, height=17, box=20.0), cfg.camera(1, width=64, height=48, box=12.0)]
    opts = mrt.RenderOptions(frame_seed=0x77, spp=2, bounces=4, rr_start=2, megakernel=megakernel)
    b = Batch(ctx, scenes)
    try:
        assert b.info()["bvh_nodes"] == sum(s.stats()["n_nodes"] for s in scenes)
        outs = [ctx.empty((c.width * c.height, 3), np.float32) for c in cams]
        _say(f"edge batch render, megakernel={megakernel}")
        b.render(list(zip(cams, outs)), _env(cfg), opts)
        for j, (s, cam, out) in enumerate(zip(scenes, cams, outs)):
            _say(f"edge single render {j}")
            ref = mrt.Pathtracer(ctx, s).render(cam, _env(cfg), opts).download()
            assert np.array_equal(_u32(out.download()), _u32(ref)), f"view {j} differs inside the batch"
        sky = np.tile(np.array(cfg.CONFIG["sky"], np.float32), (7 * 3, 1))
        assert np.array_equal(outs[1].download().reshape(-1, 3), sky)
    finally:
        b.close()
        empty.close()


def test_batch_scene_count_limit(ctx, small_batch):
    """1..24 scenes per batch; 0 and 25 are refused with an error, not truncated."""
    s = small_batch["scenes"][0]
    b = Batch(ctx, [s] * 24)
    assert b.info()["scenes"] == 24 and b.info()["triangles"] == 24 * s.n_tris
    b.close()
    for n in (0, 25):
        with pytest.raises(mrt.MrtError):
            Batch(ctx, [s] * n)
