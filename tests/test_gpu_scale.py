"""The scale workload on the GPU against the oracle: 3200 level-4 icospheres (16,384,000
triangles) in [-100, 100]^3, 3840x2160, 8 bounces (the estimator of secondaryRays.comp:64-100
with the depth of :128-129).

  * the device BVH4 equals the host build byte for byte, stays within the node limit, and passes
    the structural checks (tree, every triangle once, leaves <= 4, boxes contain and are tight);
  * the bounce-1 extension rays of one frame equal the oracle's bit for bit, and every 64th of
    them traces to the oracle's binary-BVH hits bit for bit;
  * one 8-bounce frame: wavefront, megakernel and the oracle's per-path estimator give the same
    image bits.
"""
import numpy as np
import pytest

import minotert_amd as mrt
from minotert_amd import _capi

from bvh4_check import validate_bvh4

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-scale] {msg}", flush=True)


def _env(cfg):
    return mrt.Environment(sky=tuple(cfg.CONFIG["sky"]), sun_dir=tuple(cfg.CONFIG["sun_dir"]),
                           sun_e=tuple(cfg.CONFIG["sun_e"]))


@pytest.fixture(scope="module")
def scene4(ctx, cfg, orc):
    _say("generating the 16,384,000-triangle scene")
    verts, mat, alb = mrt.generate_scene(cfg.scene_desc(4))
    assert verts.shape[0] == 16_384_000
    _say("device BVH4 build")
    s = mrt.Scene(ctx, verts, mat, alb, builder="device")
    _say("oracle binary BVH build")
    bvh = orc.BVH(verts)
    yield {"verts": verts, "mat": mat, "alb": alb, "scene": s, "bvh": bvh}
    s.close()


def test_scale_device_bvh_equals_host(ctx, scene4):
    st = scene4["scene"].stats()
    assert st["n_nodes"] < _capi.MAX_WIDE_NODES and st["depth"] <= 24
    _say(f"host BVH4 build ({st['n_nodes']} device nodes, depth {st['depth']})")
    hn, ht, hst = mrt.runtime.build_bvh_host(scene4["verts"])
    assert (hst["n_nodes"], hst["depth"], hst["leaves"]) == (st["n_nodes"], st["depth"], st["leaves"])
    dn, dt = scene4["scene"].export()
    assert np.array_equal(dn, hn) and np.array_equal(dt, ht)
    _say("structural checks of the exported BVH4")
    res = validate_bvh4(dn, dt, scene4["verts"], max_leaf=4)
    assert res["depth"] == st["depth"]


def test_scale_bounce1_rays_and_hits_match_oracle(ctx, cfg, scene4):
    cam = cfg.camera(4)
    p = cfg.path_params(4)
    pt = mrt.Pathtracer(ctx, scene4["scene"])
    _say("bounce-1 capture at 3840x2160")
    gb = pt.primary_rays((cam.width, cam.height), cam, planes=False)
    opts = mrt.RenderOptions(frame_seed=p.frame_seed, sample_base=0, spp=1, bounces=p.bounces, rr_start=p.rr_start)
    d_rays, d_pid, n = pt.capture_rays(gb, cam, _env(cfg), opts, bounce=1)
    rays = d_rays.download()[:n]
    pid = d_pid.download()[:n]
    order = np.argsort(pid, kind="stable")
    _say("oracle bounce-1 rays")
    ref_rays, ref_pix = scene4["bvh"].bounce1_rays(cam, p)
    assert n == ref_rays.shape[0] > 1_000_000
    assert np.array_equal(pid[order], ref_pix)
    got = np.ascontiguousarray(rays[order]).view(np.float32).reshape(-1, 8)
    assert np.array_equal(got.view(np.uint32), ref_rays.view(np.uint32))
    sub = np.ascontiguousarray(got[::64])
    _say(f"trace every 64th bounce-1 ray ({sub.shape[0]}) on the GPU and in the oracle")
    d_sub = ctx.to_device(sub.view(mrt.RAY_DTYPE).reshape(-1))
    d_hits = ctx.empty(sub.shape[0], mrt.HIT_DTYPE)
    scene4["scene"].trace_closest(d_sub, d_hits)
    hits = d_hits.download()
    ref = scene4["bvh"].trace_closest(sub)
    assert np.array_equal(np.ascontiguousarray(hits).view(np.uint32).reshape(-1),
                          np.ascontiguousarray(ref).view(np.uint32).reshape(-1))
    assert np.mean(hits["tri"] != mrt.MISS) > 0.05
    for a in (d_rays, d_pid, d_sub, d_hits):
        a.free()


def test_scale_frame_wavefront_megakernel_oracle(ctx, cfg, scene4):
    cam = cfg.camera(4)
    p = cfg.path_params(4, frame=2)
    pt = mrt.Pathtracer(ctx, scene4["scene"])
    o = dict(frame_seed=p.frame_seed, spp=1, bounces=p.bounces, rr_start=p.rr_start)
    assert p.bounces == 8
    _say("8-bounce 3840x2160 frame: wavefront")
    a = pt.render(cam, _env(cfg), mrt.RenderOptions(**o)).download()
    _say("8-bounce 3840x2160 frame: megakernel from vertex 1")
    b = pt.render(cam, _env(cfg), mrt.RenderOptions(megakernel=True, **o)).download()
    assert np.isfinite(a).all() and a.mean() > 0.05
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    _say("8-bounce 3840x2160 frame: oracle")
    ref = scene4["bvh"].render_path(scene4["mat"], scene4["alb"], cam, p)
    bad = np.nonzero((a.view(np.uint32) != ref.view(np.uint32)).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} pixels differ, e.g. {bad[:5]}"
