"""GPU parity of the closest-hit and any-hit traversal (trace.hip) against the CPU oracle.

Bar: triangle IDs, t, u and v bit-identical to the oracle (closest hit is independent of
the tree because the box test is conservative and ties go to the smaller ID). The allowed
tolerance for t/u/v would be 1e-5 relative; the kernels meet it exactly.
"""
import numpy as np
import pytest

import minotert_amd as mrt

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-trace] {msg}", flush=True)


def _trace(ctx, scene, rays):
    d_rays = ctx.to_device(np.ascontiguousarray(rays, np.float32))
    d_hits = ctx.empty(rays.shape[0], mrt.HIT_DTYPE)
    scene.trace_closest(d_rays, d_hits)
    return d_hits.download()


def _assert_hits_equal(got, ref, what):
    bad = np.nonzero(got["tri"] != ref["tri"])[0]
    assert bad.size == 0, f"{what}: {bad.size} triangle IDs differ, first ray {bad[:5]}: {got[bad[:5]]} vs {ref[bad[:5]]}"
    for f in ("t", "u", "v"):
        eq = got[f].view(np.uint32) == ref[f].view(np.uint32)
        assert eq.all(), f"{what}: field {f} differs at {np.nonzero(~eq)[0][:5]}"


@pytest.fixture(scope="module")
def gscene0(ctx, scene0):
    s = mrt.Scene(ctx, scene0["verts"], scene0["mat"], scene0["alb"])
    yield s
    s.close()


@pytest.fixture(scope="module")
def gscene1(ctx, scene1):
    _say("building config-1 scene (1,024,000 triangles)")
    s = mrt.Scene(ctx, scene1["verts"], scene1["mat"], scene1["alb"])
    yield s
    s.close()


def test_primary_rays_config0_vs_bruteforce(ctx, orc, cfg, scene0, gscene0):
    rays = orc.primary_rays(cfg.camera(0))
    _say(f"config-0 primary rays n={rays.shape[0]}")
    got = _trace(ctx, gscene0, rays)
    ref = orc.closest_brute(scene0["verts"], rays)
    _assert_hits_equal(got, ref, "config0 primary")
    assert np.sum(ref["tri"] != mrt.MISS) > 40000


def test_random_rays_config0_vs_bruteforce(ctx, orc, cfg, scene0, gscene0):
    rays = orc.random_rays(cfg.CONFIG["random_ray_seed"], 10.0, 1 << 16)
    _say(f"config-0 random rays n={rays.shape[0]}")
    got = _trace(ctx, gscene0, rays)
    ref = orc.closest_brute(scene0["verts"], rays)
    _assert_hits_equal(got, ref, "config0 random")


def test_bounce1_rays_config1_vs_oracle_bvh(ctx, orc, cfg, scene1, gscene1):
    rays, _ = scene1["bvh"].bounce1_rays(cfg.camera(1), cfg.path_params(1))
    _say(f"config-1 bounce-1 rays n={rays.shape[0]}")
    got = _trace(ctx, gscene1, rays)
    ref = scene1["bvh"].trace_closest(rays)
    _assert_hits_equal(got, ref, "config1 bounce1")


def test_random_rays_config1_vs_oracle_bvh(ctx, orc, cfg, scene1, gscene1):
    rays = orc.random_rays(cfg.CONFIG["random_ray_seed"] + 1, 40.0, 1 << 18)
    _say(f"config-1 random rays n={rays.shape[0]}")
    got = _trace(ctx, gscene1, rays)
    ref = scene1["bvh"].trace_closest(rays)
    _assert_hits_equal(got, ref, "config1 random")


def test_any_hit_shadow_rays(ctx, orc, cfg, scene0, gscene0):
    rays = orc.random_rays(cfg.CONFIG["random_ray_seed"] + 2, 10.0, 1 << 15)
    rays[::3, 7] = 5.0  # a third of the rays have a finite tmax
    _say(f"config-0 any-hit n={rays.shape[0]}")
    d_rays = ctx.to_device(rays)
    d_occ = ctx.empty(rays.shape[0], np.uint32)
    gscene0.trace_any(d_rays, d_occ)
    got = d_occ.download()
    ref = scene0["bvh"].trace_any(rays)
    assert np.array_equal(got.astype(np.uint8), ref)


def test_edge_cases(ctx, orc, cfg, scene0, gscene0):
    base = orc.primary_rays(cfg.camera(0, width=16, height=16))
    rays = np.concatenate([base[:37], base[37:40]])  # not a multiple of 64
    rays[0, 4:7] = 0.0                                 # zero direction: invalid, must miss
    rays[1, 4:7] = [0.0, 1.0, 0.0]                     # axis-aligned (two zero components)
    rays[2, 7] = 1e-3                                  # tmax before anything
    rays[3, 3] = 1e9                                   # tmin beyond everything
    rays[4, 0:3] = [0.0, 0.0, 0.0]                     # origin inside the scene box
    _say("edge cases n=40")
    got = _trace(ctx, gscene0, rays)
    ref = orc.closest_brute(scene0["verts"], rays)
    _assert_hits_equal(got, ref, "edge cases")
    assert got["tri"][0] == mrt.MISS and got["tri"][2] == mrt.MISS and got["tri"][3] == mrt.MISS
    # n = 0 is a no-op
    d_rays = ctx.to_device(rays)
    d_hits = ctx.empty(1, mrt.HIT_DTYPE)
    gscene0.trace_closest(d_rays, d_hits, n=0)


def test_tiny_and_empty_scenes(ctx, orc):
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    s = mrt.Scene(ctx, tri, np.zeros(1, np.uint32), np.ones((1, 3), np.float32))
    rays = np.array([[0.25, 0.25, 1, 0, 0, 0, -1, np.inf], [2, 2, 1, 0, 0, 0, -1, np.inf]], np.float32)
    _say("one-triangle scene")
    got = _trace(ctx, s, rays)
    assert got["tri"][0] == 0 and got["t"][0] == 1.0 and got["u"][0] == 0.25 and got["v"][0] == 0.25
    assert got["tri"][1] == mrt.MISS and np.isinf(got["t"][1])
    s.close()
    e = mrt.Scene(ctx, np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), np.ones((1, 3), np.float32))
    _say("empty scene")
    got = _trace(ctx, e, rays)
    assert (got["tri"] == mrt.MISS).all()
    e.close()


def test_deterministic_repeat(ctx, orc, cfg, gscene0):
    rays = orc.random_rays(7, 10.0, 50000)
    _say("determinism n=50000 x2")
    a = _trace(ctx, gscene0, rays)
    b = _trace(ctx, gscene0, rays)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_profile_counts_launches(ctx, orc, cfg, gscene0):
    rays = orc.random_rays(9, 10.0, 4096)
    d_rays = ctx.to_device(rays)
    d_hits = ctx.empty(rays.shape[0], mrt.HIT_DTYPE)
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(3):
        gscene0.trace_closest(d_rays, d_hits)
    ms, n = ctx.profile_query("trace_closest")
    ctx.profile(False)
    assert n == 3 and ms > 0.0


def test_copy_kernel(ctx):
    x = np.arange(1 << 20, dtype=np.uint32)
    a = ctx.to_device(x)
    b = ctx.empty(x.shape, np.uint32)
    ctx.copy_device(b, a)
    assert np.array_equal(b.download(), x)
    y = np.arange(1001, dtype=np.uint8)
    c = ctx.to_device(y)
    d = ctx.empty(y.shape, np.uint8)
    ctx.copy_device(d, c)
    assert np.array_equal(d.download(), y)


def test_all_variants_identical(ctx, orc, cfg, scene0, scene1, gscene0, gscene1):
    """Every launch-shape variant of both kernels returns the oracle's results."""
    names = mrt.runtime.trace_variants()
    rays0 = orc.random_rays(21, 10.0, 20000)
    ref0 = orc.closest_brute(scene0["verts"], rays0)
    rays1, _ = scene1["bvh"].bounce1_rays(cfg.camera(1, width=480, height=270), cfg.path_params(1))
    ref1 = scene1["bvh"].trace_closest(rays1)
    occ_ref = scene1["bvh"].trace_any(rays1)
    try:
        for i, name in enumerate(names):
            _say(f"variant {i} {name}")
            ctx.set_trace_variant(i, False)
            ctx.set_trace_variant(i, True)
            _assert_hits_equal(_trace(ctx, gscene0, rays0), ref0, f"variant {name} config0")
            _assert_hits_equal(_trace(ctx, gscene1, rays1), ref1, f"variant {name} config1")
            d_rays = ctx.to_device(rays1)
            d_occ = ctx.empty(rays1.shape[0], np.uint32)
            gscene1.trace_any(d_rays, d_occ)
            assert np.array_equal(d_occ.download().astype(np.uint8), occ_ref), f"variant {name} any-hit"
    finally:
        ctx.set_trace_variant(0, False)
        ctx.set_trace_variant(0, True)
