"""Edge cases of the C ABI on the GPU: zero-size calls are no-ops that leave the outputs
untouched, an empty scene misses every ray, a one-triangle scene (host or device BVH) matches
brute force, rays with zero or subnormal direction components match brute force, a 1x1 frame
matches the oracle, and tiny sorts are stable."""
import numpy as np
import pytest

import minotert_amd as mrt
from minotert_amd import runtime as rtm

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-edges] {msg}", flush=True)


def _rays(n, seed, box=3.0):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, mrt.RAY_DTYPE)
    r["o"] = rng.uniform(-box, box, (n, 3))
    d = rng.normal(size=(n, 3))
    r["d"] = d / np.linalg.norm(d, axis=1)[:, None]
    r["tmin"] = 0.0
    r["tmax"] = np.inf
    return r


def test_zero_sized_calls_leave_outputs_untouched(ctx, scene0):
    _say("zero-size trace, order, sort, accumulate, relMSE")
    s = mrt.Scene(ctx, scene0["verts"], scene0["mat"], scene0["alb"])
    try:
        rays = ctx.to_device(_rays(4, 1))
        sentinel = np.full(4, 0xA5A5A5A5, np.uint32).view(np.float32)
        hits_host = np.zeros(4, mrt.HIT_DTYPE)
        hits_host["t"] = sentinel
        hits = ctx.to_device(hits_host)
        s.trace_closest(rays, hits, n=0)
        occ = ctx.to_device(np.full(4, 7, np.uint32))
        s.trace_any(rays, occ, n=0)
        order = ctx.to_device(np.full(4, 9, np.uint32))
        rtm.ray_order(s, rays, order, n=0)
        rtm.trace_closest_ordered(s, rays, order, hits, n=0)
        keys, vals = ctx.to_device(np.arange(4, dtype=np.uint32)[::-1].copy()), ctx.to_device(np.arange(4, dtype=np.uint32))
        rtm.sort_pairs(ctx, keys, vals, n=0)
        ctx.sync()
        assert np.array_equal(hits.download()["t"].view(np.uint32), sentinel.view(np.uint32))
        assert np.all(occ.download() == 7) and np.all(order.download() == 9)
        assert np.array_equal(keys.download(), [3, 2, 1, 0])
        from minotert_amd.pathtracer import accumulate, rel_mse
        acc = ctx.to_device(np.ones((2, 3), np.float32))
        accumulate(ctx, acc, ctx.empty((0, 3), np.float32), 1)
        assert np.all(acc.download() == 1.0)
        assert rel_mse(ctx, ctx.empty((0, 3), np.float32), ctx.empty((0, 3), np.float32)) == 0.0
    finally:
        s.close()


def test_empty_scene_misses_everything(ctx, orc):
    _say("empty scene")
    s = mrt.Scene(ctx, np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), np.zeros((1, 3), np.float32))
    try:
        assert s.stats()["n_tris"] == 0
        r = _rays(1000, 2)
        hits = ctx.empty(1000, mrt.HIT_DTYPE)
        s.trace_closest(ctx.to_device(r), hits)
        h = hits.download()
        assert np.all(h["tri"] == mrt.MISS) and np.all(np.isinf(h["t"]))
        occ = ctx.to_device(np.full(1000, 5, np.uint32))
        s.trace_any(ctx.to_device(r), occ)
        assert np.all(occ.download() == 0)
    finally:
        s.close()


@pytest.mark.parametrize("megakernel", [False, True])
def test_empty_scene_frame_is_the_sky(ctx, cfg, megakernel):
    """Every path misses at vertex 0: the queues after it are empty (zero-count traversal and
    megakernel launches) and every pixel is the sky radiance exactly."""
    _say(f"empty-scene frame, megakernel={megakernel}")
    s = mrt.Scene(ctx, np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), np.zeros((1, 3), np.float32))
    try:
        sky = (0.25, 0.5, 1.0)
        env = mrt.Environment(sky=sky, sun_dir=tuple(cfg.CONFIG["sun_dir"]), sun_e=tuple(cfg.CONFIG["sun_e"]))
        opts = mrt.RenderOptions(frame_seed=3, spp=2, bounces=4, rr_start=3, megakernel=megakernel)
        img = mrt.Pathtracer(ctx, s).render(cfg.camera(0, width=64, height=48), env, opts).download()
    finally:
        s.close()
    assert np.array_equal(img.reshape(-1, 3), np.tile(np.array(sky, np.float32), (64 * 48, 1)))


@pytest.mark.parametrize("builder", ["host", "device"])
@pytest.mark.parametrize("ntri", [1, 2, 5])
def test_tiny_scenes_match_brute_force(ctx, orc, builder, ntri):
    _say(f"{ntri}-triangle scene, {builder} BVH")
    rng = np.random.default_rng(ntri)
    c = rng.uniform(-1, 1, (ntri, 1, 3))
    v = (c + rng.normal(0, 0.8, (ntri, 3, 3))).astype(np.float32).reshape(ntri, 9)
    s = mrt.Scene(ctx, v, np.zeros(ntri, np.uint32), np.full((1, 3), 0.5, np.float32), builder=builder)
    try:
        r = _rays(20000, 3 + ntri, box=2.0)
        hits = ctx.empty(r.shape[0], mrt.HIT_DTYPE)
        s.trace_closest(ctx.to_device(r), hits)
        got = hits.download()
        ref = orc.closest_brute(v, r.view(np.float32).reshape(-1, 8))
        assert np.array_equal(got["tri"], ref["tri"])
        assert np.array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32))
        assert np.any(got["tri"] != mrt.MISS)
    finally:
        s.close()


def test_axis_parallel_and_subnormal_directions_match_brute_force(ctx, orc, scene0):
    """Direction components that are zero (either sign) or so small that 1/d overflows: the box
    test must stay conservative (a slab whose plane distances cannot be formed must not cull), so
    closest hits, any hits and t equal the oracle's brute force on the 20,482-triangle scene,
    for every traversal variant and through the ordered entry point."""
    rng = np.random.default_rng(77)
    n = 30000
    r = _rays(n, 78, box=20.0)
    d = r["d"].copy()
    tiny = np.array([0.0, -0.0, 1e-39, -1e-39, 1e-45, -1e-45, 3e-39, 1e-38, -1.2e-38], np.float32)
    for axes in (1, 2):  # one or two tiny components per ray, the rest renormalized
        sel = np.arange(n)[(np.arange(n) % 2) == axes - 1]
        for i in sel:
            ks = rng.choice(3, axes, replace=False)
            d[i, ks] = tiny[rng.integers(0, len(tiny), axes)]
            rest = [k for k in range(3) if k not in ks]
            d[i, rest] /= np.linalg.norm(d[i, rest])
    r["d"] = d
    ref = orc.closest_brute(scene0["verts"], r.view(np.float32).reshape(-1, 8))
    assert np.count_nonzero(ref["tri"] != mrt.MISS) > 1000
    s = mrt.Scene(ctx, scene0["verts"], scene0["mat"], scene0["alb"])
    before = (ctx.trace_variant(False), ctx.trace_variant(True))
    try:
        d_rays = ctx.to_device(r)
        hits = ctx.empty(n, mrt.HIT_DTYPE)
        occ = ctx.empty(n, np.uint32)
        for v, name in enumerate(mrt.runtime.trace_variants()):
            _say(f"zero and subnormal direction components, variant {name}")
            ctx.set_trace_variant(v)
            ctx.set_trace_variant(v, True)
            s.trace_closest(d_rays, hits)
            got = hits.download()
            assert np.array_equal(got["tri"], ref["tri"]), name
            assert np.array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32)), name
            s.trace_any(d_rays, occ)
            assert np.array_equal(occ.download() != 0, ref["tri"] != mrt.MISS), name
        order = ctx.empty(n, np.uint32)
        rtm.ray_order(s, d_rays, order)
        rtm.trace_closest_ordered(s, d_rays, order, hits)
        assert np.array_equal(hits.download()["tri"], ref["tri"])
    finally:
        ctx.set_trace_variant(before[0], False)
        ctx.set_trace_variant(before[1], True)
        s.close()


def test_non_finite_and_inverted_rays_match_brute_force(ctx, orc, scene0):
    """Rays with NaN or infinite components, a zero direction, an empty or a single-point
    [tmin, tmax], a negative tmin or a zero tmax: the traversal accepts any floats and must give
    the brute force's answer (no hit where its comparisons fail on a NaN), for every variant and
    through the ordered entry point (the sort key clamps a NaN origin into the grid)."""
    rng = np.random.default_rng(91)
    n = 24000
    r = _rays(n, 92, box=20.0)
    ref0 = orc.closest_brute(scene0["verts"], r.view(np.float32).reshape(-1, 8))
    hit = np.flatnonzero(ref0["tri"] != mrt.MISS)
    orig = r.view(np.float32).reshape(-1, 8).copy()
    f = r.view(np.float32).reshape(-1, 8)  # o.xyz, tmin, d.xyz, tmax (a view of r)
    kinds = 12
    for i in range(n):
        k = i % kinds
        if k == 0:
            f[i, rng.integers(0, 3)] = np.nan            # origin component
        elif k == 1:
            f[i, 4 + rng.integers(0, 3)] = np.nan        # direction component
        elif k == 2:
            f[i, 3] = np.nan                             # tmin
        elif k == 3:
            f[i, 7] = np.nan                             # tmax
        elif k == 4:
            f[i, rng.integers(0, 3)] = rng.choice([np.inf, -np.inf])
        elif k == 5:
            f[i, 4 + rng.integers(0, 3)] = rng.choice([np.inf, -np.inf])
        elif k == 6:
            f[i, 4:7] = 0.0                              # no direction
        elif k == 7:
            f[i, 3], f[i, 7] = 5.0, 1.0                  # tmin > tmax
        elif k == 8:
            f[i, 3] = -np.inf                            # hits behind the origin count
        elif k == 9:
            f[i, 7] = 0.0
        elif k == 10 and len(hit):                       # [t, t] around a known hit
            j = hit[rng.integers(0, len(hit))]
            f[i] = orig[j]
            f[i, 3] = f[i, 7] = ref0["t"][j]
        # k == 11: unchanged
    ref = orc.closest_brute(scene0["verts"], f)
    assert np.count_nonzero(ref["tri"] != mrt.MISS) > 200
    s = mrt.Scene(ctx, scene0["verts"], scene0["mat"], scene0["alb"])
    before = (ctx.trace_variant(False), ctx.trace_variant(True))
    try:
        d_rays = ctx.to_device(r)
        hits = ctx.empty(n, mrt.HIT_DTYPE)
        occ = ctx.empty(n, np.uint32)
        for v, name in enumerate(mrt.runtime.trace_variants()):
            _say(f"non-finite and inverted rays, variant {name}")
            ctx.set_trace_variant(v)
            ctx.set_trace_variant(v, True)
            s.trace_closest(d_rays, hits)
            got = hits.download()
            assert np.array_equal(got["tri"], ref["tri"]), name
            assert np.array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32)), name
            s.trace_any(d_rays, occ)
            assert np.array_equal(occ.download() != 0, ref["tri"] != mrt.MISS), name
        order = ctx.empty(n, np.uint32)
        rtm.ray_order(s, d_rays, order)
        assert np.array_equal(np.sort(order.download()), np.arange(n, dtype=np.uint32))
        rtm.trace_closest_ordered(s, d_rays, order, hits)
        assert np.array_equal(hits.download()["tri"], ref["tri"])
    finally:
        ctx.set_trace_variant(before[0], False)
        ctx.set_trace_variant(before[1], True)
        s.close()


def test_one_pixel_frame_matches_oracle(ctx, orc, cfg, scene0):
    _say("1x1 path-traced frame")
    cam = cfg.camera(0, width=1, height=1)
    s = mrt.Scene(ctx, scene0["verts"], scene0["mat"], scene0["alb"])
    try:
        env = mrt.Environment(sky=tuple(cfg.CONFIG["sky"]), sun_dir=tuple(cfg.CONFIG["sun_dir"]),
                              sun_e=tuple(cfg.CONFIG["sun_e"]))
        opts = mrt.RenderOptions(frame_seed=0x77, sample_base=0, spp=1, bounces=4, rr_start=3)
        got = mrt.Pathtracer(ctx, s).render(cam, env, opts).download()
    finally:
        s.close()
    p = cfg.path_params(1)
    p.frame_seed, p.sample, p.bounces, p.rr_start = 0x77, 0, 4, 3
    p.sky, p.sun_dir, p.sun_e = tuple(cfg.CONFIG["sky"]), tuple(cfg.CONFIG["sun_dir"]), tuple(cfg.CONFIG["sun_e"])
    ref = scene0["bvh"].render_path(scene0["mat"], scene0["alb"], cam, p)
    assert np.array_equal(got.reshape(-1).view(np.uint32), ref.reshape(-1).view(np.uint32))


@pytest.mark.parametrize("n", [1, 2, 3, 65])
def test_tiny_sorts_are_stable(ctx, n):
    rng = np.random.default_rng(n)
    k = rng.integers(0, 4, n, dtype=np.uint32)
    v = np.arange(n, dtype=np.uint32)
    dk, dv = ctx.to_device(k), ctx.to_device(v)
    rtm.sort_pairs(ctx, dk, dv, bits=2)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(dk.download(), k[order]) and np.array_equal(dv.download(), v[order])


def test_render_options_out_of_range_are_refused(ctx, cfg):
    """mrt_render refuses options outside its limits before any GPU work, naming the limit: spp
    outside 1..4096, more than 2^28 paths (the count printed without 32-bit wrap), more than 60
    bounces, roulette before bounce 2, and a viewport over 2^28 pixels."""
    s = mrt.Scene(ctx, np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), np.zeros((1, 3), np.float32))
    pt = mrt.Pathtracer(ctx, s)
    env = mrt.Environment(sky=(0.25, 0.5, 1.0), sun_dir=tuple(cfg.CONFIG["sun_dir"]), sun_e=(0.0, 0.0, 0.0))
    out = ctx.empty((1, 3), np.float32)  # never written: every call below fails its checks first
    try:
        small, big = cfg.camera(0, width=8, height=8), cfg.camera(0, width=8192, height=8192)
        cases = [(small, dict(spp=0), "spp 0 not in 1..4096"),
                 (small, dict(spp=4097), "spp 4097 not in 1..4096"),
                 (big, dict(spp=5), "335544320 paths exceed 2^28"),
                 (big, dict(spp=64), "4294967296 paths exceed 2^28"),
                 (small, dict(bounces=61), "bounces 61 > 60"),
                 (small, dict(rr_start=1), "rr_start must be >= 2"),
                 (cfg.camera(0, width=16384, height=16385), dict(spp=1), "viewport 16384x16385 out of range")]
        for cam, kw, msg in cases:
            _say(f"refused options {kw} at {cam.width}x{cam.height}")
            with pytest.raises(mrt.MrtError, match=msg.replace("^", r"\^").replace(".", r"\.")):
                pt.render(cam, env, mrt.RenderOptions(**kw), out=out)
        # the same context still renders afterwards
        img = pt.render(small, env, mrt.RenderOptions(spp=1, bounces=2, rr_start=2)).download()
        assert np.array_equal(img.reshape(-1, 3), np.tile(np.array([0.25, 0.5, 1.0], np.float32), (64, 1)))
    finally:
        s.close()


def test_profile_region_times_a_stream_range(ctx, scene0):
    """mrt_profile_region: one event pair around several calls, reported as one launch whose time
    covers the kernels inside it (per-launch profiling off); a region opened twice or closed
    without being opened is refused."""
    s = mrt.Scene(ctx, scene0["verts"], scene0["mat"], scene0["alb"])
    try:
        rays = ctx.to_device(_rays(200_000, 5, box=12.0))
        hits = ctx.empty(200_000, mrt.HIT_DTYPE)
        s.trace_closest(rays, hits)
        ctx.sync()
        ctx.profile_reset()
        _say("profile region around 4 traces, then per-launch events of 4 more")
        ctx.profile_region("four_traces", True)
        for _ in range(4):
            s.trace_closest(rays, hits)
        ctx.profile_region("four_traces", False)
        ctx.profile(True)
        for _ in range(4):
            s.trace_closest(rays, hits)
        ctx.profile(False)
        region_ms, region_n = ctx.profile_query("four_traces")
        kernel_ms, kernel_n = ctx.profile_query("trace_closest")
        assert region_n == 1 and kernel_n == 4
        assert 0.0 < kernel_ms and 0.5 * kernel_ms < region_ms < 20.0 * kernel_ms
        with pytest.raises(mrt.MrtError, match="not open"):
            ctx.profile_region("never_opened", False)
        ctx.profile_region("twice", True)
        with pytest.raises(mrt.MrtError, match="already open"):
            ctx.profile_region("twice", True)
        ctx.profile_reset()  # drops the open region
        ctx.profile_region("twice", True)
        ctx.profile_region("twice", False)
        assert ctx.profile_query("twice")[1] == 1
    finally:
        s.close()
