"""Cache-terminated path tracing (paths.hip + cache.hip) against the oracle.

Bars:
  * image: bit-identical to the oracle's cache-terminated paths (oracle/oracle.cpp path_one with
    cache_mode 1) given the per-pixel cache values the GPU used;
  * those cache values: the oracle's encoding + MLP of the oracle's vertex samples, within the
    MLP forward tolerance (1e-3 relative L2; MFMA summation order);
  * training batch: samples and targets bit-identical to the oracle's plain paths
    (orc_path_suffix), path ids the frame's predicted subset of the paths that reach the vertex;
  * fallbacks: enable = 0, non-finite weights, an exceeded query budget and a loss spike give
    the plain image bit for bit.
"""
import numpy as np
import pytest

import minotert_amd as mrt
from minotert_amd import convergence as cv
from minotert_amd.cache import GRAD_WEIGHTS, TABLE, WEIGHTS
from minotert_amd.pathtracer import CachedPathtracer, CacheOptions
from oracle import oracle_cache as oc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-cached] {msg}", flush=True)


@pytest.fixture(scope="module")
def rig(ctx, scene1):
    s = mrt.Scene(ctx, scene1["verts"], scene1["mat"], scene1["alb"])
    bounds = cv.scene_bounds(scene1["verts"])
    cache = cv.make_cache(ctx, bounds)
    yield {"scene": s, "bounds": bounds, "cache": cache, "pt": mrt.Pathtracer(ctx, s),
           "cpt": CachedPathtracer(ctx, s, cache)}
    cache.close()
    s.close()


def _randomize(cache, seed):
    rng = np.random.default_rng(seed)
    W = rng.uniform(-0.25, 0.25, (6, 64, 64)).astype(np.float32)
    W[5, 3:] = 0
    W[5, :3] = np.abs(W[5, :3])  # mostly positive outputs
    T = rng.uniform(-0.5, 0.5, cache.table_floats()).astype(np.float32)
    cache.set(WEIGHTS, W)
    cache.set(TABLE, T)
    return W, T


def _grid(cache, bounds):
    o = cache.opts
    return oc.grid(o.levels, o.log2_table, o.base_res, o.max_res, bounds[0], bounds[1])


def _params(cfg, frame, vertex=1):
    p = cfg.path_params(3, frame=frame)
    p.cache_mode = 1
    p.cache_vertex = vertex
    return p


@pytest.mark.parametrize("vertex", [1, 0, 2])
def test_cached_image_matches_oracle(ctx, cfg, scene1, rig, vertex):
    cam = cfg.camera(3, width=240, height=136)
    W, T = _randomize(rig["cache"], 40 + vertex)
    p = _params(cfg, 7, vertex)
    _say(f"cached frame 240x136 vertex={vertex}, inference only")
    crgb = ctx.to_device(np.zeros((cam.width * cam.height, 3), np.float32))
    rgb = rig["cpt"].render(cam, cv.environment(), cv.frame_options(3, 7),
                            cache_opts=CacheOptions(vertex=vertex, train_samples=0), cache_rgb=crgb).download()
    cache_rgb = crgb.download()
    ref = scene1["bvh"].render_path(scene1["mat"], scene1["alb"], cam, p, cache_rgb=cache_rgb)
    bad = np.nonzero((rgb.view(np.uint32) != ref.view(np.uint32)).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} pixels differ, e.g. {bad[:5]}: {rgb[bad[:3]]} vs {ref[bad[:3]]}"
    # the cache values the GPU used: oracle encoding + MLP of the oracle's vertex samples
    pids = np.arange(cam.width * cam.height, dtype=np.uint32)
    smp, _, valid = scene1["bvh"].path_suffix(scene1["mat"], scene1["alb"], cam, p, pids)
    assert valid.any()
    from minotert_amd.cache import SAMPLE_DTYPE
    s = smp[valid].copy().view(SAMPLE_DTYPE).reshape(-1)
    x = oc.encode(_grid(rig["cache"], rig["bounds"]), T, s)
    y = np.maximum(oc.forward(W, x)[:, :3], 0)
    got = cache_rgb[valid]
    rel = np.linalg.norm(got - y) / np.linalg.norm(y)
    assert rel < 1e-3, rel
    assert np.all(cache_rgb[~valid] == 0)


def _selected(orc, valid, pids, frame_seed, sample_base, target):
    """The frame's training paths: among the paths that reach the cache vertex on a surface
    (count of them), those with pcg(pid ^ tseed) < target 2^32 / count (all when count <= target),
    tseed = pcg(frame_seed ^ pcg(sample_base ^ 0x7A1E))."""
    count = int(valid.sum())
    thr = 0xFFFFFFFF if count <= target else (target << 32) // count
    tseed = orc.pcg(frame_seed ^ orc.pcg(sample_base ^ 0x7A1E))
    h = np.array([orc.pcg(int(p) ^ tseed) for p in pids], np.uint64)
    return valid & ((thr == 0xFFFFFFFF) | (h < thr)), count


@pytest.mark.parametrize("target", [600, 100000])
def test_cached_training_batch_matches_oracle(ctx, orc, cfg, scene1, rig, target):
    """The training batch: exactly the predicted subset of the paths that reach vertex 1 on a
    surface (about `target` of them, or all when fewer), samples and targets bit-identical to the
    oracle's plain paths; the image is the oracle's given the cache values."""
    cam = cfg.camera(3, width=240, height=136)
    cache = rig["cache"]
    W, T = _randomize(cache, 50)
    steps0 = cache.stats()["steps"]
    p = _params(cfg, 9)
    npix = cam.width * cam.height
    co = CacheOptions(vertex=1, train_samples=target, train_steps=2)
    ex = rig["cpt"].export_buffers(npix)
    crgb = ctx.to_device(np.zeros((npix, 3), np.float32))
    _say(f"cached frame 240x136 with {target} training samples, 2 steps")
    opts = cv.frame_options(3, 9)
    rgb = rig["cpt"].render(cam, cv.environment(), opts, cache_opts=co, cache_rgb=crgb, export=ex).download()
    ref = scene1["bvh"].render_path(scene1["mat"], scene1["alb"], cam, p, cache_rgb=crgb.download())
    assert np.array_equal(rgb.view(np.uint32), ref.view(np.uint32))
    n = int(ex["count"].download()[0])
    pid = ex["pid"].download()[:n]
    samples = ex["samples"].download()[:n]
    targets = ex["targets"].download()[:n]
    order = np.argsort(pid)
    pid, samples, targets = pid[order], samples[order], targets[order]
    pids = np.arange(npix, dtype=np.uint32)
    smp_o, tgt_o, valid_o = scene1["bvh"].path_suffix(scene1["mat"], scene1["alb"], cam, p, pids)
    sel, count = _selected(orc, valid_o, pids, opts.frame_seed, opts.sample_base, target)
    assert np.array_equal(pid, pids[sel]), "training paths: the predicted subset of the vertex-1 surface paths"
    if count > target:
        assert abs(n - target) <= 5 * np.sqrt(target), (n, target, count)
    else:
        assert n == count
    got_s = np.ascontiguousarray(samples).view(np.float32).reshape(-1, 16)
    assert np.array_equal(got_s.view(np.uint32), smp_o[sel].view(np.uint32))
    assert np.array_equal(targets.view(np.uint32), tgt_o[sel].view(np.uint32))
    assert cache.stats()["steps"] == steps0 + 2
    assert not np.any(cache.get(GRAD_WEIGHTS))  # the steps consumed the gradients


def test_overlapped_training_same_state_as_sequential(ctx, cfg, rig):
    """overlap_training moves a frame's training steps to the start of the next cached frame with
    the same cache, on the second stream beside its traversal; that frame waits for them at the
    cache vertex, so it infers from the same parameters as without the option.
    Training is not bit-reproducible (the gradient sums use atomics, and Adam's normalized first
    steps turn last-bit gradient differences of near-zero entries into full steps), so two runs
    without the option already differ slightly after one trained frame; a large Adam epsilon here
    (1e-2: near-zero gradients then take near-zero steps) keeps that noise small. The check: frame 0
    is bit-identical with and without the option, and frame 1 with the option differs from frame 1
    without it about as much as two runs without it differ from each other (the noise level, from a
    second run), and far less than frame 1 differs from the frame the untrained cache renders (a
    frame that did not wait for the steps would infer from untrained parameters, or a mix). A read
    of the cache (stats) runs steps still left, so every cache reports every frame's steps."""
    cam = cfg.camera(3, width=240, height=136)
    env = cv.environment()
    caches = [cv.make_cache(ctx, rig["bounds"], seed=5, eps=1e-2) for _ in range(4)]
    try:
        imgs = []
        for c, overlap in zip(caches[:3], (True, False, False)):
            cpt = CachedPathtracer(ctx, rig["scene"], c)
            co = CacheOptions(vertex=1, train_samples=8192, train_steps=4, overlap_training=overlap)
            _say(f"two cached frames 240x136, overlap_training={overlap}")
            imgs.append([cpt.render(cam, env, cv.frame_options(3, 30 + f), cache_opts=co).download() for f in range(2)])
            assert c.stats()["steps"] == 8  # with the overlap, the second frame's steps run here
        untrained = CachedPathtracer(ctx, rig["scene"], caches[3]).render(
            cam, env, cv.frame_options(3, 31), cache_opts=CacheOptions(vertex=1, train_samples=0)).download()
        (a0, a1), (b0, b1), (c0, c1) = imgs
        assert np.array_equal(a0.view(np.uint32), b0.view(np.uint32))
        assert np.array_equal(b0.view(np.uint32), c0.view(np.uint32))
        d_ab = np.linalg.norm(a1 - b1) / np.linalg.norm(b1)     # with / without the option
        d_bc = np.linalg.norm(c1 - b1) / np.linalg.norm(b1)     # two runs without it: the noise
        d_bu = np.linalg.norm(b1 - untrained) / np.linalg.norm(b1)
        _say(f"frame 1: with/without overlap {d_ab:.2e}, two runs without {d_bc:.2e}, trained/untrained {d_bu:.2e}")
        assert d_ab < 0.2 * d_bu, (d_ab, d_bu)
        assert d_ab < 4.0 * d_bc + 2e-3, (d_ab, d_bc)
        assert not caches[0].stats()["nonfinite"]
    finally:
        for c in caches:
            c.close()


def test_overlapped_training_flushed_by_another_cache(ctx, cfg, rig):
    """Steps left by an overlapped frame run before a cached frame with another cache (which
    reuses the training batch's workspace), so neither cache loses a frame's training."""
    cam = cfg.camera(3, width=240, height=136)
    env = cv.environment()
    a, b = cv.make_cache(ctx, rig["bounds"], seed=6), cv.make_cache(ctx, rig["bounds"], seed=7)
    try:
        co = CacheOptions(vertex=1, train_samples=2048, train_steps=3, overlap_training=True)
        _say("overlapped frames alternating between two caches")
        for f in range(3):
            CachedPathtracer(ctx, rig["scene"], a).render(cam, env, cv.frame_options(3, 40 + f), cache_opts=co)
            CachedPathtracer(ctx, rig["scene"], b).render(cam, env, cv.frame_options(3, 50 + f), cache_opts=co)
        assert a.stats()["steps"] == 9 and b.stats()["steps"] == 9
        assert not a.stats()["nonfinite"] and not b.stats()["nonfinite"]
    finally:
        a.close()
        b.close()


def test_cached_training_steps_take_equal_shares(ctx, cfg, rig):
    """Sample j of the frame's batch goes to step j mod steps, so with n >= steps samples every step,
    the last one included, trains on a share (its loss sum is that share's, > 0). A batch smaller
    than the fixed per-step capacity used to fill only the first step."""
    cam = cfg.camera(3, width=240, height=136)
    cache = rig["cache"]
    _randomize(cache, 70)
    npix = cam.width * cam.height
    ex = rig["cpt"].export_buffers(npix)
    co = CacheOptions(vertex=1, train_samples=4096, train_steps=4)
    steps0 = cache.stats()["steps"]
    _say("cached frame 240x136, 4096 training paths, 4 steps")
    rig["cpt"].render(cam, cv.environment(), cv.frame_options(3, 21), cache_opts=co, export=ex)
    n = int(ex["count"].download()[0])
    st = cache.stats()
    assert st["steps"] == steps0 + 4
    # fewer than 4096 paths reach vertex 1 on a surface in this image: all of them train, spread
    # over the four steps (rows of the batch: 4096 + 4096 / 8 + 256)
    assert 4 <= n <= 4096 + 512 + 256, n
    assert st["loss_sum"] > 0.0, "the last step trained on no samples"


def test_cached_frame_steps_average_as_single_steps(ctx, cfg, rig):
    """A frame's training steps apply the table's running average in closed form (the first step
    moves every entry by a^S, each step adds (1 - a^(S-j+1)) times its own change). Replaying the
    frame's exported batch on a twin cache as four steps of their own (the per-step formula,
    pinned bit-exact to the oracle's Adam in test_gpu_cache.py) gives the same parameters and
    averages up to float rounding (tolerance 1e-6 of the largest magnitude; the gradient
    scatter's atomics also reorder additions). A wrong exponent in the closed form moves the
    averages by ~(1 - a) of a step's change, far outside it."""
    from minotert_amd.cache import AVG_TABLE, AVG_WEIGHTS
    cam = cfg.camera(3, width=240, height=136)
    ca = cv.make_cache(ctx, rig["bounds"], seed=11, ema_decay=0.9)
    cb = cv.make_cache(ctx, rig["bounds"], seed=11, ema_decay=0.9)
    try:
        assert np.array_equal(ca.get(TABLE), cb.get(TABLE))
        ex = rig["cpt"].export_buffers(cam.width * cam.height)
        steps = 4
        co = CacheOptions(vertex=1, train_samples=4096, train_steps=steps)
        _say("cached frame 240x136 on cache A (4 steps, closed-form average)")
        CachedPathtracer(ctx, rig["scene"], ca).render(cam, cv.environment(), cv.frame_options(3, 23), cache_opts=co,
                                                       export=ex)
        n = int(ex["count"].download()[0])
        assert n >= 4 * steps, n
        s = ex["samples"].download()[:n]
        t = ex["targets"].download()[:n]
        _say(f"replay of {n} samples on cache B as {steps} single steps")
        for i in range(steps):  # sample j trained in step j mod steps
            cb.train(ctx.to_device(np.ascontiguousarray(s[i::steps])), ctx.to_device(np.ascontiguousarray(t[i::steps])))
        assert ca.stats()["steps"] == cb.stats()["steps"] == steps
        assert not ca.stats()["nonfinite"]
        for which in (TABLE, AVG_TABLE, WEIGHTS, AVG_WEIGHTS):
            a, b = ca.get(which), cb.get(which)
            tol = 1e-6 * float(np.abs(b).max())
            err = float(np.abs(a - b).max())
            assert err <= tol, (which, err, tol)
        assert np.any(cb.get(AVG_TABLE) != 0)
    finally:
        ca.close()
        cb.close()


def test_cached_fallbacks_bit_identical_to_plain(ctx, cfg, rig):
    cam = cfg.camera(3, width=200, height=120)
    env = cv.environment()
    opts = cv.frame_options(3, 11)
    plain = rig["pt"].render(cam, env, opts).download()
    cache = rig["cache"]
    _randomize(cache, 60)
    _say("fallback: disabled cache")
    off = rig["cpt"].render(cam, env, opts, cache_opts=CacheOptions(enable=False)).download()
    assert np.array_equal(off.view(np.uint32), plain.view(np.uint32))
    on = rig["cpt"].render(cam, env, opts, cache_opts=CacheOptions(train_samples=0)).download()
    assert not np.array_equal(on.view(np.uint32), plain.view(np.uint32))
    _say("fallback: non-finite weights")
    W = cache.get(WEIGHTS)
    W[123] = np.nan
    cache.set(WEIGHTS, W)
    assert cache.stats()["nonfinite"]
    bad = rig["cpt"].render(cam, env, opts, cache_opts=CacheOptions(train_samples=512, train_steps=1)).download()
    assert np.array_equal(bad.view(np.uint32), plain.view(np.uint32))
    _randomize(cache, 61)
    assert not cache.stats()["nonfinite"]


def test_cache_budgets_fall_back_to_plain(ctx, cfg, rig):
    """The other forced fallbacks: a query budget the frame exceeds (0: no path may end into the
    cache; 1: fewer than reach the vertex) and a loss spike (the cache's last training step has a
    mean loss per sample above the budget). Each frame renders mrt_render's image bit for bit
    (paths continue to full depth); budgets the frame meets leave the cached image unchanged."""
    cam = cfg.camera(3, width=200, height=120)
    env = cv.environment()
    opts = cv.frame_options(3, 12)
    npix = cam.width * cam.height
    plain = rig["pt"].render(cam, env, opts).download()
    cache = rig["cache"]
    _randomize(cache, 62)

    def render(**kw):
        return rig["cpt"].render(cam, env, opts, cache_opts=CacheOptions(train_samples=0, **kw)).download()

    cached = render()
    assert not np.array_equal(cached.view(np.uint32), plain.view(np.uint32))
    for budget, want in ((0, plain), (1, plain), (npix, cached)):
        _say(f"query budget {budget}")
        assert np.array_equal(render(query_budget=budget).view(np.uint32), want.view(np.uint32)), budget
    _say("one training frame, then loss budgets")
    rig["cpt"].render(cam, env, cv.frame_options(3, 13), cache_opts=CacheOptions(train_samples=2048, train_steps=1))
    loss = cache.stats()["loss_sum"]
    assert loss > 0.0
    cached = render()  # the trained cache (inference-only frames leave it as it is)
    assert not np.array_equal(cached.view(np.uint32), plain.view(np.uint32))
    assert np.array_equal(render(loss_budget=1e-30).view(np.uint32), plain.view(np.uint32))
    assert np.array_equal(render(loss_budget=1e30).view(np.uint32), cached.view(np.uint32))
    both = render(loss_budget=1e30, query_budget=npix)
    assert np.array_equal(both.view(np.uint32), cached.view(np.uint32))
    _randomize(cache, 63)


def test_nonfinite_verdict_carried_through_a_frames_steps(ctx, cfg, rig):
    """A NaN in a table entry the encoding can read: the frame falls back to the plain image, and
    its training steps (four, with no samples) still run. The first step of the group reads every
    used entry and flags it; the other three do not read the entry and carry that verdict, so the
    flag is still set after the frame and the next frame falls back too. A NaN in a dense level's
    slot past its (res+1)^3 cells is never read: it neither sets the flag nor disables the cache."""
    cam = cfg.camera(3, width=200, height=120)
    env = cv.environment()
    co = CacheOptions(train_samples=2048, train_steps=4)
    cache = rig["cache"]
    o = cache.opts
    tsize = 1 << o.log2_table
    side0 = int(np.floor(o.base_res)) + 1
    assert side0 ** 3 < tsize  # level 0 is dense
    _, T = _randomize(cache, 71)
    plain = [rig["pt"].render(cam, env, cv.frame_options(3, f)).download() for f in (41, 42)]
    # unused slot of level 0: ignored
    T_bad = T.copy()
    T_bad[(side0 ** 3 + 100) * 2] = np.nan
    cache.set(TABLE, T_bad)
    assert not cache.stats()["nonfinite"]
    img = rig["cpt"].render(cam, env, cv.frame_options(3, 41), cache_opts=co).download()
    assert not np.array_equal(img.view(np.uint32), plain[0].view(np.uint32))
    assert not cache.stats()["nonfinite"]
    # used entry (the last cell of level 0): flagged, carried, both frames fall back
    _, T = _randomize(cache, 72)
    T_bad = T.copy()
    T_bad[(side0 ** 3 - 1) * 2 + 1] = np.nan
    cache.set(TABLE, T_bad)
    assert cache.stats()["nonfinite"]
    steps0 = cache.stats()["steps"]
    _say("non-finite table entry: two cached frames of four steps")
    a = rig["cpt"].render(cam, env, cv.frame_options(3, 41), cache_opts=co).download()
    assert np.array_equal(a.view(np.uint32), plain[0].view(np.uint32))
    st = cache.stats()
    assert st["steps"] == steps0 + 4 and st["nonfinite"]  # the fourth step's flag, carried from the first
    b = rig["cpt"].render(cam, env, cv.frame_options(3, 42), cache_opts=co).download()
    assert np.array_equal(b.view(np.uint32), plain[1].view(np.uint32))
    assert cache.stats()["nonfinite"]
    _randomize(cache, 73)
    assert not cache.stats()["nonfinite"]


def test_accumulate_and_rel_mse(ctx, orc):
    rng = np.random.default_rng(3)
    n = 70001
    ref = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    frames = [rng.uniform(0, 2, (n, 3)).astype(np.float32) for _ in range(3)]
    acc = ctx.to_device(np.zeros((n, 3), np.float32))
    mean = ctx.empty((n, 3), np.float32)
    for i, f in enumerate(frames):
        mrt.pathtracer.accumulate(ctx, acc, ctx.to_device(f), i + 1, mean)
    s = (frames[0] + frames[1]) + frames[2]
    assert np.array_equal(acc.download(), s)
    assert np.array_equal(mean.download(), s / np.float32(3))
    d_ref = ctx.to_device(ref)
    r = mrt.pathtracer.rel_mse(ctx, mean, d_ref, 1e-2)
    want = orc.rel_mse(s / np.float32(3), ref, 1e-2)
    assert abs(r - want) <= 1e-12 * max(1.0, want)
    assert mrt.pathtracer.rel_mse(ctx, d_ref, d_ref, 1e-2) == 0.0
