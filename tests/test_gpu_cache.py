"""Radiance cache on the GPU against the oracle (oracle/oracle_cache.cpp).

Bars (written per check below):
  * encoding: bit-exact (same FP32 operations in the same order, BF16 round-to-nearest-even);
  * forward: the MFMA sums 64 FP32 products in its own order and a BF16 rounding of a hidden
    activation can then land one ulp (2^-8 relative) apart, so outputs agree to 1e-3 in
    relative L2 norm and 1e-2 of the largest output per element;
  * gradients: the same reasons plus atomic accumulation order: 2e-3 relative L2 per layer /
    per table;
  * Adam: bit-exact given the same gradient (the update is elementwise FP32, IEEE div/sqrt).
"""
import numpy as np
import pytest

from oracle import oracle_cache as oc

pytestmark = pytest.mark.gpu

LO, HI = (-6.0, -6.0, -1.0), (6.0, 6.0, 5.0)
SMALL = dict(levels=8, log2_table=14, base_res=4.0, max_res=256.0)   # dense and hashed levels
FULL = dict(levels=16, log2_table=19, base_res=16.0, max_res=2048.0)  # the production grid


def _samples(n, seed):
    from minotert_amd.cache import SAMPLE_DTYPE
    rng = np.random.default_rng(seed)
    s = np.zeros(n, SAMPLE_DTYPE)
    lo, hi = np.array(LO), np.array(HI)
    ext = hi - lo
    s["pos"] = rng.uniform(lo - 0.05 * ext, hi + 0.05 * ext, (n, 3))  # a few outside: clamped
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
    s["dir"] = d
    m = rng.normal(size=(n, 3)); m /= np.linalg.norm(m, axis=1)[:, None]
    s["nrm"] = m
    s["alb"] = rng.uniform(0.05, 0.95, (n, 3))
    return s


def _grid(p):
    return oc.grid(p["levels"], p["log2_table"], p["base_res"], p["max_res"], LO, HI)


def _make(ctx, p, seed=3):
    from minotert_amd.cache import RadianceCache
    return RadianceCache(ctx, LO, HI, seed=seed, **p)


def _used_entries(p):
    """Mask over the table's floats of the entries Adam updates: all 2^T of a hashed level, the
    first (res+1)^3 of a dense one, the ones the encoding can read, rounded up to an even count
    (Adam works on groups of two entries; res = floor(base * b^l), b = (max/base)^(1/(L-1)))."""
    import math
    levels, tsize = p["levels"], 1 << p["log2_table"]
    b = math.exp((math.log(p["max_res"]) - math.log(p["base_res"])) / (levels - 1)) if levels > 1 else 1.0
    used = np.zeros((levels, tsize, 2), bool)
    for l in range(levels):
        side = int(math.floor(p["base_res"] * math.pow(b, l))) + 1
        used[l, :((side ** 3 + 1) // 2 * 2 if side ** 3 <= tsize else tsize)] = True
    return used.reshape(-1)


def _randomize(cache, p, seed):
    """Random weights (rows >= 3 of the last layer zero) and a random table."""
    from minotert_amd.cache import TABLE, WEIGHTS
    rng = np.random.default_rng(seed)
    W = rng.uniform(-0.3, 0.3, (6, 64, 64)).astype(np.float32)
    W[5, 3:] = 0
    T = rng.uniform(-0.5, 0.5, cache.table_floats()).astype(np.float32)
    cache.set(WEIGHTS, W)
    cache.set(TABLE, T)
    return W, T


# Batches below 1024 rows per CU use one thread per (query, level), larger ones one thread per
# query: 300,000 rows reach the second form on a 256-CU device.
@pytest.mark.parametrize("params,n", [(SMALL, 1000), (SMALL, 1), (FULL, 3000), (FULL, 300000)])
def test_cache_inputs_bit_exact(ctx, params, n):
    cache = _make(ctx, params)
    W, T = _randomize(cache, params, 11)
    s = _samples(n, 5)
    ds = ctx.to_device(s)
    out = ctx.empty((n, 64), np.float32)
    cache.inputs(ds, out)
    got = out.download()
    want = oc.encode(_grid(params), T, s)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
        f"{np.count_nonzero(got != want)} of {got.size} inputs differ"
    cache.close()


def test_cache_forward_matches_oracle(ctx):
    cache = _make(ctx, SMALL)
    W, T = _randomize(cache, SMALL, 12)
    n = 2000
    s = _samples(n, 6)
    ds = ctx.to_device(s)
    out = ctx.empty((n, 4), np.float32)
    cache.query(ds, out)
    y = out.download()
    x = oc.encode(_grid(SMALL), T, s)
    yo = oc.forward(W, x)
    assert np.all(y[:, 3] == 0)
    rel = np.linalg.norm(y[:, :3] - yo[:, :3]) / np.linalg.norm(yo[:, :3])
    assert rel < 1e-3, rel
    assert np.max(np.abs(y[:, :3] - yo[:, :3])) <= 1e-2 * np.max(np.abs(yo[:, :3])), rel
    cache.close()


def _targets(s):
    p = s["pos"]
    t = np.zeros((s.shape[0], 4), np.float32)
    t[:, 0] = 0.6 + 0.4 * np.sin(1.3 * p[:, 0]) * np.cos(0.7 * p[:, 1])
    t[:, 1] = 0.3 + 0.25 * np.cos(0.9 * p[:, 2] + p[:, 0])
    t[:, 2] = 0.1 + 0.05 * s["alb"][:, 2] + 0.2 * np.maximum(s["nrm"][:, 2], 0)
    return t


@pytest.mark.parametrize("params,n", [(SMALL, 1500), (FULL, 4000)])
def test_cache_gradients_match_oracle(ctx, params, n):
    """On the reduced grid and on the production grid (16 levels x 2^19 entries, mostly hashed)."""
    from minotert_amd.cache import GRAD_TABLE, GRAD_WEIGHTS
    cache = _make(ctx, params)
    W, T = _randomize(cache, params, 13)
    s = _samples(n, 7)
    t = _targets(s)
    cache.train(ctx.to_device(s), ctx.to_device(t), apply=False)
    gW = cache.get(GRAD_WEIGHTS).reshape(6, 64, 64)
    gT = cache.get(GRAD_TABLE)
    loss_sum = cache.stats()["loss_sum"]
    x = oc.encode(_grid(params), T, s)
    loss_o, gW_o, dx0 = oc.backward(W, x, t)
    gW_o = gW_o.reshape(6, 64, 64)
    gT_o = oc.encode_backward(_grid(params), s, dx0, T.size)
    assert abs(loss_sum / n - loss_o) <= 1e-4 * loss_o, (loss_sum / n, loss_o)
    for L in range(6):
        rel = np.linalg.norm(gW[L] - gW_o[L]) / np.linalg.norm(gW_o[L])
        assert rel < 2e-3, (L, rel)
    assert np.all(gW[5, 3:] == 0)  # padding rows of the output layer get no gradient
    rel_t = np.linalg.norm(gT - gT_o) / np.linalg.norm(gT_o)
    assert rel_t < 2e-3, rel_t
    assert np.array_equal(gT == 0, gT_o == 0) or np.count_nonzero((gT == 0) != (gT_o == 0)) < 1e-3 * gT.size
    # a second pass over the same batch accumulates: the gradient doubles
    cache.train(ctx.to_device(s), ctx.to_device(t), apply=False)
    gW2 = cache.get(GRAD_WEIGHTS).reshape(6, 64, 64)
    assert np.linalg.norm(gW2 - 2 * gW) <= 1e-5 * np.linalg.norm(gW2)
    cache.close()


@pytest.mark.parametrize("params,table_lr", [(SMALL, 0.0), (SMALL, 3e-3), (FULL, 0.0)])
def test_cache_adam_step_bit_exact(ctx, params, table_lr):
    """Adam on the weights (step lr) and on the table (step table_lr, or lr when 0), on the
    reduced and on the production grid."""
    from minotert_amd.cache import GRAD_TABLE, GRAD_WEIGHTS, TABLE, WEIGHTS, RadianceCache
    cache = RadianceCache(ctx, LO, HI, seed=3, table_lr=table_lr, **params)
    W, T = _randomize(cache, params, 14)
    s = _samples(800, 8)
    t = _targets(s)
    o = cache.opts
    m_w = np.zeros(W.size, np.float32); v_w = np.zeros(W.size, np.float32)
    m_t = np.zeros(T.size, np.float32); v_t = np.zeros(T.size, np.float32)
    Wc, Tc = W.reshape(-1).copy(), T.copy()
    for step in (1, 2, 3):
        cache.train(ctx.to_device(s), ctx.to_device(t), apply=False)
        gW, gT = cache.get(GRAD_WEIGHTS), cache.get(GRAD_TABLE)
        Wc, Tc = cache.get(WEIGHTS), cache.get(TABLE)
        cache.train(ctx.to_device(s[:0]), ctx.to_device(t[:0]), n=0, apply=True)
        Wn, m_w, v_w = oc.adam(Wc, gW, m_w, v_w, step, o.lr, o.adam_beta1, o.adam_beta2, o.adam_eps)
        Tn, m_t, v_t = oc.adam(Tc, gT, m_t, v_t, step, o.table_lr or o.lr, o.adam_beta1, o.adam_beta2, o.adam_eps,
                               sparse=True)
        assert np.array_equal(cache.get(WEIGHTS).view(np.uint32), Wn.view(np.uint32)), step
        assert np.array_equal(cache.get(TABLE).view(np.uint32), Tn.view(np.uint32)), step
        assert not np.any(cache.get(GRAD_WEIGHTS)) and not np.any(cache.get(GRAD_TABLE))
    assert cache.stats()["steps"] == 3
    cache.close()


def test_cache_running_average_for_queries(ctx):
    """ema_decay > 0: the averages follow ema = a ema + (1 - a) p after every step (bit-exact) on
    every table entry the encoding can read, and queries use the bias-corrected averages
    (forward tolerance)."""
    from minotert_amd.cache import AVG_TABLE, AVG_WEIGHTS, GRAD_TABLE, GRAD_WEIGHTS, TABLE, WEIGHTS, RadianceCache
    a = 0.9
    cache = RadianceCache(ctx, LO, HI, seed=5, ema_decay=a, **SMALL)
    W, T = _randomize(cache, SMALL, 15)
    s = _samples(600, 9)
    t = _targets(s)
    o = cache.opts
    eW = np.zeros(W.size, np.float32); eT = np.zeros(T.size, np.float32)
    m_w = np.zeros(W.size, np.float32); v_w = np.zeros(W.size, np.float32)
    m_t = np.zeros(T.size, np.float32); v_t = np.zeros(T.size, np.float32)
    fa, fb = np.float32(a), np.float32(1.0) - np.float32(a)
    used = _used_entries(SMALL)
    assert not used.all()  # the small grid has dense levels with unused slots
    for step in (1, 2, 3):
        cache.train(ctx.to_device(s), ctx.to_device(t), apply=False)
        gW, gT = cache.get(GRAD_WEIGHTS), cache.get(GRAD_TABLE)
        Wc, Tc = cache.get(WEIGHTS), cache.get(TABLE)
        cache.train(ctx.to_device(s[:0]), ctx.to_device(t[:0]), n=0, apply=True)
        Wn, m_w, v_w = oc.adam(Wc, gW, m_w, v_w, step, o.lr, o.adam_beta1, o.adam_beta2, o.adam_eps)
        Tn, m_t, v_t = oc.adam(Tc, gT, m_t, v_t, step, o.lr, o.adam_beta1, o.adam_beta2, o.adam_eps, sparse=True)
        eW = fa * eW + fb * Wn
        eT = np.where(used, fa * eT + fb * Tn, eT)  # slots no encoding reads are not averaged
        assert np.array_equal(cache.get(AVG_WEIGHTS).view(np.uint32), eW.view(np.uint32)), step
        assert np.array_equal(cache.get(AVG_TABLE).view(np.uint32), eT.view(np.uint32)), step
    scale = np.float32(1.0 / (1.0 - a ** 3))
    q = _samples(1000, 10)
    out = ctx.empty((1000, 4), np.float32)
    cache.query(ctx.to_device(q), out)
    y = out.download()[:, :3]
    yo = oc.forward(eW.reshape(6, 64, 64) * scale, oc.encode(_grid(SMALL), eT, q, float(scale)))[:, :3]
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    assert rel < 1e-3, rel
    cache.close()


def test_cache_training_fits_a_smooth_field(ctx):
    """Online training on a smooth radiance field: the relative loss falls by 10x and the fit
    on held-out samples reaches relMSE < 0.02."""
    cache = _make(ctx, SMALL, seed=21)
    rng_seeds = iter(range(1000, 2000))
    losses = []
    n = 8192
    for it in range(300):
        s = _samples(n, next(rng_seeds))
        cache.train(ctx.to_device(s), ctx.to_device(_targets(s)))
        if it % 50 == 0 or it == 299:
            losses.append(cache.stats()["loss_sum"] / n)
    st = cache.stats()
    assert not st["nonfinite"] and st["steps"] == 300
    assert losses[-1] < 0.1 * losses[0], losses
    s = _samples(4096, 77)
    out = ctx.empty((4096, 4), np.float32)
    cache.query(ctx.to_device(s), out)
    y = out.download()[:, :3]
    t = _targets(s)[:, :3]
    rel = np.mean((y - t) ** 2 / (t * t + 1e-2))
    assert rel < 0.02, (rel, losses)
    cache.close()


@pytest.mark.parametrize("which", ["table", "weights"])
def test_cache_nonfinite_flag_after_steps(ctx, which):
    """The non-finite flag: from a full check after parameters are set from the host, and from the
    Adam step itself after a step (an untouched table entry or a weight that is NaN stays NaN)."""
    from minotert_amd.cache import TABLE, WEIGHTS
    cache = _make(ctx, SMALL)
    W, T = _randomize(cache, SMALL, 21)
    s = _samples(256, 22)
    t = _targets(s)
    ds, dt = ctx.to_device(s), ctx.to_device(t)
    cache.train(ds, dt)
    assert cache.stats()["nonfinite"] is False  # flag of the step
    if which == "table":
        T = T.copy()
        T[-2] = np.nan  # last entry pair of the finest level: no sample of this batch touches it
        cache.set(TABLE, T)
    else:
        W = W.copy()
        W[5, 2, 63] = np.inf
        cache.set(WEIGHTS, W)
    assert cache.stats()["nonfinite"] is True   # full check after the host update
    cache.train(ds, dt)
    assert cache.stats()["nonfinite"] is True   # flag of the step that kept or produced it
    cache.close()


def test_cache_rejects_bad_options(ctx):
    import minotert_amd as mrt
    from minotert_amd.cache import RadianceCache, WEIGHTS
    with pytest.raises(mrt.MrtError):
        RadianceCache(ctx, LO, HI, levels=0)
    with pytest.raises(mrt.MrtError):
        RadianceCache(ctx, LO, LO, levels=4, log2_table=10)
    with pytest.raises(mrt.MrtError):
        RadianceCache(ctx, LO, HI, levels=4, log2_table=10, table_lr=-1.0)
    cache = _make(ctx, SMALL)
    with pytest.raises(mrt.MrtError):
        cache.set(WEIGHTS, np.zeros(10, np.float32))
    cache.close()
