"""Pins of the oracle's radiance cache (CPU only): BF16 rounding known answers, the hash-grid
encoding against closed forms (constant and linear fields, the published spatial hash), the
encoding gradient as the exact adjoint of the encoding, the MLP gradient against float64
finite differences of an independent numpy network, and Adam's closed form."""
import numpy as np
import pytest

from oracle import oracle_cache as oc

SAMPLE = np.dtype([("pos", "<f4", 3), ("pad0", "<f4"), ("dir", "<f4", 3), ("pad1", "<f4"),
                   ("nrm", "<f4", 3), ("pad2", "<f4"), ("alb", "<f4", 3), ("pad3", "<f4")])


def _samples(n, seed=0, lo=-5.0, hi=5.0):
    rng = np.random.default_rng(seed)
    s = np.zeros(n, SAMPLE)
    s["pos"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
    s["dir"] = d
    nr = rng.normal(size=(n, 3)); nr /= np.linalg.norm(nr, axis=1)[:, None]
    s["nrm"] = nr
    s["alb"] = rng.uniform(0.2, 0.9, (n, 3))
    return s


def test_bf16_round_known_answers():
    L = oc.lib()
    one = 1.0
    cases = {1.0: 1.0, 1.0 + 2 ** -8: 1.0, 1.0 + 3 * 2 ** -8: 1.0 + 2 ** -6, 1.0 + 2 ** -7: 1.0 + 2 ** -7,
             -2.5: -2.5, 3.0e38: 3.0e38 if False else None}
    for x, want in cases.items():
        got = L.orc_bf16_round(x)
        if want is not None:
            assert got == want, (x, got, want)
    x = np.random.default_rng(1).normal(size=10000).astype(np.float32)
    assert np.array_equal(oc.bf16(x), np.array([L.orc_bf16_round(float(v)) for v in x], np.float32))
    assert np.all(np.abs(oc.bf16(x) - x) <= np.abs(x) * 2.0 ** -8)
    assert one == 1.0


def _grid(levels=4, log2=12, base=4.0, maxr=64.0):
    return oc.grid(levels, log2, base, maxr, (-5.0, -5.0, -5.0), (5.0, 5.0, 5.0))


def _res(g, l):
    b = np.exp((np.log(g.max_res) - np.log(g.base_res)) / (g.levels - 1))
    return np.float32(np.floor(g.base_res * b ** l))


def test_encoding_constant_and_linear_fields():
    g = _grid()
    T = 1 << g.log2_table
    s = _samples(500, seed=2)
    # constant per level: every interpolation returns the constant (weights sum to one)
    table = np.zeros((g.levels, T, 2), np.float32)
    for l in range(g.levels):
        table[l, :, 0] = 0.25 * (l + 1)
        table[l, :, 1] = -0.5 * (l + 1)
    x = oc.encode(g, table, s)
    for l in range(g.levels):
        assert np.allclose(x[:, 2 * l], oc.bf16(np.float32(0.25 * (l + 1))), atol=4e-3)
        assert np.allclose(x[:, 2 * l + 1], oc.bf16(np.float32(-0.5 * (l + 1))), atol=4e-3)
    # a linear field of the grid coordinates on dense levels is reproduced by trilinear interpolation
    dense_levels = [l for l in range(g.levels) if (int(_res(g, l)) + 1) ** 3 <= T]
    assert dense_levels
    table[:] = 0
    for l in dense_levels:
        r = int(_res(g, l)); side = r + 1
        idx = np.arange(side ** 3)
        gx, gy, gz = idx % side, (idx // side) % side, idx // (side * side)
        table[l, :side ** 3, 0] = (0.01 * gx + 0.02 * gy - 0.015 * gz).astype(np.float32)
        table[l, :side ** 3, 1] = (0.005 * gx).astype(np.float32)
    x = oc.encode(g, table, s)
    u = (s["pos"].astype(np.float64) + 5.0) / 10.0
    for l in dense_levels:
        X = u * float(_res(g, l))
        want0 = 0.01 * X[:, 0] + 0.02 * X[:, 1] - 0.015 * X[:, 2]
        assert np.allclose(x[:, 2 * l], want0, atol=1e-2 * max(1.0, np.abs(want0).max()) * 2 ** -6 + 2e-2)
        assert np.allclose(x[:, 2 * l + 1], 0.005 * X[:, 0], rtol=2 ** -7, atol=1e-3)
    # the remaining inputs: direction, normal, albedo, zero padding
    assert np.array_equal(x[:, 32:35], oc.bf16(s["dir"]))
    assert np.array_equal(x[:, 35:38], oc.bf16(s["nrm"]))
    assert np.array_equal(x[:, 38:41], oc.bf16(s["alb"]))
    assert np.all(x[:, 41:] == 0) and np.all(x[:, 2 * g.levels:32] == 0)


def test_hashed_level_uses_the_published_spatial_hash():
    g = _grid(levels=2, log2=8, base=16.0, maxr=32.0)  # (17)^3 > 256: both levels hashed
    T = 1 << g.log2_table
    s = _samples(1, seed=5)
    for corner_entry in range(0, T, 37):
        table = np.zeros((g.levels, T, 2), np.float32)
        table[0, corner_entry, 0] = 1.0
        x = oc.encode(g, table, s)
        # expected: sum of the trilinear weights of the corners that hash to corner_entry
        r = float(_res(g, 0))
        X = (s["pos"][0].astype(np.float32) + np.float32(5.0)) * np.float32(0.1) * np.float32(r)
        c0 = np.floor(X).astype(np.int64)
        fr = X - c0
        w = 0.0
        for c in range(8):
            d = [(c >> k) & 1 for k in range(3)]
            gx, gy, gz = c0[0] + d[0], c0[1] + d[1], c0[2] + d[2]
            h = ((gx * 1) ^ (gy * 2654435761) ^ (gz * 805459861)) & 0xFFFFFFFF & (T - 1)
            if h == corner_entry:
                w += np.prod([fr[k] if d[k] else 1 - fr[k] for k in range(3)])
        assert abs(x[0, 0] - w) <= 2 ** -8 * max(w, 1e-30) + 1e-7


def test_encoding_backward_is_the_adjoint():
    g = _grid()
    T = 1 << g.log2_table
    s = _samples(300, seed=3)
    rng = np.random.default_rng(4)
    t = rng.normal(size=(g.levels, T, 2)).astype(np.float32)
    gx = np.zeros((300, 64), np.float32)
    gx[:, :2 * g.levels] = rng.normal(size=(300, 2 * g.levels))
    # <gx, E t> with E linear (computed in float64 without BF16 rounding) == <E^T gx, t>
    Et = np.zeros((300, 2 * g.levels))
    eye = np.zeros_like(t)
    # E t through the oracle without rounding: use exactly representable table values
    t8 = np.round(t * 8) / 8
    x = oc.encode(g, t8, s)
    gt = oc.encode_backward(g, s, gx, t.size)
    lhs = float(np.sum(gx[:, :2 * g.levels].astype(np.float64) * x[:, :2 * g.levels]))
    rhs = float(np.sum(gt.astype(np.float64) * t8.reshape(-1)))
    assert abs(lhs - rhs) <= 2 ** -7 * np.sum(np.abs(gx[:, :2 * g.levels]) * np.abs(x[:, :2 * g.levels])) + 1e-6
    assert Et.shape and eye.shape


def _ident(x):
    return x


def _bf(x):
    return oc.bf16(np.asarray(x, np.float32)).astype(np.float64)


def _np_forward(W, x, rnd=_ident):
    """float64 network of six 64x64 layers with ReLU after the first five; `rnd` is applied where
    the GPU rounds to BF16 (each hidden activation)."""
    a = [x]
    z = None
    for L in range(6):
        z = a[-1] @ W[L].T
        if L < 5:
            a.append(rnd(np.maximum(z, 0.0)))
    return z[:, :3], a


def _np_backward(W, x, t, rnd=_ident):
    """Gradient of the relative L2 loss mean((y-t)^2 / (y^2+0.01)) with the denominator held
    constant, by matrix products; `rnd` is applied where the GPU rounds deltas to BF16."""
    y, a = _np_forward(W, x, rnd)
    n = x.shape[0]
    den = y * y + 0.01
    loss = np.sum((y - t) ** 2 / den) / n
    d = np.zeros((n, 64))
    d[:, :3] = rnd(2 * (y - t) / den / n)
    dW = np.zeros_like(W)
    for L in range(5, -1, -1):
        dW[L] = d.T @ a[L]
        g = d @ W[L]
        if L > 0:
            d = rnd(g * (a[L] > 0))
    return loss, dW


def test_mlp_gradient_against_finite_differences_and_numpy():
    rng = np.random.default_rng(7)
    W = rng.uniform(-0.3, 0.3, (6, 64, 64))
    W[5, 3:] = 0
    x = np.zeros((40, 64)); x[:, :41] = rng.uniform(-1, 1, (40, 41))
    t = rng.uniform(0, 1, (40, 3))
    loss, dW = _np_backward(W, x, t)
    # the numpy network's gradient against central differences in float64 (denominator detached)
    y0, _ = _np_forward(W, x)
    den0 = y0 * y0 + 0.01
    def loss_fixed_den(Wp):
        y, _ = _np_forward(Wp, x)
        return np.sum((y - t) ** 2 / den0) / x.shape[0]
    for (L, o, i) in [(0, 3, 5), (2, 10, 20), (5, 1, 7), (4, 63, 0), (1, 0, 40), (3, 17, 33)]:
        e = 1e-6
        Wp, Wm = W.copy(), W.copy()
        Wp[L, o, i] += e; Wm[L, o, i] -= e
        fd = (loss_fixed_den(Wp) - loss_fixed_den(Wm)) / (2 * e)
        assert abs(fd - dW[L, o, i]) <= 1e-6 * max(1.0, abs(fd)), (L, o, i, fd, dW[L, o, i])
    # the oracle (BF16 weights, activations and deltas, FP32 sums) against the same numpy
    # network with BF16 rounding at those points: agreement to FP32 summation order
    Wb = _bf(W)
    xb = _bf(x)
    t4 = np.concatenate([t, np.zeros((40, 1))], 1)
    loss_o, dW_o, dx0 = oc.backward(W.astype(np.float32), xb.astype(np.float32), t4)
    loss_r, dW_r = _np_backward(Wb, xb, t, _bf)
    dW_o = dW_o.reshape(6, 64, 64)
    for L in range(6):
        err = np.linalg.norm(dW_o[L] - dW_r[L]) / np.linalg.norm(dW_r[L])
        assert err < 1e-5, (L, err)
    assert abs(loss_o - loss_r) <= 1e-5 * loss_r
    # the input gradient (used by the encoding backward) is W0^T d0, without the ReLU mask
    _, a = _np_forward(Wb, xb, _bf)
    y = (a[5] @ Wb[5].T)[:, :3]
    d = np.zeros((40, 64)); d[:, :3] = _bf(2 * (y - t) / (y * y + 0.01) / 40)
    for L in range(5, 0, -1):
        d = _bf((d @ Wb[L]) * (a[L] > 0))
    assert np.allclose(dx0, d @ Wb[0], rtol=1e-4, atol=1e-9)
    # without the BF16 rounding the gradients stay within the rounding's reach (loose)
    _, dW_n = _np_backward(Wb, xb, t)
    for L in range(6):
        assert np.linalg.norm(dW_o[L] - dW_n[L]) / np.linalg.norm(dW_n[L]) < 0.15
    # forward: bit-for-bit structure of the rounding, FP32 sums
    y_o = oc.forward(W.astype(np.float32), xb.astype(np.float32))
    y_r, _ = _np_forward(Wb, xb, _bf)
    assert np.allclose(y_o[:, :3], y_r, rtol=1e-5, atol=1e-6)


def test_adam_closed_form():
    p = np.array([1.0, -2.0, 0.5], np.float32)
    g = np.array([0.1, -3.0, 0.0], np.float32)
    p1, m1, v1 = oc.adam(p, g, np.zeros(3, np.float32), np.zeros(3, np.float32), 1, 0.01, 0.9, 0.99, 1e-8)
    # first step: m_hat = g, v_hat = g^2 -> p - lr * g / (|g| + eps)
    want = p - 0.01 * g / (np.abs(g) + 1e-8)
    assert np.allclose(p1, want, rtol=1e-6, atol=1e-7)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.01 * g * g)
    # dense: a zero gradient still moves a parameter with momentum; sparse: it stays, moments too
    p2, m2, v2 = oc.adam(p1, np.zeros(3, np.float32), m1, v1, 2, 0.01, 0.9, 0.99, 1e-8)
    assert p2[0] != p1[0] and m2[0] == np.float32(0.9) * m1[0]
    p3, m3, v3 = oc.adam(p1, np.array([0.0, 1.0, 0.0], np.float32), m1, v1, 2, 0.01, 0.9, 0.99, 1e-8, sparse=True)
    assert p3[0] == p1[0] and m3[0] == m1[0] and v3[0] == v1[0]
    assert p3[1] != p1[1]
    # a step of the dense rule from zero moments: bias correction makes it lr * sign(g) at any step
    p4, _, _ = oc.adam(np.zeros(1, np.float32), np.array([2.0], np.float32), np.zeros(1, np.float32),
                       np.zeros(1, np.float32), 1, 0.01, 0.9, 0.99, 1e-8)
    assert abs(p4[0] + 0.01) < 1e-7


def _group_average(e0, ps, a, gain_power):
    """The table average of a group of S steps as the library applies it: the first step moves
    every entry to a^S e + (1 - a^S) p_0, step j (1-based) adds (1 - a^(S-j+gain_power))
    (p_j - p_{j-1}). gain_power = 1 is the library's form; 0 plants a wrong exponent."""
    S = len(ps) - 1
    aS = np.float32(a) ** np.float32(S)
    e = aS * e0 + (np.float32(1) - aS) * ps[0]
    for j in range(1, S + 1):
        gain = np.float32(1) - np.float32(a) ** np.float32(S - j + gain_power)
        e = e + gain * (ps[j] - ps[j - 1])
    return e


def test_group_average_equals_per_step_average():
    """A frame's S Adam steps average the table in closed form (one pass over every entry on the
    first step, then only the entries each step updates); over sparse oracle Adam steps this
    equals S per-step averages e = a e + (1 - a) p to float rounding, for entries touched on some
    steps and not others. The same sum with the exponent one lower misses by ~(1 - a) of a
    step's change."""
    rng = np.random.default_rng(21)
    n, S, a = 4096, 4, 0.95
    p = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    e0 = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    ps = [p]
    e_step = e0.copy()
    for step in range(1, S + 1):
        g = rng.normal(0.0, 1e-2, n).astype(np.float32)
        g[rng.random(n) < 0.6] = 0.0  # untouched entries keep value and moments
        p, m, v = oc.adam(p, g, m, v, step, 1e-2, 0.9, 0.99, 1e-8, sparse=True)
        ps.append(p)
        e_step = np.float32(a) * e_step + np.float32(1 - a) * p
    assert any(np.any(ps[j] != ps[j - 1]) for j in range(1, S + 1))
    tol = 1e-6 * float(np.abs(e_step).max())
    assert float(np.abs(_group_average(e0, ps, a, 1) - e_step).max()) <= tol
    assert float(np.abs(_group_average(e0, ps, a, 0) - e_step).max()) > 100 * tol
