"""The device BVH builder (bvh_build.hip) against the host builder (bvh_host.cpp).

Both follow the same rules (binned SAH, leaf size, median split when all centroids coincide,
stable partitions, BFS collapse to BVH4) with order-independent arithmetic, so the bar is
byte equality of the exported wide nodes and leaf-ordered triangle records, on the generated
scenes and on inputs that reach every path of the device builder: sizes around its class limits
(64, 65, 4096, 4097), one triangle, coplanar soups (a zero-extent centroid axis), thousands of
copies of one triangle (median splits in every class), bin counts 4, 7 and 64, leaf sizes 1 and
8. Traversal over the device-built scene then equals the brute-force oracle.
"""
import numpy as np
import pytest

import minotert_amd as mrt

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-bvh] {msg}", flush=True)


def _both(ctx, verts, max_leaf=4, bins=16):
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
    n = verts.shape[0]
    mat = np.zeros(n, np.uint32)
    alb = np.full((1, 3), 0.5, np.float32)
    out = {}
    for b in ("host", "device"):
        s = mrt.Scene(ctx, verts, mat, alb, max_leaf=max_leaf, bins=bins, builder=b)
        try:
            out[b] = (s.export(), s.stats())
        finally:
            s.close()
    return out


def _check_equal(out):
    (hn, ht), hs = out["host"]
    (dn, dt), ds = out["device"]
    for k in ("n_nodes", "depth", "leaves"):
        assert hs[k] == ds[k], (k, hs[k], ds[k])
    assert np.array_equal(hn, dn), f"{int(np.sum(np.any(hn != dn, axis=1)))} of {hn.shape[0]} wide nodes differ"
    assert np.array_equal(ht, dt), f"{int(np.sum(np.any(ht != dt, axis=1)))} triangle records differ"


def _soup(n, seed, scale=10.0, size=0.3):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-scale, scale, (n, 1, 3))
    return (c + rng.normal(0.0, size, (n, 3, 3))).astype(np.float32).reshape(n, 9)


@pytest.mark.parametrize("cid", [0, 1])
def test_generated_scenes_identical(ctx, cfg, cid):
    verts, _, _ = mrt.generate_scene(cfg.scene_desc(cid))
    _say(f"config {cid}: {verts.shape[0]} triangles")
    out = _both(ctx, verts)
    _check_equal(out)
    _say(f"host {out['host'][1]['build_ms']:.1f} ms, device {out['device'][1]['build_ms']:.1f} ms")


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 65, 100, 4096, 4097, 9000, 50000])
def test_sizes_identical(ctx, n):
    _say(f"random soup of {n}")
    _check_equal(_both(ctx, _soup(n, n)))


@pytest.mark.parametrize("bins,max_leaf", [(4, 4), (7, 1), (64, 8), (16, 1), (33, 2)])
def test_options_identical(ctx, bins, max_leaf):
    _say(f"bins {bins} max_leaf {max_leaf}")
    _check_equal(_both(ctx, _soup(20000, bins * 10 + max_leaf), max_leaf=max_leaf, bins=bins))


def test_degenerate_inputs_identical(ctx):
    # coplanar: every centroid has z = 0 (zero extent on one axis)
    v = _soup(12000, 5)
    v[:, 2::3] = 0.0
    _say("coplanar")
    _check_equal(_both(ctx, v))
    # copies of one triangle: all centroids coincide, median splits in the big, mid and small classes
    one = np.array([0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0], np.float32)
    _say("duplicates")
    _check_equal(_both(ctx, np.tile(one, (9000, 1))))
    # duplicates mixed with a spread-out soup, and signed zeros in the coordinates
    v = np.concatenate([np.tile(one, (3000, 1)), _soup(3000, 6)])
    v[::7, 0] = -0.0
    _say("mixed duplicates and -0")
    _check_equal(_both(ctx, v))


def test_device_built_trace_matches_oracle(ctx, orc, cfg, scene0):
    cam = cfg.camera(0)
    rays = orc.primary_rays(cam)
    s = mrt.Scene(ctx, scene0["verts"], scene0["mat"], scene0["alb"], builder="device")
    try:
        d_rays = ctx.to_device(rays)
        d_hits = ctx.empty(rays.shape[0], mrt.HIT_DTYPE)
        s.trace_closest(d_rays, d_hits)
        got = d_hits.download()
    finally:
        s.close()
    ref = orc.closest_brute(scene0["verts"], rays)
    assert np.array_equal(got["tri"], ref["tri"])
    assert np.array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32))
