"""Pins of the CPU oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a dropped term, a wrong sign
or index, a transposed operand) fails: closed forms, invariants, an independent double-
precision formulation (Moller-Trumbore), the reference's own matrix route for the camera,
brute force, and committed golden vectors (tests/golden/) for regression.
"""
import math
import os

import numpy as np
import pytest

from refcam import reference_matrices

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- RNG (random.glsl:22-31)
def test_pcg_known_answers(orc):
    # Worked by hand from random.glsl:22-27 for v = 0:
    #   state = 0 * 747796405 + 2891336453                 = 0xAC564B05
    #   state >> 28 = 0xA, shift = 0xA + 4 = 14; state >> 14 = 0x2B159
    #   0x0002B159 ^ 0xAC564B05 = 0xAC54FA5C; then * 277803737 mod 2^32, then word ^ (word >> 22)
    state = 2891336453
    assert state == 0xAC564B05
    x = (state >> 14) ^ state
    assert x == 0xAC54FA5C
    word = (x * 277803737) & 0xFFFFFFFF
    expect0 = ((word >> 22) ^ word) & 0xFFFFFFFF
    assert orc.pcg(0) == expect0
    # the hash of 1 and of 0xFFFFFFFF follow the same steps
    for v in (1, 0xFFFFFFFF, 12345, 0x9E3779B9):
        s = (v * 747796405 + 2891336453) & 0xFFFFFFFF
        w = ((((s >> ((s >> 28) + 4)) ^ s) * 277803737) & 0xFFFFFFFF)
        assert orc.pcg(v) == ((w >> 22) ^ w)


def test_pcg_is_injective_and_uniform(orc):
    # Every stage of PCG-RXS-M-XS is a bijection of 32-bit words, so the hash must be too:
    # no collisions on 2^18 distinct inputs; output bits balanced.
    import ctypes
    lib = orc.lib()
    vals = np.array([lib.orc_pcg(i * 2654435761 & 0xFFFFFFFF) for i in range(1 << 18)], np.uint64)
    assert np.unique(vals).size == vals.size
    bits = np.unpackbits(vals.astype(np.uint32).view(np.uint8)).reshape(-1, 32)
    assert np.all(np.abs(bits.mean(axis=0) - 0.5) < 0.005)
    # randomFloat: 24-bit grid in [0, 1 - 2^-24]
    st = ctypes.c_uint32(7)
    f = np.array([lib.orc_rand_float(ctypes.byref(st)) for _ in range(100000)])
    assert f.min() >= 0.0 and f.max() <= 1.0 - 2.0 ** -24
    assert np.all((f * 2 ** 24) == np.floor(f * 2 ** 24))
    hist, _ = np.histogram(f, bins=16, range=(0, 1))
    chi2 = np.sum((hist - 6250.0) ** 2 / 6250.0)
    assert chi2 < 40.0  # 15 dof, p ~ 5e-4


# ---------------------------------------------------------------- sampling (random.glsl:10-19)
def test_sincos_turns_accuracy_and_symmetry(orc):
    xs = np.concatenate([np.arange(0, 1 << 24, 257), [0, (1 << 22), (1 << 23), 3 << 22, (1 << 24) - 1]]) / float(1 << 24)
    err = 0.0
    for x in xs[::13]:
        s, c = orc.sincos_turns(float(x))
        err = max(err, abs(s - math.sin(2 * math.pi * x)), abs(c - math.cos(2 * math.pi * x)))
    assert err < 3e-7
    # exact quarter-turn symmetry: sin(x + 1/4) == cos(x), cos(x + 1/4) == -sin(x)
    for k in (1, 77, 4095, 1 << 20):
        x = k / float(1 << 24)
        s0, c0 = orc.sincos_turns(x)
        s1, c1 = orc.sincos_turns(x + 0.25)
        assert s1 == c0 and c1 == -s0
    assert orc.sincos_turns(0.0) == (0.0, 1.0)
    assert orc.sincos_turns(0.5) == (0.0, -1.0)


def test_cosine_direction_distribution(orc):
    rng = np.random.default_rng(3)
    n = np.array([0.3, -0.5, 0.81], np.float64)
    n = (n / np.linalg.norm(n)).astype(np.float32)
    cos = []
    for _ in range(20000):
        x1, x2 = (rng.integers(0, 1 << 24, 2) / float(1 << 24)).tolist()
        d = orc.cosine_dir(n, x1, x2)
        assert abs(np.linalg.norm(d.astype(np.float64)) - 1.0) < 3e-7
        cos.append(float(np.dot(d.astype(np.float64), n)))
    cos = np.array(cos)
    assert cos.min() >= -1e-6
    # cosine-weighted: E[cos] = 2/3, E[cos^2] = 1/2
    assert abs(cos.mean() - 2.0 / 3.0) < 0.01
    assert abs((cos ** 2).mean() - 0.5) < 0.01


# ---------------------------------------------------------------- camera (camera.ixx, math.ixx)
def test_camera_rays_match_reference_matrix_route(orc, cfg):
    for cid, (W, H) in ((0, (7, 5)), (1, (16, 9))):
        cam = cfg.camera(cid, width=W, height=H)
        V, Pm = reference_matrices(cam)
        iV, iP = np.linalg.inv(V), np.linalg.inv(Pm)
        rays = orc.primary_rays(cam)
        for y in range(H):
            for x in range(W):
                uv = np.array([(x + 0.5) / W, 1.0 - (y + 0.5) / H])
                ndc = np.array([uv[0] * 2 - 1, uv[1] * 2 - 1, 1.0, 1.0])
                vd = (iP @ ndc)[:3]
                wd = (iV @ np.append(vd, 0.0))[:3]
                wd /= np.linalg.norm(wd)
                got = rays[y * W + x, 4:7].astype(np.float64)
                assert np.abs(got - wd).max() < 1e-6
                assert np.allclose(rays[y * W + x, 0:3], cam.position, atol=1e-6)
    # the centre pixel of an odd viewport looks along the camera direction
    cam = cfg.camera(0, width=7, height=5)
    b = orc.camera_basis(cam)
    assert np.abs(orc.primary_rays(cam)[2 * 7 + 3, 4:7] - b[9:12]).max() < 1e-7


# ---------------------------------------------------------------- watertight triangle test
TRI = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)


def _ray(o, d, tmin=0.0, tmax=np.inf):
    return np.array(list(o) + [tmin] + list(d) + [tmax], np.float32)


def test_triangle_known_answers(orc):
    h, r = orc.intersect(_ray((0.25, 0.25, 1), (0, 0, -1)), TRI)
    assert h and r[0] == 1.0 and r[2] == 0.25 and r[3] == 0.25      # u, v = weights of v1, v2
    h, r = orc.intersect(_ray((0.1, 0.7, -3), (0, 0, 1)), TRI)        # back side: two-sided
    assert h and r[0] == 3.0 and abs(r[2] - 0.1) < 1e-7 and abs(r[3] - 0.7) < 1e-7
    assert not orc.intersect(_ray((0.6, 0.6, 1), (0, 0, -1)), TRI)[0]  # outside the hypotenuse
    assert orc.intersect(_ray((0.5, 0.5, 1), (0, 0, -1)), TRI)[0]      # on the hypotenuse edge
    assert orc.intersect(_ray((0.0, 0.0, 1), (0, 0, -1)), TRI)[0]      # on a vertex
    assert orc.intersect(_ray((0.5, 0.0, 1), (0, 0, -1)), TRI)[0]      # on an axis edge
    assert not orc.intersect(_ray((0.25, 0.25, 1), (0, 0, -1), tmax=0.5), TRI)[0]
    assert not orc.intersect(_ray((0.25, 0.25, 1), (0, 0, -1), tmin=1.5), TRI)[0]
    assert not orc.intersect(_ray((0.25, 0.25, 1), (1, 0, 0)), TRI)[0]  # parallel
    assert not orc.intersect(_ray((0.25, 0.25, 1), (0, 0, 1)), TRI)[0]  # pointing away
    assert not orc.intersect(_ray((0.25, 0.25, 1), (0, 0, 0)), TRI)[0]  # degenerate direction
    # oblique ray, closed form: o + t d with t = 2 hits (0.3, 0.2, 0)
    d = np.array([0.1, -0.2, -0.5]); o = np.array([0.3, 0.2, 0.0]) - 2 * d
    h, r = orc.intersect(_ray(o, d), TRI)
    assert h and abs(r[0] - 2.0) < 2e-6 and abs(r[2] - 0.3) < 2e-6 and abs(r[3] - 0.2) < 2e-6


def _unit_sphere_soup(orc, cfg, level=2):
    """One icosphere of radius 1 at the origin (box 0, radius range [1, 1])."""
    d = cfg.scene_desc(0, spheres=1, level=level, box=0.0)
    d.ground = 0
    d.r_min = d.r_max = 1.0
    v, m, a = orc.generate_scene(d)
    return v


def test_watertight_no_leaks_through_edges_and_vertices(orc, cfg):
    v = _unit_sphere_soup(orc, cfg)
    tris = v.reshape(-1, 3, 3).astype(np.float64)
    c = tris.reshape(-1, 3).mean(axis=0)
    targets = np.concatenate([tris.reshape(-1, 3), 0.5 * (tris[:, 0] + tris[:, 1]), 0.5 * (tris[:, 1] + tris[:, 2])])
    rays = []
    for p in targets[::3]:
        dirn = p - c
        o = c + 10.0 * dirn / np.linalg.norm(dirn) + np.array([1e-3, -2e-3, 0.5e-3])
        dd = p - o
        rays.append(_ray(o, dd / np.linalg.norm(dd)))
    hits = orc.closest_brute(v, np.array(rays))
    assert np.all(hits["tri"] != 0xFFFFFFFF), "ray aimed at a shared vertex/edge leaked through the soup"


def _moller_trumbore_f64(verts, rays):
    """Independent closest-hit in float64 (Moller & Trumbore 1997), ties to the smaller ID."""
    v = verts.reshape(-1, 3, 3).astype(np.float64)
    v0, e1, e2 = v[:, 0], v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    out_t = np.full(rays.shape[0], np.inf); out_id = np.full(rays.shape[0], 0xFFFFFFFF, np.uint64)
    for i, r in enumerate(rays.astype(np.float64)):
        o, d = r[0:3], r[4:7]
        p = np.cross(d, e2)
        det = np.einsum("ij,ij->i", e1, p)
        ok = np.abs(det) > 1e-300
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        tv = o - v0
        u = np.einsum("ij,ij->i", tv, p) * inv
        q = np.cross(tv, e1)
        w = (q @ d) * inv
        t = np.einsum("ij,ij->i", e2, q) * inv
        m = ok & (u >= 0) & (w >= 0) & (u + w <= 1) & (t >= r[3]) & (t <= r[7])
        if m.any():
            idx = np.nonzero(m)[0]
            j = idx[np.argmin(t[idx])]
            out_t[i], out_id[i] = t[j], j
    return out_t, out_id


def test_closest_hit_vs_independent_double_precision(orc, cfg, scene0):
    rays = np.concatenate([orc.primary_rays(cfg.camera(0, width=16, height=16)),
                           orc.random_rays(11, 10.0, 200)])
    got = orc.closest_brute(scene0["verts"], rays)
    t64, id64 = _moller_trumbore_f64(scene0["verts"], rays)
    hit = id64 != 0xFFFFFFFF
    assert np.array_equal(got["tri"] != 0xFFFFFFFF, hit)
    same = got["tri"][hit] == id64[hit]
    # a different ID is only allowed at a near-tie of t (shared edges / coplanar neighbours)
    t_got = got["t"][hit].astype(np.float64)
    assert np.all(same | (np.abs(t_got - t64[hit]) <= 1e-5 * t64[hit]))
    assert np.all(np.abs(t_got - t64[hit]) <= 1e-5 * np.maximum(t64[hit], 1.0))
    assert same.mean() > 0.99


# ---------------------------------------------------------------- scene generator
def test_scene_counts_and_geometry(orc, cfg):
    for cid, expect in ((0, 20482), (1, 1024000)):
        assert cfg.scene_desc(cid).tri_count() == expect
    d = cfg.scene_desc(4)
    assert d.tri_count() == 16384000
    v, m, a = orc.generate_scene(cfg.scene_desc(0))
    assert v.shape == (20482, 9) and m.max() == 64 and a.shape == (65, 3)
    assert np.all((a[:64] >= 0.2) & (a[:64] < 0.9)) and np.all(a[64] == 0.5)
    pts = v[:-2].reshape(64, -1, 3).astype(np.float64)
    cen = 0.5 * (pts.min(axis=1) + pts.max(axis=1))
    rad = np.linalg.norm(pts - cen[:, None, :], axis=2)
    assert np.all(np.abs(cen) <= 10.0 + 1e-4)
    assert np.all((rad.mean(axis=1) >= 0.5 - 1e-4) & (rad.mean(axis=1) <= 2.0 + 1e-4))
    assert np.all(rad.std(axis=1) / rad.mean(axis=1) < 1e-6)  # every vertex on its sphere


def test_scene_checksums_golden(orc, cfg):
    import hashlib
    import json
    with open(os.path.join(GOLDEN, "scene_checksums.json")) as f:
        gold = json.load(f)
    for cid in ("0", "1"):
        v, m, a = orc.generate_scene(cfg.scene_desc(int(cid)))
        h = hashlib.sha256(v.tobytes() + m.tobytes() + a.tobytes()).hexdigest()
        assert h == gold[cid], f"config {cid} soup changed"


# ---------------------------------------------------------------- BVH
def test_bvh_structure(orc, scene0, scene1):
    for s in (scene0, scene1):
        assert s["bvh"].validate() == 0
        info = s["bvh"].info()
        assert info["max_leaf"] <= 4 and info["leaves"] >= s["verts"].shape[0] // 4
        assert math.isfinite(info["sah"]) and info["sah"] > 1.0


def test_bvh_traversal_equals_bruteforce(orc, cfg, scene0):
    rays = np.concatenate([orc.primary_rays(cfg.camera(0, width=96, height=96)),
                           orc.random_rays(cfg.CONFIG["random_ray_seed"], 10.0, 8192)])
    a = scene0["bvh"].trace_closest(rays)
    b = orc.closest_brute(scene0["verts"], rays)
    assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))
    occ = scene0["bvh"].trace_any(rays)
    assert np.array_equal(occ.astype(bool), b["tri"] != 0xFFFFFFFF)


def test_traversal_independent_of_thread_count(orc, cfg, scene1):
    rays = orc.random_rays(5, 40.0, 20000)
    a = scene1["bvh"].trace_closest(rays, threads=1)
    b = scene1["bvh"].trace_closest(rays, threads=7)
    assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


# ---------------------------------------------------------------- shading / estimator
def test_path_tracer_convex_object_closed_form(orc, cfg):
    """A diffuse ray leaving a convex object never hits it again, so with a constant sky L and
    the directional light E every hit pixel is a*max(0, n.l)*E + a*L and every miss L."""
    v = _unit_sphere_soup(orc, cfg, level=3)
    mat = np.zeros(v.shape[0], np.uint32)
    alb = np.array([[0.6, 0.3, 0.9]], np.float32)
    bvh = orc.BVH(v)
    cam = cfg.camera(0, width=48, height=48)
    cam.position = (0.0, -4.0, 0.3)
    cam.yaw, cam.pitch = math.pi / 2, -0.07
    for sun_e in ((0.0, 0.0, 0.0), (2.0, 2.0, 2.0)):
        p = cfg.path_params(1, frame=2)
        p.sky = (0.8, 0.9, 1.0)
        p.sun_e = sun_e
        rgb = bvh.render_path(mat, alb, cam, p)
        rays = orc.primary_rays(cam)
        hits = bvh.trace_closest(rays)
        hit = hits["tri"] != 0xFFFFFFFF
        assert hit.mean() > 0.1
        np.testing.assert_array_equal(rgb[~hit], np.tile(np.float32(p.sky), (int((~hit).sum()), 1)))
        tv = v.reshape(-1, 3, 3).astype(np.float64)[hits["tri"][hit]]
        n = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0]); n /= np.linalg.norm(n, axis=1)[:, None]
        n *= -np.sign(np.sum(n * rays[hit, 4:7], axis=1))[:, None]
        cosl = np.maximum(0.0, n @ np.array(p.sun_dir))
        expect = alb[0][None, :] * cosl[:, None] * np.array(sun_e)[None, :] + alb[0][None, :] * np.array(p.sky)[None, :]
        np.testing.assert_allclose(rgb[hit], expect, rtol=2e-6, atol=2e-7)
        if sun_e == (0.0, 0.0, 0.0):
            np.testing.assert_array_equal(rgb[hit], np.tile(alb[0] * np.float32(p.sky), (int(hit.sum()), 1)))


def test_path_tracer_roulette_unbiased_in_expectation(orc, cfg):
    """Two planes facing each other: roulette (rr_start = 2) must match the no-roulette mean."""
    quad = lambda z, flip: ([-2, -2, z, 2, -2, z, 2, 2, z], [-2, -2, z, 2, 2, z, -2, 2, z]) if not flip \
        else ([-2, -2, z, 2, 2, z, 2, -2, z], [-2, -2, z, -2, 2, z, 2, 2, z])
    v = np.array(quad(0.0, False) + quad(1.0, True), np.float32).reshape(-1, 9)
    mat = np.zeros(4, np.uint32); alb = np.array([[0.7, 0.7, 0.7]], np.float32)
    bvh = orc.BVH(v)
    cam = cfg.camera(0, width=16, height=16)
    cam.position = (0.0, 0.0, 0.5); cam.yaw = 0.3; cam.pitch = -0.5
    means = {}
    for rr in (2, 99):
        acc = np.zeros((256, 3))
        for s in range(192):
            p = cfg.path_params(1, frame=0, sample=s)
            p.bounces, p.rr_start, p.sky, p.sun_e = 6, rr, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
            acc += bvh.render_path(mat, alb, cam, p)
        means[rr] = acc.mean() / 192
    assert abs(means[2] - means[99]) < 0.03 * means[99]


def test_lambert_matches_closed_form(orc, cfg, scene0):
    cam = cfg.camera(0, width=64, height=64)
    l, e = cfg.CONFIG["sun_dir"], cfg.CONFIG["lambert_light_e"]
    hits, rgb = scene0["bvh"].render_lambert(scene0["mat"], scene0["alb"], cam, l, e)
    rays = orc.primary_rays(cam)
    hit = hits["tri"] != 0xFFFFFFFF
    assert np.all(rgb[~hit] == 0)
    tv = scene0["verts"].reshape(-1, 3, 3).astype(np.float64)[hits["tri"][hit]]
    n = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0]); n /= np.linalg.norm(n, axis=1)[:, None]
    n *= -np.sign(np.sum(n * rays[hit, 4:7], axis=1))[:, None]
    cosl = n @ np.array(l)
    a = scene0["alb"][scene0["mat"][hits["tri"][hit]]].astype(np.float64)
    lit = rgb[hit].sum(axis=1) > 0
    expect = a * np.maximum(cosl, 0)[:, None] * np.array(e)[None, :]
    np.testing.assert_allclose(rgb[hit][lit], expect[lit], rtol=1e-5, atol=1e-7)
    # dark pixels are back-facing or shadowed: check shadowing by an independent float64 test
    dark = np.nonzero(hit)[0][~lit & (cosl > 1e-3)]
    assert dark.size > 0
    u = hits["u"][dark].astype(np.float64); w = hits["v"][dark].astype(np.float64)
    tv = scene0["verts"].reshape(-1, 3, 3).astype(np.float64)[hits["tri"][dark]]
    p = tv[:, 0] + u[:, None] * (tv[:, 1] - tv[:, 0]) + w[:, None] * (tv[:, 2] - tv[:, 0])
    sh = np.array([_ray(pi + 1e-3 * np.array(l), l) for pi in p[:60]])
    t64, _ = _moller_trumbore_f64(scene0["verts"], sh)
    assert np.all(np.isfinite(t64))


@pytest.mark.parametrize("vertex", [0, 1, 2])
def test_path_suffix_targets_reproduce_the_plain_paths(orc, cfg, scene1, vertex):
    """Cache training data (orc_path_suffix): ending every path at the vertex with its own
    training target as the cache value gives back the plain path's radiance (up to the rounding
    of (L - L_before) / thr * thr), and the sample is that vertex's position, incoming
    direction, normal and albedo (the next plain vertex is found along the recorded direction)."""
    cam = cfg.camera(1, width=96, height=54)
    p = cfg.path_params(1, frame=3)
    p.cache_vertex = vertex
    npix = cam.width * cam.height
    plain = scene1["bvh"].render_path(scene1["mat"], scene1["alb"], cam, p)
    smp, tgt, valid = scene1["bvh"].path_suffix(scene1["mat"], scene1["alb"], cam, p, np.arange(npix, dtype=np.uint32))
    assert valid.sum() > 20
    p.cache_mode = 1
    cached = scene1["bvh"].render_path(scene1["mat"], scene1["alb"], cam, p, cache_rgb=tgt[:, :3])
    assert np.allclose(cached, plain, rtol=2e-6, atol=1e-6)
    assert np.array_equal(cached[~valid], plain[~valid])  # no vertex there: the path is the plain one
    # a planted mistake (targets without the division by the throughput) is caught
    bad = scene1["bvh"].render_path(scene1["mat"], scene1["alb"], cam, p, cache_rgb=tgt[:, :3] * 0.5)
    assert not np.allclose(bad, plain, rtol=1e-3, atol=1e-3)
    s = smp[valid]
    assert np.allclose(np.linalg.norm(s[:, 4:7], axis=1), 1.0, atol=1e-5)   # incoming direction
    assert np.allclose(np.linalg.norm(s[:, 8:11], axis=1), 1.0, atol=1e-5)  # normal
    assert np.all(np.einsum("ij,ij->i", s[:, 4:7], s[:, 8:11]) <= 1e-6)     # normal faces the ray
    assert np.all((s[:, 12:15] >= 0.2) & (s[:, 12:15] <= 0.9))              # sphere albedo range
    assert np.all(s[:, [3, 7, 11, 15]] == 0)


def test_rel_mse_closed_form(orc):
    ref = np.array([1.0, 2.0, 0.0, 4.0], np.float32)
    img = np.array([1.5, 2.0, 0.1, 3.0], np.float32)
    x, r = img.astype(np.float64), ref.astype(np.float64)
    expect = np.mean((x - r) ** 2 / (r ** 2 + 1e-2))
    assert abs(expect - np.mean([0.25 / 1.01, 0.0, 0.01 / 0.01, 1.0 / 16.01])) < 1e-7
    assert abs(orc.rel_mse(img, ref, 1e-2) - expect) < 1e-12


# ---------------------------------------------------------------- golden regression vectors
def test_golden_vectors(orc, cfg, scene0):
    g = np.load(os.path.join(GOLDEN, "oracle_vectors.npz"))
    assert np.array_equal(np.array([orc.pcg(int(x)) for x in g["pcg_in"]], np.uint32), g["pcg_out"])
    rays = orc.primary_rays(cfg.camera(0, width=32, height=32))
    assert np.array_equal(rays.view(np.uint32), g["primary_rays_32"].view(np.uint32))
    h = orc.closest_brute(scene0["verts"], g["probe_rays"])
    assert np.array_equal(np.ascontiguousarray(h).view(np.uint32).reshape(-1, 4), g["probe_hits"])
