"""GPU parity of the render stages (render.hip) against the CPU oracle.

Bar: bit-identical hits, rays and radiance (the wavefront kernels evaluate the oracle's
operation sequence per path; compaction order cannot change a path's result). The
north-star tolerance for radiance is 1e-5 relative.
"""
import numpy as np
import pytest

import minotert_amd as mrt

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-render] {msg}", flush=True)


def _env(cfg):
    return mrt.Environment(sky=tuple(cfg.CONFIG["sky"]), sun_dir=tuple(cfg.CONFIG["sun_dir"]),
                           sun_e=tuple(cfg.CONFIG["sun_e"]))


@pytest.fixture(scope="module")
def pt0(ctx, scene0):
    s = mrt.Scene(ctx, scene0["verts"], scene0["mat"], scene0["alb"])
    yield mrt.Pathtracer(ctx, s)
    s.close()


@pytest.fixture(scope="module")
def pt1(ctx, scene1):
    s = mrt.Scene(ctx, scene1["verts"], scene1["mat"], scene1["alb"])
    yield mrt.Pathtracer(ctx, s)
    s.close()


def test_gbuffer_config0(ctx, orc, cfg, scene0, pt0):
    cam = cfg.camera(0)
    prev = cfg.camera(0)
    prev.yaw += 0.01
    _say("config-0 primary rays + G-buffer planes")
    gb = pt0.primary_rays((cam.width, cam.height), cam, prev)
    hits = gb.hits.download()
    ref = orc.closest_brute(scene0["verts"], orc.primary_rays(cam))
    assert np.array_equal(np.ascontiguousarray(hits).view(np.uint32).reshape(-1),
                          np.ascontiguousarray(ref).view(np.uint32).reshape(-1))
    # planes against a float64 restatement of primaryRay.comp:60-75 from the same hits
    basis = orc.camera_basis(cam).astype(np.float64)
    pos, right, up, fwd = basis[0:3], basis[3:6], basis[6:9], basis[9:12]
    rays = orc.primary_rays(cam).astype(np.float64)
    hit = ref["tri"] != mrt.MISS
    v = scene0["verts"].astype(np.float64)[ref["tri"][hit]]
    u_, v_ = ref["u"][hit].astype(np.float64), ref["v"][hit].astype(np.float64)
    p = v[:, 0:3] + u_[:, None] * (v[:, 3:6] - v[:, 0:3]) + v_[:, None] * (v[:, 6:9] - v[:, 0:3])
    q = p - pos
    vz = q @ fwd
    depth = gb.depth.download()
    np.testing.assert_allclose(depth[hit], cam.near_plane / vz, rtol=1e-5)
    assert (depth[~hit] == 0).all()
    nrm = gb.normal.download()
    n = np.cross(v[:, 3:6] - v[:, 0:3], v[:, 6:9] - v[:, 0:3])
    n /= np.linalg.norm(n, axis=1)[:, None]
    flip = np.sum(n * rays[hit, 4:7], axis=1) > 0
    n[flip] *= -1
    np.testing.assert_allclose(nrm[hit, :3], n, atol=2e-6)
    np.testing.assert_array_equal(nrm[~hit, :3], orc.primary_rays(cam)[~hit, 4:7])
    mot = gb.motion.download()
    assert np.abs(mot[hit]).max() > 0.1 and (mot[~hit] == 0).all()


@pytest.mark.parametrize("zoom", [False, True])
def test_motion_vectors_closed_form(ctx, cfg, scene0, pt0, zoom):
    """Motion plane against the reference's matrix route in float64 (primaryRay.comp:73-75 with
    the prevView of pathtracer.ixx:91,102): a non-square viewport, and a previous camera that
    differs in yaw, pitch and position (and field of view, which the reference ignores: both
    positions are projected with the current projection)."""
    from refcam import hit_positions, motion_close, reference_motion
    cam = cfg.camera(0, width=320, height=200)
    prev = cfg.camera(0, width=320, height=200)
    prev.yaw += 0.013
    prev.pitch -= 0.007
    prev.position = tuple(np.array(prev.position) + np.array([0.21, -0.13, 0.08]))
    if zoom:
        prev.vertical_fov *= 0.8
    _say(f"config-0 motion vectors 320x200 zoom={zoom}")
    gb = pt0.primary_rays((cam.width, cam.height), cam, prev)
    hits = gb.hits.download()
    hit, p = hit_positions(scene0["verts"], hits)
    mot = gb.motion.download()
    want = reference_motion(cam, prev, p)
    assert hit.sum() > 1000 and np.abs(want).max() > 2.0
    assert motion_close(mot[hit], want), np.abs(mot[hit] - want).max()
    assert (mot[~hit] == 0).all()


def test_lambert_config0(ctx, orc, cfg, scene0, pt0):
    cam = cfg.camera(0)
    _say("config-0 Lambert + shadow")
    gb = pt0.primary_rays((cam.width, cam.height), cam, planes=False)
    rgb = pt0.shade_lambert(gb, cam, cfg.CONFIG["sun_dir"], cfg.CONFIG["lambert_light_e"]).download()
    ref_hits, ref_rgb = scene0["bvh"].render_lambert(scene0["mat"], scene0["alb"], cam, cfg.CONFIG["sun_dir"],
                                                     cfg.CONFIG["lambert_light_e"])
    assert np.array_equal(rgb.view(np.uint32), ref_rgb.view(np.uint32))
    assert rgb.max() > 0.5 and (rgb == 0).any()


def test_capture_bounce1_rays_config1(ctx, orc, cfg, scene1, pt1):
    cam = cfg.camera(1)
    p = cfg.path_params(1)
    _say("config-1 bounce-1 capture at 1920x1080")
    gb = pt1.primary_rays((cam.width, cam.height), cam, planes=False)
    opts = mrt.RenderOptions(frame_seed=p.frame_seed, sample_base=0, spp=1, bounces=p.bounces, rr_start=p.rr_start)
    d_rays, d_pid, n = pt1.capture_rays(gb, cam, _env(cfg), opts, bounce=1)
    rays = d_rays.download()[:n]
    pid = d_pid.download()[:n]
    order = np.argsort(pid)
    ref_rays, ref_pix = scene1["bvh"].bounce1_rays(cam, p)
    assert n == ref_rays.shape[0]
    assert np.array_equal(pid[order], ref_pix)
    got = rays[order].view(np.float32).reshape(-1, 8)
    assert np.array_equal(got.view(np.uint32), ref_rays.view(np.uint32))


@pytest.mark.parametrize("sample", [0, 5])
def test_path_tracing_config1_reduced(ctx, orc, cfg, scene1, pt1, sample):
    cam = cfg.camera(1, width=320, height=180)
    p = cfg.path_params(1, frame=3, sample=sample)
    _say(f"config-1 path tracing 320x180 sample={sample}")
    opts = mrt.RenderOptions(frame_seed=p.frame_seed, sample_base=sample, spp=1, bounces=p.bounces,
                             rr_start=p.rr_start)
    rgb = pt1.render(cam, _env(cfg), opts).download()
    ref = scene1["bvh"].render_path(scene1["mat"], scene1["alb"], cam, p)
    bad = np.nonzero((rgb.view(np.uint32) != ref.view(np.uint32)).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} pixels differ, e.g. {bad[:5]}: {rgb[bad[:3]]} vs {ref[bad[:3]]}"
    assert rgb.mean() > 0.05


def test_spp_resolve_and_determinism(ctx, orc, cfg, scene1, pt1):
    cam = cfg.camera(1, width=160, height=90)
    p = cfg.path_params(1, frame=1)
    opts = mrt.RenderOptions(frame_seed=p.frame_seed, sample_base=0, spp=4, bounces=p.bounces, rr_start=p.rr_start)
    _say("config-1 spp=4 at 160x90, twice")
    a = pt1.render(cam, _env(cfg), opts).download()
    b = pt1.render(cam, _env(cfg), opts).download()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    samples = [scene1["bvh"].render_path(scene1["mat"], scene1["alb"], cam, cfg.path_params(1, frame=1, sample=j))
               for j in range(4)]
    acc = samples[0].copy()
    for j in range(1, 4):
        acc = acc + samples[j]
    ref = acc / np.float32(4.0)
    assert np.array_equal(a.view(np.uint32), ref.view(np.uint32))


def test_invalid_arguments_rejected(ctx, cfg, pt0):
    cam = cfg.camera(0)
    gb = pt0.primary_rays((cam.width, cam.height), cam, planes=False)
    with pytest.raises(mrt.MrtError):
        pt0.secondary_rays(gb, cam, mrt.Environment(), mrt.RenderOptions(spp=0))
    with pytest.raises(mrt.MrtError):
        pt0.secondary_rays(gb, cam, mrt.Environment(), mrt.RenderOptions(rr_start=1))
    bad = cfg.camera(0, width=128, height=128)
    with pytest.raises(mrt.MrtError):
        pt0.secondary_rays(gb, bad, mrt.Environment(), mrt.RenderOptions())


@pytest.mark.parametrize("bounces,spp", [(4, 2), (8, 1)])
def test_megakernel_matches_wavefront_and_oracle(ctx, cfg, scene1, pt1, bounces, spp):
    """One lane per path after vertex 0 (MRT_RENDER_MEGAKERNEL) gives the wavefront's image bit
    for bit; at 1 spp both equal the oracle."""
    cam = cfg.camera(1, width=320, height=180)
    p = cfg.path_params(1, frame=5, sample=2)
    p.bounces = bounces
    _say(f"megakernel vs wavefront, bounces={bounces} spp={spp}")
    base = dict(frame_seed=p.frame_seed, sample_base=2, spp=spp, bounces=bounces, rr_start=p.rr_start)
    a = pt1.render(cam, _env(cfg), mrt.RenderOptions(**base)).download()
    b = pt1.render(cam, _env(cfg), mrt.RenderOptions(megakernel=True, **base)).download()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    if spp == 1:
        ref = scene1["bvh"].render_path(scene1["mat"], scene1["alb"], cam, p)
        assert np.array_equal(b.view(np.uint32), ref.view(np.uint32))
