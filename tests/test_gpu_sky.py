"""The physical sky on the GPU (sky.hip) against the oracle (oracle/oracle_sky.cpp).

Both sides evaluate exp, sin, cos, acos and x^1.5 with the same basic-operation methods
(csrc/transc.h, oracle/oracle_transc.hpp) and every other operation separately rounded in the
same order, so the bar is bit-exact: each LUT (checked on the GPU's own inputs: the GPU
transmittance LUT feeds the oracle's multiple-scattering and sky-view LUTs), the miss radiance
of 50k directions (around the sun disc too), and a path-traced image with the physical sky
(no next-event estimation, sun disc on: the reference's estimator).
"""
import numpy as np
import pytest

import minotert_amd as mrt
from minotert_amd.sky import MULTISCATTERING, TRANSMITTANCE, VIEW, PhysicalSky, SUN_DIRECTION, SUN_ILLUMINANCE
from oracle import oracle_sky as osk

pytestmark = pytest.mark.gpu
PROBE = (0.0, 0.0, 0.5)


def _errs(a, b, floor=1e-5):
    """|a - b| / (|b| + floor * max|b|) per element (the floor covers texels near zero)."""
    return np.abs(a - b) / (np.abs(b) + floor * np.max(np.abs(b)))


def _rel(a, b, floor=1e-5):
    return float(np.max(_errs(a, b, floor)))


def _q(a, b, q):
    """q-quantile over texels / directions of the worst channel's error."""
    return float(np.quantile(_errs(a, b).reshape(-1, 3).max(axis=1), q))


@pytest.fixture(scope="module")
def sky(ctx):
    s = PhysicalSky(ctx)
    s.update_view(PROBE, SUN_DIRECTION, SUN_ILLUMINANCE, sun_disc=True)
    yield s
    s.close()


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def test_sky_luts_match_oracle(ctx, sky):
    a = osk.earth()
    T = sky.get(TRANSMITTANCE)[..., :3]
    To = osk.transmittance(a)
    M = sky.get(MULTISCATTERING)[..., :3]
    Mo = osk.multiscattering(a, T)
    V = sky.get(VIEW)[..., :3]
    Vo = osk.view(a, T, M, PROBE, SUN_DIRECTION, SUN_ILLUMINANCE)
    print(f"[gpu-sky] max rel: transmittance {_rel(T, To):.2e} multiscattering {_rel(M, Mo):.2e} view {_rel(V, Vo):.2e}",
          flush=True)
    assert _same(T, To) and _same(M, Mo) and _same(V, Vo)
    assert V[..., 2].mean() > V[..., 0].mean() > 0  # a blue sky


def test_sky_radiance_matches_oracle(ctx, sky):
    rng = np.random.default_rng(4)
    d = rng.normal(size=(50000, 3))
    sun = np.array(SUN_DIRECTION)
    d[:200] = sun + rng.normal(scale=0.002, size=(200, 3))  # around the sun disc
    d = (d / np.linalg.norm(d, axis=1)[:, None]).astype(np.float32)
    dd = ctx.to_device(d)
    out = ctx.empty((d.shape[0], 3), np.float32)
    sky.radiance(dd, out)
    got = out.download()
    a = osk.earth()
    ref = osk.radiance(a, sky.get(TRANSMITTANCE)[..., :3], sky.get(VIEW)[..., :3], PROBE, SUN_DIRECTION,
                       SUN_ILLUMINANCE, True, d)
    assert np.all(np.isfinite(got))
    print(f"[gpu-sky] radiance max rel {_rel(got, ref):.2e}", flush=True)
    assert _same(got, ref)
    assert (got.max(axis=1) > 1e4).sum() > 10  # some directions see the sun disc


def test_path_tracing_with_the_physical_sky(ctx, cfg, scene1, sky):
    """The reference's estimator (skyColor on escape, no NEE) through the wavefront."""
    s = mrt.Scene(ctx, scene1["verts"], scene1["mat"], scene1["alb"])
    pt = mrt.Pathtracer(ctx, s)
    cam = cfg.camera(1, width=200, height=112)
    p = cfg.path_params(1, frame=6)
    p.sun_e = (0.0, 0.0, 0.0)
    env = mrt.Environment(sky=(1.0, 1.0, 1.0), sun_dir=tuple(SUN_DIRECTION), sun_e=(0.0, 0.0, 0.0), physical_sky=sky)
    opts = mrt.RenderOptions(frame_seed=p.frame_seed, sample_base=0, spp=1, bounces=p.bounces, rr_start=p.rr_start)
    rgb = pt.render(cam, env, opts).download()
    mk = pt.render(cam, env, mrt.RenderOptions(frame_seed=p.frame_seed, sample_base=0, spp=1, bounces=p.bounces,
                                               rr_start=p.rr_start, megakernel=True)).download()
    assert np.array_equal(rgb.view(np.uint32), mk.view(np.uint32))
    st = osk.SkyState(osk.earth(), sky.get(TRANSMITTANCE)[..., :3], sky.get(VIEW)[..., :3], PROBE, SUN_DIRECTION,
                      SUN_ILLUMINANCE, True)
    p.sky_state = st
    ref = scene1["bvh"].render_path(scene1["mat"], scene1["alb"], cam, p)
    assert np.all(np.isfinite(rgb))
    print(f"[gpu-sky] image max rel {_rel(rgb, ref):.2e}", flush=True)
    assert _same(rgb, ref)
    s.close()


def test_sky_view_is_required(ctx, cfg):
    s = PhysicalSky(ctx)
    with pytest.raises(mrt.MrtError):
        s.radiance(ctx.empty((1, 3), np.float32), ctx.empty((1, 3), np.float32))
    s.close()
