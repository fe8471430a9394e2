"""Pins of the oracle's physical sky (oracle/oracle_sky.cpp), CPU only: phase functions
normalized over the sphere, LUT parameterizations inverted by their counterparts, the zenith
transmittance against the closed-form optical depth of the exponential and tent-shaped
densities, a non-scattering atmosphere giving no in-scattered light, and the sun disc."""
import math

import numpy as np
import pytest

from oracle import oracle_sky as osk


def _ulp_err(got, want):
    want32 = np.float32(want)
    spacing = np.spacing(np.abs(want32).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / spacing


def test_transcendentals_against_libm():
    """exp, sin, cos, acos and x^1.5 built from basic operations agree with libm in double to
    within 2 ulp on the ranges the sky uses (and exp keeps 2 ulp down to the denormal range)."""
    rng = np.random.default_rng(9)
    x = rng.uniform(-80.0, 80.0, 200000).astype(np.float32)
    assert np.max(_ulp_err(osk.tx("exp", x), np.exp(x.astype(np.float64)))) <= 2.0
    x = rng.uniform(0.0, 6.5, 200000).astype(np.float32)
    assert np.max(_ulp_err(osk.tx("sin", x), np.sin(x.astype(np.float64))) *
                  np.minimum(1.0, np.abs(np.sin(x.astype(np.float64))) / 1e-3)) <= 2.0
    assert np.max(np.abs(osk.tx("sin", x) - np.sin(x.astype(np.float64)))) < 2e-7
    assert np.max(np.abs(osk.tx("cos", x) - np.cos(x.astype(np.float64)))) < 2e-7
    x = rng.uniform(-1.0, 1.0, 200000).astype(np.float32)
    x[:4] = [-1.0, 1.0, 0.5, -0.5]
    assert np.max(_ulp_err(osk.tx("acos", x), np.arccos(x.astype(np.float64)))) <= 2.0
    x = rng.uniform(0.0, 4.0, 100000).astype(np.float32)
    assert np.max(_ulp_err(osk.tx("pow15", x), x.astype(np.float64) ** 1.5)) <= 1.5
    x = np.exp(rng.uniform(-30.0, 30.0, 200000)).astype(np.float32)
    assert np.max(np.abs(osk.tx("log", x) - np.log(x.astype(np.float64)))) < 4e-6  # absolute, |ln x| <= 30
    x = rng.uniform(1e-6, 64.0, 200000).astype(np.float32)
    assert np.max(_ulp_err(osk.tx("pow_inv2.4", x), x.astype(np.float64) ** (1.0 / 2.4))) <= 3.0
    assert osk.tx("exp", np.float32([0.0]))[0] == 1.0 and osk.tx("acos", np.float32([1.0]))[0] == 0.0
    assert osk.tx("log", np.float32([1.0]))[0] == 0.0
    assert osk.tx("exp", np.float32([-200.0]))[0] == 0.0


def test_phase_functions_integrate_to_one():
    c = np.linspace(-1.0, 1.0, 400001)
    for mie, g in [(False, 0.0), (True, 0.0), (True, 0.5), (True, 0.8)]:
        p = np.array([osk.phase(mie, g, float(x)) for x in c[::20]])
        integral = 2.0 * math.pi * np.trapezoid(p, c[::20])
        assert abs(integral - 1.0) < (2e-4 if g < 0.7 else 2e-3), (mie, g, integral)
    # Rayleigh: 3/(16 pi) (1 + c^2); forward-peaked Mie for g > 0 (evaluated with -cos, as the shader does)
    assert abs(osk.phase(False, 0.0, 1.0) - 6.0 / (16.0 * math.pi)) < 1e-7
    assert osk.phase(True, 0.8, -1.0) > 10 * osk.phase(True, 0.8, 1.0)


def test_transmittance_parameterization_round_trip():
    a = osk.earth()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        # v near 0 is the ground, where h - bottom ~ rho^2 / (2 bottom) is below float resolution
        u, v = rng.uniform(0.001, 0.999), rng.uniform(0.05, 0.999)
        h, mu = osk.trans_params(float(u), float(v), a.bottom_radius, a.top_radius)
        assert a.bottom_radius <= h <= a.top_radius + 1e-3 and -1.0 <= mu <= 1.0
        u2, v2 = osk.trans_uv(h, mu, a.bottom_radius, a.top_radius)
        assert abs(u2 - u) < 2e-3 and abs(v2 - v) < 1e-3, (u, v, u2, v2)


def test_view_parameterization_round_trip():
    a = osk.earth()
    vh = a.bottom_radius + 0.5
    beta = math.acos(math.sqrt(vh * vh - a.bottom_radius ** 2) / vh)
    rng = np.random.default_rng(2)
    for _ in range(2000):
        u, v = rng.uniform(0.02, 0.98, 2)
        vz, lv = osk.view_params(float(u), float(v), 192, 108, vh, a.bottom_radius)
        ground = math.acos(max(-1.0, min(1.0, vz))) > math.pi - beta
        u2, v2 = osk.view_uv(ground, vz, lv, 192, 108, vh, a.bottom_radius)
        assert abs(u2 - u) < 2e-3 and abs(v2 - v) < 2e-3, (u, v, u2, v2)


def _extinction(a, h):
    ozo = np.clip(np.where(h < a.absorption_density0_layer_width,
                           a.absorption_density0_linear * h + a.absorption_density0_constant,
                           a.absorption_density1_linear * h + a.absorption_density1_constant), 0.0, 1.0)
    return (np.outer(np.exp(a.rayleigh_density_exp_scale * h), a.rayleigh_scattering[:])
            + np.outer(np.exp(a.mie_density_exp_scale * h), a.mie_extinction[:])
            + np.outer(ozo, a.absorption_extinction[:]))


def test_transmittance_against_float64_quadrature_and_closed_form():
    """Texels against the LUT's 40-step rule (t_s = tmax (s + 0.3) / 40, the first step 0.3 long)
    restated in float64 numpy; the zenith texel also against the exact optical depth of the
    exponential densities beta H (1 - e^{-100/H}) and the ozone tent (15 km), which the 40-step
    rule under-estimates by ~15% (it starts its first sample 0.3 steps above the ground)."""
    a = osk.earth()
    T = osk.transmittance(a)
    for (x, y) in [(0, 0), (10, 5), (128, 32), (255, 63), (200, 1), (3, 60)]:
        h0, mu = osk.trans_params((x + 0.5) / 256, (y + 0.5) / 64, a.bottom_radius, a.top_radius)
        pos = np.array([0.0, 0.0, h0]); d = np.array([0.0, math.sqrt(max(0.0, 1 - mu * mu)), mu])
        b = 2 * pos @ d; c = pos @ pos - a.top_radius ** 2
        tmax = (-b + math.sqrt(b * b - 4 * c)) / 2
        ts = tmax * (np.arange(40) + 0.3) / 40
        dts = np.diff(np.concatenate([[0.0], ts]))
        h = np.linalg.norm(pos[None, :] + ts[:, None] * d[None, :], axis=1) - a.bottom_radius
        tau = (_extinction(a, h) * dts[:, None]).sum(axis=0)
        got = -np.log(T[y, x].astype(np.float64))
        assert np.allclose(got, tau, rtol=2e-3), ((x, y), got, tau)
    Hr, Hm = 8.0, 1.2
    exact = (np.array(a.rayleigh_scattering[:]) * Hr * (1 - math.exp(-100 / Hr))
             + np.array(a.mie_extinction[:]) * Hm * (1 - math.exp(-100 / Hm))
             + np.array(a.absorption_extinction[:]) * 15.0)
    got = -np.log(T[0, 0].astype(np.float64))
    assert np.all(got < exact) and np.allclose(got, exact, rtol=0.2), (got, exact)
    # longer paths through the atmosphere transmit less: T falls from the zenith (u = 0) to the horizon (u = 1)
    assert np.all(np.diff(T[0, :, 2]) <= 1e-6)
    assert np.all((T > 0) & (T <= 1.0))


def test_non_scattering_atmosphere_scatters_nothing():
    a = osk.earth()
    for k in range(3):
        a.rayleigh_scattering[k] = 0.0
        a.mie_scattering[k] = 0.0
        a.mie_extinction[k] = 1e-6
    T = osk.transmittance(a)
    ms = osk.multiscattering(a, T)
    assert np.all(ms == 0)
    view = osk.view(a, T, ms, (0, 0, 0.5), (0, 0.8, 0.6), (8, 8, 8))
    assert np.all(view == 0)


def test_earth_sky_is_blue_and_the_sun_disc_is_where_the_sun_is():
    a = osk.earth()
    T = osk.transmittance(a)
    ms = osk.multiscattering(a, T)
    sun = np.array([-0.435286462, 0.818654716, 0.374606609], np.float32)
    view = osk.view(a, T, ms, (0, 0, 0.5), sun, (8, 8, 8))
    zen = osk.radiance(a, T, view, (0, 0, 0.5), sun, (8, 8, 8), False, np.array([[0.0, 0.001, 0.9999995]], np.float32))[0]
    assert zen[2] > zen[1] > zen[0] > 0
    # the sun disc (radius 0.2525 deg) adds 120000 / E times the transmitted sun luminance
    ax = np.cross(sun, [0, 0, 1]).astype(np.float64); ax /= np.linalg.norm(ax)
    def rot(ang):
        c, s = math.cos(ang), math.sin(ang)
        return (sun * c + np.cross(ax, sun) * s + ax * np.dot(ax, sun) * (1 - c)).astype(np.float32)
    dirs = np.stack([sun, rot(math.radians(0.1)), rot(math.radians(0.4))])
    on = osk.radiance(a, T, view, (0, 0, 0.5), sun, (8, 8, 8), True, dirs)
    off = osk.radiance(a, T, view, (0, 0, 0.5), sun, (8, 8, 8), False, dirs)
    assert np.all(on[0] > 1000 * off[0]) and np.all(on[1] > 1000 * off[1]) and on[0][0] > on[1][0]
    assert np.array_equal(on[2], off[2])
