"""Pins of the oracle's post-processing (oracle/oracle_post.cpp), CPU only.

Tonemappers: the sRGB encode against its float64 closed form and the 8-bit values every sRGB
table lists; each operator at the points its design fixes (Reinhard's white point maps to 1,
Hable's normalisation W = 11.2 and zero, ACES's gray axis and saturation, Uchimura's identity
segment and shoulder limit, the AMD curve's midIn -> midOut and hdrMax -> 1).
Denoiser: a float64 brute force of the spatial term with scipy's bilinear interpolation
(map_coordinates, order 1, edge-clamped), flat images staying flat, depth edges between columns
not leaking colour, sky pixels passing through, and the dither's PCG seed.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as orc
from oracle import oracle_post as op

LINEAR, REINHARD, HABLE, ACES, UCHIMURA, AMD = range(6)


def srgb64(c):
    c = np.asarray(c, np.float64)
    return np.where(c < 0.0031308, 12.92 * c, 1.055 * np.abs(c) ** (1.0 / 2.4) - 0.055)


def tm(op_id, rgb, **kw):
    return op.tonemap(np.asarray(rgb, np.float32).reshape(-1, 3), op.tonemap_params(op=op_id, **kw))


def test_defaults_match_reference():
    p = op.tonemap_params()
    assert p.op == AMD and p.exposure == 1.0 and p.reinhard_max == 8.0
    assert np.allclose(list(p.uchimura), [1.0, 1.0, 0.22, 0.4, 1.33, 0.0])
    assert np.allclose(list(p.amd), [16.0, 2.0, 1.0, 0.18, 0.18])


def test_linear_is_srgb_encode():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(0, 0.004, 3000), rng.uniform(0, 1, 30000), rng.uniform(1, 50, 3000)])
    x = x.astype(np.float32)
    out, q = tm(LINEAR, np.repeat(x, 3))
    want = srgb64(np.repeat(x, 3).astype(np.float64)).reshape(-1, 3)
    assert np.max(np.abs(out - want) / np.maximum(np.abs(want), 1e-3)) < 2e-6
    # 8-bit sRGB of linear 0.5, 0.2140 (sRGB 128) and 0.0031308 (the linear/power junction)
    _, q = tm(LINEAR, [0.5, 0.5, 0.5, 0.21404114, 0.21404114, 0.21404114, 0.0, 1.0, 2.0])
    assert q[0, :3].tolist() == [188, 188, 188] and q[1, 0] == 128
    assert q[2].tolist() == [0, 255, 255, 255]
    # exposure multiplies the input
    a, _ = tm(LINEAR, [0.25, 0.5, 1.0], exposure=2.0)
    b, _ = tm(LINEAR, [0.5, 1.0, 2.0])
    assert np.array_equal(a, b)


def test_reinhard_white_point_and_toe():
    out, _ = tm(REINHARD, [8.0, 8.0, 8.0, 3.0, 3.0, 3.0], reinhard_max=8.0)
    assert abs(out[0, 0] - 1.0) < 1e-6  # v = maxWhite maps to 1
    # extended Reinhard with an unbounded white is x / (1 + x)
    out, _ = tm(REINHARD, [3.0, 0.5, 0.01], reinhard_max=1e18)
    assert np.allclose(out[0], srgb64([0.75, 0.5 / 1.5, 0.01 / 1.01]), rtol=2e-6)


def test_hable_normalisation():
    out, _ = tm(HABLE, [5.6, 0.0, 100.0])  # 2 * 5.6 = W = 11.2 maps to 1; 0 to 0; monotone above
    assert abs(out[0, 0] - 1.0) < 2e-6 and abs(out[0, 1]) < 1e-6 and out[0, 2] > 1.0


def test_aces_gray_axis_and_saturation():
    g = np.array([0.02, 0.18, 0.5, 1.0, 3.0], np.float32)
    out, _ = tm(ACES, np.repeat(g, 3))
    # both matrices have unit row sums (of the rows a row vector meets): gray stays gray
    assert np.max(np.abs(out - out[:, :1])) < 2e-4
    assert np.all(np.diff(out[:, 0]) > 0)
    out, q = tm(ACES, [1e4, 1e4, 1e4, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    assert q[0, :3].tolist() == [255, 255, 255] and q[1, :3].tolist() == [0, 0, 0]
    # pure red stays reddest: a transposed input matrix would let green and blue catch up
    assert out[2, 0] > 3 * out[2, 1] and out[2, 0] > 3 * out[2, 2]


def test_uchimura_identity_segment_and_shoulder():
    # defaults P=1, a=1, m=0.22, l=0.4: l0 = (P - m) l / a = 0.312, identity on [m, m + l0)
    x = np.linspace(0.2201, 0.5319, 97, dtype=np.float32)
    out, _ = tm(UCHIMURA, np.repeat(x, 3))
    assert np.max(np.abs(out[:, 0] - srgb64(x))) < 2e-6
    out, _ = tm(UCHIMURA, [0.0, 60.0, 0.532])  # toe at 0 gives the pedestal 0; shoulder tends to P
    assert abs(out[0, 0]) < 1e-7 and abs(out[0, 1] - 1.0) < 2e-6
    assert abs(out[0, 2] - srgb64(0.532)) < 2e-6  # shoulder starts at S1 = m + a l0 (continuity)


def test_amd_curve_anchors():
    out, _ = tm(AMD, [0.18, 0.18, 0.18, 16.0, 16.0, 16.0])
    assert np.max(np.abs(out[0] - srgb64(0.18))) < 2e-5  # midIn -> midOut
    assert np.max(np.abs(out[1] - 1.0)) < 2e-5           # hdrMax -> 1
    out, _ = tm(AMD, [0.18, 0.18, 0.18, 0.18, 0.18, 0.18], amd=(16.0, 2.0, 1.0, 0.18, 0.3))
    assert np.max(np.abs(out[0] - srgb64(0.3))) < 2e-5
    out, _ = tm(AMD, [0.0, 0.0, 0.0, 0.5, 0.1, 0.0])
    assert np.all(out[0] < 1e-6) and out[1, 0] > out[1, 1] > out[1, 2]


# ---- denoiser ----------------------------------------------------------------------------
def dither(w, h, frame):
    y, x = np.mgrid[0:h, 0:w]
    seed = (x * 709 + y * 1153 + frame * 1361) & 0xFFFFFFFF
    return np.vectorize(lambda s: (orc.pcg(int(s)) & 0xFFFFFF) / 16777216.0 * 0.005)(seed).astype(np.float64)


def taps(radius):
    """(dx, dy) of the reference's loops: dx integral, dy from -sqrt(r^2 - dx^2) in unit steps."""
    out = []
    r = float(radius)
    dx = -r
    while dx <= r:
        pt = np.float32(math.sqrt(np.float32(r * r - dx * dx)))
        dy = -pt
        while dy <= pt:
            out.append((dx, float(dy)))
            dy = np.float32(dy + np.float32(1.0))
        dx += 1.0
    return np.array(out)


def flat_gbuffer(w, h, depth=0.05, n=(0.0, 0.0, 1.0)):
    d = np.full((h, w), depth, np.float32)
    nrm = np.zeros((h, w, 4), np.float32)
    nrm[..., :3] = n
    return d, nrm


def test_spatial_term_against_scipy_bilinear():
    """On a flat surface the range weight is one constant, so the filter is the normalised
    Gaussian sum of bilinear samples: checked in float64 with map_coordinates."""
    rng = np.random.default_rng(3)
    w, h, sigma, ks = 23, 17, 2.0, 2.0
    rgb = rng.uniform(0, 4, (h, w, 3)).astype(np.float32)
    d, nrm = flat_gbuffer(w, h)
    got = op.denoise_bilateral(rgb, d, nrm, 0.1, sigma=sigma, k_sigma=ks, threshold=0.12, frame=5)
    T = taps(round(ks * sigma))
    wgt = np.exp(-(T[:, 0] ** 2 + T[:, 1] ** 2) / (2 * sigma * sigma))
    want = np.zeros((h, w, 3))
    for y in range(h):
        for x in range(w):
            coords = np.stack([y + T[:, 1], x + T[:, 0]])
            for c in range(3):
                s = ndimage.map_coordinates(rgb[..., c].astype(np.float64), coords, order=1, mode="nearest")
                want[y, x, c] = np.sum(wgt * s) / np.sum(wgt)
    want += dither(w, h, 5)[..., None]
    assert np.max(np.abs(got - want) / want) < 2e-5


def test_flat_image_stays_flat_and_sky_passes_through():
    w, h = 31, 9
    rgb = np.full((h, w, 3), [0.3, 1.5, 7.0], np.float32)
    d, nrm = flat_gbuffer(w, h)
    d[4, 10] = 0.0  # a sky pixel: passes through, still dithered
    rgb[4, 10] = [9.0, 9.0, 9.0]
    got = op.denoise_bilateral(rgb, d, nrm, 0.1, frame=2)
    want = rgb.astype(np.float64) + dither(w, h, 2)[..., None]
    mask = np.ones((h, w), bool)
    mask[4, 10] = False
    # neighbours of the sky pixel: its depth 0 gives an infinite distance, so its weight is the
    # minimum exp(0) against exp(1 / (2 threshold^2)) for the surface (relative 3e-16)
    assert np.max(np.abs(got[mask] - want[mask]) / want[mask]) < 1e-5
    assert np.array_equal(got[4, 10], (rgb[4, 10] + np.float32(dither(w, h, 2)[4, 10])).astype(np.float32))


def test_depth_edge_between_columns_does_not_leak():
    """Horizontal taps land on texel centres, so across a vertical depth edge every sample of the
    other side has distance term >= 1 and weight exp(0) = 1 against exp(34.7) on the own side."""
    w, h = 40, 12
    rgb = np.zeros((h, w, 3), np.float32)
    rgb[:, :20] = [1.0, 0.0, 0.0]
    rgb[:, 20:] = [0.0, 0.0, 1.0]
    d, nrm = flat_gbuffer(w, h)
    d[:, 20:] = 0.01  # reverse-Z: distance 10 against 2 (near 0.1)
    got = op.denoise_bilateral(rgb, d, nrm, 0.1, frame=0) - dither(w, h, 0)[..., None]
    assert np.max(np.abs(got[:, :20] - [1.0, 0.0, 0.0])) < 1e-6
    assert np.max(np.abs(got[:, 20:] - [0.0, 0.0, 1.0])) < 1e-6
    # a normal crease at equal depth: dot = 0 on the far side, weight exp(0) against exp(34.7)
    d, nrm = flat_gbuffer(w, h)
    nrm[:, 20:, :3] = [1.0, 0.0, 0.0]
    got = op.denoise_bilateral(rgb, d, nrm, 0.1, frame=0) - dither(w, h, 0)[..., None]
    assert np.max(np.abs(got[:, :20] - [1.0, 0.0, 0.0])) < 1e-6
    # without the edge the colours mix
    d, nrm = flat_gbuffer(w, h)
    got = op.denoise_bilateral(rgb, d, nrm, 0.1, frame=0)
    assert got[5, 19, 2] > 0.2


@pytest.mark.parametrize("frame", [0, 7])
def test_dither_seed(frame):
    w, h = 5, 4
    rgb = np.zeros((h, w, 3), np.float32)
    d = np.zeros((h, w), np.float32)  # all sky: output is the dither alone
    nrm = np.zeros((h, w, 4), np.float32)
    got = op.denoise_bilateral(rgb, d, nrm, 0.1, frame=frame)
    assert np.allclose(got[..., 0], dither(w, h, frame), rtol=0, atol=1e-9)
    assert np.array_equal(got[..., 0], got[..., 2])
