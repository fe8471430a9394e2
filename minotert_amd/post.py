"""Post-processing on the device: the reference's bilateral denoiser and tonemappers.

    Denoiser(ctx).bilateral(color, depth, normal, w, h, near_plane, out, frame=0)
        Denoiser::bilateral (src/gfx/modules/denoiser.ixx:34-97; src/gpu/denoise/bilateral.comp)
    Tonemapper(ctx).amd(color, out_rgb=None, out_rgba8=None)   (also linear, reinhard, hable,
        aces, uchimura): Tonemapper (src/gfx/modules/tonemapper.ixx:57-373; src/gpu/tonemap/*.comp)

All images are DeviceArrays, row-major: colour W*H*3 float32, depth W*H, normal W*H*4, RGBA8
W*H*4 uint8.
"""
from __future__ import annotations

import ctypes as C
import struct
import zlib
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import lib
from .runtime import Context, DeviceArray

LINEAR, REINHARD, HABLE, ACES, UCHIMURA, AMD = range(6)
OPERATORS = {"linear": LINEAR, "reinhard": REINHARD, "hable": HABLE, "aces": ACES, "uchimura": UCHIMURA, "amd": AMD}


@dataclass
class BilateralParams:
    """Denoiser::BilateralParams (denoiser.ixx:20-31) defaults."""
    sigma: float = 5.0
    k_sigma: float = 2.0
    threshold: float = 0.12


@dataclass
class TonemapParams:
    """Tonemapper parameters (tonemapper.ixx:19-55); the renderer's default operator is AMD."""
    op: int = AMD
    exposure: float = 1.0
    reinhard_max: float = 8.0
    uchimura: tuple = (1.0, 1.0, 0.22, 0.4, 1.33, 0.0)  # maxBrightness contrast linearStart linearLength blackTightness pedestal
    amd: tuple = (16.0, 2.0, 1.0, 0.18, 0.18)           # hdrMax contrast shoulder midIn midOut

    def struct(self) -> _capi.mrt_tonemap_params:
        p = _capi.mrt_tonemap_params()
        p.op, p.exposure, p.reinhard_max = int(self.op), float(self.exposure), float(self.reinhard_max)
        for k, v in enumerate(self.uchimura):
            p.uchimura[k] = float(v)
        for k, v in enumerate(self.amd):
            p.amd[k] = float(v)
        return p


def tonemap_defaults() -> TonemapParams:
    """The library's defaults (mrt_tonemap_defaults)."""
    p = _capi.mrt_tonemap_params()
    _capi.check(lib().mrt_tonemap_defaults(C.byref(p)))
    return TonemapParams(p.op, p.exposure, p.reinhard_max, tuple(p.uchimura), tuple(p.amd))


class Denoiser:
    def __init__(self, ctx: Context):
        self.ctx = ctx

    def bilateral(self, color: DeviceArray, depth: DeviceArray, normal: DeviceArray, width: int, height: int,
                  near_plane: float, out: DeviceArray, params: BilateralParams | None = None, frame: int = 0,
                  km_per_unit: float = 1.0) -> None:
        """Edge-aware bilateral filter + 0.005 dither; pixels with depth 0 (no hit) pass through.
        near_plane is in scene units; km_per_unit converts it to the reference's kilometres (its
        depth term weighs distance differences in km), e.g. 1e-3 for a scene modelled in metres."""
        params = params or BilateralParams()
        p = _capi.mrt_bilateral_params(params.sigma, params.k_sigma, params.threshold, near_plane * km_per_unit, frame)
        self.ctx.check(lib().mrt_denoise_bilateral(self.ctx.handle, color.ptr, depth.ptr, normal.ptr, width, height,
                                                   C.byref(p), out.ptr))


class Tonemapper:
    def __init__(self, ctx: Context):
        self.ctx = ctx

    def apply(self, color: DeviceArray, params: TonemapParams, out_rgb: DeviceArray | None = None,
              out_rgba8: DeviceArray | None = None) -> None:
        """Exposure, the operator, then sRGB encode; float and/or RGBA8 output."""
        n = color.nbytes // 12
        p = params.struct()
        self.ctx.check(lib().mrt_tonemap(self.ctx.handle, color.ptr, n, C.byref(p),
                                         out_rgb.ptr if out_rgb is not None else None,
                                         out_rgba8.ptr if out_rgba8 is not None else None))

    def _op(self, op, color, exposure, out_rgb, out_rgba8, **kw):
        p = TonemapParams(op=op, exposure=exposure, **kw)
        self.apply(color, p, out_rgb, out_rgba8)

    def linear(self, color, exposure=1.0, out_rgb=None, out_rgba8=None):
        self._op(LINEAR, color, exposure, out_rgb, out_rgba8)

    def reinhard(self, color, exposure=1.0, hdr_max=8.0, out_rgb=None, out_rgba8=None):
        self._op(REINHARD, color, exposure, out_rgb, out_rgba8, reinhard_max=hdr_max)

    def hable(self, color, exposure=1.0, out_rgb=None, out_rgba8=None):
        self._op(HABLE, color, exposure, out_rgb, out_rgba8)

    def aces(self, color, exposure=1.0, out_rgb=None, out_rgba8=None):
        self._op(ACES, color, exposure, out_rgb, out_rgba8)

    def uchimura(self, color, exposure=1.0, params=TonemapParams.uchimura, out_rgb=None, out_rgba8=None):
        self._op(UCHIMURA, color, exposure, out_rgb, out_rgba8, uchimura=tuple(params))

    def amd(self, color, exposure=1.0, params=TonemapParams.amd, out_rgb=None, out_rgba8=None):
        self._op(AMD, color, exposure, out_rgb, out_rgba8, amd=tuple(params))


def write_png(path: str, rgba8: np.ndarray, width: int, height: int) -> None:
    """RGBA8 pixels (row-major, top row first: the camera's y = 0 row) as an 8-bit RGBA PNG."""
    px = np.ascontiguousarray(rgba8, np.uint8).reshape(height, width, 4)
    raw = b"".join(b"\x00" + px[y].tobytes() for y in range(height))  # filter type 0 per row

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", width, height, 8, 6, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))
