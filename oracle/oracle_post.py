"""ctypes binding of the oracle's post-processing (oracle/oracle_post.cpp).

TEST INFRASTRUCTURE ONLY: the checker of the GPU denoiser and tonemappers
(minotert_amd/csrc/post.hip). Images are numpy float32 arrays: colour (h, w, 3), depth (h, w),
normal (h, w, 4).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .oracle import lib as _lib

P = C.c_void_p


class orc_bilateral(C.Structure):
    _fields_ = [("sigma", C.c_float), ("k_sigma", C.c_float), ("threshold", C.c_float), ("near_plane", C.c_float),
                ("frame", C.c_uint32)]


class orc_tonemap_params(C.Structure):
    _fields_ = [("op", C.c_uint32), ("exposure", C.c_float), ("reinhard_max", C.c_float),
                ("uchimura", C.c_float * 6), ("amd", C.c_float * 5)]


_BOUND = False


def lib():
    global _BOUND
    L = _lib()
    if not _BOUND:
        L.orc_denoise_bilateral.argtypes = [P, P, P, C.c_uint32, C.c_uint32, C.POINTER(orc_bilateral), P]
        L.orc_tonemap_defaults.argtypes = [C.POINTER(orc_tonemap_params)]
        L.orc_tonemap.argtypes = [C.POINTER(orc_tonemap_params), P, C.c_uint32, P, P]
        _BOUND = True
    return L


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    if shape is not None:
        assert a.shape == shape, (a.shape, shape)
    return a


def denoise_bilateral(rgb, depth, normal, near_plane: float, sigma=5.0, k_sigma=2.0, threshold=0.12,
                      frame: int = 0) -> np.ndarray:
    rgb = _f32(rgb)
    h, w = rgb.shape[:2]
    depth, normal = _f32(depth, (h, w)), _f32(normal, (h, w, 4))
    out = np.empty((h, w, 3), np.float32)
    p = orc_bilateral(sigma, k_sigma, threshold, near_plane, frame)
    lib().orc_denoise_bilateral(rgb.ctypes.data, depth.ctypes.data, normal.ctypes.data, w, h, C.byref(p),
                                out.ctypes.data)
    return out


def tonemap_params(op=None, exposure=None, reinhard_max=None, uchimura=None, amd=None) -> orc_tonemap_params:
    """The oracle's defaults with the given fields replaced."""
    p = orc_tonemap_params()
    lib().orc_tonemap_defaults(C.byref(p))
    if op is not None:
        p.op = op
    if exposure is not None:
        p.exposure = exposure
    if reinhard_max is not None:
        p.reinhard_max = reinhard_max
    if uchimura is not None:
        for k, v in enumerate(uchimura):
            p.uchimura[k] = v
    if amd is not None:
        for k, v in enumerate(amd):
            p.amd[k] = v
    return p


def tonemap(rgb, p: orc_tonemap_params):
    """(sRGB float (n, 3), RGBA8 (n, 4)) of n linear RGB pixels."""
    rgb = _f32(rgb).reshape(-1, 3)
    n = rgb.shape[0]
    out = np.empty((n, 3), np.float32)
    q = np.empty((n, 4), np.uint8)
    lib().orc_tonemap(C.byref(p), rgb.ctypes.data, n, out.ctypes.data, q.ctypes.data)
    return out, q
