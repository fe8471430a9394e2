"""ctypes binding of the oracle's physical sky (oracle/oracle_sky.cpp).

TEST INFRASTRUCTURE ONLY: the checker of the GPU sky kernels (minotert_amd/csrc/sky.hip).
LUTs are numpy arrays of shape (h, w, 3).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .oracle import lib as _lib

P = C.c_void_p
F3 = C.c_float * 3


class orc_atmosphere(C.Structure):
    _fields_ = [("bottom_radius", C.c_float), ("top_radius", C.c_float), ("rayleigh_density_exp_scale", C.c_float),
                ("rayleigh_scattering", F3), ("mie_density_exp_scale", C.c_float), ("mie_scattering", F3),
                ("mie_extinction", F3), ("mie_absorption", F3), ("mie_phase_g", C.c_float),
                ("absorption_density0_layer_width", C.c_float), ("absorption_density0_constant", C.c_float),
                ("absorption_density0_linear", C.c_float), ("absorption_density1_constant", C.c_float),
                ("absorption_density1_linear", C.c_float), ("absorption_extinction", F3), ("ground_albedo", F3)]


class orc_sky_state(C.Structure):
    _fields_ = [("atmo", orc_atmosphere), ("trans", P), ("tw", C.c_uint32), ("th", C.c_uint32), ("view", P),
                ("vw", C.c_uint32), ("vh", C.c_uint32), ("probe", F3), ("sun_dir", F3), ("illum", F3),
                ("sun_disc", C.c_int)]


_BOUND = False


def lib():
    global _BOUND
    L = _lib()
    if not _BOUND:
        L.orc_sky_earth.argtypes = [C.POINTER(orc_atmosphere)]
        L.orc_sky_transmittance.argtypes = [C.POINTER(orc_atmosphere), C.c_uint32, C.c_uint32, P]
        L.orc_sky_multiscattering.argtypes = [C.POINTER(orc_atmosphere), P, C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.c_uint32, P]
        L.orc_sky_view.argtypes = [C.POINTER(orc_atmosphere), P, C.c_uint32, C.c_uint32, P, C.c_uint32, C.c_uint32, P, P,
                                   P, C.c_uint32, C.c_uint32, P]
        L.orc_sky_radiance.argtypes = [C.POINTER(orc_atmosphere), P, C.c_uint32, C.c_uint32, P, C.c_uint32, C.c_uint32,
                                       P, P, P, C.c_int, P, C.c_uint32, P]
        L.orc_tx_eval.argtypes = [C.c_int, P, C.c_uint32, P]
        L.orc_sky_phase.restype = C.c_float
        L.orc_sky_phase.argtypes = [C.c_int, C.c_float, C.c_float]
        F2 = C.c_float * 2
        L.orc_sky_trans_uv.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, F2]
        L.orc_sky_trans_params.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, F2]
        L.orc_sky_view_uv.argtypes = [C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, C.c_float, C.c_float, F2]
        L.orc_sky_view_params.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int, C.c_float, C.c_float, F2]
        _BOUND = True
    return L


TX = {"exp": 0, "sin": 1, "cos": 2, "acos": 3, "pow15": 4, "log": 5, "pow_inv2.4": 6}


def tx(fn: str, x) -> np.ndarray:
    """The sky's basic-operation transcendental `fn` on float32 inputs."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().orc_tx_eval(TX[fn], _p(x), x.size, _p(out))
    return out


def _call2(fn, *args):
    out = (C.c_float * 2)()
    fn(*args, out)
    return out[0], out[1]


def phase(mie: bool, g: float, cos_t: float) -> float:
    return lib().orc_sky_phase(1 if mie else 0, g, cos_t)


def trans_uv(h, mu, bottom, top):
    return _call2(lib().orc_sky_trans_uv, h, mu, bottom, top)


def trans_params(u, v, bottom, top):
    return _call2(lib().orc_sky_trans_params, u, v, bottom, top)


def view_uv(ground, vz, lv, w, h, view_h, bottom):
    return _call2(lib().orc_sky_view_uv, 1 if ground else 0, vz, lv, w, h, view_h, bottom)


def view_params(u, v, w, h, view_h, bottom):
    return _call2(lib().orc_sky_view_params, u, v, w, h, view_h, bottom)


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _f3(v):
    return np.ascontiguousarray(v, np.float32).reshape(3)


def earth() -> orc_atmosphere:
    a = orc_atmosphere()
    lib().orc_sky_earth(C.byref(a))
    return a


def transmittance(a, w=256, h=64) -> np.ndarray:
    out = np.empty((h, w, 3), np.float32)
    lib().orc_sky_transmittance(C.byref(a), w, h, _p(out))
    return out


def multiscattering(a, trans, w=32, h=32) -> np.ndarray:
    trans = np.ascontiguousarray(trans, np.float32)
    out = np.empty((h, w, 3), np.float32)
    lib().orc_sky_multiscattering(C.byref(a), _p(trans), trans.shape[1], trans.shape[0], w, h, _p(out))
    return out


def view(a, trans, ms, probe, sun_dir, illum, w=192, h=108) -> np.ndarray:
    trans = np.ascontiguousarray(trans, np.float32)
    ms = np.ascontiguousarray(ms, np.float32)
    pr, sd, il = _f3(probe), _f3(sun_dir), _f3(illum)
    out = np.empty((h, w, 3), np.float32)
    lib().orc_sky_view(C.byref(a), _p(trans), trans.shape[1], trans.shape[0], _p(ms), ms.shape[1], ms.shape[0], _p(pr),
                       _p(sd), _p(il), w, h, _p(out))
    return out


def radiance(a, trans, vw_lut, probe, sun_dir, illum, sun_disc, dirs) -> np.ndarray:
    trans = np.ascontiguousarray(trans, np.float32)
    vw_lut = np.ascontiguousarray(vw_lut, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    pr, sd, il = _f3(probe), _f3(sun_dir), _f3(illum)
    out = np.empty((dirs.shape[0], 3), np.float32)
    lib().orc_sky_radiance(C.byref(a), _p(trans), trans.shape[1], trans.shape[0], _p(vw_lut), vw_lut.shape[1],
                           vw_lut.shape[0], _p(pr), _p(sd), _p(il), 1 if sun_disc else 0, _p(dirs), dirs.shape[0],
                           _p(out))
    return out


class SkyState:
    """A sky for the oracle's path tracer (PathParams.sky_state); keeps its LUT arrays alive."""

    def __init__(self, a, trans, vw_lut, probe, sun_dir, illum, sun_disc):
        self.trans = np.ascontiguousarray(trans, np.float32)
        self.view = np.ascontiguousarray(vw_lut, np.float32)
        s = orc_sky_state()
        s.atmo = a
        s.trans, s.tw, s.th = self.trans.ctypes.data, self.trans.shape[1], self.trans.shape[0]
        s.view, s.vw, s.vh = self.view.ctypes.data, self.view.shape[1], self.view.shape[0]
        for k in range(3):
            s.probe[k], s.sun_dir[k], s.illum[k] = float(probe[k]), float(sun_dir[k]), float(illum[k])
        s.sun_disc = 1 if sun_disc else 0
        self.struct = s
