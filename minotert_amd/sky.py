"""The reference's physical sky on the device (mrt_sky_*): Atmosphere / Sky of
src/gfx/modules/sky.ixx:18-266 with the LUT shaders src/gpu/sky/gen*.comp.

    sky = PhysicalSky(ctx)                                  # transmittance + multiple scattering
    sky.update_view(probe_km, sun_dir, sun_illuminance)     # the sky-view LUT (per camera position)
    env = Environment(physical_sky=sky, sun_e=(0, 0, 0))    # misses take skyColor(dir)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import lib
from .runtime import Context, DeviceArray

TRANSMITTANCE, MULTISCATTERING, VIEW = 0, 1, 2
SIZES = {TRANSMITTANCE: (256, 64), MULTISCATTERING: (32, 32), VIEW: (192, 108)}
# Sky defaults of the reference (sky.ixx: sunDirection, sunIlluminance)
SUN_DIRECTION = (-0.435286462, 0.818654716, 0.374606609)
SUN_ILLUMINANCE = (8.0, 8.0, 8.0)


def earth() -> _capi.mrt_atmosphere_params:
    p = _capi.mrt_atmosphere_params()
    _capi.check(lib().mrt_sky_earth(C.byref(p)))
    return p


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


class PhysicalSky:
    def __init__(self, ctx: Context, params: _capi.mrt_atmosphere_params | None = None):
        self.ctx = ctx
        self.params = params if params is not None else earth()
        self._h = C.c_void_p()
        ctx.check(lib().mrt_sky_create(ctx.handle, C.byref(self.params), C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def update_view(self, probe_km=(0.0, 0.0, 0.5), sun_dir=SUN_DIRECTION, sun_illuminance=SUN_ILLUMINANCE,
                    sun_disc: bool = True) -> None:
        self.probe, self.sun_dir, self.illum, self.sun_disc = tuple(probe_km), tuple(sun_dir), tuple(sun_illuminance), sun_disc
        self.ctx.check(lib().mrt_sky_update_view(self.ctx.handle, self._h, _f3(probe_km), _f3(sun_dir),
                                                 _f3(sun_illuminance), 1 if sun_disc else 0))

    def radiance(self, dirs: DeviceArray, out: DeviceArray, n: int | None = None) -> None:
        n = dirs.shape[0] if n is None else n
        self.ctx.check(lib().mrt_sky_radiance(self.ctx.handle, self._h, dirs.ptr, n, out.ptr))

    def get(self, which: int) -> np.ndarray:
        w, h = SIZES[which]
        out = np.empty((h, w, 4), np.float32)
        self.ctx.check(lib().mrt_sky_get(self.ctx.handle, self._h, which, out.ctypes.data_as(_capi.F32P), out.size))
        return out

    def close(self) -> None:
        if self._h and self.ctx.handle:
            lib().mrt_sky_destroy(self.ctx.handle, self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
