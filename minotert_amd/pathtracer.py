"""Render-stage API mirroring the reference's Pathtracer (src/gfx/modules/pathtracer.ixx:18-200).

    gbuffer = pt.primary_rays(size, camera, prev_camera)          # pathtracer.ixx:29
    color   = pt.secondary_rays(gbuffer, camera, environment)     # pathtracer.ixx:118-120

The reference's G-buffer holds visibility (R32Uint), depth (R16F), normal (RGBA16F) and
motion (RG16F) images (pathtracer.ixx:42-69); here the planes are device arrays: the hit
records (visibility = hit["tri"], plus t and barycentrics), depth, normal and motion in FP32.
The reference's sky LUTs (Atmosphere) are replaced by the constant sky + directional light of
Environment. All work runs through the C ABI on the context's stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from .configs import CameraDesc
from .runtime import HIT_DTYPE, RAY_DTYPE, Context, DeviceArray, Scene, camera_struct


@dataclass
class Environment:
    """Miss radiance and the directional light. sky is a constant radiance; with physical_sky
    (a minotert_amd.sky.PhysicalSky with a current view) escaped rays take the reference's
    skyColor (secondaryRays.comp:36-58) instead. sun_e = 0 turns next-event estimation off."""
    sky: tuple = (1.0, 1.0, 1.0)
    sun_dir: tuple = (-0.435286462, 0.818654716, 0.374606609)
    sun_e: tuple = (2.0, 2.0, 2.0)
    physical_sky: object = None


@dataclass
class GBuffer:
    hits: DeviceArray
    depth: DeviceArray | None
    normal: DeviceArray | None
    motion: DeviceArray | None
    width: int
    height: int

    def struct(self) -> _capi.mrt_gbuffer:
        g = _capi.mrt_gbuffer()
        g.hits_dev = self.hits.ptr
        g.depth_dev = self.depth.ptr if self.depth is not None else None
        g.normal_dev = self.normal.ptr if self.normal is not None else None
        g.motion_dev = self.motion.ptr if self.motion is not None else None
        g.width, g.height = self.width, self.height
        return g

    @property
    def visibility(self) -> np.ndarray:
        return self.hits.download()["tri"].reshape(self.height, self.width)


@dataclass
class RenderOptions:
    frame_seed: int = 0
    sample_base: int = 0
    spp: int = 1
    bounces: int = 4
    rr_start: int = 3
    sort_rays: bool = False
    megakernel: bool = False

    def struct(self, env: Environment) -> _capi.mrt_render_opts:
        o = _capi.mrt_render_opts()
        o.frame_seed, o.sample_base, o.spp = self.frame_seed, self.sample_base, self.spp
        o.bounces, o.rr_start = self.bounces, self.rr_start
        o.flags = (_capi.MRT_RENDER_SORT_RAYS if self.sort_rays else 0) | (
            _capi.MRT_RENDER_MEGAKERNEL if self.megakernel else 0)
        for k in range(3):
            o.sky[k], o.sun_dir[k], o.sun_e[k] = env.sky[k], env.sun_dir[k], env.sun_e[k]
        if env.physical_sky is not None:
            o.sky_lut = env.physical_sky.handle
        return o


class Pathtracer:
    """Two-stage render of one scene on one context."""

    def __init__(self, ctx: Context, scene: Scene):
        self.ctx = ctx
        self.scene = scene

    def primary_rays(self, size, camera: CameraDesc, prev_camera: CameraDesc | None = None,
                     planes: bool = True) -> GBuffer:
        w, h = int(size[0]), int(size[1])
        if (w, h) != (camera.width, camera.height):
            raise ValueError("size must match the camera viewport")
        n = w * h
        gb = GBuffer(hits=self.ctx.empty(n, HIT_DTYPE),
                     depth=self.ctx.empty(n, np.float32) if planes else None,
                     normal=self.ctx.empty((n, 4), np.float32) if planes else None,
                     motion=self.ctx.empty((n, 2), np.float32) if planes else None, width=w, height=h)
        cam = camera_struct(camera)
        prev = camera_struct(prev_camera if prev_camera is not None else camera)
        g = gb.struct()
        self.ctx.check(_capi.lib().mrt_primary_rays(self.ctx.handle, self.scene.handle, C.byref(cam), C.byref(prev),
                                                    C.byref(g)))
        return gb

    def secondary_rays(self, gbuffer: GBuffer, camera: CameraDesc, env: Environment,
                       opts: RenderOptions | None = None, out: DeviceArray | None = None) -> DeviceArray:
        opts = opts or RenderOptions()
        rgb = out if out is not None else self.ctx.empty((gbuffer.width * gbuffer.height, 3), np.float32)
        cam = camera_struct(camera)
        g = gbuffer.struct()
        o = opts.struct(env)
        self.ctx.check(_capi.lib().mrt_secondary_rays(self.ctx.handle, self.scene.handle, C.byref(cam), C.byref(g),
                                                      C.byref(o), rgb.ptr))
        return rgb

    def capture_rays(self, gbuffer: GBuffer, camera: CameraDesc, env: Environment, opts: RenderOptions,
                     bounce: int = 1):
        """Extension rays of the given bounce (path ids = pixel + sample * W * H), unordered."""
        cap = gbuffer.width * gbuffer.height * opts.spp
        rays = self.ctx.empty(cap, RAY_DTYPE)
        pid = self.ctx.empty(cap, np.uint32)
        cnt = self.ctx.empty(1, np.uint32)
        cam = camera_struct(camera)
        g = gbuffer.struct()
        o = opts.struct(env)
        o.capture_bounce = bounce
        o.capture_rays_dev, o.capture_pid_dev, o.capture_count_dev = rays.ptr, pid.ptr, cnt.ptr
        self.ctx.check(_capi.lib().mrt_secondary_rays(self.ctx.handle, self.scene.handle, C.byref(cam), C.byref(g),
                                                      C.byref(o), None))
        n = int(cnt.download()[0])
        return rays, pid, n

    def shade_lambert(self, gbuffer: GBuffer, camera: CameraDesc, light_dir, light_e) -> DeviceArray:
        rgb = self.ctx.empty((gbuffer.width * gbuffer.height, 3), np.float32)
        cam = camera_struct(camera)
        g = gbuffer.struct()
        l = (C.c_float * 3)(*light_dir)
        e = (C.c_float * 3)(*light_e)
        self.ctx.check(_capi.lib().mrt_shade_lambert(self.ctx.handle, self.scene.handle, C.byref(cam), C.byref(g),
                                                     l, e, rgb.ptr))
        return rgb

    def render(self, camera: CameraDesc, env: Environment, opts: RenderOptions | None = None,
               out: DeviceArray | None = None) -> DeviceArray:
        opts = opts or RenderOptions()
        rgb = out if out is not None else self.ctx.empty((camera.width * camera.height, 3), np.float32)
        cam = camera_struct(camera)
        o = opts.struct(env)
        self.ctx.check(_capi.lib().mrt_render(self.ctx.handle, self.scene.handle, C.byref(cam), C.byref(o), rgb.ptr))
        return rgb


@dataclass
class CacheOptions:
    """Cache termination of a frame (mrt_cache_render_opts): paths end at `vertex` into the
    cache; about `train_samples` of the paths that reach that vertex on a surface are traced in
    full and train it in `train_steps` steps (mrt_cache_render_opts). overlap_training: the steps
    run at the start of the next cached frame with this cache, beside its traversal; any read of
    the cache runs them first. query_budget / loss_budget (None: unchecked): a frame in which more
    paths reach the vertex on a surface, or after a training step whose mean loss per sample is
    higher, renders plain."""
    enable: bool = True
    vertex: int = 1
    train_samples: int = 65536
    train_steps: int = 4
    overlap_training: bool = False
    query_budget: int | None = None
    loss_budget: float | None = None

    def struct(self, cache_rgb=None, export=None) -> _capi.mrt_cache_render_opts:
        o = _capi.mrt_cache_render_opts()
        o.enable, o.vertex, o.train_samples, o.train_steps = int(self.enable), self.vertex, self.train_samples, self.train_steps
        o.overlap_training = int(self.overlap_training)
        if self.query_budget is not None:
            o.limits |= _capi.CACHE_LIMIT_QUERIES
            o.query_budget = int(self.query_budget)
        if self.loss_budget is not None:
            o.limits |= _capi.CACHE_LIMIT_LOSS
            o.loss_budget = float(self.loss_budget)
        if cache_rgb is not None:
            o.cache_rgb_dev = cache_rgb.ptr
        if export is not None:
            o.train_samples_dev, o.train_targets_dev, o.train_pid_dev, o.train_count_dev = (
                export["samples"].ptr, export["targets"].ptr, export["pid"].ptr, export["count"].ptr)
        return o


class CachedPathtracer(Pathtracer):
    """Path tracer whose paths end into a radiance cache trained online on its own full paths.
    With opts.enable false, or a cache whose weights are not finite, frames are bit-identical
    to Pathtracer.render."""

    def __init__(self, ctx: Context, scene: Scene, cache):
        super().__init__(ctx, scene)
        self.cache = cache

    def export_buffers(self, capacity: int) -> dict:
        from .cache import SAMPLE_DTYPE
        return {"samples": self.ctx.empty(capacity, SAMPLE_DTYPE), "targets": self.ctx.empty((capacity, 4), np.float32),
                "pid": self.ctx.empty(capacity, np.uint32), "count": self.ctx.empty(1, np.uint32)}

    def render(self, camera: CameraDesc, env: Environment, opts: RenderOptions | None = None,
               out: DeviceArray | None = None, cache_opts: CacheOptions | None = None, cache_rgb: DeviceArray | None = None,
               export: dict | None = None) -> DeviceArray:
        opts = opts or RenderOptions()
        cache_opts = cache_opts or CacheOptions()
        rgb = out if out is not None else self.ctx.empty((camera.width * camera.height, 3), np.float32)
        cam = camera_struct(camera)
        o = opts.struct(env)
        co = cache_opts.struct(cache_rgb, export)
        self.ctx.check(_capi.lib().mrt_render_cached(self.ctx.handle, self.scene.handle, self.cache.handle, C.byref(cam),
                                                     C.byref(o), C.byref(co), rgb.ptr))
        return rgb


def accumulate(ctx: Context, acc: DeviceArray, rgb: DeviceArray, frames: int, mean: DeviceArray | None = None) -> None:
    """acc += rgb; mean = acc / frames (device, float32)."""
    n = int(np.prod(rgb.shape))
    ctx.check(_capi.lib().mrt_accumulate(ctx.handle, acc.ptr, rgb.ptr, n, frames, mean.ptr if mean is not None else None))


def rel_mse(ctx: Context, img: DeviceArray, ref: DeviceArray, eps: float = 1e-2) -> float:
    """Mean over values of (x - r)^2 / (r^2 + eps), summed in double on the device."""
    out = C.c_double()
    ctx.check(_capi.lib().mrt_rel_mse(ctx.handle, img.ptr, ref.ptr, int(np.prod(img.shape)), eps, C.byref(out)))
    return out.value


def _views(views):
    arr = (_capi.mrt_batch_view * len(views))()
    for i, (cam, out) in enumerate(views):
        arr[i].camera = camera_struct(cam)
        arr[i].rgb_dev = out.ptr
    return arr


class Batch:
    """Several scenes rendered in one frame: concatenated BVHs, one traversal launch per bounce."""

    def __init__(self, ctx: Context, scenes):
        self.ctx = ctx
        handles = (C.c_void_p * len(scenes))(*[s.handle for s in scenes])
        self._h = C.c_void_p()
        ctx.check(_capi.lib().mrt_batch_create(ctx.handle, handles, len(scenes), C.byref(self._h)))
        self.n = len(scenes)

    def info(self) -> dict:
        v = (C.c_uint64 * 4)()
        self.ctx.check(_capi.lib().mrt_batch_info(self.ctx.handle, self._h, v))
        return {"scenes": v[0], "triangles": v[1], "bvh_nodes": v[2], "device_bytes": v[3]}

    def render(self, views, env: Environment, opts: RenderOptions) -> None:
        """views: list of (CameraDesc, DeviceArray rgb[W*H, 3]) in scene order."""
        arr = _views(views)
        o = opts.struct(env)
        self.ctx.check(_capi.lib().mrt_batch_render(self.ctx.handle, self._h, arr, len(views), C.byref(o)))

    def graph(self, views, env: Environment, opts: RenderOptions) -> "FrameGraph":
        return FrameGraph(self, views, env, opts)

    def close(self) -> None:
        if self._h and self.ctx.handle:
            _capi.lib().mrt_batch_destroy(self.ctx.handle, self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FrameGraph:
    """A batch frame captured into a HIP graph; each launch renders the next frame index."""

    def __init__(self, batch: Batch, views, env: Environment, opts: RenderOptions):
        self.batch = batch
        self.ctx = batch.ctx
        arr = _views(views)
        o = opts.struct(env)
        self._h = C.c_void_p()
        self.ctx.check(_capi.lib().mrt_graph_create(self.ctx.handle, batch._h, arr, len(views), C.byref(o),
                                                    C.byref(self._h)))

    def set_frame(self, frame: int) -> None:
        self.ctx.check(_capi.lib().mrt_graph_set_frame(self.ctx.handle, self._h, frame))

    def launch(self) -> None:
        self.ctx.check(_capi.lib().mrt_graph_launch(self.ctx.handle, self._h))

    def close(self) -> None:
        if self._h and self.ctx.handle:
            _capi.lib().mrt_graph_destroy(self.ctx.handle, self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
