"""Cache-vs-plain convergence protocol on the GPU.

A reference image is the mean of many plain samples (a different seed family from the frames
being measured). Each method then renders 1-spp frames that are accumulated into a running mean;
after every frame the relative MSE  mean((x - r)^2 / (r^2 + eps))  against the reference is
taken (mrt_rel_mse, not timed), and the time of the frames themselves (render + accumulate,
synchronized) is summed until the target is reached. The cached method trains the cache inside
its frames, so its time includes training.

No reference counterpart: the reference renders 8 spp per frame with no accumulation or error
metric (secondaryRays.comp:124-134).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import configs
from .cache import RadianceCache
from .configs import CameraDesc
from .pathtracer import (CachedPathtracer, CacheOptions, Environment, Pathtracer, RenderOptions, accumulate,
                         rel_mse)
from .runtime import Context, DeviceArray

REFERENCE_SEED_OFFSET = 1 << 20  # reference frames use frame_seed(REFERENCE_SEED_OFFSET + i)


def environment() -> Environment:
    c = configs.CONFIG
    return Environment(sky=tuple(c["sky"]), sun_dir=tuple(c["sun_dir"]), sun_e=tuple(c["sun_e"]))


def scene_bounds(verts: np.ndarray, margin: float = 1e-3):
    """Axis-aligned bounds of a triangle soup (n x 9 floats), widened by a relative margin."""
    p = np.asarray(verts, np.float64).reshape(-1, 3)
    lo, hi = p.min(axis=0), p.max(axis=0)
    pad = (hi - lo) * margin
    return tuple((lo - pad).tolist()), tuple((hi + pad).tolist())


def make_cache(ctx: Context, bounds, cid: int = 3, seed: int | None = None, **overrides) -> RadianceCache:
    """The radiance cache with the hyper-parameters of the cache workload."""
    c = configs.config(cid)
    kw = dict(levels=int(c["hash_levels"]), log2_table=int(c["hash_log2_table"]), base_res=float(c["hash_base_res"]),
              max_res=float(c["hash_max_res"]), lr=float(c["adam_lr"]), beta1=float(c["adam_beta1"]),
              beta2=float(c["adam_beta2"]), eps=float(c["adam_eps"]), ema_decay=float(c.get("ema_decay", 0.0)),
              table_lr=float(c.get("table_lr", 0.0)),
              seed=int(c["cache_seed"] if seed is None else seed))
    kw.update(overrides)
    return RadianceCache(ctx, bounds[0], bounds[1], **kw)


def cache_options(cid: int = 3, **overrides) -> CacheOptions:
    c = configs.config(cid)
    o = CacheOptions(enable=True, vertex=int(c["cache_vertex"]), train_samples=int(c["train_suffixes"]),
                     train_steps=int(c["train_steps"]), overlap_training=bool(c.get("overlap_training", False)))
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


def frame_options(cid: int, frame: int, spp: int = 1, sample_base: int = 0) -> RenderOptions:
    c = configs.config(cid)
    return RenderOptions(frame_seed=configs.frame_seed(frame), sample_base=sample_base, spp=spp,
                         bounces=int(c["bounces"]), rr_start=int(c["rr_start"]))


def render_reference(pt: Pathtracer, cam: CameraDesc, env: Environment, spp: int, cid: int = 3,
                     chunk: int = 16) -> DeviceArray:
    """Mean of `spp` plain samples per pixel (chunks of `chunk` samples per launch)."""
    ctx = pt.ctx
    n = cam.width * cam.height
    acc = ctx.to_device(np.zeros((n, 3), np.float32))
    mean = ctx.empty((n, 3), np.float32)
    rgb = ctx.empty((n, 3), np.float32)
    chunk = max(1, min(chunk, spp))
    calls = max(1, spp // chunk)
    for i in range(calls):
        pt.render(cam, env, frame_options(cid, REFERENCE_SEED_OFFSET + i, spp=chunk), out=rgb)
        accumulate(ctx, acc, rgb, i + 1, mean)
    ctx.sync()
    acc.free()
    rgb.free()
    return mean


@dataclass
class Convergence:
    frames: int = 0
    ms: float = 0.0                   # summed frame time until the target (or the frame limit)
    reached: bool = False
    relmse: list = field(default_factory=list)
    frame_ms: list = field(default_factory=list)

    @property
    def ms_per_frame(self) -> float:
        return self.ms / max(1, self.frames)


def converge(ctx: Context, render_frame, ref: DeviceArray, n: int, target: float = 0.01, max_frames: int = 512,
             eps: float = 1e-2, warmup: int = 0, max_ms: float | None = None) -> Convergence:
    """Runs render_frame(f, out) for f = 0, 1, ... and accumulates frames >= warmup into a running
    mean until relMSE(mean, ref) <= target. Warm-up frames are timed but not accumulated."""
    acc = ctx.to_device(np.zeros((n, 3), np.float32))
    mean = ctx.empty((n, 3), np.float32)
    rgb = ctx.empty((n, 3), np.float32)
    out = Convergence()
    ctx.sync()
    for f in range(max_frames):
        t0 = time.perf_counter()
        render_frame(f, rgb)
        if f >= warmup:
            accumulate(ctx, acc, rgb, f - warmup + 1, mean)
        ctx.sync()
        dt = (time.perf_counter() - t0) * 1e3
        out.frames += 1
        out.ms += dt
        out.frame_ms.append(dt)
        if f >= warmup:
            r = rel_mse(ctx, mean, ref, eps)
            out.relmse.append(r)
            if r <= target:
                out.reached = True
                break
        if max_ms is not None and out.ms >= max_ms:
            break
    acc.free()
    mean.free()
    rgb.free()
    return out


def plain_frames(pt: Pathtracer, cam: CameraDesc, env: Environment, cid: int = 3):
    def render(f, out):
        pt.render(cam, env, frame_options(cid, f), out=out)
    return render


def cached_frames(cpt: CachedPathtracer, cam: CameraDesc, env: Environment, copts: CacheOptions, cid: int = 3):
    def render(f, out):
        cpt.render(cam, env, frame_options(cid, f), out=out, cache_opts=copts)
    return render
