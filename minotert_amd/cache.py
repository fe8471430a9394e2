"""Radiance cache handle over the C ABI (hash-grid encoding + MLP, trained online).

No reference counterpart: the reference renderer has no learned component. The cache predicts
the outgoing radiance at a path vertex from its position, incoming direction, normal and albedo;
the path tracer ends paths into it after the first bounce and trains it every frame on path
suffixes it traces itself.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import lib
from .runtime import Context, DeviceArray

SAMPLE_DTYPE = np.dtype([("pos", "<f4", 3), ("pad0", "<f4"), ("dir", "<f4", 3), ("pad1", "<f4"),
                         ("nrm", "<f4", 3), ("pad2", "<f4"), ("alb", "<f4", 3), ("pad3", "<f4")])
WEIGHTS, TABLE, GRAD_WEIGHTS, GRAD_TABLE, AVG_WEIGHTS, AVG_TABLE = 0, 1, 2, 3, 4, 5
LAYERS, WIDTH = 6, 64


class RadianceCache:
    def __init__(self, ctx: Context, lo, hi, levels: int = 16, log2_table: int = 19, base_res: float = 16.0,
                 max_res: float = 2048.0, lr: float = 1e-2, beta1: float = 0.9, beta2: float = 0.99,
                 eps: float = 1e-8, seed: int = 1, ema_decay: float = 0.0, table_lr: float = 0.0):
        self.ctx = ctx
        o = _capi.mrt_cache_opts()
        o.levels, o.log2_table, o.base_res, o.max_res = levels, log2_table, base_res, max_res
        for k in range(3):
            o.lo[k], o.hi[k] = lo[k], hi[k]
        o.lr, o.adam_beta1, o.adam_beta2, o.adam_eps, o.seed = lr, beta1, beta2, eps, seed
        o.ema_decay = ema_decay
        o.table_lr = table_lr
        self.opts = o
        self.levels, self.log2_table = levels, log2_table
        self._h = C.c_void_p()
        ctx.check(lib().mrt_cache_create(ctx.handle, C.byref(o), C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def table_floats(self) -> int:
        return self.levels * (1 << self.log2_table) * 2

    def query(self, samples: DeviceArray, out: DeviceArray, n: int | None = None) -> None:
        n = samples.shape[0] if n is None else n
        self.ctx.check(lib().mrt_cache_query(self.ctx.handle, self._h, samples.ptr, n, out.ptr))

    def inputs(self, samples: DeviceArray, out: DeviceArray, n: int | None = None) -> None:
        n = samples.shape[0] if n is None else n
        self.ctx.check(lib().mrt_cache_inputs(self.ctx.handle, self._h, samples.ptr, n, out.ptr))

    def train(self, samples: DeviceArray, targets: DeviceArray, n: int | None = None, apply: bool = True) -> None:
        n = samples.shape[0] if n is None else n
        self.ctx.check(lib().mrt_cache_train(self.ctx.handle, self._h, samples.ptr, targets.ptr, n, 1 if apply else 0))

    def stats(self) -> dict:
        loss = C.c_float()
        steps = C.c_uint32()
        bad = C.c_uint32()
        self.ctx.check(lib().mrt_cache_stats(self.ctx.handle, self._h, C.byref(loss), C.byref(steps), C.byref(bad)))
        return {"loss_sum": loss.value, "steps": steps.value, "nonfinite": bool(bad.value)}

    def get(self, which: int) -> np.ndarray:
        n = LAYERS * WIDTH * WIDTH if which in (WEIGHTS, GRAD_WEIGHTS, AVG_WEIGHTS) else self.table_floats()
        out = np.empty(n, np.float32)
        self.ctx.check(lib().mrt_cache_get(self.ctx.handle, self._h, which, out.ctypes.data_as(_capi.F32P), n))
        return out

    def set(self, which: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, np.float32).reshape(-1)
        self.ctx.check(lib().mrt_cache_set(self.ctx.handle, self._h, which, v.ctypes.data_as(_capi.F32P), v.size))

    def close(self) -> None:
        if self._h and self.ctx.handle:
            lib().mrt_cache_destroy(self.ctx.handle, self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
