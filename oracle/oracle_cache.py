"""ctypes binding of the oracle's radiance-cache functions (oracle/oracle_cache.cpp).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the GPU cache kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .oracle import lib as _lib

P = C.c_void_p


class orc_grid(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("log2_table", C.c_uint32), ("base_res", C.c_float), ("max_res", C.c_float),
                ("lo", C.c_float * 3), ("hi", C.c_float * 3)]


_BOUND = False


def lib():
    global _BOUND
    L = _lib()
    if not _BOUND:
        L.orc_bf16_round.restype = C.c_float
        L.orc_bf16_round.argtypes = [C.c_float]
        L.orc_cache_encode.argtypes = [C.POINTER(orc_grid), P, P, P, P, P, C.c_uint32, C.c_float, P]
        L.orc_mlp_forward.argtypes = [P, P, C.c_uint32, P, P]
        L.orc_mlp_backward.restype = C.c_double
        L.orc_mlp_backward.argtypes = [P, P, C.c_uint32, P, P, P]
        L.orc_cache_encode_backward.argtypes = [C.POINTER(orc_grid), P, P, C.c_uint32, P]
        L.orc_adam.argtypes = [P, P, P, P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]
        _BOUND = True
    return L


def _p(a):
    return C.c_void_p(a.ctypes.data)


def grid(levels, log2_table, base_res, max_res, lo, hi) -> orc_grid:
    g = orc_grid()
    g.levels, g.log2_table, g.base_res, g.max_res = levels, log2_table, base_res, max_res
    for k in range(3):
        g.lo[k], g.hi[k] = lo[k], hi[k]
    return g


def bf16(x) -> np.ndarray:
    """Round-to-nearest-even to BF16, returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounded = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    special = (u & 0x7F800000) == 0x7F800000
    r = np.where(special, u & 0xFFFF0000, rounded).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def encode(g, table, samples, scale: float = 1.0) -> np.ndarray:
    """MLP inputs: bf16(scale * trilinear features) per level, then direction, normal, albedo."""
    n = samples.shape[0]
    pos = np.ascontiguousarray(samples["pos"], np.float32)
    dr = np.ascontiguousarray(samples["dir"], np.float32)
    nr = np.ascontiguousarray(samples["nrm"], np.float32)
    al = np.ascontiguousarray(samples["alb"], np.float32)
    t = np.ascontiguousarray(table, np.float32)
    out = np.empty((n, 64), np.float32)
    lib().orc_cache_encode(C.byref(g), _p(t), _p(pos), _p(dr), _p(nr), _p(al), n, scale, _p(out))
    return out


def forward(W, x0, with_acts=False):
    W = np.ascontiguousarray(W, np.float32)
    x0 = np.ascontiguousarray(x0, np.float32)
    n = x0.shape[0]
    y = np.empty((n, 4), np.float32)
    acts = np.empty((5, n, 64), np.float32) if with_acts else None
    lib().orc_mlp_forward(_p(W), _p(x0), n, _p(y), _p(acts) if with_acts else None)
    return (y, acts) if with_acts else y


def backward(W, x0, target):
    W = np.ascontiguousarray(W, np.float32)
    x0 = np.ascontiguousarray(x0, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    n = x0.shape[0]
    dW = np.empty(W.size, np.float32)
    dx0 = np.empty((n, 64), np.float32)
    loss = lib().orc_mlp_backward(_p(W), _p(x0), n, _p(target), _p(dW), _p(dx0))
    return loss, dW, dx0


def encode_backward(g, samples, dx0, table_floats) -> np.ndarray:
    pos = np.ascontiguousarray(samples["pos"], np.float32)
    dx0 = np.ascontiguousarray(dx0, np.float32)
    out = np.zeros(table_floats, np.float32)
    lib().orc_cache_encode_backward(C.byref(g), _p(pos), _p(dx0), pos.shape[0], _p(out))
    return out


def adam(p, g, m, v, step, lr, b1, b2, eps, sparse=False):
    p = np.ascontiguousarray(p, np.float32).copy()
    m = np.ascontiguousarray(m, np.float32).copy()
    v = np.ascontiguousarray(v, np.float32).copy()
    g = np.ascontiguousarray(g, np.float32)
    lib().orc_adam(_p(p), _p(g), _p(m), _p(v), p.size, step, lr, b1, b2, eps, 1 if sparse else 0)
    return p, m, v
