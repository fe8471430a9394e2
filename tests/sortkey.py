"""Closed form of the ray-order key: direction octant (3 bits: x, y, z negative -> 4, 2, 1) above
a 27-bit Morton code of the origin quantized to 512 cells per axis inside the scene bounds
(bit 3k+2 = bit k of the x cell, 3k+1 of y, 3k of z). Cells are trunc(clamp((o - lo) * s, 0,
511)) with s = float32(512 / float32(hi - lo)), in float32 as the device evaluates them."""
from __future__ import annotations

import numpy as np


def scene_bounds(verts: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    v = np.asarray(verts, np.float32).reshape(-1, 3)
    return v.min(axis=0), v.max(axis=0)


def cells(o: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    ext = (hi - lo).astype(np.float32)
    scale = np.where(ext > 0, np.float32(512.0) / np.where(ext > 0, ext, np.float32(1)), np.float32(0)).astype(np.float32)
    q = (o.astype(np.float32) - lo.astype(np.float32)) * scale
    q = np.minimum(np.maximum(q, np.float32(0)), np.float32(511))
    return q.astype(np.uint32)


def morton3(cx, cy, cz, bits: int = 9, order=(2, 1, 0)) -> np.ndarray:
    m = np.zeros(cx.shape, np.uint64)
    for k in range(bits):
        for c, sh in zip((cx, cy, cz), order):
            m |= ((c.astype(np.uint64) >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k + sh)
    return m


def octant(d: np.ndarray, weights=(4, 2, 1)) -> np.ndarray:
    return ((d[:, 0] < 0) * weights[0] + (d[:, 1] < 0) * weights[1] + (d[:, 2] < 0) * weights[2]).astype(np.uint64)


def ray_keys(rays: np.ndarray, lo, hi, plant: str | None = None) -> np.ndarray:
    """rays: (n, 8) float32 {o, tmin, d, tmax}. `plant` swaps in a plausible mistake (for the
    checks that the order test catches each one): 'xz' interleave, 'octant' bit order, 'bits10'
    quantization, 'no_octant'."""
    c = cells(rays[:, 0:3], np.asarray(lo, np.float32), np.asarray(hi, np.float32))
    order = (0, 1, 2) if plant == "xz" else (2, 1, 0)
    m = morton3(c[:, 0], c[:, 1], c[:, 2], order=order)
    if plant == "bits10":
        c10 = np.minimum((c.astype(np.uint64) << np.uint64(1)), np.uint64(1023))
        m = morton3(c10[:, 0], c10[:, 1], c10[:, 2], bits=9)
    w = (1, 2, 4) if plant == "octant" else (4, 2, 1)
    oc = octant(rays[:, 4:7], w)
    if plant == "no_octant":
        oc = np.zeros_like(oc)
    return (oc << np.uint64(27)) | m


def expected_order(rays: np.ndarray, lo, hi, plant: str | None = None) -> np.ndarray:
    return np.argsort(ray_keys(rays, lo, hi, plant), kind="stable").astype(np.uint32)


def hand_made_rays(lo: np.ndarray, hi: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Rays at chosen cells and octants, in an order unlike their keys, with the key each must
    have: the centre of cell (cx, cy, cz) and a direction in octant o."""
    cell_list = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (511, 511, 511), (3, 5, 6), (256, 0, 1), (1, 256, 0),
                 (0, 1, 256), (510, 2, 77)]
    ext = (hi - lo).astype(np.float64)
    rows, keys = [], []
    for o in range(8):
        d = np.array([-1.0 if o & 4 else 1.0, -1.0 if o & 2 else 1.0, -1.0 if o & 1 else 1.0]) / np.sqrt(3.0)
        for cx, cy, cz in cell_list:
            p = lo.astype(np.float64) + (np.array([cx, cy, cz]) + 0.5) * ext / 512.0
            rows.append(np.r_[p, 0.0, d, np.inf].astype(np.float32))
            m = 0
            for k in range(9):
                m |= ((cx >> k) & 1) << (3 * k + 2) | ((cy >> k) & 1) << (3 * k + 1) | ((cz >> k) & 1) << (3 * k)
            keys.append((o << 27) | m)
    rows = np.array(rows, np.float32)
    keys = np.array(keys, np.uint64)
    perm = np.random.default_rng(5).permutation(len(rows))
    return rows[perm], keys[perm]
