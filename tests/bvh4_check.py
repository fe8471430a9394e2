"""Structural validation of an exported BVH4 (numpy, level by level from the root).

Node layout (128 B, 32 words): six float rows lo.x, hi.x, lo.y, hi.y, lo.z, hi.z over the four
child slots, four child words, four spare words. Child word: internal child = node index, leaf =
0x80000000 | first << 3 | (count - 1), empty slot = 0xFFFFFFFF. Leaf records are 12 words
{v0, id}, {v1, 0}, {v2, 0}.

Checked:
  * every node is reached from the root exactly once (a tree: no sharing, nothing unreachable);
  * every leaf-record index is covered by exactly one leaf, and the record ids are a permutation
    of 0..n-1 whose vertices equal the soup's triangle bit for bit;
  * leaves hold 1..max_leaf triangles;
  * containment: a child node's slot boxes lie inside its slot box in the parent, and a leaf's
    triangles lie inside the leaf's slot box;
  * tightness (builder-specific): each slot box is the exact min/max of what it holds.
"""
from __future__ import annotations

import numpy as np

EMPTY = 0xFFFFFFFF
LEAF = 0x80000000


def _slot_boxes(nodes: np.ndarray) -> np.ndarray:
    """(N, 4, 2, 3) float32: [node, slot, lo/hi, axis]."""
    rows = nodes[:, :24].view(np.float32).reshape(-1, 6, 4)  # lox hix loy hiy loz hiz
    lo = np.stack([rows[:, 0], rows[:, 2], rows[:, 4]], axis=-1)  # (N, 4, 3)
    hi = np.stack([rows[:, 1], rows[:, 3], rows[:, 5]], axis=-1)
    return np.stack([lo, hi], axis=2)


def validate_bvh4(nodes: np.ndarray, tris: np.ndarray, verts: np.ndarray, max_leaf: int = 4,
                  tight: bool = True) -> dict:
    """Raises AssertionError naming the first violated invariant; returns counts on success."""
    nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1, 32)
    tris = np.ascontiguousarray(tris, np.uint32).reshape(-1, 12)
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
    n_nodes, n_tris = nodes.shape[0], verts.shape[0]
    assert tris.shape[0] == n_tris, "one leaf record per triangle"
    assert n_nodes >= 1, "a BVH4 has a root node"
    boxes = _slot_boxes(nodes)
    child = nodes[:, 24:28]
    visits = np.zeros(n_nodes, np.int64)
    cover = np.zeros(n_tris, np.int64)
    rec_v = tris.view(np.float32).reshape(-1, 3, 4)[:, :, :3]  # (n, 3 vertices, xyz)
    frontier = np.array([0], np.int64)
    depth = 0
    leaves = 0
    while frontier.size:
        depth += 1
        visits[frontier] += 1
        assert np.all(visits[frontier] == 1), "a node is reached twice (not a tree)"
        cw = child[frontier]                      # (F, 4)
        pb = boxes[frontier]                      # (F, 4, 2, 3)
        used = cw != EMPTY
        is_leaf = used & ((cw & LEAF) != 0)
        is_int = used & ((cw & LEAF) == 0)
        # ---- internal children ----
        fi, si = np.nonzero(is_int)
        kids = cw[fi, si].astype(np.int64)
        assert np.all(kids < n_nodes), "a child index points past the node array"
        if kids.size:
            cb = boxes[kids]                      # (K, 4, 2, 3)
            kused = child[kids] != EMPTY          # (K, 4)
            assert np.all(kused.any(axis=1)), "an internal node without children"
            plo = pb[fi, si, 0][:, None, :]
            phi = pb[fi, si, 1][:, None, :]
            inside = (cb[:, :, 0] >= plo).all(-1) & (cb[:, :, 1] <= phi).all(-1)
            assert np.all(inside | ~kused), "a child node's box is not inside its slot box in the parent"
            if tight:
                big = np.float32(np.inf)
                ulo = np.where(kused[..., None], cb[:, :, 0], big).min(axis=1)
                uhi = np.where(kused[..., None], cb[:, :, 1], -big).max(axis=1)
                assert np.array_equal(ulo, pb[fi, si, 0]) and np.array_equal(uhi, pb[fi, si, 1]), \
                    "a slot box is not the union of its child node's boxes"
        # ---- leaves ----
        fl, sl = np.nonzero(is_leaf)
        lw = cw[fl, sl]
        first = ((lw >> 3) & 0x0FFFFFFF).astype(np.int64)
        count = (lw & 7).astype(np.int64) + 1
        assert np.all(count <= max_leaf), f"a leaf holds more than {max_leaf} triangles"
        assert np.all(first + count <= n_tris), "a leaf range points past the triangle records"
        leaves += lw.size
        for k in range(int(count.max()) if count.size else 0):
            m = count > k
            rec = first[m] + k
            np.add.at(cover, rec, 1)
            v = rec_v[rec]                        # (L, 3, 3)
            lo = pb[fl[m], sl[m], 0][:, None, :]
            hi = pb[fl[m], sl[m], 1][:, None, :]
            assert np.all((v >= lo) & (v <= hi)), "a leaf triangle lies outside the leaf's slot box"
        if tight and lw.size:
            big = np.float32(np.inf)
            ulo = np.full((lw.size, 3), big, np.float32)
            uhi = np.full((lw.size, 3), -big, np.float32)
            for k in range(int(count.max())):
                m = count > k
                v = rec_v[first[m] + k]
                ulo[m] = np.minimum(ulo[m], v.min(axis=1))
                uhi[m] = np.maximum(uhi[m], v.max(axis=1))
            assert np.array_equal(ulo, pb[fl, sl, 0]) and np.array_equal(uhi, pb[fl, sl, 1]), \
                "a leaf slot box is not the bounding box of its triangles"
        frontier = kids
    assert np.all(visits == 1), f"{int(np.sum(visits == 0))} nodes are unreachable from the root"
    assert np.all(cover == 1), "a leaf record is covered by no leaf or by several"
    ids = tris[:, 3].astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(n_tris)), "leaf record ids are not a permutation of the triangles"
    want = verts.reshape(-1, 3, 3)[ids]
    assert np.array_equal(rec_v.view(np.uint32), np.ascontiguousarray(want).view(np.uint32)), \
        "a leaf record's vertices differ from its triangle's"
    for w in (1, 2):
        assert not np.any(tris.reshape(-1, 3, 4)[:, w, 3]), "the spare word of a record is not zero"
    return {"nodes": n_nodes, "depth": depth, "leaves": leaves}
