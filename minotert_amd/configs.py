"""Workload recipes: scene descriptions, cameras and render parameters of the five configs.

All values come from ``configs.json`` next to this file. The camera follows the reference's
Camera fields (src/gfx/camera.ixx:8-22): position, yaw, pitch, vertical FoV, viewport; it
is placed at position_over_box * box and aimed at the origin (yaw/pitch from the look
direction, Z up as in camera.ixx:37).
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

_HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(_HERE, "configs.json"), "r", encoding="utf-8") as _f:
    CONFIG = json.load(_f)


def _pcg(v: int) -> int:
    """32-bit PCG hash (random.glsl:22-27); used here only to derive batch layouts."""
    state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return ((word >> 22) ^ word) & 0xFFFFFFFF


@dataclass
class SceneDesc:
    seed: int
    n_spheres: int
    level: int
    ground: int
    box: float
    r_min: float
    r_max: float
    alb_min: float
    alb_max: float
    ground_z: float = 0.0
    ground_half: float = 0.0
    ground_albedo: float = 0.0

    def tri_count(self) -> int:
        return self.n_spheres * 20 * 4 ** self.level + (2 if self.ground else 0)

    def as_tuple(self):
        return (self.seed, self.n_spheres, self.level, self.ground, self.box, self.r_min, self.r_max,
                self.alb_min, self.alb_max, self.ground_z, self.ground_half, self.ground_albedo)


@dataclass
class CameraDesc:
    position: tuple
    yaw: float
    pitch: float
    vertical_fov: float
    width: int
    height: int
    near_plane: float = 0.001


@dataclass
class PathParams:
    frame_seed: int
    sample: int
    bounces: int
    rr_start: int
    sky: tuple = (1.0, 1.0, 1.0)
    sun_dir: tuple = (0.0, 0.0, 1.0)
    sun_e: tuple = (0.0, 0.0, 0.0)
    cache_mode: int = 0
    cache_vertex: int = 1


def config(cid: int) -> dict:
    for c in CONFIG["configs"]:
        if c["id"] == cid:
            return c
    raise KeyError(f"no config {cid}")


def scene_desc(cid: int, *, spheres: int | None = None, box: float | None = None, seed: int | None = None,
               level: int | None = None) -> SceneDesc:
    c = config(cid)
    sp = CONFIG["sphere"]
    b = float(box if box is not None else c["box"])
    ground = bool(c.get("ground", False))
    g = CONFIG["ground"]
    return SceneDesc(
        seed=int(seed if seed is not None else CONFIG["scene_seed_base"] + cid),
        n_spheres=int(spheres if spheres is not None else c["spheres"]),
        level=int(level if level is not None else c["level"]),
        ground=1 if ground else 0,
        box=b, r_min=sp["r_min"], r_max=sp["r_max"], alb_min=sp["alb_min"], alb_max=sp["alb_max"],
        ground_z=(-b - g["z_offset_below_box"]) if ground else 0.0,
        ground_half=(g["half_extent_over_box"] * b) if ground else 0.0,
        ground_albedo=g["albedo"] if ground else 0.0,
    )


def camera(cid: int, *, width: int | None = None, height: int | None = None, box: float | None = None) -> CameraDesc:
    c = config(cid)
    b = float(box if box is not None else c["box"])
    cam = CONFIG["camera"]
    pos = [cam["position_over_box"][k] * b for k in range(3)]
    tgt = cam["target"]
    d = [tgt[k] - pos[k] for k in range(3)]
    yaw = math.atan2(d[1], d[0])
    pitch = math.atan2(d[2], math.hypot(d[0], d[1]))
    return CameraDesc(position=tuple(pos), yaw=yaw, pitch=pitch,
                      vertical_fov=math.radians(cam["vertical_fov_deg"]),
                      width=int(width if width is not None else c["width"]),
                      height=int(height if height is not None else c["height"]))


def frame_seed(frame: int) -> int:
    return (CONFIG["path_seed_base"] ^ frame) & 0xFFFFFFFF


def path_params(cid: int, frame: int = 0, sample: int = 0, cache_mode: int = 0) -> PathParams:
    c = config(cid)
    return PathParams(frame_seed=frame_seed(frame), sample=sample, bounces=int(c.get("bounces", 1)),
                      rr_start=int(c.get("rr_start", 3)), sky=tuple(CONFIG["sky"]),
                      sun_dir=tuple(CONFIG["sun_dir"]), sun_e=tuple(CONFIG["sun_e"]), cache_mode=cache_mode)


@dataclass
class BatchEntry:
    scene: SceneDesc
    width: int
    height: int


def batch_layout(cid: int = 2) -> list:
    """The 16 scenes of the batched workload: triangle counts log-uniform in [tri_min, tri_max]
    (level-3 icospheres, 1280 triangles each), box half-width scaled to keep the sphere
    density of the incoherent workload, resolution drawn from the listed set."""
    c = config(cid)
    base = CONFIG["scene_seed_base"] + cid
    per = 20 * 4 ** c["level"]
    out = []
    lo, hi = math.log(c["tri_min"]), math.log(c["tri_max"])
    for j in range(c["scenes"]):
        u = _pcg(j ^ _pcg(base)) / 4294967296.0
        tris = math.exp(lo + u * (hi - lo))
        spheres = max(1, int(round(tris / per)))
        box = 40.0 * (spheres / 800.0) ** (1.0 / 3.0)
        res = c["resolutions"][_pcg(j ^ _pcg(base ^ 0xA5A5A5A5)) % len(c["resolutions"])]
        sd = scene_desc(1, spheres=spheres, box=box, seed=base + 0x100 * (j + 1), level=c["level"])
        out.append(BatchEntry(scene=sd, width=int(res[0]), height=int(res[1])))
    return out
