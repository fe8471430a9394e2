"""CPU checks that the closed forms the GPU tests compare against have the power to catch
plausible mistakes: each planted mistake (in a numpy restatement, on the CPU) must fail the
same comparison the GPU test makes.

  * motion vectors (primaryRay.comp:73-75): the kernel's formula, restated in float64 from the
    float32 camera basis, passes the float64 matrix-route comparison; a flipped sign, swapped
    pitches, the current view for the previous frame, the previous camera's own projection, or
    a flipped y axis each fail it;
  * ray-order keys: hand-worked keys equal the closed form; a swapped interleave, a reversed
    octant bit order, a 10-bit quantization or a missing octant each change the order of a
    random ray set."""
import math

import numpy as np
import pytest

import minotert_amd as mrt
from refcam import hit_positions, motion_close, reference_motion
from sortkey import expected_order, hand_made_rays, ray_keys, scene_bounds


def _kernel_motion(orc, cam, prev, p, plant=None):
    """k_gbuffer_planes' formula: view coordinates from each camera's float32 basis, NDC with
    the current camera's proj_w = h H / W, proj_h = h, difference over the pixel pitch."""
    bn = orc.camera_basis(cam).astype(np.float64)
    bp = orc.camera_basis(prev if plant != "current_view" else cam).astype(np.float64)

    def proj(b, c):
        h = 1.0 / math.tan(0.5 * c.vertical_fov)
        w = h * c.height / c.width
        q = p - b[0:3]
        vx, vy, vz = q @ b[3:6], q @ b[6:9], q @ b[9:12]
        return w * vx / vz, (-1.0 if plant == "flip_y" else 1.0) * h * vy / vz

    x0, y0 = proj(bn, cam)
    x1, y1 = proj(bp, prev if plant == "prev_projection" else cam)
    m = np.stack([x0 - x1, y0 - y1], axis=1)
    if plant == "sign":
        m = -m
    scale = np.array([cam.height, cam.width] if plant == "pitch" else [cam.width, cam.height], np.float64)
    return m * scale


@pytest.fixture(scope="module")
def motion_case(orc, cfg):
    verts, _, _ = mrt.generate_scene(cfg.scene_desc(0))
    cam = cfg.camera(0, width=120, height=80)
    prev = cfg.camera(0, width=120, height=80)
    prev.yaw += 0.013
    prev.pitch -= 0.007
    prev.position = tuple(np.array(prev.position) + np.array([0.21, -0.13, 0.08]))
    prev.vertical_fov *= 0.8
    hits = orc.closest_brute(verts, orc.primary_rays(cam))
    hit, p = hit_positions(verts, hits)
    return cam, prev, p, reference_motion(cam, prev, p)


def test_kernel_motion_formula_matches_reference(orc, motion_case):
    cam, prev, p, want = motion_case
    assert p.shape[0] > 500 and np.abs(want).max() > 2.0
    assert motion_close(_kernel_motion(orc, cam, prev, p).astype(np.float32), want)


@pytest.mark.parametrize("plant", ["sign", "pitch", "current_view", "prev_projection", "flip_y"])
def test_planted_motion_mistake_is_caught(orc, motion_case, plant):
    cam, prev, p, want = motion_case
    assert not motion_close(_kernel_motion(orc, cam, prev, p, plant).astype(np.float32), want)


def test_hand_worked_keys_equal_closed_form(cfg):
    verts, _, _ = mrt.generate_scene(cfg.scene_desc(1))
    lo, hi = scene_bounds(verts)
    rows, keys = hand_made_rays(lo, hi)
    assert np.array_equal(ray_keys(rows, lo, hi), keys)
    assert len(np.unique(keys)) == len(keys)


@pytest.mark.parametrize("plant", ["xz", "octant", "bits10", "no_octant"])
def test_planted_key_mistake_changes_the_order(orc, cfg, plant):
    verts, _, _ = mrt.generate_scene(cfg.scene_desc(1))
    lo, hi = scene_bounds(verts)
    rays = orc.random_rays(91, 44.0, 200000)
    assert not np.array_equal(expected_order(rays, lo, hi, plant), expected_order(rays, lo, hi))
    rows, keys = hand_made_rays(lo, hi)
    assert not np.array_equal(ray_keys(rows, lo, hi, plant), keys)
