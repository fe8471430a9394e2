"""The reference's camera matrices in float64 and the motion vectors they give.

view and projection exactly as look() and perspective() (src/stx/math.ixx:822-860) with the
camera direction of src/gfx/camera.ixx:26-44; column-major mat[c][r] -> numpy M[r, c].
Motion (src/gpu/primaryRay.comp:73-75): (ndc(P V p) - ndc(P V_prev p)).xy / OutPitch, with
OutPitch = (1/W, 1/H), the current projection for both, V_prev = prevCamera.view()
(src/gfx/modules/pathtracer.ixx:91,102)."""
from __future__ import annotations

import math

import numpy as np


def reference_matrices(cam):
    d = np.array([math.cos(cam.pitch) * math.cos(cam.yaw), math.cos(cam.pitch) * math.sin(cam.yaw),
                  math.sin(cam.pitch)])
    up = np.array([0.0, 0.0, 1.0])
    s = np.cross(up, d); s /= np.linalg.norm(s)
    u = np.cross(d, s)
    pos = np.array(cam.position, np.float64)
    V = np.eye(4)
    V[0, :3] = -s; V[1, :3] = u; V[2, :3] = d
    V[0, 3] = s @ pos; V[1, 3] = -(u @ pos); V[2, 3] = -(d @ pos)
    h = 1.0 / math.tan(0.5 * cam.vertical_fov)
    w = h * cam.height / cam.width
    Pm = np.zeros((4, 4))
    Pm[0, 0] = w; Pm[1, 1] = h; Pm[3, 2] = 1.0; Pm[2, 3] = getattr(cam, "near_plane", 0.001)
    return V, Pm


def ndc(M: np.ndarray, p: np.ndarray) -> np.ndarray:
    """(n, 3) positions -> (n, 3) NDC of clip = M [p, 1] divided by w."""
    c = np.c_[p, np.ones(p.shape[0])] @ M.T
    return c[:, :3] / c[:, 3:4]


def reference_motion(cam, prev, p: np.ndarray) -> np.ndarray:
    """Motion in pixels of world positions p (n, 3) between prev and cam (float64)."""
    V, Pm = reference_matrices(cam)
    Vp, _ = reference_matrices(prev)
    a = ndc(Pm @ V, p)
    b = ndc(Pm @ Vp, p)
    return (a[:, :2] - b[:, :2]) * np.array([cam.width, cam.height], np.float64)


def hit_positions(verts: np.ndarray, hits) -> tuple[np.ndarray, np.ndarray]:
    """(hit mask, float64 positions v0 + u (v1 - v0) + v (v2 - v0) of the hit pixels)."""
    hit = hits["tri"] != 0xFFFFFFFF
    v = verts.astype(np.float64)[hits["tri"][hit]]
    u_, v_ = hits["u"][hit].astype(np.float64), hits["v"][hit].astype(np.float64)
    p = v[:, 0:3] + u_[:, None] * (v[:, 3:6] - v[:, 0:3]) + v_[:, None] * (v[:, 6:9] - v[:, 0:3])
    return hit, p


def motion_close(got: np.ndarray, want: np.ndarray) -> bool:
    """The GPU's float32 motion against the float64 reference: 2e-3 pixels + 1e-4 relative."""
    return bool(np.all(np.abs(got - want) <= 2e-3 + 1e-4 * np.abs(want)))
