"""ctypes binding of the scalar CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg, as the checker. The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MRT_ORACLE_LIB selects another build of the same sources (the sanitizer build of `make SAN=1`).
LIB_PATH = os.environ.get("MRT_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")


class orc_camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("yaw", C.c_float), ("pitch", C.c_float), ("vfov", C.c_float),
                ("width", C.c_uint32), ("height", C.c_uint32)]


class orc_scene_desc(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("n_spheres", C.c_uint32), ("level", C.c_uint32), ("ground", C.c_uint32),
                ("box", C.c_float), ("r_min", C.c_float), ("r_max", C.c_float), ("alb_min", C.c_float),
                ("alb_max", C.c_float), ("ground_z", C.c_float), ("ground_half", C.c_float),
                ("ground_albedo", C.c_float)]


class orc_path_params(C.Structure):
    _fields_ = [("frame_seed", C.c_uint32), ("sample", C.c_uint32), ("bounces", C.c_uint32),
                ("rr_start", C.c_uint32), ("sky", C.c_float * 3), ("sun_dir", C.c_float * 3),
                ("sun_e", C.c_float * 3), ("cache_mode", C.c_uint32), ("cache_vertex", C.c_uint32),
                ("sky_state", C.c_void_p)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        _LIB = C.CDLL(LIB_PATH)
        L = _LIB
        P = C.c_void_p
        L.orc_pcg.restype = C.c_uint32; L.orc_pcg.argtypes = [C.c_uint32]
        L.orc_rand_float.restype = C.c_float; L.orc_rand_float.argtypes = [C.POINTER(C.c_uint32)]
        L.orc_sincos_turns.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_cosine_dir.argtypes = [P, C.c_float, C.c_float, P]
        L.orc_camera_basis.argtypes = [C.POINTER(orc_camera), P]
        L.orc_primary_rays.argtypes = [C.POINTER(orc_camera), P]
        L.orc_offset_ray.argtypes = [P, P, P]
        L.orc_scene_tri_count.restype = C.c_uint32; L.orc_scene_tri_count.argtypes = [C.POINTER(orc_scene_desc)]
        L.orc_scene_generate.restype = C.c_int
        L.orc_scene_generate.argtypes = [C.POINTER(orc_scene_desc), P, P, P]
        L.orc_intersect.restype = C.c_int; L.orc_intersect.argtypes = [P, P, C.c_uint32, P]
        L.orc_closest_brute.argtypes = [P, C.c_uint32, P, C.c_uint32, P, C.c_int]
        L.orc_bvh_build.restype = P; L.orc_bvh_build.argtypes = [P, C.c_uint32]
        L.orc_bvh_free.argtypes = [P]
        L.orc_bvh_info.argtypes = [P, P, C.POINTER(C.c_double)]
        L.orc_bvh_validate.restype = C.c_int; L.orc_bvh_validate.argtypes = [P]
        L.orc_trace_closest.argtypes = [P, P, C.c_uint32, P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_int]
        L.orc_trace_any.argtypes = [P, P, C.c_uint32, P, C.c_int]
        L.orc_render_lambert.argtypes = [P, P, P, C.POINTER(orc_camera), P, P, P, P, C.c_int]
        L.orc_render_path.argtypes = [P, P, P, C.POINTER(orc_camera), C.POINTER(orc_path_params), P, P, C.c_int]
        L.orc_bounce1_rays.restype = C.c_uint32
        L.orc_path_suffix.argtypes = [P, P, P, C.POINTER(orc_camera), C.POINTER(orc_path_params), P, C.c_uint32, P, P, P,
                                      C.c_int]
        L.orc_bounce1_rays.argtypes = [P, C.POINTER(orc_camera), C.POINTER(orc_path_params), P, P, C.c_int]
        L.orc_random_rays.argtypes = [C.c_uint32, C.c_float, C.c_uint32, P]
        L.orc_rel_mse.restype = C.c_double; L.orc_rel_mse.argtypes = [P, P, C.c_size_t, C.c_double]
        L.orc_hardware_threads.restype = C.c_int
    return _LIB


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def f32(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.float32)


def camera(c) -> orc_camera:
    o = orc_camera()
    for k in range(3):
        o.pos[k] = c.position[k]
    o.yaw, o.pitch, o.vfov, o.width, o.height = c.yaw, c.pitch, c.vertical_fov, c.width, c.height
    return o


def scene_desc(d) -> orc_scene_desc:
    return orc_scene_desc(*d.as_tuple())


def path_params(p) -> orc_path_params:
    o = orc_path_params()
    o.frame_seed, o.sample, o.bounces, o.rr_start, o.cache_mode = p.frame_seed, p.sample, p.bounces, p.rr_start, p.cache_mode
    o.cache_vertex = getattr(p, "cache_vertex", 1)
    st = getattr(p, "sky_state", None)  # an oracle_sky.SkyState kept alive by the caller
    o.sky_state = C.addressof(st.struct) if st is not None else None
    for k in range(3):
        o.sky[k], o.sun_dir[k], o.sun_e[k] = p.sky[k], p.sun_dir[k], p.sun_e[k]
    return o


def pcg(v: int) -> int:
    return lib().orc_pcg(v & 0xFFFFFFFF)


def sincos_turns(x: float):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos_turns(x, C.byref(s), C.byref(c))
    return s.value, c.value


def cosine_dir(n, xi1, xi2) -> np.ndarray:
    n = f32(n); out = np.empty(3, np.float32)
    lib().orc_cosine_dir(_p(n), xi1, xi2, _p(out))
    return out


def camera_basis(c) -> np.ndarray:
    out = np.empty(14, np.float32)
    cam = camera(c)
    lib().orc_camera_basis(C.byref(cam), _p(out))
    return out


def primary_rays(c) -> np.ndarray:
    out = np.empty((c.width * c.height, 8), np.float32)
    cam = camera(c)
    lib().orc_primary_rays(C.byref(cam), _p(out))
    return out


def offset_ray(p, n) -> np.ndarray:
    p = f32(p); n = f32(n); out = np.empty(3, np.float32)
    lib().orc_offset_ray(_p(p), _p(n), _p(out))
    return out


def generate_scene(d):
    sd = scene_desc(d)
    n = lib().orc_scene_tri_count(C.byref(sd))
    nm = d.n_spheres + (1 if d.ground else 0)
    verts = np.empty((n, 9), np.float32)
    mat = np.empty(n, np.uint32)
    alb = np.empty((max(nm, 1), 3), np.float32)
    lib().orc_scene_generate(C.byref(sd), _p(verts), _p(mat), _p(alb))
    return verts, mat, alb[:nm]


def intersect(ray, tri, tid: int = 0):
    ray = f32(ray); tri = f32(tri); hit = np.zeros(4, np.float32)
    h = lib().orc_intersect(_p(ray), _p(tri), tid, _p(hit))
    return bool(h), hit


def hits_view(raw: np.ndarray) -> np.ndarray:
    """(n,4) u32 hit words -> structured (t, tri, u, v)."""
    dt = np.dtype([("t", "<f4"), ("tri", "<u4"), ("u", "<f4"), ("v", "<f4")])
    return np.ascontiguousarray(raw).view(dt).reshape(-1)


def closest_brute(verts, rays, threads: int = 0) -> np.ndarray:
    verts = f32(verts).reshape(-1, 9); rays = f32(rays).reshape(-1, 8)
    out = np.empty((rays.shape[0], 4), np.uint32)
    lib().orc_closest_brute(_p(verts), verts.shape[0], _p(rays), rays.shape[0], _p(out), threads)
    return hits_view(out)


class BVH:
    def __init__(self, verts):
        self.verts = f32(verts).reshape(-1, 9)
        self._h = C.c_void_p(lib().orc_bvh_build(_p(self.verts), self.verts.shape[0]))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_bvh_free(self._h)
            self._h = None

    def info(self) -> dict:
        info = np.zeros(4, np.uint32); sah = C.c_double()
        lib().orc_bvh_info(self._h, _p(info), C.byref(sah))
        return {"nodes": int(info[0]), "leaves": int(info[1]), "max_depth": int(info[2]),
                "max_leaf": int(info[3]), "sah": sah.value}

    def validate(self) -> int:
        return lib().orc_bvh_validate(self._h)

    def trace_closest(self, rays, threads: int = 0, with_stats: bool = False):
        rays = f32(rays).reshape(-1, 8)
        out = np.empty((rays.shape[0], 4), np.uint32)
        nb, nt = C.c_uint64(), C.c_uint64()
        lib().orc_trace_closest(self._h, _p(rays), rays.shape[0], _p(out), C.byref(nb), C.byref(nt), threads)
        h = hits_view(out)
        if with_stats:
            return h, nb.value, nt.value
        return h

    def trace_any(self, rays, threads: int = 0) -> np.ndarray:
        rays = f32(rays).reshape(-1, 8)
        out = np.empty(rays.shape[0], np.uint8)
        lib().orc_trace_any(self._h, _p(rays), rays.shape[0], _p(out), threads)
        return out

    def render_lambert(self, mat, albedo, cam, light_dir, light_e, threads: int = 0):
        mat = np.ascontiguousarray(mat, np.uint32); albedo = f32(albedo)
        ldir = f32(light_dir); le = f32(light_e)  # kept alive for the duration of the call
        n = cam.width * cam.height
        hits = np.empty((n, 4), np.uint32); rgb = np.empty((n, 3), np.float32)
        oc = camera(cam)
        lib().orc_render_lambert(self._h, _p(mat), _p(albedo), C.byref(oc), _p(ldir), _p(le),
                                 _p(hits), _p(rgb), threads)
        return hits_view(hits), rgb

    def render_path(self, mat, albedo, cam, params, cache_rgb=None, threads: int = 0) -> np.ndarray:
        mat = np.ascontiguousarray(mat, np.uint32); albedo = f32(albedo)
        n = cam.width * cam.height
        rgb = np.empty((n, 3), np.float32)
        oc = camera(cam); pp = path_params(params)
        cr = None if cache_rgb is None else f32(cache_rgb)
        lib().orc_render_path(self._h, _p(mat), _p(albedo), C.byref(oc), C.byref(pp),
                              None if cr is None else _p(cr), _p(rgb), threads)
        return rgb

    def path_suffix(self, mat, albedo, cam, params, pids, threads: int = 0):
        """Cache samples (n x 16), targets (n x 4) and validity of plain paths at params.cache_vertex."""
        mat = np.ascontiguousarray(mat, np.uint32); albedo = f32(albedo)
        pids = np.ascontiguousarray(pids, np.uint32)
        n = pids.size
        samples = np.zeros((n, 16), np.float32); targets = np.zeros((n, 4), np.float32); valid = np.zeros(n, np.uint8)
        oc = camera(cam); pp = path_params(params)
        lib().orc_path_suffix(self._h, _p(mat), _p(albedo), C.byref(oc), C.byref(pp), _p(pids), n, _p(samples),
                              _p(targets), _p(valid), threads)
        return samples, targets, valid.astype(bool)

    def bounce1_rays(self, cam, params, threads: int = 0):
        n = cam.width * cam.height
        rays = np.empty((n, 8), np.float32); pix = np.empty(n, np.uint32)
        oc = camera(cam); pp = path_params(params)
        cnt = lib().orc_bounce1_rays(self._h, C.byref(oc), C.byref(pp), _p(rays), _p(pix), threads)
        return rays[:cnt].copy(), pix[:cnt].copy()


def random_rays(seed: int, box: float, n: int) -> np.ndarray:
    out = np.empty((n, 8), np.float32)
    lib().orc_random_rays(seed, box, n, _p(out))
    return out


def rel_mse(img, ref, eps: float = 1e-2) -> float:
    img = f32(img).reshape(-1); ref = f32(ref).reshape(-1)
    return lib().orc_rel_mse(_p(img), _p(ref), img.size, eps)


def hardware_threads() -> int:
    return lib().orc_hardware_threads()
