"""Python host-side handles over the C ABI: device context, device arrays and scenes.

These mirror the ownership rules of include/minotert.h: a Context owns the HIP stream;
DeviceArray owns one mrt_alloc buffer; Scene owns the uploaded triangles and the BVH.
Errors raise MrtError with the library's message (the reference raises C++ exceptions,
src/stx/except.ixx:12-20).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import check, lib
from .configs import CameraDesc, SceneDesc

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("tri", "<u4"), ("u", "<f4"), ("v", "<f4")])
MISS = 0xFFFFFFFF


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def scene_desc_struct(d: SceneDesc) -> _capi.mrt_scene_desc:
    return _capi.mrt_scene_desc(*d.as_tuple())


def camera_struct(c: CameraDesc) -> _capi.mrt_camera:
    cam = _capi.mrt_camera()
    for k in range(3):
        cam.position[k] = c.position[k]
    cam.yaw, cam.pitch, cam.vertical_fov = c.yaw, c.pitch, c.vertical_fov
    cam.near_plane = getattr(c, "near_plane", 0.001)
    cam.width, cam.height = c.width, c.height
    return cam


def generate_scene(desc: SceneDesc):
    """Host-side scene generator of the library: (verts [n,9] f32, material [n] u32, albedo [m,3] f32)."""
    s = scene_desc_struct(desc)
    n = lib().mrt_scene_desc_tri_count(C.byref(s))
    n_mat = desc.n_spheres + (1 if desc.ground else 0)
    verts = np.empty((n, 9), np.float32)
    mat = np.empty(n, np.uint32)
    alb = np.empty((max(n_mat, 1), 3), np.float32)
    check(lib().mrt_scene_generate(C.byref(s), _ptr(verts), _ptr(mat), _ptr(alb)))
    return verts, mat, alb[:n_mat]


def check_size(n_tris: int, n_nodes: int = 0) -> None:
    """Raises MrtError (MRT_ERR_INVALID) if a scene or batch of this size cannot be traced."""
    check(lib().mrt_check_size(int(n_tris), int(n_nodes)))


def build_bvh_host(verts: np.ndarray, max_leaf: int = 4, bins: int = 16):
    """The library's host BVH4 build with no device: (wide nodes as an (n_nodes, 32) uint32
    array, leaf-ordered triangle records as an (n_tris, 12) uint32 array, stats dict); the same
    bytes as Scene(...).export() for either builder."""
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
    n = verts.shape[0]
    nodes = np.empty((max(n, 1), 32), np.uint32)
    tris = np.empty((n, 12), np.uint32)
    st = _capi.mrt_scene_stats()
    opts = _capi.mrt_bvh_opts(max_leaf, bins, _capi.BVH_HOST)
    check(lib().mrt_bvh_build_host(_ptr(verts), n, C.byref(opts), nodes.ctypes.data, nodes.nbytes,
                                   tris.ctypes.data, tris.nbytes, C.byref(st)))
    return nodes[:st.n_nodes].copy(), tris, {"n_tris": st.n_tris, "n_nodes": st.n_nodes, "depth": st.depth,
                                            "leaves": st.leaves, "build_ms": st.build_ms}


def trace_variants() -> list:
    """Names of the traversal launch-shape variants (index = variant id)."""
    return [lib().mrt_trace_variant_name(i).decode() for i in range(lib().mrt_trace_variant_count())]


def frame_variants() -> list:
    """Names of the frame launch-shape variants (shading block, megakernel shape)."""
    return [lib().mrt_frame_variant_name(i).decode() for i in range(lib().mrt_frame_variant_count())]


def camera_basis(c: CameraDesc) -> np.ndarray:
    out = np.empty(14, np.float32)
    cam = camera_struct(c)
    check(lib().mrt_camera_basis(C.byref(cam), out.ctypes.data_as(_capi.F32P)))
    return out


class Context:
    """A device + stream (replaces the Vulkan context of src/sys/vulkan.ixx:106-115)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        opts = _capi.mrt_opts(device, 0)
        check(lib().mrt_create(C.byref(self._h), C.byref(opts)))

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def check(self, status: int) -> None:
        check(status, self._h)

    def close(self) -> None:
        if self._h:
            lib().mrt_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        self.check(lib().mrt_sync(self._h))

    def device_info(self) -> dict:
        info = (C.c_uint32 * 4)()
        self.check(lib().mrt_device_info(self._h, info))
        return {"compute_units": info[0], "wavefront": info[1], "lds_per_block": info[2], "device": info[3]}

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        self.check(lib().mrt_device_name(self._h, buf, 256))
        return buf.value.decode()

    def trace_variant(self, any_hit: bool = False) -> int:
        v = C.c_uint32()
        self.check(lib().mrt_trace_variant_get(self._h, 1 if any_hit else 0, C.byref(v)))
        return v.value

    def frame_variant(self) -> int:
        v = C.c_uint32()
        self.check(lib().mrt_frame_variant_get(self._h, C.byref(v)))
        return v.value

    def set_frame_variant(self, v: int) -> None:
        self.check(lib().mrt_frame_variant_set(self._h, v))

    def set_trace_variant(self, v: int, any_hit: bool = False) -> None:
        self.check(lib().mrt_trace_variant_set(self._h, 1 if any_hit else 0, int(v)))

    def empty(self, shape, dtype) -> "DeviceArray":
        return DeviceArray(self, shape, dtype)

    def to_device(self, host: np.ndarray) -> "DeviceArray":
        host = np.ascontiguousarray(host)
        d = DeviceArray(self, host.shape, host.dtype)
        d.upload(host)
        return d

    def profile(self, on: bool = True) -> None:
        self.check(lib().mrt_profile_enable(self._h, 1 if on else 0))

    def profile_reset(self) -> None:
        self.check(lib().mrt_profile_reset(self._h))

    def profile_query(self, kernel: str):
        ms = C.c_double()
        n = C.c_uint64()
        self.check(lib().mrt_profile_query(self._h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def profile_region(self, name: str, begin: bool) -> None:
        """Starts (begin) or ends a named region of the context stream timed by one pair of HIP
        events, independent of profile(); profile_query(name) reports it."""
        self.check(lib().mrt_profile_region(self._h, name.encode(), 1 if begin else 0))

    def copy_device(self, dst: "DeviceArray", src: "DeviceArray", nbytes: int | None = None) -> None:
        nb = min(dst.nbytes, src.nbytes) if nbytes is None else nbytes
        self.check(lib().mrt_copy_device(self._h, dst.ptr, src.ptr, nb))


class DeviceArray:
    """One device allocation with a numpy shape and dtype."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        ctx.check(lib().mrt_alloc(ctx.handle, max(self.nbytes, 1), C.byref(self.ptr)))

    def upload(self, host: np.ndarray) -> None:
        host = np.ascontiguousarray(host)
        if host.nbytes != self.nbytes:
            raise ValueError(f"upload of {host.nbytes} bytes into {self.nbytes}")
        self.ctx.check(lib().mrt_upload(self.ctx.handle, self.ptr, _ptr(host), self.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.ctx.check(lib().mrt_download(self.ctx.handle, _ptr(out), self.ptr, self.nbytes))
        return out

    def free(self) -> None:
        if self.ptr and self.ctx.handle:
            lib().mrt_free(self.ctx.handle, self.ptr)
        self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Scene:
    """Triangles + materials on the device with their BVH (built at construction, on the host or,
    with builder="device", on the GPU; both give the same bytes)."""

    def __init__(self, ctx: Context, verts: np.ndarray, material: np.ndarray, albedo: np.ndarray,
                 max_leaf: int = 4, bins: int = 16, builder: str = "host"):
        self.ctx = ctx
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
        material = np.ascontiguousarray(material, np.uint32).reshape(-1)
        albedo = np.ascontiguousarray(albedo, np.float32).reshape(-1, 3)
        if material.shape[0] != verts.shape[0]:
            raise ValueError("one material index per triangle")
        self.n_tris = verts.shape[0]
        self._h = C.c_void_p()
        if builder not in ("host", "device"):
            raise ValueError("builder must be 'host' or 'device'")
        opts = _capi.mrt_bvh_opts(max_leaf, bins, _capi.BVH_DEVICE if builder == "device" else _capi.BVH_HOST)
        ctx.check(lib().mrt_scene_create(ctx.handle, _ptr(verts), self.n_tris, _ptr(material), _ptr(albedo),
                                         albedo.shape[0], C.byref(opts), C.byref(self._h)))

    @classmethod
    def from_desc(cls, ctx: Context, desc: SceneDesc, **kw) -> "Scene":
        v, m, a = generate_scene(desc)
        return cls(ctx, v, m, a, **kw)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def stats(self) -> dict:
        s = _capi.mrt_scene_stats()
        self.ctx.check(lib().mrt_scene_stats_get(self.ctx.handle, self._h, C.byref(s)))
        return {"n_tris": s.n_tris, "n_nodes": s.n_nodes, "depth": s.depth, "leaves": s.leaves,
                "build_ms": s.build_ms, "device_bytes": s.device_bytes}

    def export(self):
        """(wide nodes as an (n_nodes, 32) uint32 array, leaf-ordered triangle records as an
        (n_tris, 12) uint32 array): the BVH4 bytes on the device."""
        st = self.stats()
        nodes = np.empty((st["n_nodes"], 32), np.uint32)
        tris = np.empty((self.n_tris, 12), np.uint32)
        self.ctx.check(lib().mrt_scene_export(self.ctx.handle, self._h, nodes.ctypes.data, nodes.nbytes,
                                              tris.ctypes.data, tris.nbytes))
        return nodes, tris

    def close(self) -> None:
        if self._h and self.ctx.handle:
            lib().mrt_scene_destroy(self.ctx.handle, self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace_closest(self, rays: DeviceArray, hits: DeviceArray, n: int | None = None) -> None:
        n = rays.shape[0] if n is None else n
        self.ctx.check(lib().mrt_trace_closest(self.ctx.handle, self._h, rays.ptr, n, hits.ptr))

    def trace_any(self, rays: DeviceArray, occluded: DeviceArray, n: int | None = None) -> None:
        n = rays.shape[0] if n is None else n
        self.ctx.check(lib().mrt_trace_any(self.ctx.handle, self._h, rays.ptr, n, occluded.ptr))


def ray_order(scene: Scene, rays: DeviceArray, order: DeviceArray, n: int | None = None) -> None:
    """Coherence order of the rays (octant, then origin Morton code) into `order`."""
    n = rays.shape[0] if n is None else n
    scene.ctx.check(lib().mrt_ray_order(scene.ctx.handle, scene.handle, rays.ptr, n, order.ptr))


def trace_closest_ordered(scene: Scene, rays: DeviceArray, order: DeviceArray, hits: DeviceArray,
                          n: int | None = None) -> None:
    n = rays.shape[0] if n is None else n
    scene.ctx.check(lib().mrt_trace_closest_ordered(scene.ctx.handle, scene.handle, rays.ptr, order.ptr, n, hits.ptr))


def sort_pairs(ctx: Context, keys: DeviceArray, values: DeviceArray, n: int | None = None, bits: int = 32) -> None:
    """Stable radix sort of (key, value) u32 pairs on the device by the low `bits` key bits."""
    n = keys.shape[0] if n is None else n
    ctx.check(lib().mrt_sort_pairs(ctx.handle, keys.ptr, values.ptr, n, bits))
