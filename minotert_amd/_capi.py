"""ctypes binding of the C ABI declared in include/minotert.h (libminotert.so).

The library is built in-tree into minotert_amd/_lib/. Importing this module without the
built library raises ImportError: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libminotert.so")

MRT_OK = 0
STATUS_NAMES = {0: "MRT_OK", 1: "MRT_ERR_INVALID", 2: "MRT_ERR_HIP", 3: "MRT_ERR_OOM",
                4: "MRT_ERR_UNSUPPORTED", 5: "MRT_ERR_STATE"}


class mrt_ray(C.Structure):
    _fields_ = [("o", C.c_float * 3), ("tmin", C.c_float), ("d", C.c_float * 3), ("tmax", C.c_float)]


class mrt_hit(C.Structure):
    _fields_ = [("t", C.c_float), ("tri", C.c_uint32), ("u", C.c_float), ("v", C.c_float)]


class mrt_camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("yaw", C.c_float), ("pitch", C.c_float),
                ("vertical_fov", C.c_float), ("near_plane", C.c_float), ("width", C.c_uint32),
                ("height", C.c_uint32)]


class mrt_gbuffer(C.Structure):
    _fields_ = [("hits_dev", C.c_void_p), ("depth_dev", C.c_void_p), ("normal_dev", C.c_void_p),
                ("motion_dev", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32)]


class mrt_render_opts(C.Structure):
    _fields_ = [("frame_seed", C.c_uint32), ("sample_base", C.c_uint32), ("spp", C.c_uint32),
                ("bounces", C.c_uint32), ("rr_start", C.c_uint32), ("flags", C.c_uint32),
                ("sky", C.c_float * 3), ("sun_dir", C.c_float * 3), ("sun_e", C.c_float * 3),
                ("capture_bounce", C.c_uint32), ("capture_rays_dev", C.c_void_p),
                ("capture_pid_dev", C.c_void_p), ("capture_count_dev", C.c_void_p), ("sky_lut", C.c_void_p)]


MRT_RENDER_SORT_RAYS = 1
MRT_RENDER_MEGAKERNEL = 2


class mrt_batch_view(C.Structure):
    _fields_ = [("camera", mrt_camera), ("rgb_dev", C.c_void_p)]


class mrt_cache_opts(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("log2_table", C.c_uint32), ("base_res", C.c_float), ("max_res", C.c_float),
                ("lo", C.c_float * 3), ("hi", C.c_float * 3), ("lr", C.c_float), ("adam_beta1", C.c_float),
                ("adam_beta2", C.c_float), ("adam_eps", C.c_float), ("seed", C.c_uint32), ("ema_decay", C.c_float),
                ("table_lr", C.c_float)]


class mrt_opts(C.Structure):
    _fields_ = [("device", C.c_int), ("flags", C.c_uint32)]


class mrt_scene_desc(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("n_spheres", C.c_uint32), ("level", C.c_uint32), ("ground", C.c_uint32),
                ("box", C.c_float), ("r_min", C.c_float), ("r_max", C.c_float), ("alb_min", C.c_float),
                ("alb_max", C.c_float), ("ground_z", C.c_float), ("ground_half", C.c_float),
                ("ground_albedo", C.c_float)]


class mrt_bvh_opts(C.Structure):
    _fields_ = [("max_leaf", C.c_uint32), ("bins", C.c_uint32), ("builder", C.c_uint32)]


BVH_HOST, BVH_DEVICE = 0, 1
MAX_WIDE_NODES = 1 << 25          # MRT_MAX_WIDE_NODES
MAX_TRIANGLES = (1 << 28) - 1     # MRT_MAX_TRIANGLES
RAY_ORDER_NOALLOC_MAX = 1 << 21   # MRT_RAY_ORDER_NOALLOC_MAX
CACHE_LIMIT_QUERIES = 1           # MRT_CACHE_LIMIT_QUERIES
CACHE_LIMIT_LOSS = 2              # MRT_CACHE_LIMIT_LOSS


class mrt_scene_stats(C.Structure):
    _fields_ = [("n_tris", C.c_uint32), ("n_nodes", C.c_uint32), ("depth", C.c_uint32), ("leaves", C.c_uint32),
                ("build_ms", C.c_double), ("device_bytes", C.c_uint64)]


class mrt_atmosphere_params(C.Structure):
    _fields_ = [("bottom_radius", C.c_float), ("top_radius", C.c_float), ("rayleigh_density_exp_scale", C.c_float),
                ("rayleigh_scattering", C.c_float * 3), ("mie_density_exp_scale", C.c_float),
                ("mie_scattering", C.c_float * 3), ("mie_extinction", C.c_float * 3), ("mie_absorption", C.c_float * 3),
                ("mie_phase_g", C.c_float), ("absorption_density0_layer_width", C.c_float),
                ("absorption_density0_constant", C.c_float), ("absorption_density0_linear", C.c_float),
                ("absorption_density1_constant", C.c_float), ("absorption_density1_linear", C.c_float),
                ("absorption_extinction", C.c_float * 3), ("ground_albedo", C.c_float * 3)]


class mrt_cache_render_opts(C.Structure):
    _fields_ = [("enable", C.c_uint32), ("vertex", C.c_uint32), ("train_samples", C.c_uint32),
                ("train_steps", C.c_uint32), ("cache_rgb_dev", C.c_void_p), ("train_samples_dev", C.c_void_p),
                ("train_targets_dev", C.c_void_p), ("train_pid_dev", C.c_void_p), ("train_count_dev", C.c_void_p),
                ("overlap_training", C.c_uint32), ("limits", C.c_uint32), ("query_budget", C.c_uint32),
                ("loss_budget", C.c_float)]



class mrt_bilateral_params(C.Structure):
    _fields_ = [("sigma", C.c_float), ("k_sigma", C.c_float), ("threshold", C.c_float), ("near_plane", C.c_float),
                ("frame", C.c_uint32)]


class mrt_tonemap_params(C.Structure):
    _fields_ = [("op", C.c_uint32), ("exposure", C.c_float), ("reinhard_max", C.c_float),
                ("uchimura", C.c_float * 6), ("amd", C.c_float * 5)]


P = C.c_void_p
U32 = C.c_uint32
SZ = C.c_size_t
F32P = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)

# name -> (restype, argtypes)
PROTOTYPES = {
    "mrt_abi_version": (U32, []),
    "mrt_create": (C.c_int, [C.POINTER(P), C.POINTER(mrt_opts)]),
    "mrt_destroy": (C.c_int, [P]),
    "mrt_last_error": (C.c_char_p, [P]),
    "mrt_sync": (C.c_int, [P]),
    "mrt_device_info": (C.c_int, [P, U32P]),
    "mrt_device_name": (C.c_int, [P, C.c_char_p, SZ]),
    "mrt_alloc": (C.c_int, [P, SZ, C.POINTER(P)]),
    "mrt_free": (C.c_int, [P, P]),
    "mrt_upload": (C.c_int, [P, P, P, SZ]),
    "mrt_download": (C.c_int, [P, P, P, SZ]),
    "mrt_copy_device": (C.c_int, [P, P, P, SZ]),
    "mrt_scene_desc_tri_count": (U32, [C.POINTER(mrt_scene_desc)]),
    "mrt_scene_generate": (C.c_int, [C.POINTER(mrt_scene_desc), P, P, P]),
    "mrt_camera_basis": (C.c_int, [C.POINTER(mrt_camera), F32P]),
    "mrt_check_size": (C.c_int, [C.c_uint64, C.c_uint64]),
    "mrt_bvh_build_host": (C.c_int, [P, U32, C.POINTER(mrt_bvh_opts), P, SZ, P, SZ, C.POINTER(mrt_scene_stats)]),
    "mrt_scene_create": (C.c_int, [P, P, U32, P, P, U32, C.POINTER(mrt_bvh_opts), C.POINTER(P)]),
    "mrt_scene_destroy": (C.c_int, [P, P]),
    "mrt_scene_stats_get": (C.c_int, [P, P, C.POINTER(mrt_scene_stats)]),
    "mrt_scene_export": (C.c_int, [P, P, P, SZ, P, SZ]),
    "mrt_trace_closest": (C.c_int, [P, P, P, U32, P]),
    "mrt_trace_any": (C.c_int, [P, P, P, U32, P]),
    "mrt_batch_create": (C.c_int, [P, C.POINTER(P), U32, C.POINTER(P)]),
    "mrt_batch_destroy": (C.c_int, [P, P]),
    "mrt_batch_info": (C.c_int, [P, P, C.POINTER(C.c_uint64)]),
    "mrt_batch_render": (C.c_int, [P, P, P, U32, C.POINTER(mrt_render_opts)]),
    "mrt_graph_create": (C.c_int, [P, P, P, U32, C.POINTER(mrt_render_opts), C.POINTER(P)]),
    "mrt_graph_set_frame": (C.c_int, [P, P, U32]),
    "mrt_graph_launch": (C.c_int, [P, P]),
    "mrt_graph_destroy": (C.c_int, [P, P]),
    "mrt_cache_create": (C.c_int, [P, C.POINTER(mrt_cache_opts), C.POINTER(P)]),
    "mrt_cache_destroy": (C.c_int, [P, P]),
    "mrt_cache_query": (C.c_int, [P, P, P, U32, P]),
    "mrt_cache_inputs": (C.c_int, [P, P, P, U32, P]),
    "mrt_cache_train": (C.c_int, [P, P, P, P, U32, C.c_int]),
    "mrt_cache_stats": (C.c_int, [P, P, C.POINTER(C.c_float), U32P, U32P]),
    "mrt_cache_get": (C.c_int, [P, P, U32, F32P, SZ]),
    "mrt_cache_set": (C.c_int, [P, P, U32, F32P, SZ]),
    "mrt_render_cached": (C.c_int, [P, P, P, C.POINTER(mrt_camera), C.POINTER(mrt_render_opts),
                                    C.POINTER(mrt_cache_render_opts), P]),
    "mrt_accumulate": (C.c_int, [P, P, P, SZ, U32, P]),
    "mrt_sky_earth": (C.c_int, [C.POINTER(mrt_atmosphere_params)]),
    "mrt_sky_create": (C.c_int, [P, C.POINTER(mrt_atmosphere_params), C.POINTER(P)]),
    "mrt_sky_update_view": (C.c_int, [P, P, F32P, F32P, F32P, C.c_int]),
    "mrt_sky_radiance": (C.c_int, [P, P, P, U32, P]),
    "mrt_sky_get": (C.c_int, [P, P, U32, F32P, SZ]),
    "mrt_sky_destroy": (C.c_int, [P, P]),
    "mrt_rel_mse": (C.c_int, [P, P, P, SZ, C.c_double, C.POINTER(C.c_double)]),
    "mrt_denoise_bilateral": (C.c_int, [P, P, P, P, U32, U32, C.POINTER(mrt_bilateral_params), P]),
    "mrt_tonemap_defaults": (C.c_int, [C.POINTER(mrt_tonemap_params)]),
    "mrt_tonemap": (C.c_int, [P, P, U32, C.POINTER(mrt_tonemap_params), P, P]),
    "mrt_ray_order": (C.c_int, [P, P, P, U32, P]),
    "mrt_ray_order_noalloc": (C.c_int, [P, P, P, U32, P]),
    "mrt_trace_closest_ordered": (C.c_int, [P, P, P, P, U32, P]),
    "mrt_sort_pairs": (C.c_int, [P, P, P, U32, U32]),
    "mrt_primary_rays":(C.c_int, [P, P, C.POINTER(mrt_camera), C.POINTER(mrt_camera), C.POINTER(mrt_gbuffer)]),
    "mrt_secondary_rays": (C.c_int, [P, P, C.POINTER(mrt_camera), C.POINTER(mrt_gbuffer),
                                     C.POINTER(mrt_render_opts), P]),
    "mrt_render": (C.c_int, [P, P, C.POINTER(mrt_camera), C.POINTER(mrt_render_opts), P]),
    "mrt_shade_lambert": (C.c_int, [P, P, C.POINTER(mrt_camera), C.POINTER(mrt_gbuffer), F32P, F32P, P]),
    "mrt_frame_variant_count": (U32, []),
    "mrt_frame_variant_name": (C.c_char_p, [U32]),
    "mrt_frame_variant_set": (C.c_int, [P, U32]),
    "mrt_frame_variant_get": (C.c_int, [P, U32P]),
    "mrt_trace_variant_count": (U32, []),
    "mrt_trace_variant_name": (C.c_char_p, [U32]),
    "mrt_trace_variant_set": (C.c_int, [P, C.c_int, U32]),
    "mrt_trace_variant_get": (C.c_int, [P, C.c_int, U32P]),
    "mrt_profile_enable": (C.c_int, [P, C.c_int]),
    "mrt_profile_reset": (C.c_int, [P]),
    "mrt_profile_query": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "mrt_profile_region": (C.c_int, [P, C.c_char_p, C.c_int]),
}


def load(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"minotert HIP library not built: {path} is missing "
                          f"(run __graft_entry__.build() or make -C minotert_amd/csrc)")
    lib = C.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


class MrtError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def check(status: int, ctx=None) -> None:
    if status != MRT_OK:
        msg = lib().mrt_last_error(ctx)
        raise MrtError(status, msg.decode() if msg else "")
