"""minotert_amd — MI355X-native ray-tracing hot path of MinoteRT.

HIP kernels for gfx950 behind the C ABI of include/minotert.h (libminotert.so, built
in-tree into minotert_amd/_lib/). This package is the Python host side: ctypes binding,
device handles and the render-stage API mirroring the reference's Pathtracer
(src/gfx/modules/pathtracer.ixx:18-200).
"""
from . import configs  # noqa: F401
from ._capi import LIB_PATH, MrtError  # noqa: F401
from .pathtracer import Environment, GBuffer, Pathtracer, RenderOptions  # noqa: F401
from .runtime import (HIT_DTYPE, MISS, RAY_DTYPE, Context, DeviceArray, Scene,  # noqa: F401
                      camera_basis, generate_scene)

__all__ = ["configs", "LIB_PATH", "MrtError", "Environment", "GBuffer", "Pathtracer", "RenderOptions",
           "HIT_DTYPE", "MISS", "RAY_DTYPE", "Context", "DeviceArray", "Scene", "camera_basis", "generate_scene"]
