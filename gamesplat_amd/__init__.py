"""gamesplat_amd: MI355X (gfx950) Gaussian-splatting frame pipeline.

The product is libgamesplat_hip.so (C ABI in include/gs_api.h); this package
holds its sources (csrc/), the build recipe (build.py), the headless C++
renderer shim (renderer/) and a thin ctypes binding used by tests and the
benchmark.
"""
from ._lib import (  # noqa: F401
    Context,
    Denoiser,
    GsCameraUniforms,
    GsError,
    GsSplatEffectParams,
    GsTaaUniforms,
    camera_from_pose,
    generate_scene,
    load,
    look_at_origin_pose,
    taa_jitter,
    taa_uniforms,
)
from . import _lib as abi  # noqa: F401
from .scenes import orbit_camera, random_view_camera, scene_config  # noqa: F401
