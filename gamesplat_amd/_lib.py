"""ctypes binding of libgamesplat_hip.so (the C ABI declared in include/gs_api.h).

The shared library is built in-tree by ``python -m gamesplat_amd.build``.  There
is no fallback: if the library is missing or a HIP device is absent, calls
fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# GS_LIB_PATH selects another build of the same library (e.g. to compare two
# builds on one machine); by default the in-tree build is loaded.
LIB_PATH = Path(os.environ.get("GS_LIB_PATH", str(Path(__file__).resolve().parent / "libgamesplat_hip.so")))

GS_OK = 0
GS_ERR_INVALID = -1
GS_ERR_OOM = -2
GS_ERR_HIP = -3
GS_ERR_STATE = -4
GS_ERR_LIMIT = -5

GS_IMAGE_ALPHA_RGBA16F = 0
GS_IMAGE_GPS_SAMPLE_RGBA32F = 1
GS_IMAGE_GPS_DEPTH_R32F = 2
GS_IMAGE_ACCUM_RGBA32F = 3
GS_IMAGE_FRAME_AVG_RGBA32F = 4
GS_IMAGE_TAA_RGBA32F = 5
GS_IMAGE_TAA_XYZ_RGBA32F = 6
GS_IMAGE_DISPLAY_RGBA32F = 7

GS_BUFFER_SORTED_KEYS = 0
GS_BUFFER_SORTED_IDS = 1
GS_BUFFER_VISIBLE_IDS = 2
GS_BUFFER_VISIBLE_KEYS = 3
GS_BUFFER_PROJECTED = 4
GS_BUFFER_GPS_PACKED = 5

GS_OPT_GPS_KEEP_CELLS = 0
GS_OPT_TILE_SIZE = 1
GS_OPT_QUADRANT_CULL = 2
GS_OPT_TIGHT_RECTS = 3
GS_OPT_RADIX_BITS = 4
GS_OPT_BLOCKS_PER_CU = 5
GS_OPT_GRAPH = 6
GS_OPT_REFERENCE_BLEND = 7
GS_OPT_WIDE_INDEX = 8
GS_OPT_TILE_KEY16 = 9
GS_SORT_STATIC_MAX_PAIRS = 1 << 22
MAX_VIEWS = 64

STAGES = ("cull", "sort", "bin", "composite", "gps_expand", "gps_splat", "gps_resolve", "cull_kernel")

GAUSSIAN_FLOATS = 59  # GsGaussian: pos 0-2, scale 3-5, rot 6-9 (w,x,y,z), dc 10-12, opacity 13, rest 14-58


class GsCameraUniforms(C.Structure):
    _fields_ = [
        ("view", C.c_float * 16),
        ("proj", C.c_float * 16),
        ("viewport", C.c_float * 2),
        ("orthographic", C.c_float),
        ("persp_focal", C.c_float),
        ("ortho_focal", C.c_float),
        ("clip_y_sign", C.c_float),
        ("clip_z_01", C.c_float),
        ("pad", C.c_float * 1),
    ]

    def as_array(self) -> np.ndarray:
        return np.frombuffer(bytes(self), dtype=np.float32).copy()

    @classmethod
    def from_array(cls, a) -> "GsCameraUniforms":
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.size == 40
        return cls.from_buffer_copy(a.tobytes())


class GsTaaUniforms(C.Structure):
    """TaaAccumUBO (shaders/accum.glsl:64-75); matrices column-major."""
    _fields_ = [
        ("inv_view", C.c_float * 16),
        ("prev_view", C.c_float * 16),
        ("viewport_orthographic", C.c_float * 4),
        ("focal_clip", C.c_float * 4),
        ("prev_focal_clip", C.c_float * 4),
        ("proj_z", C.c_float * 4),
        ("history_valid", C.c_float),
        ("view_changed", C.c_float),
        ("pad", C.c_float * 2),
    ]

    def as_array(self) -> np.ndarray:
        return np.frombuffer(bytes(self), dtype=np.float32).copy()

    @classmethod
    def from_array(cls, a) -> "GsTaaUniforms":
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.size == 52
        return cls.from_buffer_copy(a.tobytes())


class GsAnimation(C.Structure):
    """GSAnim appearance data (include/gs_api.h)."""
    _fields_ = [
        ("rows", C.c_uint32),
        ("state_count", C.c_uint32),
        ("deltas", C.c_void_p),
        ("scales", C.c_float * 48),
        ("layer_count", C.c_uint32),
        ("layer_data", C.c_void_p),
        ("lut_size", C.c_uint32),
        ("display_lut", C.c_void_p),
        ("log2_min", C.c_float),
        ("log2_max", C.c_float),
    ]


class GsPlyLayout(C.Structure):
    _fields_ = [("stride", C.c_uint32), ("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32),
                ("scale", C.c_int32 * 3), ("rot", C.c_int32 * 4), ("opacity", C.c_int32), ("dc", C.c_int32 * 3),
                ("rest", C.c_int32 * 45)]


class GsSogData(C.Structure):
    _fields_ = [("count", C.c_uint32), ("means_mins", C.c_float * 3), ("means_maxs", C.c_float * 3),
                ("scales_codebook", C.c_float * 256), ("sh0_codebook", C.c_float * 256),
                ("means_l", C.c_void_p), ("means_u", C.c_void_p), ("scales", C.c_void_p), ("quats", C.c_void_p),
                ("sh0", C.c_void_p), ("shN_bands", C.c_int32), ("shN_count", C.c_int32),
                ("shN_codebook", C.c_float * 256), ("shN_centroids", C.c_void_p), ("centroids_width", C.c_int32),
                ("centroids_height", C.c_int32), ("shN_labels", C.c_void_p)]


class GsSplatEffectParams(C.Structure):
    _fields_ = [("center_radius", C.c_float * 4), ("params", C.c_float * 4), ("color", C.c_float * 4)]


class GsCullOptions(C.Structure):
    _fields_ = [("sh_degree", C.c_int32), ("sh_lod_distance", C.c_float), ("collect_stats", C.c_int32)]


class GsHbmRates(C.Structure):
    _fields_ = [("read_gbs", C.c_double), ("planes_read_gbs", C.c_double), ("write_gbs", C.c_double),
                ("copy_gbs", C.c_double)]


class GsStats(C.Structure):
    _fields_ = [
        ("gaussian_count", C.c_uint32),
        ("visible_count", C.c_uint32),
        ("tile_entries", C.c_uint32),
        ("gps_requested_work_items", C.c_uint32),
        ("gps_work_budget_exhausted", C.c_uint32),
        ("gps_dropped_work_items", C.c_uint32),
        ("gps_max_work_items", C.c_uint32),
        ("diagnostics", C.c_uint32 * 8),
        ("diagnostics_valid", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "diagnostics"}
        d["diagnostics"] = list(self.diagnostics)
        return d


class GsFrameInfo(C.Structure):
    _fields_ = [
        ("views", C.c_uint32),
        ("total_tiles", C.c_uint32),
        ("tile_size", C.c_int32),
        ("radix_bits", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("graph", C.c_int32),
        ("graph_captures", C.c_uint32),
        ("graph_replays", C.c_uint32),
        ("overflow_reruns", C.c_uint32),
        ("pad", C.c_uint32),
        ("held_bytes", C.c_uint64),
    ]


class GsTuneResult(C.Structure):
    _fields_ = [
        ("tile_size", C.c_int32),
        ("radix_bits", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("configs", C.c_int32),
        ("best_ms", C.c_double),
        ("default_ms", C.c_double),
    ]


EXPORTS = {
    "gs_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "gs_destroy": (None, [C.c_void_p]),
    "gs_last_error": (C.c_char_p, []),
    "gs_device_info": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "gs_hbm_probe": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(GsHbmRates)]),
    "gs_set_memory_budget": (C.c_int, [C.c_void_p, C.c_size_t]),
    "gs_set_option": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "gs_upload_gaussians": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "gs_set_gps_budget": (C.c_int, [C.c_void_p, C.c_uint32]),
    "gs_cull": (C.c_int, [C.c_void_p, C.POINTER(GsCameraUniforms), C.POINTER(GsSplatEffectParams),
                          C.POINTER(GsCullOptions)]),
    "gs_sort": (C.c_int, [C.c_void_p]),
    "gs_sort_pairs_static": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "gs_composite_alpha": (C.c_int, [C.c_void_p, C.c_int]),
    "gs_gps_sample": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_float]),
    "gs_accumulate": (C.c_int, [C.c_void_p, C.c_uint32]),
    "gs_stochastic_sample": (C.c_int, [C.c_void_p, C.c_uint32, C.c_float, C.POINTER(C.c_float)]),
    "gs_taa_jitter": (None, [C.c_uint32, C.POINTER(C.c_float)]),
    "gs_decode_ply": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(GsPlyLayout), C.c_void_p]),
    "gs_decode_sog": (C.c_int, [C.POINTER(GsSogData), C.c_void_p]),
    "gs_planes_from_gaussians": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "gs_decode_ply_planes": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(GsPlyLayout), C.c_void_p]),
    "gs_upload_planes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "gs_upload_ply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(GsPlyLayout)]),
    "gs_upload_sog": (C.c_int, [C.c_void_p, C.POINTER(GsSogData)]),
    "gs_alloc_pinned": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "gs_free_pinned": (None, [C.c_void_p]),
    "gs_upload_animation": (C.c_int, [C.c_void_p, C.POINTER(GsAnimation)]),
    "gs_set_animation_frame": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_uint32,
                                         C.POINTER(C.c_float)]),
    "gs_display_layers": (C.c_int, [C.c_void_p, C.c_int, C.c_float]),
    "gs_frame_average": (C.c_int, [C.c_void_p, C.c_uint32]),
    "gs_taa_uniforms": (C.c_int, [C.POINTER(GsCameraUniforms), C.POINTER(GsCameraUniforms), C.c_int,
                                  C.POINTER(GsTaaUniforms)]),
    "gs_taa_accumulate": (C.c_int, [C.c_void_p, C.POINTER(GsTaaUniforms), C.c_int]),
    "gs_render_alpha": (C.c_int, [C.c_void_p, C.POINTER(GsCameraUniforms), C.POINTER(GsSplatEffectParams),
                                  C.POINTER(GsCullOptions), C.c_int]),
    "gs_render_views": (C.c_int, [C.c_void_p, C.POINTER(GsCameraUniforms), C.c_uint32,
                                  C.POINTER(GsSplatEffectParams), C.POINTER(GsCullOptions), C.c_int]),
    "gs_read_view_image": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t]),
    "gs_frame_info": (C.c_int, [C.c_void_p, C.POINTER(GsFrameInfo)]),
    "gs_autotune": (C.c_int, [C.c_void_p, C.POINTER(GsCameraUniforms), C.c_uint32, C.POINTER(GsCullOptions),
                              C.c_int, C.POINTER(GsTuneResult)]),
    "gs_denoiser_param_count": (C.c_uint32, []),
    "gs_denoiser_create": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "gs_denoiser_destroy": (None, [C.c_void_p]),
    "gs_denoiser_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "gs_denoiser_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "gs_denoiser_forward_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                               C.c_uint32, C.POINTER(C.c_float), C.c_void_p, C.c_void_p]),
    "gs_denoiser_adam_step": (C.c_int, [C.c_void_p, C.c_float]),
    "gs_denoiser_capture": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64]),
    "gs_denoiser_pool_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "gs_denoiser_train": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, C.c_void_p]),
    "gs_denoise": (C.c_int, [C.c_void_p]),
    "gs_denoiser_judge": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float)]),
    "gs_denoiser_read_image": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "gs_read_image": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "gs_read_buffer": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "gs_stats": (C.c_int, [C.c_void_p, C.POINTER(GsStats)]),
    "gs_synchronize": (C.c_int, [C.c_void_p]),
    "gs_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "gs_stage_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint32)]),
    "gs_reset_stage_times": (C.c_int, [C.c_void_p]),
    "gs_stream": (C.c_void_p, [C.c_void_p]),
    "gs_generate_scene": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_float, C.c_void_p]),
    "gs_camera_from_pose": (C.c_int, [C.POINTER(C.c_float), C.c_float, C.c_float, C.c_float,
                                      C.POINTER(GsCameraUniforms)]),
    "gs_look_at_origin_pose": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float)]),
}

_lib = None


class GsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gs error {code}: {msg}")
        self.code = code


def load() -> C.CDLL:
    """Load libgamesplat_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gamesplat_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != GS_OK:
        raise GsError(code, load().gs_last_error().decode())


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def taa_jitter(frame_seed: int) -> tuple[float, float]:
    """Halton (2, 3) jitter of a StochasticSplats sample, in pixels."""
    out = (C.c_float * 2)()
    load().gs_taa_jitter(frame_seed, out)
    return float(out[0]), float(out[1])


def taa_uniforms(cam: GsCameraUniforms, prev: GsCameraUniforms | None = None,
                 view_changed: bool = False) -> GsTaaUniforms:
    """TAA uniforms for cam with prev's history (None = no history).  A view
    that cannot be inverted raises GsError."""
    u = GsTaaUniforms()
    check(load().gs_taa_uniforms(C.byref(cam), None if prev is None else C.byref(prev), int(view_changed),
                                 C.byref(u)))
    return u


def generate_scene(seed: int, count: int, clusters: int, opacity_logit_mean: float = 0.0) -> np.ndarray:
    """Seeded synthetic scene, (count, 59) float32 GsGaussian rows."""
    out = np.empty((count, GAUSSIAN_FLOATS), dtype=np.float32)
    check(load().gs_generate_scene(seed, count, clusters, opacity_logit_mean, out.ctypes.data_as(C.c_void_p)))
    return out


def look_at_origin_pose(position) -> np.ndarray:
    pos = np.ascontiguousarray(position, dtype=np.float32)
    out = np.empty(16, dtype=np.float32)
    check(load().gs_look_at_origin_pose(_fptr(pos), _fptr(out)))
    return out


def camera_from_pose(pose, focal_over_height: float, width: int, height: int) -> GsCameraUniforms:
    p = np.ascontiguousarray(pose, dtype=np.float32).reshape(16)
    cam = GsCameraUniforms()
    check(load().gs_camera_from_pose(_fptr(p), focal_over_height, float(width), float(height), C.byref(cam)))
    return cam


class Context:
    """One gs_ctx: one HIP stream and the device memory of one renderer."""

    def __init__(self, device: int = 0):
        lib = load()
        self._lib = lib
        h = C.c_void_p()
        check(lib.gs_create(C.byref(h), device))
        self._h = h
        self.n = 0
        self.width = 0
        self.height = 0

    def close(self):
        if self._h:
            self._lib.gs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def device_info(self) -> str:
        buf = C.create_string_buffer(512)
        check(self._lib.gs_device_info(self._h, buf, 512))
        return buf.value.decode()

    def hbm_probe(self, nbytes: int = 1 << 30) -> dict:
        """Read, 15-plane read, write and copy rates (GB/s) of probe kernels
        over buffers of nbytes."""
        r = GsHbmRates()
        check(self._lib.gs_hbm_probe(self._h, nbytes, C.byref(r)))
        return {k: round(getattr(r, k), 1) for k, _ in GsHbmRates._fields_}

    def set_memory_budget(self, nbytes: int):
        check(self._lib.gs_set_memory_budget(self._h, nbytes))

    def set_option(self, option: int, value: int):
        check(self._lib.gs_set_option(self._h, option, value))

    def upload(self, gaussians: np.ndarray):
        g = np.ascontiguousarray(gaussians, dtype=np.float32)
        assert g.ndim == 2 and g.shape[1] == GAUSSIAN_FLOATS
        check(self._lib.gs_upload_gaussians(self._h, g.ctypes.data_as(C.c_void_p), g.shape[0]))
        self.n = g.shape[0]

    def set_gps_budget(self, items: int):
        check(self._lib.gs_set_gps_budget(self._h, items))

    def cull(self, cam: GsCameraUniforms, effect: GsSplatEffectParams | None = None, sh_degree: int = 3,
             sh_lod_distance: float = 0.0, collect_stats: bool = False):
        opts = GsCullOptions(sh_degree, sh_lod_distance, int(collect_stats))
        check(self._lib.gs_cull(self._h, C.byref(cam), C.byref(effect) if effect is not None else None,
                                C.byref(opts)))
        self.width = int(cam.viewport[0])
        self.height = int(cam.viewport[1])
        self._views = [(self.width, self.height)]

    def sort(self):
        check(self._lib.gs_sort(self._h))

    def composite(self, tile: int = 0):
        check(self._lib.gs_composite_alpha(self._h, tile))

    def render_alpha(self, cam: GsCameraUniforms, sh_degree: int = 3, tile: int = 0):
        opts = GsCullOptions(sh_degree, 0.0, 0)
        check(self._lib.gs_render_alpha(self._h, C.byref(cam), None, C.byref(opts), tile))
        self.width = int(cam.viewport[0])
        self.height = int(cam.viewport[1])
        self._views = [(self.width, self.height)]

    def render_views(self, cams, sh_degree: int = 3, tile: int = 0):
        """One batched sorted-alpha frame over up to 64 cameras."""
        arr = (GsCameraUniforms * len(cams))(*cams)
        opts = GsCullOptions(sh_degree, 0.0, 0)
        check(self._lib.gs_render_views(self._h, arr, len(cams), None, C.byref(opts), tile))
        self._views = [(int(c.viewport[0]), int(c.viewport[1])) for c in cams]
        self.width, self.height = self._views[0]

    def read_view_image(self, view: int) -> np.ndarray:
        views = getattr(self, "_views", [])
        w, h = views[view] if 0 <= view < len(views) else (0, 0)  # the library reports a bad view
        out = np.empty((h, w, 4), dtype=np.float16)
        check(self._lib.gs_read_view_image(self._h, view, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def frame_info(self) -> dict:
        fi = GsFrameInfo()
        check(self._lib.gs_frame_info(self._h, C.byref(fi)))
        return {k: getattr(fi, k) for k, _ in fi._fields_}

    def autotune(self, cams, frames: int = 3, sh_degree: int = 3) -> dict:
        arr = (GsCameraUniforms * len(cams))(*cams)
        opts = GsCullOptions(sh_degree, 0.0, 0)
        res = GsTuneResult()
        check(self._lib.gs_autotune(self._h, arr, len(cams), C.byref(opts), frames, C.byref(res)))
        return {k: getattr(res, k) for k, _ in res._fields_}

    def gps_sample(self, frame_seed: int, supersample: int = 1, linear_color: bool = False,
                   background: float = 0.0):
        check(self._lib.gs_gps_sample(self._h, frame_seed, supersample, int(linear_color), background))

    def stochastic_sample(self, frame_seed: int, background: float = 0.0, jitter=None):
        j = None if jitter is None else (C.c_float * 2)(float(jitter[0]), float(jitter[1]))
        check(self._lib.gs_stochastic_sample(self._h, frame_seed, background, j))

    def upload_animation(self, deltas=None, scales=None, layers=None, lut=None, log2_min=0.0, log2_max=1.0):
        """GSAnim appearance data: deltas (states, rows, 48) int8 with scales (48,),
        or layers (layers, rows, 48) half bits; optional LUT (size, size, size, 3)
        uint16 indexed [b][g][r].  None for everything drops the animation."""
        if deltas is None and layers is None:
            check(self._lib.gs_upload_animation(self._h, None))
            return
        a = GsAnimation()
        keep = []
        if deltas is not None:
            d = np.ascontiguousarray(deltas, np.int8)
            a.rows, a.state_count = d.shape[1], d.shape[0]
            a.deltas = d.ctypes.data
            a.scales[:] = [float(x) for x in np.asarray(scales, np.float32)]
            keep.append(d)
        else:
            lay = np.ascontiguousarray(layers, np.uint16)
            a.rows, a.layer_count = lay.shape[1], lay.shape[0]
            a.layer_data = lay.ctypes.data
            keep.append(lay)
        if lut is not None:
            t = np.ascontiguousarray(lut, np.uint16)
            a.lut_size = t.shape[0]
            a.display_lut = t.ctypes.data
            a.log2_min, a.log2_max = log2_min, log2_max
            keep.append(t)
        check(self._lib.gs_upload_animation(self._h, C.byref(a)))

    def set_animation_frame(self, state_a: int = 0, state_b: int = 0, weight: float = 0.0, layer_weights=None):
        if layer_weights is None:
            check(self._lib.gs_set_animation_frame(self._h, state_a, state_b, weight, 0, None))
            return
        w = np.ascontiguousarray(layer_weights, np.float32).reshape(-1, 3)
        check(self._lib.gs_set_animation_frame(self._h, state_a, state_b, weight, w.shape[0],
                                               w.ctypes.data_as(C.POINTER(C.c_float))))

    def display_layers(self, source: int, exposure_ev: float = 0.0):
        check(self._lib.gs_display_layers(self._h, source, exposure_ev))

    def frame_average(self, sample_index: int):
        check(self._lib.gs_frame_average(self._h, sample_index))

    def taa_accumulate(self, uniforms: GsTaaUniforms, use_frame_average: bool = False):
        check(self._lib.gs_taa_accumulate(self._h, C.byref(uniforms), int(use_frame_average)))

    def accumulate(self, sample_count: int):
        check(self._lib.gs_accumulate(self._h, sample_count))

    def synchronize(self):
        check(self._lib.gs_synchronize(self._h))

    def read_image(self, which: int, shape=None) -> np.ndarray:
        w, h = self.width, self.height
        if which == GS_IMAGE_ALPHA_RGBA16F:
            out = np.empty((h, w, 4), dtype=np.float16)
        elif which == GS_IMAGE_GPS_DEPTH_R32F:
            out = np.empty((h, w), dtype=np.float32)
        else:
            out = np.empty((h, w, 4), dtype=np.float32)
        check(self._lib.gs_read_image(self._h, which, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def read_buffer(self, which: int, count: int | None = None, shape=None) -> np.ndarray:
        if which in (GS_BUFFER_SORTED_KEYS, GS_BUFFER_SORTED_IDS):
            out = np.empty(self.n, dtype=np.uint32)
        elif which in (GS_BUFFER_VISIBLE_IDS, GS_BUFFER_VISIBLE_KEYS):
            v = self.stats()["visible_count"] if count is None else count
            out = np.empty(v, dtype=np.uint32)
        elif which == GS_BUFFER_PROJECTED:
            out = np.empty((self.n, 16), dtype=np.float32)
        elif which == GS_BUFFER_GPS_PACKED:
            out = np.empty(shape, dtype=np.uint64)
        else:
            raise ValueError(which)
        check(self._lib.gs_read_buffer(self._h, which, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def stats(self) -> dict:
        s = GsStats()
        check(self._lib.gs_stats(self._h, C.byref(s)))
        return s.as_dict()

    def set_profiling(self, on: bool):
        check(self._lib.gs_set_profiling(self._h, int(on)))

    def stage_times(self) -> dict:
        ms = (C.c_double * len(STAGES))()
        cnt = (C.c_uint32 * len(STAGES))()
        check(self._lib.gs_stage_times(self._h, ms, cnt))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(STAGES)}

    def reset_stage_times(self):
        check(self._lib.gs_reset_stage_times(self._h))

    def stream(self) -> int:
        return self._lib.gs_stream(self._h) or 0


class Denoiser:
    """The learned denoiser of a render context (gs_denoiser_*)."""

    def __init__(self, ctx: "Context", seed: int = 0):
        self._lib = ctx._lib
        self._ctx = ctx
        h = C.c_void_p()
        check(self._lib.gs_denoiser_create(ctx._h, seed, C.byref(h)))
        self._h = h
        self.P = int(self._lib.gs_denoiser_param_count())

    def close(self):
        if self._h:
            self._lib.gs_denoiser_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def get_params(self) -> np.ndarray:
        out = np.empty(self.P, np.float32)
        check(self._lib.gs_denoiser_get_params(self._h, out.ctypes.data_as(C.c_void_p), self.P))
        return out

    def set_params(self, flat: np.ndarray):
        f = np.ascontiguousarray(flat, np.float32)
        assert f.shape == (self.P,)
        check(self._lib.gs_denoiser_set_params(self._h, f.ctypes.data_as(C.c_void_p), self.P))

    def forward_backward(self, inp: np.ndarray, target: np.ndarray, want_out: bool = True, want_grad: bool = True):
        x = np.ascontiguousarray(inp, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        n, h, w, c = x.shape
        assert c == 8 and t.shape == (n, h, w, 4)
        loss = C.c_float()
        out = np.empty((n, h, w, 4), np.float32) if want_out else None
        grad = np.empty(self.P, np.float32) if want_grad else None
        check(self._lib.gs_denoiser_forward_backward(
            self._h, x.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p), n, h, w, C.byref(loss),
            out.ctypes.data_as(C.c_void_p) if out is not None else None,
            grad.ctypes.data_as(C.c_void_p) if grad is not None else None))
        return loss.value, out, grad

    def adam_step(self, lr: float):
        check(self._lib.gs_denoiser_adam_step(self._h, lr))

    def capture(self, crops: int, seed: int):
        check(self._lib.gs_denoiser_capture(self._h, crops, seed))

    def pool_size(self) -> int:
        n = C.c_uint32()
        check(self._lib.gs_denoiser_pool_size(self._h, C.byref(n)))
        return n.value

    def train(self, steps: int, batch: int, lr: float, seed: int) -> np.ndarray:
        losses = np.empty(steps, np.float32)
        check(self._lib.gs_denoiser_train(self._h, steps, batch, lr, seed, losses.ctypes.data_as(C.c_void_p)))
        return losses

    def judge(self) -> tuple[float, float]:
        """(PSNR dB, max abs error) of the last denoised image against the
        context's sorted-alpha image of the same cull."""
        p, m = C.c_double(), C.c_float()
        check(self._lib.gs_denoiser_judge(self._h, C.byref(p), C.byref(m)))
        return p.value, m.value

    def denoise(self):
        check(self._lib.gs_denoise(self._h))

    def read_image(self, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), np.float16)
        check(self._lib.gs_denoiser_read_image(self._h, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out
