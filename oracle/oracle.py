"""ctypes binding of the CPU oracle (oracle/_build/libgs_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by gamesplat_amd/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libgs_oracle.so"

_lib = None

_SIGS = {
    "or_log_d": (C.c_double, [C.c_double]),
    "or_exp_d": (C.c_double, [C.c_double]),
    "or_sin_d": (C.c_double, [C.c_double]),
    "or_cos_d": (C.c_double, [C.c_double]),
    "or_logf": (C.c_float, [C.c_float]),
    "or_expf": (C.c_float, [C.c_float]),
    "or_sinf": (C.c_float, [C.c_float]),
    "or_cosf": (C.c_float, [C.c_float]),
    "or_dilog": (C.c_float, [C.c_float]),
    "or_inv_dilog": (C.c_float, [C.c_float]),
    "or_logf_fma": (C.c_float, [C.c_float]),
    "or_logf_fma_n": (None, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "or_gps_sincosf_n": (None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "or_hash01_bits": (C.c_uint32, [C.c_uint32]),
    "or_hash01": (C.c_float, [C.c_uint32]),
    "or_pcg2d": (None, [C.POINTER(C.c_uint32)]),
    "or_f32_to_f16": (C.c_uint16, [C.c_float]),
    "or_f16_to_f32": (C.c_float, [C.c_uint16]),
    "or_cull_anim": (None, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_int,
                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "or_cull": (None, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_int,
                       C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.c_void_p, C.c_int]),
    "or_radix_sort": (None, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "or_clipped": (C.c_int, [C.c_float, C.c_float]),
    "or_composite": (None, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                            C.c_void_p, C.c_int]),
    "or_quad_fixed": (None, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p]),
    "or_gps_expand": (None, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_int,
                             C.c_void_p, C.c_void_p, C.c_void_p]),
    "or_gps_splat": (None, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_uint32, C.c_float,
                            C.c_void_p]),
    "or_gps_resolve": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_void_p,
                              C.c_void_p, C.c_void_p]),
    "or_accumulate": (None, [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_uint32]),
    "or_hash13": (C.c_float, [C.c_float, C.c_float, C.c_float]),
    "or_depth24": (C.c_uint32, [C.c_float, C.c_float]),
    "or_stochastic_splats": (None, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_float, C.c_float,
                                    C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_int]),
    "or_halton": (C.c_float, [C.c_uint32, C.c_uint32]),
    "or_layers_display": (None, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float,
                                 C.c_void_p]),
    "or_taa_accumulate": (None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                 C.c_void_p, C.c_void_p]),
}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def threads_default() -> int:
    return int(os.environ.get("GS_ORACLE_THREADS", "0"))


class OrAnimation(C.Structure):
    _fields_ = [("count", C.c_uint32), ("state_a", C.c_int32), ("state_b", C.c_int32), ("weight", C.c_float),
                ("scales", C.c_float * 48), ("layer_count", C.c_int32), ("words", C.c_void_p),
                ("layer_weights", C.c_void_p)]


def pack_animation_words(deltas=None, layers=None) -> np.ndarray:
    """The animation buffer as the renderer packs it: int8 deltas four per
    little-endian word, or half coefficients two per word (zero padded)."""
    raw = (np.ascontiguousarray(deltas, np.int8).view(np.uint8).ravel() if layers is None
           else np.ascontiguousarray(layers, np.uint16).view(np.uint8).ravel())
    pad = (-raw.size) % 4
    return np.concatenate([raw, np.zeros(pad, np.uint8)]).view("<u4").copy()


def cull(gaussians: np.ndarray, cam, effect=None, sh_degree=3, sh_lod_distance=0.0, collect_stats=False,
         threads=None, animation=None):
    """Returns (keys, ids, projected (n,16), visible_count, diag[8]) with reference semantics.

    animation: None, or a dict with "deltas" (states, rows, 48) int8, "scales"
    (48,), "state_a", "state_b", "weight"; or with "layers" (layers, rows, 48)
    half bits and "layer_weights" (layers, 3), optionally "layer_count"."""
    lib = load()
    g = np.ascontiguousarray(gaussians, dtype=np.float32)
    n = g.shape[0]
    keys = np.empty(n, np.uint32)
    ids = np.empty(n, np.uint32)
    proj = np.empty((n, 16), np.float32)
    vis = C.c_uint32(0)
    diag = np.zeros(8, np.uint32)
    an = None
    keep = []
    if animation is not None:
        an = OrAnimation()
        an.count = n
        if "layers" in animation:
            words = pack_animation_words(layers=animation["layers"])
            lw = np.ascontiguousarray(animation["layer_weights"], np.float32).ravel()
            an.layer_count = int(animation.get("layer_count", len(animation["layers"])))
            an.layer_weights = lw.ctypes.data
            keep.append(lw)
        else:
            words = pack_animation_words(deltas=animation["deltas"])
            an.state_a = int(animation.get("state_a", 0))
            an.state_b = int(animation.get("state_b", 0))
            an.weight = float(animation.get("weight", 0.0))
            an.scales[:] = [float(x) for x in np.asarray(animation["scales"], np.float32)]
        an.words = words.ctypes.data
        keep.append(words)
    lib.or_cull_anim(_p(g), n, C.byref(cam), C.byref(effect) if effect is not None else None, sh_degree,
                     sh_lod_distance, int(collect_stats), C.byref(an) if an is not None else None, _p(keys),
                     _p(ids), _p(proj), C.byref(vis), _p(diag), threads_default() if threads is None else threads)
    return keys, ids, proj, vis.value, diag


def radix_sort(keys: np.ndarray, ids: np.ndarray):
    lib = load()
    k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    i = np.ascontiguousarray(ids, dtype=np.uint32).copy()
    lib.or_radix_sort(_p(k), _p(i), k.size)
    return k, i


def composite(projected: np.ndarray, sorted_ids: np.ndarray, width: int, height: int, clip_y_sign: float = 1.0,
              mode: int = 0, threads=None, clip_z_01: float = 0.0) -> np.ndarray:
    """mode 0 fp64 back-to-front, 1 fp64 front-to-back, 2 fp16-per-blend. (H, W, 4) float64, top row first.
    Splats whose ndc depth lies outside the clip volume ([-1, 1], [0, 1] with clip_z_01) are not drawn."""
    lib = load()
    p = np.ascontiguousarray(projected, dtype=np.float32)
    s = np.ascontiguousarray(sorted_ids, dtype=np.uint32)
    out = np.empty((height, width, 4), np.float64)
    lib.or_composite(_p(p), _p(s), s.size, width, height, clip_y_sign, clip_z_01, mode, _p(out),
                     threads_default() if threads is None else threads)
    return out


def quad_fixed(rec: np.ndarray, width: float, height: float, jitter=(0.0, 0.0)) -> np.ndarray:
    lib = load()
    r = np.ascontiguousarray(rec, dtype=np.float32)
    out = np.empty(4, np.int32)
    lib.or_quad_fixed(_p(r), width, height, float(jitter[0]), float(jitter[1]), _p(out))
    return out


def gps_expand(projected, ids, max_work_items, ss=1.0, clip_z_01=0.0, linear_color=False):
    lib = load()
    p = np.ascontiguousarray(projected, dtype=np.float32)
    i = np.ascontiguousarray(ids, dtype=np.uint32)
    n = i.size
    prepared = np.zeros((n, 12), np.uint32)
    work = np.zeros((max_work_items, 2), np.uint32)
    stats = np.zeros(2, np.uint32)
    lib.or_gps_expand(_p(p), _p(i), n, max_work_items, ss, clip_z_01, int(linear_color), _p(prepared), _p(work),
                      _p(stats))
    return prepared, work, int(stats[0]), int(stats[1])


def gps_splat(prepared, work, work_count, gps_width, gps_height, frame_seed, ss=1.0, cells=None):
    lib = load()
    if cells is None:
        cells = np.full(gps_width * gps_height, 0xFFFFFFFF00000000, dtype=np.uint64)
    lib.or_gps_splat(_p(prepared), _p(work), work_count, gps_width, gps_height, frame_seed, ss, _p(cells))
    return cells


def gps_resolve(cells, width, height, ss=1, background=0.0, linear_color=False, projected=None):
    lib = load()
    color = np.empty((height, width, 4), np.float32)
    depth = np.empty((height, width), np.float32)
    proj = np.ascontiguousarray(projected, dtype=np.float32) if projected is not None else np.zeros(16, np.float32)
    lib.or_gps_resolve(_p(cells), width, height, ss, background, int(linear_color), _p(proj), _p(color), _p(depth))
    return color, depth


def stochastic_splats(projected, ids, width, height, frame_seed, clip_z_01=0.0, background=0.0, jitter=(0.0, 0.0),
                      threads=None):
    """StochasticSplats sample: (color (H, W, 4), depth (H, W)) in framebuffer row order."""
    lib = load()
    p = np.ascontiguousarray(projected, dtype=np.float32)
    i = np.ascontiguousarray(ids, dtype=np.uint32)
    color = np.empty((height, width, 4), np.float32)
    depth = np.empty((height, width), np.float32)
    lib.or_stochastic_splats(_p(p), _p(i), i.size, width, height, clip_z_01, float(frame_seed), background,
                             float(jitter[0]), float(jitter[1]), _p(color), _p(depth),
                             threads_default() if threads is None else threads)
    return color, depth


def layers_display(image, lut, log2_min, log2_max, exposure_ev=0.0):
    """layers_display_fs of a linear RGBA image (..., 4); lut (size, size, size, 3) uint16 [b][g][r]."""
    lib = load()
    src = np.ascontiguousarray(image, dtype=np.float32)
    t = np.ascontiguousarray(lut, dtype=np.uint16)
    out = np.empty(src.shape[:-1] + (4,), np.float32)
    lib.or_layers_display(_p(src), src.size // 4, _p(t), t.shape[0], log2_min, log2_max, exposure_ev, _p(out))
    return out


def taa_accumulate(cur_color, cur_depth, hist_color, hist_xyz, uniforms):
    """taa_accum_fs over framebuffer-row images; uniforms: 52 floats (TaaAccumUBO).  Returns (color, xyz)."""
    lib = load()
    cc = np.ascontiguousarray(cur_color, dtype=np.float32)
    cd = np.ascontiguousarray(cur_depth, dtype=np.float32)
    hc = np.ascontiguousarray(hist_color, dtype=np.float32)
    hx = np.ascontiguousarray(hist_xyz, dtype=np.float32)
    u = np.ascontiguousarray(uniforms, dtype=np.float32)
    assert u.size == 52
    h, w = cd.shape
    oc = np.empty((h, w, 4), np.float32)
    ox = np.empty((h, w, 4), np.float32)
    lib.or_taa_accumulate(_p(cc), _p(cd), _p(hc), _p(hx), w, h, _p(u), _p(oc), _p(ox))
    return oc, ox


def accumulate(history, sample, n):
    lib = load()
    h = np.ascontiguousarray(history, dtype=np.float32)
    s = np.ascontiguousarray(sample, dtype=np.float32)
    out = np.empty_like(h)
    lib.or_accumulate(_p(h), _p(s), float(n), _p(out), h.size)
    return out


def logf_fma(x: np.ndarray) -> np.ndarray:
    """The binary32 FMA log of each element (positive normal floats): the GPS
    sampler's and the StochasticSplats cutoff's."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    load().or_logf_fma_n(_p(x), _p(y), x.size)
    return y


def gps_sincosf(x: np.ndarray):
    """The GPS sampler's binary32 (sin, cos) of each element."""
    x = np.ascontiguousarray(x, np.float32)
    s, c = np.empty_like(x), np.empty_like(x)
    load().or_gps_sincosf_n(_p(x), _p(s), _p(c), x.size)
    return s, c


def scalar(name, *args):
    return getattr(load(), name)(*args)


def pcg2d(x: int, y: int):
    v = (C.c_uint32 * 2)(x, y)
    load().or_pcg2d(v)
    return v[0], v[1]
