"""Workload definitions: seeded synthetic scenes and cameras.

Configurations (sizes, seeds and camera rules):
  c1  100k Gaussians, 16 clusters, 1280x720, one camera at radius 15 (seed 0x5EED0001)
  c2  1M Gaussians, 64 clusters, 1920x1080, 120-frame orbit, 3 deg/frame (seed 0x5EED0002)
  c3  6M Gaussians, 256 clusters, opacity logit mean 1, 1920x1080, static camera (seed 0x5EED0003)
  c5  3M Gaussians, 64 clusters, 64 views at radius 10-20 of mixed sizes (seed 0x5EED0005)
Cameras are pinhole, vertical FOV 60 deg, looking at the origin with +Z up,
built through the same path as the reference's baked views.
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import camera_from_pose, generate_scene, look_at_origin_pose

FOCAL_OVER_HEIGHT = 0.5 / math.tan(math.radians(30.0))  # 60 deg vertical FOV

CONFIGS = {
    "c1": dict(seed=0x5EED0001, count=100_000, clusters=16, opacity_mean=0.0, width=1280, height=720),
    "c2": dict(seed=0x5EED0002, count=1_000_000, clusters=64, opacity_mean=0.0, width=1920, height=1080),
    "c3": dict(seed=0x5EED0003, count=6_000_000, clusters=256, opacity_mean=1.0, width=1920, height=1080),
    "c5": dict(seed=0x5EED0005, count=3_000_000, clusters=64, opacity_mean=0.0, width=1920, height=1080),
}


def scene_config(name: str, count: int | None = None) -> tuple[np.ndarray, dict]:
    cfg = dict(CONFIGS[name])
    if count is not None:
        cfg["count"] = count
    g = generate_scene(cfg["seed"], cfg["count"], cfg["clusters"], cfg["opacity_mean"])
    return g, cfg


def _direction(seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    d = rng.standard_normal(3)
    return d / np.linalg.norm(d)


def random_view_camera(seed: int, width: int, height: int, radius: float = 15.0):
    pos = (_direction(seed) * radius).astype(np.float32)
    return camera_from_pose(look_at_origin_pose(pos), FOCAL_OVER_HEIGHT, width, height)


def orbit_camera(seed: int, frame: int, width: int, height: int, radius: float = 15.0, step_deg: float = 3.0):
    p0 = _direction(seed) * radius
    a = math.radians(step_deg * frame)
    c, s = math.cos(a), math.sin(a)
    pos = np.array([c * p0[0] - s * p0[1], s * p0[0] + c * p0[1], p0[2]], dtype=np.float32)
    return camera_from_pose(look_at_origin_pose(pos), FOCAL_OVER_HEIGHT, width, height)


def multiview_cameras(seed: int = 0x5EED0005):
    """64 views: 16 each of 256x256, 640x480, 1280x720, 1920x1080; radius U[10, 20]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = [(256, 256)] * 16 + [(640, 480)] * 16 + [(1280, 720)] * 16 + [(1920, 1080)] * 16
    cams = []
    for w, h in sizes:
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        r = rng.uniform(10.0, 20.0)
        cams.append(camera_from_pose(look_at_origin_pose((d * r).astype(np.float32)), FOCAL_OVER_HEIGHT, w, h))
    return cams
