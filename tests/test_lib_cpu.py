"""The C ABI library without a GPU: it loads, exports every function that
include/gs_api.h declares, and its host-only entry points (scene generator,
camera builders) behave as specified."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import gamesplat_amd as gs
from gamesplat_amd import abi

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(gs_[a-z0-9_]+)\s*\(", text))
    return names


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(abi.LIB_PATH))
    declared = _declared()
    assert len(declared) >= 25
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing
    assert declared <= set(abi.EXPORTS)


def test_no_gpu_fails_loudly():
    try:
        ctx = gs.Context(0)
    except abi.GsError as e:
        assert e.code == abi.GS_ERR_HIP
        return
    ctx.close()
    pytest.skip("a GPU is present")


def test_generator_deterministic_and_prefix_stable():
    a = gs.generate_scene(42, 70000, 16, 0.0)
    b = gs.generate_scene(42, 70000, 16, 0.0)
    np.testing.assert_array_equal(a, b)
    c = gs.generate_scene(42, 1000, 16, 0.0)
    np.testing.assert_array_equal(a[:1000], c)
    d = gs.generate_scene(43, 1000, 16, 0.0)
    assert not np.array_equal(c, d)


def test_generator_activation_and_distributions():
    g = gs.generate_scene(0x5EED0001, 200000, 16, 0.0)
    q = g[:, 6:10]
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    ls = np.log(g[:, 3:6].astype(np.float64))
    assert abs(ls.mean() + 4.5) < 0.01 and abs(ls.std() - 0.8) < 0.01
    o = g[:, 13].astype(np.float64)
    logit = np.log(o / (1 - o))
    assert abs(logit.mean()) < 0.02 and abs(logit.std() - 2.0) < 0.02
    assert abs(g[:, 10:13].std() - 0.5) < 0.01
    rest = g[:, 14:].reshape(-1, 15, 3)  # RGB triples per coefficient
    for k, l in ((0, 1), (5, 2), (12, 3)):
        assert abs(rest[:, k, :].std() - 0.1 / (l + 1)) < 0.002
    pos = g[:, 0:3]
    assert np.abs(pos).max() < 8 + 1.5 * 7


def test_camera_from_pose_matches_baked_view_rules():
    pos = np.array([3.0, -12.0, 8.0], np.float32)
    pose = gs.look_at_origin_pose(pos).reshape(4, 4)
    np.testing.assert_allclose(pose[:3, 3], pos)
    np.testing.assert_allclose(pose[:3, 2], pos / np.linalg.norm(pos), atol=1e-6)  # back axis
    assert abs(pose[2, 0]) < 1e-6  # right axis horizontal (Z up)
    cam = gs.camera_from_pose(pose.reshape(16), 0.8660254, 1920, 1080)
    assert abs(cam.persp_focal - 0.8660254 * 1080) < 1e-2
    assert list(cam.viewport) == [1920.0, 1080.0]
    with pytest.raises(abi.GsError):
        gs.camera_from_pose(pose.reshape(16), 0.0, 1920, 1080)
