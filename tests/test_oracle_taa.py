"""Pins of the StochasticSplats TAA oracle (or_taa_accumulate, or_halton,
jittered or_stochastic_splats) and of the library's host-only TAA helpers
(gs_taa_jitter, gs_taa_uniforms), CPU only.

- halton against its closed forms (radical inverses);
- jitter: moving every record centre by a dyadic offset equals jittering by
  it (quad corners and fragment centre move together, with the same sign);
- world reconstruction against a numpy FP64 projection of known points, for
  perspective and orthographic cameras and both clip_y_sign values;
- an unchanged view blends with the running-mean weights, n = min(alpha, 128);
- a view moved by exactly one pixel (orthographic) reads the neighbouring
  history texel, with the sign the projection fixes;
- every invalid-history case (no history, behind the camera, outside the
  previous view, empty history, a moved position beyond 100 units) yields
  the current colour with alpha 1.
"""
import numpy as np
import pytest

import gamesplat_amd as gs
from gamesplat_amd import scenes
from oracle import oracle as orc


def test_halton_closed_forms():
    assert orc.scalar("or_halton", 1, 2) == 0.5
    assert orc.scalar("or_halton", 2, 2) == 0.25
    assert orc.scalar("or_halton", 3, 2) == 0.75
    assert orc.scalar("or_halton", 5, 2) == 0.625  # 101b -> 0.101b
    assert np.isclose(orc.scalar("or_halton", 1, 3), 1 / 3, rtol=1e-7)
    assert np.isclose(orc.scalar("or_halton", 2, 3), 2 / 3, rtol=1e-7)
    assert np.isclose(orc.scalar("or_halton", 4, 3), 1 / 3 + 1 / 9, rtol=1e-6)  # 11 in base 3
    assert orc.scalar("or_halton", 0, 2) == 0.0


def test_library_jitter_matches_oracle_halton():
    for seed in [0, 1, 2, 7, 1023, 1024, 1025, 123456789, 2**32 - 1]:
        idx = seed % 1024 + 1
        want = (orc.scalar("or_halton", idx, 2) - 0.5, orc.scalar("or_halton", idx, 3) - 0.5)
        got = gs.taa_jitter(seed)
        assert np.float32(got[0]) == np.float32(want[0]) and np.float32(got[1]) == np.float32(want[1]), seed
        assert -0.5 <= got[0] < 0.5 and -0.5 <= got[1] < 0.5


def _record(cx, cy, rgb, opacity, ndc_z, sigma):
    a = c = sigma * sigma + 0.25
    det = a * c
    rad = float(np.ceil(3.0 * np.sqrt(a)))
    return np.array([*rgb, opacity, cx, cy, rad, rad, c / det, 0.0, a / det, ndc_z, a, 0.0, c, det], np.float32)


def test_jitter_equals_moved_centres():
    W, H = 40, 32
    recs = np.stack([_record(12.25, 10.5, (1, 0, 0), 0.9, 0.2, 3.0),
                     _record(20.0, 15.75, (0, 1, 0), 0.8, 0.1, 4.0),
                     _record(27.5, 20.25, (0, 0, 1), 0.95, 0.4, 2.5)])
    ids = np.arange(3, dtype=np.uint32)
    jit = (0.375, -0.25)
    moved = recs.copy()
    moved[:, 4] += np.float32(jit[0])
    moved[:, 5] += np.float32(jit[1])
    for seed in (1, 9):
        a, da = orc.stochastic_splats(recs, ids, W, H, seed, jitter=jit)
        b, db = orc.stochastic_splats(moved, ids, W, H, seed)
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(da, db)
        c0, _ = orc.stochastic_splats(recs, ids, W, H, seed)
        assert not np.array_equal(a, c0)  # the jitter changes coverage
    q0 = orc.quad_fixed(recs[0], W, H)
    q1 = orc.quad_fixed(recs[0], W, H, jitter=jit)
    np.testing.assert_array_equal(q1 - q0, [96, 96, -64, -64])  # 0.375 * 256, -0.25 * 256


# --- TAA ------------------------------------------------------------------

def _persp(fovy_focal, W, H, near=0.1, far=100.0):
    """A camera looking down -z from the origin: view = identity."""
    cam = gs.GsCameraUniforms()
    view = np.eye(4, dtype=np.float32)
    proj = np.zeros((4, 4), np.float32)
    proj[0, 0] = 2 * fovy_focal / W
    proj[1, 1] = 2 * fovy_focal / H
    proj[2, 2] = -(far + near) / (far - near)
    proj[2, 3] = -2 * far * near / (far - near)
    proj[3, 2] = -1.0
    cam.view[:] = view.T.flatten()
    cam.proj[:] = proj.T.flatten()
    cam.viewport[:] = [W, H]
    cam.orthographic = 0.0
    cam.persp_focal = fovy_focal
    cam.ortho_focal = 10.0
    cam.clip_y_sign = 1.0
    cam.clip_z_01 = 0.0
    return cam, proj


def _u(cam, prev=None, changed=False):
    return gs.taa_uniforms(cam, prev, changed).as_array()


def _depth_of(view_z, proj):
    """Window depth of a view-space z (GL clip, [0, 1] window)."""
    clip_z = proj[2, 2] * view_z + proj[2, 3]
    clip_w = -view_z
    return 0.5 * clip_z / clip_w + 0.5


def _images(rng, W, H, alpha=4.0):
    cur = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    hc = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    hc[..., 3] = alpha
    return cur, hc


@pytest.mark.parametrize("ysign", [1.0, -1.0])
def test_world_reconstruction_perspective(ysign):
    W, H, f = 32, 24, 30.0
    cam, proj = _persp(f, W, H)
    cam.clip_y_sign = ysign
    rng = np.random.default_rng(1)
    vz = rng.uniform(-20.0, -2.0, (H, W))
    depth = _depth_of(vz, proj.astype(np.float64)).astype(np.float32)
    cur, hc = _images(rng, W, H)
    oc, ox = orc.taa_accumulate(cur, depth, hc, np.zeros_like(hc), _u(cam))
    # closed form: pixel centre (i + .5, j + .5), rows flipped when clip_y_sign = -1
    i = np.arange(W)[None, :] + 0.5
    j = np.arange(H)[:, None] + 0.5
    py = j if ysign > 0 else H - j
    z = -proj[2, 3] / ((depth.astype(np.float64) * 2 - 1) + proj[2, 2])
    np.testing.assert_allclose(ox[..., 2], z, rtol=1e-5)
    np.testing.assert_allclose(ox[..., 0], (i - W / 2) * z / f, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(ox[..., 1], (py - H / 2) * z / f, rtol=1e-4, atol=1e-5)
    # no history: current colour, alpha 1
    np.testing.assert_array_equal(oc[..., :3], cur[..., :3])
    assert (oc[..., 3] == 1.0).all() and (ox[..., 3] == 1.0).all()


def test_world_reconstruction_orthographic_and_camera_motion():
    W, H = 32, 24
    cam, proj = _persp(30.0, W, H)
    cam.orthographic = 1.0
    cam.ortho_focal = 4.0
    # a camera moved to (3, -2, 5) and turned 30 degrees about y
    th = np.deg2rad(30.0)
    rot = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    eye = np.array([3.0, -2.0, 5.0])
    view = np.eye(4)
    view[:3, :3] = rot.T
    view[:3, 3] = -rot.T @ eye
    cam.view[:] = view.T.flatten().astype(np.float32)
    depth = np.full((H, W), 0.3, np.float32)
    rng = np.random.default_rng(2)
    cur, hc = _images(rng, W, H)
    _, ox = orc.taa_accumulate(cur, depth, hc, np.zeros_like(hc), _u(cam))
    z = -proj[2, 3] / ((0.3 * 2 - 1) + proj[2, 2])
    i = np.arange(W)[None, :] + 0.5
    j = np.arange(H)[:, None] + 0.5
    vx = (i - W / 2) / -4.0 + 0 * j
    vy = (j - H / 2) / -4.0 + 0 * i
    pv = np.stack([vx, vy, np.full_like(vx, z), np.ones_like(vx)], -1)
    world = pv @ np.linalg.inv(view).T
    np.testing.assert_allclose(ox[..., :3], world[..., :3], rtol=1e-4, atol=1e-4)


def test_unchanged_view_is_a_running_mean():
    W, H = 20, 16
    cam, proj = _persp(25.0, W, H)
    rng = np.random.default_rng(3)
    depth = _depth_of(rng.uniform(-15, -3, (H, W)), proj).astype(np.float32)
    for alpha, n in [(1.0, 1.0), (5.0, 5.0), (200.0, 128.0)]:
        cur, hc = _images(rng, W, H, alpha)
        hx = rng.uniform(-1, 1, (H, W, 4)).astype(np.float32)
        oc, ox = orc.taa_accumulate(cur, depth, hc, hx, _u(cam, cam))
        want = (hc[..., :3].astype(np.float64) * n + cur[..., :3]) / (n + 1)
        np.testing.assert_allclose(oc[..., :3], want, atol=2e-5)
        assert (oc[..., 3] == min(n + 1, 128.0)).all()
        _, world = orc.taa_accumulate(cur, depth, hc, hx, _u(cam))
        np.testing.assert_allclose(ox[..., :3], (hx[..., :3] * n + world[..., :3]) / (n + 1), atol=2e-5)


def test_one_pixel_move_reads_the_neighbour():
    """Orthographic, focal 4 px/unit: the previous view sat 0.25 units to the
    right (+x), so a world point lies one pixel further left in it."""
    W, H = 24, 16
    cam, proj = _persp(25.0, W, H)
    cam.orthographic = 1.0
    cam.ortho_focal = 4.0
    prev = gs.GsCameraUniforms.from_array(cam.as_array())
    pview = np.eye(4, dtype=np.float32)
    pview[0, 3] = -0.25  # view = translate(-eye), eye.x = +0.25
    prev.view[:] = pview.T.flatten()
    rng = np.random.default_rng(4)
    depth = np.full((H, W), 0.4, np.float32)
    cur, hc = _images(rng, W, H, 3.0)
    oc, _ = orc.taa_accumulate(cur, depth, hc, np.zeros_like(hc), _u(cam, prev))
    # ortho: screen x = -focal * view x + W/2, so view x grows to the left:
    # view_prev.x = view.x - 0.25 -> screen x grows by one pixel
    nb = np.concatenate([hc[:, 1:], hc[:, -1:]], axis=1)  # texel i + 1 (clamped at the edge)
    want = (nb[..., :3].astype(np.float64) * 3 + cur[..., :3]) / 4
    np.testing.assert_allclose(oc[:, :-1, :3], want[:, :-1], atol=1e-4)
    wrong = (np.concatenate([hc[:, :1], hc[:, :-1]], axis=1)[..., :3] * 3 + cur[..., :3]) / 4
    assert np.abs(oc[:, 1:-1, :3] - wrong[:, 1:-1]).max() > 0.1


def test_invalid_history_cases():
    W, H = 16, 12
    cam, proj = _persp(20.0, W, H)
    rng = np.random.default_rng(5)
    depth = _depth_of(np.full((H, W), -6.0), proj).astype(np.float32)
    cur, hc = _images(rng, W, H, 2.0)
    _, world = orc.taa_accumulate(cur, depth, hc, np.zeros_like(hc), _u(cam))

    def is_current(oc):
        return np.array_equal(oc[..., :3], cur[..., :3]) and (oc[..., 3] == 1.0).all()

    # valid reference case for contrast
    oc, _ = orc.taa_accumulate(cur, depth, hc, world, _u(cam, cam, True))
    assert not is_current(oc)
    # no history
    assert is_current(orc.taa_accumulate(cur, depth, hc, world, _u(cam))[0])
    # empty history (alpha 0)
    h0 = hc.copy()
    h0[..., 3] = 0.0
    assert is_current(orc.taa_accumulate(cur, depth, h0, world, _u(cam, cam))[0])
    # position moved beyond 100 units while the view changed; not checked when it did not
    far = world.copy()
    far[..., 0] += 150.0
    assert is_current(orc.taa_accumulate(cur, depth, hc, far, _u(cam, cam, True))[0])
    assert not is_current(orc.taa_accumulate(cur, depth, hc, far, _u(cam, cam, False))[0])
    near = world.copy()
    near[..., 0] += 50.0
    assert not is_current(orc.taa_accumulate(cur, depth, hc, near, _u(cam, cam, True))[0])
    # previous camera turned away: points behind it
    back = gs.GsCameraUniforms.from_array(cam.as_array())
    bview = np.diag([-1.0, 1.0, -1.0, 1.0]).astype(np.float32)
    back.view[:] = bview.T.flatten()
    assert is_current(orc.taa_accumulate(cur, depth, hc, world, _u(cam, back, True))[0])
    # previous camera moved far sideways: outside its view
    side = gs.GsCameraUniforms.from_array(cam.as_array())
    sview = np.eye(4, dtype=np.float32)
    sview[0, 3] = -50.0
    side.view[:] = sview.T.flatten()
    assert is_current(orc.taa_accumulate(cur, depth, hc, world, _u(cam, side, True))[0])


def test_taa_uniforms_fields():
    cam = scenes.random_view_camera(7, 64, 48)
    prev = scenes.random_view_camera(8, 64, 48)
    prev.orthographic = 1.0
    prev.clip_y_sign = -1.0
    u = gs.taa_uniforms(cam, prev, True)
    view = np.array(cam.view, np.float64).reshape(4, 4).T
    inv = np.array(u.inv_view, np.float64).reshape(4, 4).T
    np.testing.assert_allclose(inv, np.linalg.inv(view), rtol=1e-5, atol=1e-6)
    assert list(u.prev_view) == list(prev.view)
    assert list(u.viewport_orthographic) == [64.0, 48.0, cam.orthographic, 1.0]
    assert list(u.focal_clip) == [cam.persp_focal, cam.ortho_focal, cam.clip_y_sign, cam.clip_z_01]
    assert list(u.prev_focal_clip)[:3] == [prev.persp_focal, prev.ortho_focal, -1.0]
    assert list(u.proj_z)[:2] == [cam.proj[10], cam.proj[14]]
    assert u.history_valid == 1.0 and u.view_changed == 1.0
    u0 = gs.taa_uniforms(cam)
    assert u0.history_valid == 0.0 and list(u0.prev_view) == list(cam.view)
    bad = gs.GsCameraUniforms.from_array(cam.as_array())
    bad.view[:] = [0.0] * 16
    with pytest.raises(gs.GsError):
        gs.taa_uniforms(bad, prev)
