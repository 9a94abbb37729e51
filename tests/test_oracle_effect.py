"""Pins two pieces of the oracle's cull without retyping its float code:

- the shockwave effect branch (shaders/cull.glsl:183-207): inactive or with
  zero attack/release it leaves every record bit-identical to the plain
  cull; at full strength the effect cull equals the plain cull of a scene
  rewritten in float64 from the effect's definition (exp(-band^2) wave,
  displacement along the normalised jittered direction, scale x (1 + 0.75 w),
  opacity x (1 - 0.2 w), tint + rgb w a);
- the GPS point diagnostics diag[4..7] (cull.glsl:296-318): the point
  estimate 2 pi sqrt(det) Li2(o) of every visible splat against scipy's
  dilogarithm, summed, maximised and counted over 1024 and 16384.
"""
import numpy as np
import pytest
from scipy.special import spence

import gamesplat_amd as gs
from gamesplat_amd import scenes

W, H = 640, 480
CENTER = np.array([0.5, -0.2, 0.3])
RADIUS = 9.0
TINT = (1.0, 0.55, 0.18, 0.4)


def _effect(progress, push=0.12, active=1.0):
    duration = 1.5
    return gs.GsSplatEffectParams((*CENTER, RADIUS), (progress * duration, duration, push, active), TINT)


def _scene(seed, n, progress):
    """Gaussians around the effect centre at distances spanning the wave's
    band (wave radius progress * 1.15 R, band width 0.085 R)."""
    g = gs.generate_scene(seed, n, 4, 1.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    wr, bw = progress * RADIUS * 1.15, RADIUS * 0.085
    dist = wr + bw * rng.uniform(-3.0, 3.0, n)
    g[:, 0:3] = (CENTER[None, :] + d * dist[:, None]).astype(np.float32)
    return g


def _cam(seed):
    return scenes.random_view_camera(seed, W, H, radius=18.0)


def _hash11_f64(p):
    """hash11 (shaders/cull.glsl:101-103): fract(sin(p * 127.1) * 43758.5453123),
    the product p * 127.1 rounded to float as the shader forms it, the rest in
    float64 (the shader's float sin and product move the fraction by a few 1e-3)."""
    arg = (np.float32(p) * np.float32(127.1)).astype(np.float64)
    v = np.sin(arg) * 43758.5453123
    return v - np.floor(v)


def _rewritten(g, progress, push):
    """The scene the effect describes, formed in float64 (attack = release = 1)."""
    pos = g[:, 0:3].astype(np.float64)
    to = pos - CENTER[None, :]
    dist = np.linalg.norm(to, axis=1)
    band = (dist - progress * RADIUS * 1.15) / (RADIUS * 0.085)
    w = np.exp(-band * band)
    idx = np.arange(len(g), dtype=np.float64)
    jit = np.stack([_hash11_f64(idx + k) * 2.0 - 1.0 for k in (1, 2, 3)], axis=1) * (RADIUS * 0.015)
    d = to + jit + np.array([0.0, RADIUS * 0.01, 0.0])[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ripple = np.sin(progress * 18.849556 + dist * 8.0 / RADIUS)
    out = g.copy()
    out[:, 0:3] = (pos + d * (w * RADIUS * (push + 0.012 * ripple))[:, None]).astype(np.float32)
    out[:, 3:6] = (g[:, 3:6].astype(np.float64) * (1.0 + 0.75 * w)[:, None]).astype(np.float32)
    out[:, 13] = (g[:, 13].astype(np.float64) * (1.0 - 0.2 * w)).astype(np.float32)
    return out, w


@pytest.mark.parametrize("progress,active", [(0.0, 1.0), (1.0, 1.0), (0.5, 0.0)])
def test_effect_without_strength_is_the_plain_cull(orc, progress, active):
    """progress 0 (attack = smoothstep(0, 0.08, 0) = 0), progress 1 (release
    = 1 - smoothstep(0.82, 1, 1) = 0) and an inactive effect (params.w <=
    0.5) change no bit of any record."""
    g = _scene(3, 3000, 0.5)
    cam = _cam(41)
    plain = orc.cull(g, cam)
    fx = orc.cull(g, cam, effect=_effect(progress, active=active))
    for a, b in zip(plain[:3], fx[:3]):
        np.testing.assert_array_equal(a, b)
    assert plain[3] == fx[3] > 100


@pytest.mark.parametrize("progress,push", [(0.3, 0.12), (0.5, 0.0), (0.7, 0.25)])
def test_effect_equals_rewritten_scene(orc, progress, push):
    g = _scene(5, 3000, progress)
    cam = _cam(43)
    fk, fi, fp, fv, _ = orc.cull(g, cam, effect=_effect(progress, push), sh_degree=0)
    rg, w = _rewritten(g, progress, push)
    rk, ri, rp, rv, _ = orc.cull(rg, cam, sh_degree=0)
    both = (fi != 0xFFFFFFFF) & (ri != 0xFFFFFFFF)
    assert both.sum() > 1000 and w[both].max() > 0.99 and w[both].min() < 0.01
    # raster centre: the float64 displacement differs from the shader's by the
    # hash's float rounding (about 1e-3 of the jitter) and float rounding
    dc = np.abs(fp[both, 4:6].astype(np.float64) - rp[both, 4:6]).max()
    assert dc < 0.03, dc
    # opacity and the projected covariance (scale growth)
    np.testing.assert_allclose(fp[both, 3], rp[both, 3], rtol=2e-6)
    np.testing.assert_allclose(fp[both, 12:15], rp[both, 12:15], rtol=2e-3, atol=1e-4)
    # tint: colour(effect) - colour(plain) = rgb * w * a
    tint = np.array(TINT[:3])[None, :] * (w[both] * TINT[3])[:, None]
    np.testing.assert_allclose(fp[both, 0:3] - rp[both, 0:3], tint, atol=2e-6)
    # the effect moved the splats (the comparison is not vacuous)
    pk, pi, pp, pv, _ = orc.cull(g, cam, sh_degree=0)
    moved = both & (pi != 0xFFFFFFFF)
    assert np.abs(pp[moved, 4:6] - fp[moved, 4:6]).max() > 1.0


def _li2(x):
    return spence(1.0 - np.asarray(x, np.float64))


def test_gps_point_diagnostics_against_scipy(orc):
    """diag[4] = sum ceil(points / 1024), diag[5] = max points, diag[6] / [7] =
    splats over 1024 / 16384 points, with points = ceil(2 pi sqrt(det) Li2(o))
    (det the dilated 2D covariance determinant, o the opacity).  The shader
    uses a fitted Li2, so counts are bounded by the float64 values 0.2% on
    either side of each threshold."""
    g = gs.generate_scene(17, 6000, 8, 0.0)
    rng = np.random.Generator(np.random.PCG64(17))
    g[:, 0:3] *= 0.5
    g[:, 3:6] = np.exp(rng.uniform(np.log(0.01), np.log(3.0), (len(g), 1))).astype(np.float32) * \
        rng.uniform(0.5, 1.0, (len(g), 3)).astype(np.float32)
    g[:, 13] = rng.uniform(0.01, 1.0, len(g)).astype(np.float32)
    cam = _cam(47)
    keys, ids, proj, vis, diag = orc.cull(g, cam, collect_stats=True)
    v = ids != 0xFFFFFFFF
    pts = 2.0 * np.pi * np.sqrt(proj[v, 15].astype(np.float64)) * _li2(np.clip(proj[v, 3], 0.0, 1.0))
    d = 2e-3
    lo, hi = pts * (1 - d), pts * (1 + d)
    assert (pts > 16384).sum() > 20 and (pts < 1024).sum() > 100
    assert np.ceil(lo).max() <= diag[5] <= np.ceil(hi).max() + 1
    for k, th in ((6, 1024), (7, 16384)):
        assert (lo > th).sum() <= diag[k] <= (hi > th).sum(), (k, diag[k])
    s_lo = (np.ceil(lo).astype(np.int64) + 1023) >> 10
    s_hi = (np.ceil(hi).astype(np.int64) + 1 + 1023) >> 10
    assert s_lo.sum() <= diag[4] <= s_hi.sum()
    # a splat of opacity 1 collects pi^2 / 6 of the 2 pi sqrt(det) points
    one = gs.generate_scene(19, 1, 1, 0.0)
    one[0, 0:3] = 0.0
    one[0, 3:6] = 0.5
    one[0, 13] = 1.0
    _, _, p1, v1, d1 = orc.cull(one, cam, collect_stats=True)
    assert v1 == 1
    exact = 2.0 * np.pi * np.sqrt(float(p1[0, 15])) * np.pi ** 2 / 6.0
    assert abs(int(d1[5]) - exact) <= 1e-3 * exact + 1
