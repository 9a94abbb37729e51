"""Pins of the GSAnim appearance branches of the cull oracle (fetch_vec4 /
animation_coefficient and the layer sum, shaders/cull.glsl:125-164,353-418)
and of the LUT display (layers_display_fs, shaders/accum.glsl:197-240), CPU
only.

- Deltas: the animated cull equals the plain cull of a scene whose SH
  coefficients numpy rewrote itself (base + q * scale per coefficient, with
  the channel-major coefficient index 3 + 15 c + k), for one state and for a
  lerp of two; state 0 is the scene itself.
- Layers: each colour channel equals an FP64 evaluation of
  sum_l w_l * max(SH_l(dir), 0) (no +0.5), the sum over layers equals the
  float sum of single-layer culls bit for bit, and a layer count below the
  uploaded one uses the first layers only.
- Display: an identity lattice returns the log2-domain coordinate s in
  closed form; a constant lattice returns its constant; the result is
  monotone in radiance.
"""
import numpy as np

import gamesplat_amd as gs
from gamesplat_amd import scenes
from oracle import oracle as orc

C0 = 0.28209479177387814
C1 = 0.4886025119029199
C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396]
C3 = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
      1.445305721320277, -0.5900435899266435]


def _scene(n=3000, seed=11):
    g = gs.generate_scene(seed, n, 8, 0.5)
    cam = scenes.random_view_camera(seed, 160, 120, radius=12.0)
    return g, cam


def _deltas(states, rows, seed):
    rng = np.random.default_rng(seed)
    d = rng.integers(-128, 128, (states, rows, 48), dtype=np.int16).astype(np.int8)
    scales = rng.uniform(0.001, 0.02, 48).astype(np.float32)
    return d, scales


def _coef_index(r):
    """Channel-major coefficient of rest value r = 3 k + c."""
    return 3 + (r % 3) * 15 + r // 3


def _apply_state(g, d, scales, state):
    """Coefficients of `state` (0 = the scene): base + q * scale in float32."""
    out = g.copy()
    if state == 0:
        return out
    q = d[state - 1].astype(np.float32)
    for c in range(3):
        out[:, 10 + c] = g[:, 10 + c] + q[:, c] * scales[c]
    for r in range(45):
        j = _coef_index(r)
        out[:, 14 + r] = g[:, 14 + r] + q[:, j] * scales[j]
    return out


def test_delta_state_zero_is_the_scene():
    g, cam = _scene()
    d, scales = _deltas(3, len(g), 1)
    plain = orc.cull(g, cam)[2]
    anim = orc.cull(g, cam, animation={"deltas": d, "scales": scales, "state_a": 0, "state_b": 0,
                                       "weight": 0.0})[2]
    np.testing.assert_array_equal(anim.view(np.uint32), plain.view(np.uint32))


def test_delta_state_equals_rewritten_scene():
    g, cam = _scene()
    d, scales = _deltas(3, len(g), 2)
    for state in (1, 3):
        anim = orc.cull(g, cam, animation={"deltas": d, "scales": scales, "state_a": state, "weight": 0.0})[2]
        want = orc.cull(_apply_state(g, d, scales, state), cam)[2]
        np.testing.assert_array_equal(anim.view(np.uint32), want.view(np.uint32))
        assert not np.array_equal(anim, orc.cull(g, cam)[2])


def test_delta_lerp_equals_rewritten_scene():
    g, cam = _scene()
    d, scales = _deltas(4, len(g), 3)
    for a, b, w in [(1, 3, 0.25), (0, 2, 0.5), (4, 0, 0.75)]:
        anim = orc.cull(g, cam, animation={"deltas": d, "scales": scales, "state_a": a, "state_b": b,
                                           "weight": w})[2]
        ga, gb = _apply_state(g, d, scales, a), _apply_state(g, d, scales, b)
        wf = np.float32(w)
        mixed = g.copy()
        cols = list(range(10, 13)) + list(range(14, 59))
        mixed[:, cols] = (np.float32(1.0) - wf) * ga[:, cols] + wf * gb[:, cols]
        want = orc.cull(mixed, cam)[2]
        np.testing.assert_array_equal(anim.view(np.uint32), want.view(np.uint32))


def _layers(count, rows, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-0.4, 0.4, (count, rows, 48)).astype(np.float32)
    v[:, :, :3] = rng.uniform(-1.0, 2.5, (count, rows, 3))  # DC, some layers negative
    return v.astype(np.float16).view(np.uint16), v.astype(np.float16).astype(np.float64)


def _basis(d):
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
    return np.stack([np.full_like(x, C0), -C1 * y, C1 * z, -C1 * x,
                     C2[0] * xy, C2[1] * yz, C2[2] * (2 * zz - xx - yy), C2[3] * xz, C2[4] * (xx - yy),
                     C3[0] * y * (3 * xx - yy), C3[1] * xy * z, C3[2] * y * (4 * zz - xx - yy),
                     C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy), C3[5] * z * (xx - yy),
                     C3[6] * x * (xx - 3 * yy)], 1)


def test_layers_match_fp64_sum_of_clamped_layers():
    g, cam = _scene(seed=12)
    bits, vals = _layers(3, len(g), 4)
    weights = np.array([[1.0, 0.5, 0.25], [0.3, 1.2, 0.0], [2.0, 0.1, 0.7]], np.float32)
    _, ids, proj, _, _ = orc.cull(g, cam, animation={"layers": bits, "layer_weights": weights})
    vis = ids != 0xFFFFFFFF
    view = np.array(cam.view, np.float64).reshape(4, 4).T
    campos = -view[:3, :3].T @ view[:3, 3]
    d = g[vis, 0:3].astype(np.float64) - campos
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    B = _basis(d)  # (v, 16)
    want = np.zeros((vis.sum(), 3))
    clamped = 0
    for l in range(3):
        for c in range(3):
            coef = np.concatenate([vals[l][vis][:, c:c + 1],
                                   vals[l][vis][:, [3 + c * 15 + k for k in range(15)]]], 1)
            res = (B * coef).sum(1)
            clamped += (res < 0).sum()
            want[:, c] += np.maximum(res, 0.0) * weights[l, c]
    np.testing.assert_allclose(proj[vis, 0:3], want, rtol=2e-5, atol=2e-5)
    assert clamped > 100  # the per-layer clamp is exercised


def test_layer_sum_is_the_sum_of_single_layers():
    g, cam = _scene(seed=13)
    bits, _ = _layers(2, len(g), 5)
    w = np.array([[0.75, 1.5, 0.5], [1.25, 0.25, 2.0]], np.float32)
    both = orc.cull(g, cam, animation={"layers": bits, "layer_weights": w})[2]
    one = orc.cull(g, cam, animation={"layers": bits[:1], "layer_weights": w[:1]})[2]
    two = orc.cull(g, cam, animation={"layers": bits[1:], "layer_weights": w[1:]})[2]
    np.testing.assert_array_equal(both[:, 0:3], one[:, 0:3] + two[:, 0:3])
    np.testing.assert_array_equal(both[:, 3:], one[:, 3:])  # geometry untouched
    first = orc.cull(g, cam, animation={"layers": bits, "layer_weights": w, "layer_count": 1})[2]
    np.testing.assert_array_equal(first, one)


def test_display_identity_and_constant_lattice():
    size = 17
    idx = np.arange(size, dtype=np.float64) / (size - 1)
    ident = np.zeros((size, size, size, 3), np.uint16)
    ident[..., 0] = np.round(idx[None, None, :] * 65535)
    ident[..., 1] = np.round(idx[None, :, None] * 65535)
    ident[..., 2] = np.round(idx[:, None, None] * 65535)
    rng = np.random.default_rng(6)
    img = np.exp2(rng.uniform(-12.0, 6.0, (64, 64, 4))).astype(np.float32)
    out = orc.layers_display(img, ident, -10.0, 4.0, exposure_ev=0.5)
    rad = img[..., :3].astype(np.float64) * 2.0 ** 0.5
    s = np.clip((np.log2(np.maximum(rad, 2.0 ** -10)) + 10.0) / 14.0, 0.0, 1.0)
    # lattice entries are rounded to 1/65535: half a step (7.6e-6) plus float error
    np.testing.assert_allclose(out[..., :3], s, atol=8e-6)
    assert (out[..., 3] == 1.0).all()
    const = np.full((5, 5, 5, 3), 12345, np.uint16)
    np.testing.assert_allclose(orc.layers_display(img, const, -10.0, 4.0)[..., :3], 12345 / 65535, atol=1e-7)
    ramp = np.exp2(np.linspace(-12, 6, 200)).astype(np.float32)
    col = np.zeros((200, 4), np.float32)
    col[:, 0] = ramp
    r = orc.layers_display(col, ident, -10.0, 4.0)[:, 0]
    assert (np.diff(r) >= 0).all()
