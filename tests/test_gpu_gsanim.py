"""GSAnim appearance in the HIP cull against the CPU oracle: delta states
(one state, lerps, state 0) and coefficient layers (all, first only), the
projected records bit-exact; animated multi-view frames equal their
single-view renders; the LUT display against its restatement; argument and
state errors."""
import numpy as np
import pytest

from gamesplat_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _scene(gs, n, seed):
    g = gs.generate_scene(seed, n, 16, 0.5)
    cam = scenes.random_view_camera(seed, 320, 240, radius=12.0)
    return g, cam


def _deltas(states, rows, seed):
    rng = np.random.default_rng(seed)
    d = rng.integers(-128, 128, (states, rows, 48), dtype=np.int16).astype(np.int8)
    return d, rng.uniform(0.001, 0.02, 48).astype(np.float32)


def _layers(count, rows, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-0.4, 0.4, (count, rows, 48)).astype(np.float32)
    v[:, :, :3] = rng.uniform(-1.0, 2.5, (count, rows, 3))
    return v.astype(np.float16).view(np.uint16)


@pytest.mark.parametrize("a,b,w", [(2, 0, 0.0), (1, 3, 0.25), (0, 2, 0.5), (0, 0, 0.0)])
@pytest.mark.parametrize("deg", [3, 1])
def test_delta_states_bit_exact(gs, orc, a, b, w, deg):
    g, cam = _scene(gs, 30000, 0x6A01)
    d, scales = _deltas(3, len(g), 7)
    print(f"case: GSAnim deltas state_a={a} state_b={b} weight={w} sh_degree={deg}", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.upload_animation(deltas=d, scales=scales)
        c.set_animation_frame(a, b, w)
        c.cull(cam, sh_degree=deg)
        proj = c.read_buffer(abi.GS_BUFFER_PROJECTED)
        nvis = c.stats()["visible_count"]
    _, _, want, ov, _ = orc.cull(g, cam, sh_degree=deg, animation={"deltas": d, "scales": scales, "state_a": a,
                                                                  "state_b": b, "weight": w})
    assert nvis == ov
    np.testing.assert_array_equal(proj.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("count", [3, 1])
def test_layers_bit_exact(gs, orc, count):
    g, cam = _scene(gs, 30000, 0x6A02)
    bits = _layers(3, len(g), 8)
    w = np.array([[1.0, 0.5, 0.25], [0.3, 1.2, 0.0], [2.0, 0.1, 0.7]], np.float32)
    print(f"case: GSAnim layers, {count} summed", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.upload_animation(layers=bits)
        c.set_animation_frame(layer_weights=w[:count])
        c.cull(cam)
        proj = c.read_buffer(abi.GS_BUFFER_PROJECTED)
    want = orc.cull(g, cam, animation={"layers": bits, "layer_weights": w, "layer_count": count})[2]
    np.testing.assert_array_equal(proj.view(np.uint32), want.view(np.uint32))


def test_animated_views_equal_single_renders(gs):
    g, _ = _scene(gs, 40000, 0x6A03)
    d, scales = _deltas(2, len(g), 9)
    cams = [scenes.random_view_camera(0x6A03 + k, 160 + 32 * k, 120, radius=12.0) for k in range(3)]
    print("case: GSAnim deltas over three batched views", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.upload_animation(deltas=d, scales=scales)
        c.set_animation_frame(1, 2, 0.4)
        c.render_views(cams)
        batched = [c.read_view_image(v) for v in range(3)]
        for v, cam in enumerate(cams):
            c.render_alpha(cam)
            np.testing.assert_array_equal(c.read_image(abi.GS_IMAGE_ALPHA_RGBA16F), batched[v])


def test_layer_display_matches_restatement(gs, orc):
    g, cam = _scene(gs, 30000, 0x6A04)
    bits = _layers(2, len(g), 10)
    rng = np.random.default_rng(11)
    lut = rng.integers(0, 65536, (9, 9, 9, 3)).astype(np.uint16)
    print("case: GSAnim layer frame through the display LUT", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.upload_animation(layers=bits, lut=lut, log2_min=-8.0, log2_max=3.0)
        c.set_animation_frame(layer_weights=[[1.0, 0.8, 0.6], [0.4, 0.5, 0.9]])
        c.render_alpha(cam)
        lin = c.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).astype(np.float32)
        c.display_layers(abi.GS_IMAGE_ALPHA_RGBA16F, 0.75)
        out = c.read_image(abi.GS_IMAGE_DISPLAY_RGBA32F)
    want = orc.layers_display(lin, lut, -8.0, 3.0, 0.75)
    # exp2 / log2 differ by an ulp between the device and libm
    np.testing.assert_allclose(out, want, atol=2e-4)
    assert np.abs(out - want).mean() < 1e-6


def test_animation_errors(gs):
    print("case: GSAnim argument and state errors", flush=True)
    g, cam = _scene(gs, 2000, 0x6A05)
    d, scales = _deltas(2, len(g), 12)
    with gs.Context(0) as c:
        with pytest.raises(gs.GsError):
            c.upload_animation(deltas=d, scales=scales)  # no scene yet
        c.upload(g)
        with pytest.raises(gs.GsError):
            c.upload_animation(deltas=d[:, :100], scales=scales)  # rows differ from the scene
        with pytest.raises(gs.GsError):
            c.set_animation_frame(1, 0, 0.0)  # no states uploaded
        c.upload_animation(deltas=d, scales=scales)
        with pytest.raises(gs.GsError):
            c.set_animation_frame(3, 0, 0.0)  # state beyond the two uploaded
        with pytest.raises(gs.GsError):
            c.set_animation_frame(1, 2, float("nan"))
        with pytest.raises(gs.GsError):
            c.set_animation_frame(layer_weights=[[1, 1, 1]])  # deltas, not layers
        with pytest.raises(gs.GsError):
            c.display_layers(abi.GS_IMAGE_ALPHA_RGBA16F)  # no LUT
        c.set_animation_frame(2, 1, 0.5)
        c.render_alpha(cam)
        c.upload(g)  # a new scene drops the animation
        with pytest.raises(gs.GsError):
            c.set_animation_frame(1, 0, 0.0)
        c.upload_animation(None)
