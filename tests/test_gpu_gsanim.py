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


def _sample(keys, duration, version, linear, t):
    """Playback state at time t: (state_a, state_b, weight) from (time, state,
    linear) keys, written from the asset format's rules."""
    u = t % duration
    past = [i for i, k in enumerate(keys) if k[0] <= u]
    i = past[-1] if past else 0
    lin = keys[i][2] if version == 2 else linear
    if not lin or len(keys) == 1:
        return keys[i][1], keys[i][1], 0.0
    j = (i + 1) % len(keys)
    end = keys[j][0] if j else duration
    return keys[i][1], keys[j][1], float(np.float32((u - keys[i][0]) / (end - keys[i][0])))


def _layer_weights(layer_keys, duration, t):
    u = t % duration
    out = []
    for keys in layer_keys:
        past = [i for i, k in enumerate(keys) if k[0] <= u]
        i = past[-1] if past else 0
        j = (i + 1) % len(keys)
        end = keys[j][0] if j else duration
        w = (u - keys[i][0]) / (end - keys[i][0]) if keys[i][4] and len(keys) > 1 else 0.0
        # the key weights are floats, the blend runs in double
        wi = [float(np.float32(keys[i][1 + c])) for c in range(3)]
        wj = [float(np.float32(keys[j][1 + c])) for c in range(3)]
        out.append([np.float32((1.0 - w) * wi[c] + w * wj[c]) for c in range(3)])
    return np.array(out, np.float32)


def _headless(args, tmp_path):
    import subprocess

    from headless_util import HEADLESS

    r = subprocess.run([str(HEADLESS), *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def _headless_camera(gs, seed, W, H):
    from headless_util import FOCAL_OVER_HEIGHT, headless_position

    return gs.camera_from_pose(gs.look_at_origin_pose(headless_position(seed, 0)), FOCAL_OVER_HEIGHT, W, H)


def test_renderer_mirror_drives_delta_states(gs, tmp_path):
    """gs_headless with a delta GSAnim asset: the playback state it samples
    at the last frame's time, replayed through the binding, gives the same
    image bit for bit."""
    W, H, n, seed = 192, 144, 20000, 0x6A06
    d, scales = _deltas(3, n, 13)
    d.tofile(tmp_path / "d.i8")
    scales.tofile(tmp_path / "s.f32")
    keys = [(0.0, 1, 1), (0.4, 3, 0), (0.7, 2, 1)]
    t = 0.25 + 0.3 * 2  # anim_time + step * last frame
    print("case: headless GSAnim deltas, version 2 keys", flush=True)
    _headless([f"gaussians={n}", f"seed={seed}", f"width={W}", f"height={H}", "warmup=0", "frames=3",
               f"anim_deltas={tmp_path / 'd.i8'}", f"anim_scales={tmp_path / 's.f32'}", "anim_states=3",
               "anim_keys=" + ",".join(f"{k[0]}:{k[1]}:{k[2]}" for k in keys), "anim_version=2",
               "anim_duration=1.0", "anim_time=0.25", "anim_time_step=0.3",
               f"linear_capture={tmp_path / 'out.f32'}"], tmp_path)
    got = np.fromfile(tmp_path / "out.f32", np.float32).reshape(H, W, 4)
    a, b, w = _sample(keys, 1.0, 2, False, t)
    assert (a, b) == (2, 1) and abs(w - 0.5) < 1e-6
    with gs.Context(0) as c:
        c.upload(gs.generate_scene(seed, n, 16, 0.0))
        c.upload_animation(deltas=d, scales=scales)
        c.set_animation_frame(a, b, w)
        c.render_alpha(_headless_camera(gs, seed, W, H))
        want = c.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).astype(np.float32)
        c.set_animation_frame(0, 0, 0.0)
        c.render_alpha(_headless_camera(gs, seed, W, H))
        still = c.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).astype(np.float32)
    np.testing.assert_array_equal(got, want)
    assert not np.array_equal(got, still)


def test_renderer_mirror_drives_layers_and_display(gs, tmp_path):
    W, H, n, seed = 160, 120, 20000, 0x6A07
    bits = _layers(2, n, 14)
    bits.tofile(tmp_path / "l.u16")
    rng = np.random.default_rng(15)
    lut = rng.integers(0, 65536, (9, 9, 9, 3)).astype(np.uint16)
    lut.tofile(tmp_path / "lut.u16")
    layer_keys = [[(0.0, 1.0, 0.5, 0.25, 1), (0.5, 0.2, 1.0, 2.0, 0)], [(0.0, 0.3, 0.3, 0.3, 0)]]
    spec = ";".join(",".join(f"{k[0]}:{k[1]}:{k[2]}:{k[3]}:{k[4]}" for k in ks) for ks in layer_keys)
    t = 0.2 + 0.1 * 1
    print("case: headless GSAnim layers + LUT display", flush=True)
    _headless([f"gaussians={n}", f"seed={seed}", f"width={W}", f"height={H}", "warmup=0", "frames=2",
               f"anim_layers={tmp_path / 'l.u16'}", "anim_layer_count=2", f"anim_layer_keys={spec}",
               "anim_duration=1.0", "anim_time=0.2", "anim_time_step=0.1", f"anim_lut={tmp_path / 'lut.u16'}",
               "anim_lut_size=9", "anim_log2_min=-8", "anim_log2_max=3", "exposure_ev=0.5",
               f"linear_capture={tmp_path / 'lin.f32'}", f"display_capture={tmp_path / 'disp.f32'}"], tmp_path)
    lin = np.fromfile(tmp_path / "lin.f32", np.float32).reshape(H, W, 4)
    disp = np.fromfile(tmp_path / "disp.f32", np.float32).reshape(H, W, 4)
    weights = _layer_weights(layer_keys, 1.0, t)
    with gs.Context(0) as c:
        c.upload(gs.generate_scene(seed, n, 16, 0.0))
        c.upload_animation(layers=bits, lut=lut, log2_min=-8.0, log2_max=3.0)
        c.set_animation_frame(layer_weights=weights)
        c.render_alpha(_headless_camera(gs, seed, W, H))
        want = c.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).astype(np.float32)
        c.display_layers(abi.GS_IMAGE_ALPHA_RGBA16F, 0.5)
        want_disp = c.read_image(abi.GS_IMAGE_DISPLAY_RGBA32F)
    np.testing.assert_array_equal(lin, want)
    np.testing.assert_array_equal(disp, want_disp)
    assert (disp[..., 3] == 1.0).all() and lin[..., :3].max() > 0
