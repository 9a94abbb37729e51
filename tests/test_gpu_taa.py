"""StochasticSplats with TAA on the HIP library against the CPU oracle:
Halton-jittered samples, the per-frame average and the TAA history
(colour and world position) over a moving camera, all bit-exact; the
double-buffered history survives a tile-list overflow re-run; argument and
state errors."""
import numpy as np
import pytest

import gamesplat_amd as gsm
from gamesplat_amd import abi, scenes
from headless_util import headless_position

pytestmark = pytest.mark.gpu


def _oracle_frame(orc, g, cam, seeds, bg=0.0):
    """Jittered samples of one frame and their running mean (framebuffer rows)."""
    _, ids, proj, _, _ = orc.cull(g, cam)
    W, H = int(cam.viewport[0]), int(cam.viewport[1])
    avg = depth = None
    for k, seed in enumerate(seeds):
        c, depth = orc.stochastic_splats(proj, ids, W, H, seed, background=bg, jitter=gsm.taa_jitter(seed))
        avg = c.copy() if k == 0 else orc.accumulate(avg, c, k + 1)
    return c, avg, depth


@pytest.mark.parametrize("seed", [3, 700])
def test_jittered_sample_bit_exact(gs, orc, seed):
    g = gs.generate_scene(0x7AA0001, 30000, 16, 0.0)
    cam = scenes.random_view_camera(0x7AA0001, 192, 144)
    jit = gs.taa_jitter(seed)
    print(f"case: jittered stochastic sample seed={seed} jitter={jit}", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.cull(cam)
        c.stochastic_sample(seed, 0.0, jit)
        col = c.read_image(abi.GS_IMAGE_GPS_SAMPLE_RGBA32F)
        dep = c.read_image(abi.GS_IMAGE_GPS_DEPTH_R32F)
    rc, _, rd = _oracle_frame(orc, g, cam, [seed])
    np.testing.assert_array_equal(col, rc[::-1])
    np.testing.assert_array_equal(dep, rd[::-1])


def test_taa_sequence_bit_exact(gs, orc):
    """Four frames of an orbiting camera, two jittered samples per frame,
    frame average -> TAA; colour and position history checked every frame."""
    g = gs.generate_scene(0x7AA0002, 30000, 16, 0.5)
    W, H = 160, 120
    cams = [scenes.orbit_camera(0x7AA0002, f, W, H, step_deg=2.0) for f in range(4)]
    hist_c = hist_x = None
    seed = 40
    with gs.Context(0) as c:
        c.upload(g)
        for f, cam in enumerate(cams):
            print(f"case: TAA frame {f}", flush=True)
            seeds = [seed + 1, seed + 2]
            seed += 2
            c.cull(cam)
            for k, s in enumerate(seeds):
                c.stochastic_sample(s, 0.0, gs.taa_jitter(s))
                c.frame_average(k)
            prev = cams[f - 1] if f else None
            u = gs.taa_uniforms(cam, prev, f > 0)
            c.taa_accumulate(u, use_frame_average=True)
            avg_gpu = c.read_image(abi.GS_IMAGE_FRAME_AVG_RGBA32F)
            tc = c.read_image(abi.GS_IMAGE_TAA_RGBA32F)
            tx = c.read_image(abi.GS_IMAGE_TAA_XYZ_RGBA32F)

            _, avg, depth = _oracle_frame(orc, g, cam, seeds)
            if hist_c is None:
                hist_c = np.zeros((H, W, 4), np.float32)
                hist_x = np.zeros((H, W, 4), np.float32)
            hist_c, hist_x = orc.taa_accumulate(avg, depth, hist_c, hist_x, u.as_array())
            np.testing.assert_array_equal(avg_gpu, avg[::-1])
            np.testing.assert_array_equal(tc, hist_c[::-1])
            np.testing.assert_array_equal(tx, hist_x[::-1])
            if f:
                blended = (hist_c[..., 3] > 1.0).mean()
                print(f"frame {f}: {blended:.3f} of pixels blended with history", flush=True)
                assert blended > 0.5
    # one TAA pass per frame: the history length reaches the frame count
    # (bilinear weights can leave it a few ulps off)
    assert 3.99 < hist_c[..., 3].max() < 4.01


def test_taa_single_sample_and_restart(gs, orc):
    """Without a frame average the last sample is the current colour; a
    resize needs a restart (history_valid 0)."""
    g = gs.generate_scene(0x7AA0003, 20000, 8, 0.0)
    cam = scenes.random_view_camera(0x7AA0003, 128, 96)
    print("case: TAA from the single sample, unchanged view", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.cull(cam)
        hist = None
        for f in range(3):
            s = 10 + f
            c.stochastic_sample(s, 0.0, gs.taa_jitter(s))
            u = gs.taa_uniforms(cam, cam if f else None, False)
            c.taa_accumulate(u)
            rc, _, rd = _oracle_frame(orc, g, cam, [s])
            if hist is None:
                hist = (np.zeros((96, 128, 4), np.float32),) * 2
            hist = orc.taa_accumulate(rc, rd, hist[0], hist[1], u.as_array())
        np.testing.assert_array_equal(c.read_image(abi.GS_IMAGE_TAA_RGBA32F), hist[0][::-1])
        np.testing.assert_array_equal(c.read_image(abi.GS_IMAGE_TAA_XYZ_RGBA32F), hist[1][::-1])
        # a new size: history must restart
        small = scenes.random_view_camera(0x7AA0003, 64, 48)
        c.cull(small)
        c.stochastic_sample(20, 0.0, gs.taa_jitter(20))
        with pytest.raises(gs.GsError):
            c.taa_accumulate(gs.taa_uniforms(small, cam, True))
        c.taa_accumulate(gs.taa_uniforms(small))
        assert c.read_image(abi.GS_IMAGE_TAA_RGBA32F).shape == (48, 64, 4)


def test_jittered_overflow_rerun(gs, orc):
    """Large splats overflow the first tile-list capacity; the re-run keeps
    the jitter, so the averaged frame still matches."""
    g = gs.generate_scene(17, 3000, 4, 0.0)
    g[:, 3:6] *= 40.0
    cam = scenes.random_view_camera(17, 320, 240)
    print("case: jittered sample tile-pair overflow and re-run", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.cull(cam)
        for k, s in enumerate([5, 6]):
            c.stochastic_sample(s, 0.0, gs.taa_jitter(s))
            c.frame_average(k)
        avg = c.read_image(abi.GS_IMAGE_FRAME_AVG_RGBA32F)
    _, ravg, _ = _oracle_frame(orc, g, cam, [5, 6])
    np.testing.assert_array_equal(avg, ravg[::-1])


def test_taa_errors(gs):
    print("case: TAA argument and state errors", flush=True)
    g = gs.generate_scene(1, 1000, 2, 0.0)
    cam = scenes.random_view_camera(1, 64, 64)
    with gs.Context(0) as c:
        with pytest.raises(gs.GsError):
            c.taa_accumulate(gs.taa_uniforms(cam))  # no sample
        with pytest.raises(gs.GsError):
            c.frame_average(0)
        c.upload(g)
        c.cull(cam)
        with pytest.raises(gs.GsError):
            c.stochastic_sample(1, 0.0, (float("nan"), 0.0))
        c.stochastic_sample(1, 0.0, (0.25, 0.25))
        with pytest.raises(gs.GsError):
            c.frame_average(1)  # must start at 0
        with pytest.raises(gs.GsError):
            c.taa_accumulate(gs.taa_uniforms(cam), use_frame_average=True)  # no frame average
        with pytest.raises(gs.GsError):
            c.taa_accumulate(gs.taa_uniforms(cam, cam))  # history requested, none yet
        other = scenes.random_view_camera(1, 32, 64)
        with pytest.raises(gs.GsError):
            c.taa_accumulate(gs.taa_uniforms(other))  # viewport does not match the sample
        with pytest.raises(gs.GsError):
            c.read_image(abi.GS_IMAGE_TAA_RGBA32F)
        c.taa_accumulate(gs.taa_uniforms(cam))
        c.read_image(abi.GS_IMAGE_TAA_RGBA32F)


def test_renderer_mirror_taa_matches_library_sequence(gs, tmp_path):
    """renderer_draw_frame's StochasticSplats + TAA glue (headless driver,
    taa=1): the same frame seeds, jitters, frame averages and TAA uniforms
    through the Python binding give the same final history, bit for bit."""
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parents[1] / "gamesplat_amd" / "gs_headless"
    W, H, n, seed, spp, warm, frames = 160, 120, 20000, 0x5EED0101, 2, 1, 3
    cap = tmp_path / "taa.f32"
    print("case: headless StochasticSplats + TAA vs library sequence", flush=True)
    r = subprocess.run([str(exe), f"gaussians={n}", f"seed={seed}", f"width={W}", f"height={H}",
                        "render_mode=stochastic", "taa=1", f"taa_samples={spp}", "orbit_deg=2",
                        f"warmup={warm}", f"frames={frames}", f"linear_capture={cap}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(cap, np.float32).reshape(H, W, 4)

    g = gs.generate_scene(seed, n, 16, 0.0)
    frame_seed = 0
    prev = None
    with gs.Context(0) as c:
        c.upload(g)
        for f in range(warm + frames):
            pose = gs.look_at_origin_pose(headless_position(seed, f, orbit_deg=2.0))
            cam = gs.camera_from_pose(pose, np.float32(0.8660254), W, H)
            changed = prev is not None and bytes(prev) != bytes(cam)
            u = gs.taa_uniforms(cam, prev, changed)
            c.cull(cam)
            for i in range(spp):
                frame_seed += 1
                c.stochastic_sample(frame_seed, 0.0, gs.taa_jitter(frame_seed))
                c.frame_average(i)
            c.taa_accumulate(u, use_frame_average=True)
            prev = cam
        want = c.read_image(abi.GS_IMAGE_TAA_RGBA32F)
    assert want[..., 3].max() > 3.5  # the history spans the frames
    np.testing.assert_array_equal(got, want)


def test_jitter_extremes_and_list_reuse(gs, orc):
    """Samples of one cull share tile lists: unjittered, then jitters at the
    half-pixel limits (lists grown by a pixel), a jitter beyond it (lists
    moved exactly), and back; every sample bit-exact, near the image edges
    too (a wide view of a large scene)."""
    g = gs.generate_scene(0x7AA0005, 40000, 32, 0.0)
    cam = scenes.random_view_camera(0x7AA0005, 200, 136, radius=9.0)
    _, ids, proj, _, _ = orc.cull(g, cam)
    jitters = [(0.0, 0.0), (-0.5, -0.5), (0.5, 0.5), (0.5, -0.5), (1.5, -2.25), (0.0, 0.0), (-0.25, 0.5)]
    with gs.Context(0) as c:
        c.upload(g)
        c.cull(cam)
        for k, jit in enumerate(jitters):
            print(f"case: sample {k} jitter {jit}", flush=True)
            c.stochastic_sample(100 + k, 0.0, jit)
            col = c.read_image(abi.GS_IMAGE_GPS_SAMPLE_RGBA32F)
            rc, _ = orc.stochastic_splats(proj, ids, 200, 136, 100 + k, jitter=jit)
            np.testing.assert_array_equal(col, rc[::-1])
            # covered pixels on the image border
            assert (rc[0, :, 3] > 0).any() and (rc[:, -1, 3] > 0).any()
