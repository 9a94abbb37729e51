"""StochasticSplats samples of the HIP library against the CPU oracle
(or_stochastic_splats): colour and depth bit-exact (same fragment
arithmetic, deterministic exp/log, float hash13, D24 depth, index-order ties),
the running mean of several samples bit-exact, and the tile-list overflow
re-run giving the same sample."""
import numpy as np
import pytest

from gamesplat_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _oracle_sample(orc, g, cam, seed, bg=0.0):
    keys, ids, proj, vis, _ = orc.cull(g, cam)
    W, H = int(cam.viewport[0]), int(cam.viewport[1])
    color, depth = orc.stochastic_splats(proj, ids, W, H, seed, background=bg)
    return color[::-1], depth[::-1]  # top row first, as the library reads back


@pytest.mark.parametrize("seed,bg", [(1, 0.0), (2, 0.0), (173, 0.1)])
def test_stochastic_sample_bit_exact(gs, orc, seed, bg):
    g = gs.generate_scene(0x5EED5701, 40000, 16, 0.0)
    cam = scenes.random_view_camera(0x5EED5701, 320, 240)
    print(f"case: stochastic sample seed={seed} bg={bg}", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.cull(cam)
        c.stochastic_sample(seed, bg)
        col = c.read_image(abi.GS_IMAGE_GPS_SAMPLE_RGBA32F)
        dep = c.read_image(abi.GS_IMAGE_GPS_DEPTH_R32F)
    rc, rd = _oracle_sample(orc, g, cam, seed, bg)
    diff = np.argwhere((col != rc).any(axis=2))
    print(f"pixels differing: {len(diff)} of {col.shape[0] * col.shape[1]}", flush=True)
    np.testing.assert_array_equal(col, rc)
    np.testing.assert_array_equal(dep, rd)
    assert (rc[..., 3] == 1.0).mean() > 0.2


def test_stochastic_running_mean(gs, orc):
    g = gs.generate_scene(0x5EED5702, 30000, 8, 1.0)
    cam = scenes.random_view_camera(0x5EED5702, 256, 192)
    print("case: stochastic running mean of 8 samples", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.cull(cam)
        for f in range(1, 9):
            c.stochastic_sample(f)
            c.accumulate(f)
        acc = c.read_image(abi.GS_IMAGE_ACCUM_RGBA32F)
    hist = None
    for f in range(1, 9):
        s, _ = _oracle_sample(orc, g, cam, f)
        hist = s.copy() if f == 1 else orc.accumulate(hist, s, f)
    np.testing.assert_array_equal(acc, hist)


def test_stochastic_tile_overflow_rerun(gs, orc):
    """Large splats: far more (tile, splat) pairs than the first capacity; the
    sample is redone with grown tile lists before it is read or averaged."""
    g = gs.generate_scene(13, 3000, 4, 0.0)
    g[:, 3:6] *= 40.0
    cam = scenes.random_view_camera(13, 640, 480)
    print("case: stochastic tile-pair overflow and re-run", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.cull(cam)
        c.stochastic_sample(5)
        c.accumulate(1)
        acc = c.read_image(abi.GS_IMAGE_ACCUM_RGBA32F)
        col = c.read_image(abi.GS_IMAGE_GPS_SAMPLE_RGBA32F)
    rc, _ = _oracle_sample(orc, g, cam, 5)
    np.testing.assert_array_equal(col, rc)
    np.testing.assert_array_equal(acc, rc)


def test_stochastic_needs_cull(gs):
    print("case: stochastic sample errors", flush=True)
    g = gs.generate_scene(1, 1000, 2, 0.0)
    cam = scenes.random_view_camera(1, 64, 64)
    with gs.Context(0) as c:
        with pytest.raises(gs.GsError):
            c.stochastic_sample(1)
        c.upload(g)
        with pytest.raises(gs.GsError):
            c.stochastic_sample(1)
        c.render_views([cam, cam])
        with pytest.raises(gs.GsError):
            c.stochastic_sample(1)  # needs a single-view cull
        c.cull(cam)
        c.stochastic_sample(1)
