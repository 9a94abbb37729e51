"""Batched multi-view frames, hipGraph replay and the autotuner, through the
C ABI.

Bar: every view of a batched frame is bit-identical to a single-view frame of
the same camera (the same kernels see the same tile lists in the same order),
for every tile size, digit width and composite launch shape; graph replays are
bit-identical to direct launches; the 4-bit-digit sort keeps the oracle's
order exactly.
"""
import numpy as np
import pytest

from gamesplat_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _cams(seed=7):
    sizes = [(256, 256), (200, 150), (640, 480), (333, 97), (1280, 720)]
    rng = np.random.default_rng(seed)
    out = []
    for w, h in sizes:
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        r = rng.uniform(10.0, 20.0)
        pose = scenes.look_at_origin_pose((d * r).astype(np.float32))
        out.append(scenes.camera_from_pose(pose, scenes.FOCAL_OVER_HEIGHT, w, h))
    return out


@pytest.fixture(scope="module")
def scene(gs):
    return gs.generate_scene(0x5EED0B01, 40000, 16, 0.0)


@pytest.fixture(scope="module")
def singles(gs, scene):
    """Single-view reference images per (tile, camera)."""
    cams = _cams()
    ref = {}
    with gs.Context(0) as c:
        c.upload(scene)
        for tile in (8, 16, 32):
            for k, cam in enumerate(cams):
                c.render_alpha(cam, tile=tile)
                ref[tile, k] = c.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).view(np.uint16).copy()
    return cams, ref


def _set(c, **opts):
    names = {"bits": abi.GS_OPT_RADIX_BITS, "bpc": abi.GS_OPT_BLOCKS_PER_CU, "graph": abi.GS_OPT_GRAPH,
             "quad": abi.GS_OPT_QUADRANT_CULL}
    for k, v in opts.items():
        c.set_option(names[k], v)


@pytest.mark.parametrize("tile,bits,bpc,graph", [(16, 8, 0, 0), (16, 8, 2, 0), (16, 4, 0, 1), (8, 8, 1, 0),
                                                 (32, 4, 4, 1), (8, 4, 0, 1)])
def test_views_equal_single_view(gs, scene, singles, tile, bits, bpc, graph):
    cams, ref = singles
    print(f"case: batched views tile={tile} bits={bits} blocks_per_cu={bpc} graph={graph}", flush=True)
    with gs.Context(0) as c:
        c.upload(scene)
        _set(c, bits=bits, bpc=bpc, graph=graph)
        for rep in range(2):
            c.render_views(cams, tile=tile)
            for k in range(len(cams)):
                img = c.read_view_image(k).view(np.uint16)
                np.testing.assert_array_equal(img, ref[tile, k], err_msg=f"view {k} rep {rep}")
        fi = c.frame_info()
        assert fi["views"] == len(cams) and fi["tile_size"] == tile
        if graph:
            assert fi["graph_captures"] >= 1 and fi["graph_replays"] >= 1


def test_views_reversed_order(gs, scene, singles):
    """A view's image does not depend on its slot in the batch."""
    cams, ref = singles
    print("case: batched views in reverse order", flush=True)
    with gs.Context(0) as c:
        c.upload(scene)
        c.render_views(cams[::-1], tile=16)
        for k in range(len(cams)):
            img = c.read_view_image(len(cams) - 1 - k).view(np.uint16)
            np.testing.assert_array_equal(img, ref[16, k])


def test_sixty_four_views(gs, scene, singles):
    """All 64 visibility bits, including view 63."""
    cams, ref = singles
    print("case: 64 views", flush=True)
    many = [cams[k % len(cams)] for k in range(64)]
    with gs.Context(0) as c:
        c.upload(scene)
        c.render_views(many, tile=16)
        for v in (0, 1, 17, 40, 62, 63):
            img = c.read_view_image(v).view(np.uint16)
            np.testing.assert_array_equal(img, ref[16, v % len(cams)], err_msg=f"view {v}")
        with pytest.raises(gs.GsError):
            c.render_views(many + [cams[0]], tile=16)


def test_graph_replays_match_eager(gs, scene, singles):
    """Moving cameras replay the captured frame; eager frames and readbacks in
    between; every replay equals the direct launch."""
    cams, ref = singles
    print("case: graph replay vs eager, moving camera", flush=True)
    orbit = [scenes.orbit_camera(0x5EED0001, f, 320, 240) for f in range(4)]
    with gs.Context(0) as eager, gs.Context(0) as g:
        eager.upload(scene)
        g.upload(scene)
        _set(g, graph=1)
        for rep in range(3):
            for f, cam in enumerate(orbit):
                eager.render_alpha(cam, tile=16)
                want = eager.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).view(np.uint16).copy()
                g.render_alpha(cam, tile=16)
                got = g.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).view(np.uint16)
                np.testing.assert_array_equal(got, want, err_msg=f"rep {rep} frame {f}")
                # eager work on the graph's context between replays
                g.cull(cam)
                g.sort()
                g.synchronize()
        fi = g.frame_info()
        print(f"graph captures {fi['graph_captures']} replays {fi['graph_replays']}", flush=True)
        assert fi["graph_replays"] >= 8
        assert fi["graph_captures"] <= 4


def test_radix4_sort_matches_oracle(gs, orc, c1_scene):
    g, cfg, cam = c1_scene
    print("case: 4-bit digit depth sort vs oracle", flush=True)
    with gs.Context(0) as c:
        c.upload(g)
        c.set_option(abi.GS_OPT_RADIX_BITS, 4)
        c.cull(cam)
        c.sort()
        ids = c.read_buffer(abi.GS_BUFFER_SORTED_IDS)
        keys = c.read_buffer(abi.GS_BUFFER_SORTED_KEYS)
    ok, oi, op, ov, od = orc.cull(g, cam)
    sk, si = orc.radix_sort(ok, oi)
    np.testing.assert_array_equal(keys, sk)
    np.testing.assert_array_equal(ids, si)


def test_autotune_keeps_images(gs, scene, singles):
    cams, ref = singles
    print("case: autotune over 3 views", flush=True)
    with gs.Context(0) as c:
        c.upload(scene)
        res = c.autotune(cams[:3], frames=1)
        print(f"autotune {res}", flush=True)
        assert res["configs"] == 30
        assert res["tile_size"] in (8, 16, 32) and res["radix_bits"] in (4, 8)
        assert res["blocks_per_cu"] in (0, 1, 2, 4, 8)
        assert 0.0 < res["best_ms"] <= res["default_ms"]
        c.render_views(cams[:3])  # tile 0 = the tuned default
        assert c.frame_info()["tile_size"] == res["tile_size"]
        for k in range(3):
            np.testing.assert_array_equal(c.read_view_image(k).view(np.uint16), ref[res["tile_size"], k])


def test_batch_errors(gs, scene):
    cam = _cams()[0]
    print("case: batch argument errors", flush=True)
    with gs.Context(0) as c:
        with pytest.raises(gs.GsError):
            c.render_views([cam])  # no scene
        c.upload(scene)
        with pytest.raises(gs.GsError):
            c.set_option(abi.GS_OPT_RADIX_BITS, 6)
        with pytest.raises(gs.GsError):
            c.set_option(abi.GS_OPT_BLOCKS_PER_CU, 65)
        c.render_views([cam, cam])
        with pytest.raises(gs.GsError):
            c.read_view_image(2)
        with pytest.raises(gs.GsError):
            c.gps_sample(1)  # GPS needs a single-view cull
        with pytest.raises(gs.GsError):
            c.read_buffer(abi.GS_BUFFER_SORTED_IDS)


def test_many_count_blocks(gs):
    """A batch whose bin has more than 32768 count blocks (3 views of a
    3,000,001-Gaussian scene: 35,157 blocks, not a multiple of the 8 counts a
    thread scans) takes the scan over blocks of 8192 counts, tail included;
    each view equals its single-view frame."""
    print("case: many count blocks (3 views x 3000001 Gaussians)", flush=True)
    g = gs.generate_scene(0x5EED0B02, 3_000_001, 64, 0.0)
    cams = _cams(11)[:3]
    with gs.Context(0) as c:
        c.upload(g)
        ref = []
        for cam in cams:
            c.render_alpha(cam, tile=16)
            ref.append(c.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).view(np.uint16).copy())
        c.render_views(cams, tile=16)
        for k in range(3):
            np.testing.assert_array_equal(c.read_view_image(k).view(np.uint16), ref[k])
        assert any(r.any() for r in ref)

