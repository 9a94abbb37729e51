"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: integer results (visible set, depth keys, sorted order, GPS cells)
bit-exact; projected splats bit-exact (same IEEE operations, deterministic
log/exp/sin/cos); the RGBA16F composite within 1/1024 of the FP64 oracle.
"""
import numpy as np
import pytest

from gamesplat_amd import abi, scenes

pytestmark = pytest.mark.gpu

TOL_RGBA16F = 1.0 / 1024.0


def _cull_pair(ctx, orc, g, cam, **kw):
    ctx.upload(g)
    ctx.cull(cam, **kw)
    ok, oi, op, ov, od = orc.cull(g, cam, sh_degree=kw.get("sh_degree", 3),
                                  sh_lod_distance=kw.get("sh_lod_distance", 0.0),
                                  collect_stats=kw.get("collect_stats", False), effect=kw.get("effect"))
    return ok, oi, op, ov, od


def test_cull_bit_exact(ctx, orc, c1_scene):
    g, cfg, cam = c1_scene
    print(f"case: cull c1 n={len(g)}", flush=True)
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam, collect_stats=True)
    st = ctx.stats()
    assert st["visible_count"] == ov
    vis_ids = ctx.read_buffer(abi.GS_BUFFER_VISIBLE_IDS)
    vis_keys = ctx.read_buffer(abi.GS_BUFFER_VISIBLE_KEYS)
    exp_ids = np.nonzero(oi != 0xFFFFFFFF)[0].astype(np.uint32)
    np.testing.assert_array_equal(vis_ids, exp_ids)
    np.testing.assert_array_equal(vis_keys, ok[exp_ids])
    proj = ctx.read_buffer(abi.GS_BUFFER_PROJECTED)
    np.testing.assert_array_equal(proj.view(np.uint32), op.view(np.uint32))
    assert st["diagnostics_valid"] == 1
    np.testing.assert_array_equal(np.array(st["diagnostics"], np.uint32), od)


@pytest.mark.parametrize("deg", [0, 1, 2])
def test_cull_sh_degrees(ctx, orc, c1_scene, deg):
    g, cfg, cam = c1_scene
    print(f"case: cull sh_degree={deg}", flush=True)
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g[:20000], cam, sh_degree=deg)
    proj = ctx.read_buffer(abi.GS_BUFFER_PROJECTED)
    np.testing.assert_array_equal(proj.view(np.uint32), op.view(np.uint32))


def test_cull_lod_and_effect(ctx, orc, c1_scene):
    g, cfg, cam = c1_scene
    print("case: cull sh_lod + shockwave effect", flush=True)
    fx = abi.GsSplatEffectParams((0.5, -0.2, 0.3, 9.0), (0.6, 1.5, 0.12, 1.0), (1.0, 0.55, 0.18, 0.4))
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g[:30000], cam, sh_lod_distance=15.0, effect=fx)
    assert ctx.stats()["visible_count"] == ov
    proj = ctx.read_buffer(abi.GS_BUFFER_PROJECTED)
    np.testing.assert_array_equal(proj.view(np.uint32), op.view(np.uint32))


def test_sort_bit_exact(ctx, orc, c1_scene):
    g, cfg, cam = c1_scene
    print("case: sort c1", flush=True)
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam)
    ctx.sort()
    sk, si = orc.radix_sort(ok, oi)
    np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_SORTED_KEYS), sk)
    np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_SORTED_IDS), si)


def test_sort_ties_and_sizes(ctx, orc, gs):
    # many Gaussians at identical depth: ties must resolve by index
    print("case: sort ties", flush=True)
    for n in (1, 255, 3840, 3841, 50_000):
        g = gs.generate_scene(7, n, 4, 0.0)
        g[:, 0:3] = np.round(g[:, 0:3])  # coarse positions -> many equal depths
        cam = scenes.random_view_camera(11, 640, 480)
        ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam)
        ctx.sort()
        sk, si = orc.radix_sort(ok, oi)
        np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_SORTED_IDS), si, err_msg=f"n={n}")


@pytest.mark.parametrize("spread,passes", [(0.0, 1), (2e-4, 1), (1e-3, 2), (0.3, 3), (None, 3), (1e10, 4)])
def test_sort_key_range_passes(ctx, orc, gs, spread, passes):
    # The depth sort runs only the 9-bit passes that the visible keys' range
    # (max - min) needs, and launches only those the box of the scene's
    # centres allows.  Splats on the view axis at depths 15 + U[0, spread)
    # give a range of each pass count; 1e10 puts one splat far out (range
    # above 2^27: the box then allows, and the keys need, all four passes).
    print(f"case: sort key range spread={spread}", flush=True)
    cam = scenes.random_view_camera(21, 640, 480)
    d = scenes._direction(21).astype(np.float64)
    g = gs.generate_scene(9, 20000, 4, 0.0)
    if spread == 1e10:
        g[0, 0:3] = d * (15.0 - 1e10)
        g[1, 0:3] = d * (15.0 - 0.25)
    elif spread is not None:
        t = np.random.Generator(np.random.PCG64(5)).random(len(g)) * spread
        g[:, 0:3] = (d[None, :] * -t[:, None]).astype(np.float32)
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam)
    vis = ok[oi != 0xFFFFFFFF].astype(np.int64)
    assert vis.size > 1
    span = int(vis.max() - vis.min())
    need = max(1, -(-span.bit_length() // 9))
    assert need == passes, f"key range {span:#x} needs {need} passes"
    ctx.sort()
    sk, si = orc.radix_sort(ok, oi)
    np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_SORTED_KEYS), sk)
    np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_SORTED_IDS), si)


def test_all_culled(ctx, orc, gs):
    print("case: all culled", flush=True)
    g = gs.generate_scene(3, 1000, 2, 0.0)
    # the cull has no far plane: put every splat behind the camera instead
    g[:, 0:3] = g[:, 0:3] * 0.01 + scenes._direction(5).astype(np.float32) * 45.0
    cam = scenes.random_view_camera(5, 320, 240)
    ctx.upload(g)
    ctx.cull(cam)
    assert ctx.stats()["visible_count"] == 0
    ctx.sort()
    assert (ctx.read_buffer(abi.GS_BUFFER_SORTED_IDS) == 0xFFFFFFFF).all()
    ctx.composite()
    img = ctx.read_image(abi.GS_IMAGE_ALPHA_RGBA16F)
    assert not img.any()


@pytest.mark.parametrize("tile,quad", [(8, 1), (16, 1), (16, 2), (16, 0), (32, 1)])
def test_composite_vs_fp64(ctx, orc, c1_scene, tile, quad):
    g, cfg, cam = c1_scene
    print(f"case: composite c1 tile={tile} quadrant_cull={quad}", flush=True)
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam)
    ctx.sort()
    ctx.set_option(abi.GS_OPT_QUADRANT_CULL, quad)
    try:
        ctx.composite(tile)
    finally:
        ctx.set_option(abi.GS_OPT_QUADRANT_CULL, 1)
    img = ctx.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).astype(np.float64)
    sk, si = orc.radix_sort(ok, oi)
    ref = orc.composite(op, si, cfg["width"], cfg["height"], mode=0)
    err = np.abs(img - ref)
    print(f"max abs err vs fp64 {err.max():.3e}, mean {err.mean():.3e}")
    assert err.max() <= TOL_RGBA16F
    st = ctx.stats()
    assert st["tile_entries"] > 0


@pytest.mark.parametrize("path,tile", [("calls", 16), ("frame", 16), ("frame", 32)])
def test_composite_wide_view(ctx, orc, gs, path, tile):
    """A 4352-pixel-wide view is 272 tiles of 16: more than a packed 8-bit
    tile coordinate holds, so the bin gathers pixel rectangles; at tile 32
    (136 tiles) the depth sort carries packed tile rectangles.  Splats on a
    shell around the camera cover the whole width."""
    W, H = 4352, 256
    print(f"case: composite wide view {W}x{H} path={path} tile={tile}", flush=True)
    g = gs.generate_scene(13, 60000, 4, 0.0)
    pos = scenes._direction(13) * 15.0
    rng = np.random.Generator(np.random.PCG64(3))
    d = rng.standard_normal((len(g), 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    g[:, 0:3] = (pos[None, :] + d * rng.uniform(2.0, 10.0, (len(g), 1))).astype(np.float32)
    cam = scenes.random_view_camera(13, W, H)
    if path == "calls":
        ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam)
        ctx.sort()
        ctx.composite(tile)
    else:
        ctx.upload(g)
        ctx.render_alpha(cam, tile=tile)
        ok, oi, op, ov, od = orc.cull(g, cam, sh_degree=3, sh_lod_distance=0.0, collect_stats=False, effect=None)
    img = ctx.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).astype(np.float64)
    sk, si = orc.radix_sort(ok, oi)
    ref = orc.composite(op, si, W, H, mode=0)
    assert (ref[:, 16 * 255:, 3] > 0).mean() > 0.5  # the columns past tile 255 are covered
    err = np.abs(img - ref)
    print(f"max abs err vs fp64 {err.max():.3e}, mean {err.mean():.3e}")
    assert err.max() <= TOL_RGBA16F


def _needle_scene(gs, seed=21, n=20000):
    """Elongated, rotated splats with opacities down to just above 1/255:
    the cases where the opacity-aware rectangle cut has the least slack."""
    g = gs.generate_scene(seed, n, 8, 0.0)
    rng = np.random.default_rng(seed)
    g[:, 3] *= rng.uniform(1.0, 40.0, n).astype(np.float32)
    g[:, 4] *= rng.uniform(0.5, 4.0, n).astype(np.float32)
    g[:, 13] = rng.choice(np.array([0.0040, 0.0045, 0.01, 0.05, 0.3, 0.9, 0.999], np.float32), n)
    return g


@pytest.mark.parametrize("scene,tile", [("c1", 16), ("needles", 16), ("c1", 32), ("needles", 32)])
def test_quadrant_ellipse_cull_identical_image(ctx, gs, c1_scene, scene, tile):
    """Skipping (splat, quadrant) pairs whose alpha >= 1/255 ellipse misses
    the quadrant must not change a single pixel (16x16 tiles: four 8x8
    quadrants, 32x32 tiles: sixteen)."""
    g, cfg, cam = c1_scene
    if scene == "needles":
        g = _needle_scene(gs)
    print(f"case: quadrant ellipse cull scene={scene} tile={tile}", flush=True)
    ctx.upload(g)
    out = {}
    try:
        for quad in (2, 1):
            ctx.set_option(abi.GS_OPT_QUADRANT_CULL, quad)
            ctx.render_alpha(cam, tile=tile)
            out[quad] = ctx.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).view(np.uint16).copy()
    finally:
        ctx.set_option(abi.GS_OPT_QUADRANT_CULL, 1)
    np.testing.assert_array_equal(out[1], out[2])
    assert out[1].any()


@pytest.mark.parametrize("scene", ["c1", "needles"])
def test_tight_rects_identical_image(ctx, gs, c1_scene, scene):
    """Cutting each splat's pixel rectangle to the box of its alpha >= 1/255
    ellipse must not change a single pixel, only the (tile, splat) pairs."""
    g, cfg, cam = c1_scene
    if scene == "needles":
        g = _needle_scene(gs)
    ctx.upload(g)
    fewer = False
    try:
        for tile, quad in [(16, 1), (16, 2), (16, 0), (8, 1), (32, 1)]:
            print(f"case: tight rects scene={scene} tile={tile} quadrant_cull={quad}", flush=True)
            ctx.set_option(abi.GS_OPT_QUADRANT_CULL, quad)
            out = {}
            for tight in (0, 1):
                ctx.set_option(abi.GS_OPT_TIGHT_RECTS, tight)
                ctx.render_alpha(cam, tile=tile)
                out[tight] = (ctx.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).view(np.uint16).copy(),
                              ctx.stats()["tile_entries"])
            print(f"tile entries full {out[0][1]} tight {out[1][1]}", flush=True)
            assert out[1][1] <= out[0][1]
            fewer |= out[1][1] < out[0][1]
            np.testing.assert_array_equal(out[1][0], out[0][0])
            assert out[0][0].any()
    finally:
        ctx.set_option(abi.GS_OPT_QUADRANT_CULL, 1)
        ctx.set_option(abi.GS_OPT_TIGHT_RECTS, 1)
    assert fewer


def test_composite_tile_overflow_rerun(orc, gs):
    # large splats: far more (tile, splat) pairs than the initial capacity
    print("case: composite tile-pair overflow and re-run", flush=True)
    g = gs.generate_scene(13, 3000, 4, 0.0)
    g[:, 3:6] *= 40.0
    cam = scenes.random_view_camera(13, 640, 480)
    with gs.Context(0) as c:
        c.upload(g)
        c.render_alpha(cam, tile=8)
        img = c.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).astype(np.float64)
        st = c.stats()
        assert st["tile_entries"] > 3 * len(g)
    keys, ids, proj, vis, _ = orc.cull(g, cam)
    sk, si = orc.radix_sort(keys, ids)
    ref = orc.composite(proj, si, 640, 480, mode=0)
    assert np.abs(img - ref).max() <= TOL_RGBA16F


def test_gps_bit_exact(ctx, orc, c1_scene):
    g, cfg, cam = c1_scene
    W, H = cfg["width"], cfg["height"]
    print("case: gps c1 ss=1 keep cells", flush=True)
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam)
    ctx.set_option(abi.GS_OPT_GPS_KEEP_CELLS, 1)
    try:
        budget = 1 << 22
        ctx.set_gps_budget(budget)
        ctx.gps_sample(frame_seed=1, supersample=1)
        cells = ctx.read_buffer(abi.GS_BUFFER_GPS_PACKED, shape=(H * W,))
        prepared, work, requested, overflow = orc.gps_expand(op, oi, budget, 1.0, 0.0)
        ref = orc.gps_splat(prepared, work, min(requested, budget), W, H, 1, 1.0)
        st = ctx.stats()
        assert st["gps_requested_work_items"] == requested
        assert st["gps_work_budget_exhausted"] == overflow
        np.testing.assert_array_equal(cells, ref)
        color, depth = orc.gps_resolve(ref, W, H, 1, 0.0)
        sample = ctx.read_image(abi.GS_IMAGE_GPS_SAMPLE_RGBA32F)
        np.testing.assert_array_equal(sample, color[::-1])
        np.testing.assert_array_equal(ctx.read_image(abi.GS_IMAGE_GPS_DEPTH_R32F), depth[::-1])
    finally:
        ctx.set_option(abi.GS_OPT_GPS_KEEP_CELLS, 0)
        ctx.set_gps_budget(16 * 1024 * 1024)


def test_gps_budget_overflow(ctx, orc, c1_scene):
    g, cfg, cam = c1_scene
    W, H = cfg["width"], cfg["height"]
    print("case: gps tiny budget overflow", flush=True)
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam)
    budget = 1000
    ctx.set_gps_budget(budget)
    ctx.set_option(abi.GS_OPT_GPS_KEEP_CELLS, 1)
    try:
        ctx.gps_sample(frame_seed=3, supersample=2)
        prepared, work, requested, overflow = orc.gps_expand(op, oi, budget, 2.0, 0.0)
        st = ctx.stats()
        assert overflow == 1 and st["gps_work_budget_exhausted"] == 1
        assert st["gps_requested_work_items"] == requested
        assert st["gps_dropped_work_items"] == requested - budget
        cells = ctx.read_buffer(abi.GS_BUFFER_GPS_PACKED, shape=(H * 2 * W * 2,))
        ref = orc.gps_splat(prepared, work, budget, W * 2, H * 2, 3, 2.0)
        np.testing.assert_array_equal(cells, ref)
    finally:
        ctx.set_option(abi.GS_OPT_GPS_KEEP_CELLS, 0)
        ctx.set_gps_budget(16 * 1024 * 1024)


def test_gps_budget_sequence(ctx, orc, c1_scene):
    """Budgets that rise and fall between frames: each count block writes its
    prepared records in the count pass or, for kept splats only, in the
    reserve pass, as the frame before suggests (wrongly after each change of
    budget); the cells equal the oracle's every time."""
    g, cfg, cam = c1_scene
    W, H = cfg["width"], cfg["height"]
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam)
    _, _, requested, _ = orc.gps_expand(op, oi, 1 << 30, 1.0, 0.0)
    budgets = [requested // 4, requested // 4, requested // 2, requested + 5, requested // 3, 1 << 22, requested // 8]
    ctx.set_option(abi.GS_OPT_GPS_KEEP_CELLS, 1)
    try:
        for k, budget in enumerate(budgets):
            print(f"case: gps budget sequence step {k} budget {budget} of {requested}", flush=True)
            ctx.set_gps_budget(budget)
            ctx.gps_sample(frame_seed=k + 1, supersample=1)
            prepared, work, req, overflow = orc.gps_expand(op, oi, budget, 1.0, 0.0)
            cells = ctx.read_buffer(abi.GS_BUFFER_GPS_PACKED, shape=(H * W,))
            ref = orc.gps_splat(prepared, work, min(req, budget), W, H, k + 1, 1.0)
            assert ctx.stats()["gps_work_budget_exhausted"] == overflow
            np.testing.assert_array_equal(cells, ref, err_msg=f"step {k}")
    finally:
        ctx.set_option(abi.GS_OPT_GPS_KEEP_CELLS, 0)
        ctx.set_gps_budget(16 * 1024 * 1024)


def test_gps_then_sort_and_readbacks(ctx, orc, c1_scene):
    """One cull serves a GPS sample (which compacts the visible ids alone),
    the depth sort (which then compacts in its first pass), the visible-list
    readbacks (which compact keys and ids) and a second GPS sample."""
    g, cfg, cam = c1_scene
    W, H = cfg["width"], cfg["height"]
    print("case: gps, sort, visible readbacks and gps again on one cull", flush=True)
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam)
    ctx.set_option(abi.GS_OPT_GPS_KEEP_CELLS, 1)
    try:
        ctx.gps_sample(frame_seed=2, supersample=1)
        first = ctx.read_buffer(abi.GS_BUFFER_GPS_PACKED, shape=(H * W,)).copy()
        ctx.sort()
        sk, si = orc.radix_sort(ok, oi)
        np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_SORTED_IDS), si)
        np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_SORTED_KEYS), sk)
        vis = np.nonzero(oi != 0xFFFFFFFF)[0].astype(np.uint32)
        np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_VISIBLE_IDS), vis)
        np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_VISIBLE_KEYS), ok[vis])
        ctx.gps_sample(frame_seed=2, supersample=1)
        np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_GPS_PACKED, shape=(H * W,)), first)
        prepared, work, requested, _ = orc.gps_expand(op, oi, 16 << 20, 1.0, 0.0)
        np.testing.assert_array_equal(first, orc.gps_splat(prepared, work, min(requested, 16 << 20), W, H, 2, 1.0))
    finally:
        ctx.set_option(abi.GS_OPT_GPS_KEEP_CELLS, 0)


def test_gps_accumulate(ctx, orc, c1_scene):
    g, cfg, cam = c1_scene
    W, H = cfg["width"], cfg["height"]
    print("case: gps accumulate 8 frames", flush=True)
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam)
    prepared, work, requested, overflow = orc.gps_expand(op, oi, 16 << 20, 1.0, 0.0)
    hist = None
    for f in range(1, 9):
        ctx.gps_sample(frame_seed=f, supersample=1)
        ctx.accumulate(f)
        cells = orc.gps_splat(prepared, work, min(requested, 16 << 20), W, H, f, 1.0)
        color, _ = orc.gps_resolve(cells, W, H, 1, 0.0)
        hist = orc.accumulate(color if hist is None else hist, color, f)
    acc = ctx.read_image(abi.GS_IMAGE_ACCUM_RGBA32F)
    np.testing.assert_array_equal(acc, hist[::-1])


def test_wide_index_variants(ctx, orc, c1_scene):
    """The 64-bit-index variants of the radix sorts and the bin (taken by
    arrays of 2^30 or more elements) give the same sorted order and image as
    the 32-bit-offset ones; forced on at the C1 size."""
    g, cfg, cam = c1_scene
    print("case: wide index variants c1", flush=True)
    ctx.upload(g)
    out = {}
    try:
        for wide in (0, 1):
            ctx.set_option(abi.GS_OPT_WIDE_INDEX, wide)
            ctx.cull(cam)
            ctx.sort()
            ids = ctx.read_buffer(abi.GS_BUFFER_SORTED_IDS)
            ctx.composite(16)
            img = ctx.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).view(np.uint16).copy()
            ctx.render_alpha(cam, tile=8)
            img8 = ctx.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).view(np.uint16).copy()
            out[wide] = (ids, img, img8)
    finally:
        ctx.set_option(abi.GS_OPT_WIDE_INDEX, 0)
    ok, oi, op, ov, od = orc.cull(g, cam)
    sk, si = orc.radix_sort(ok, oi)
    np.testing.assert_array_equal(out[1][0], si)
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][2], out[0][2])
    assert out[1][1].any()


@pytest.mark.parametrize("tile,size", [(16, (1280, 720)), (8, (1280, 720)), (32, (1280, 720)), (8, (2048, 2048))])
def test_tile_key_widths(ctx, gs, c1_scene, tile, size):
    """16-bit tile keys (frames of at most 65536 tiles; 2048x2048 at 8-pixel
    tiles has exactly 65536, ids up to 0xFFFF) give the same image as
    32-bit ones."""
    g, cfg, _ = c1_scene
    W, H = size
    print(f"case: tile key widths tile={tile} size={W}x{H}", flush=True)
    cam = scenes.random_view_camera(cfg["seed"], W, H)
    ctx.upload(g)
    out = {}
    try:
        for k16 in (1, 0):
            ctx.set_option(abi.GS_OPT_TILE_KEY16, k16)
            ctx.render_alpha(cam, tile=tile)
            out[k16] = ctx.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).view(np.uint16).copy()
    finally:
        ctx.set_option(abi.GS_OPT_TILE_KEY16, 1)
    np.testing.assert_array_equal(out[1], out[0])
    assert out[1].any()


def test_visible_list_after_frame(ctx, orc, c1_scene):
    """A frame sorts from the cull's dense keys and reuses the shared block
    offsets for its bin; the visible list built afterwards on demand (for a
    readback or GPS) still comes from the cull's own counts."""
    g, cfg, cam = c1_scene
    print("case: visible list after a frame c1", flush=True)
    ctx.upload(g)
    ctx.render_alpha(cam)
    ok, oi, op, ov, od = orc.cull(g, cam)
    exp_ids = np.nonzero(oi != 0xFFFFFFFF)[0].astype(np.uint32)
    np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_VISIBLE_IDS), exp_ids)
    np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_VISIBLE_KEYS), ok[exp_ids])
    sk, si = orc.radix_sort(ok, oi)
    np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_SORTED_IDS), si)


def test_sort_far_splat_under_effect(ctx, orc, gs):
    """With the shockwave effect active the centres move, so the scene's box
    no longer bounds the depths and all depth passes are launched; a splat
    far out (key range above 2^27) still sorts exactly."""
    print("case: sort far splat under the effect", flush=True)
    cam = scenes.random_view_camera(21, 640, 480)
    d = scenes._direction(21).astype(np.float64)
    g = gs.generate_scene(9, 20000, 4, 0.0)
    g[0, 0:3] = d * (15.0 - 1e10)
    fx = abi.GsSplatEffectParams((0.5, -0.2, 0.3, 9.0), (0.6, 1.5, 0.12, 1.0), (1.0, 0.55, 0.18, 0.4))
    ok, oi, op, ov, od = _cull_pair(ctx, orc, g, cam, effect=fx)
    vis = ok[oi != 0xFFFFFFFF].astype(np.int64)
    assert int(vis.max() - vis.min()).bit_length() > 27
    ctx.sort()
    sk, si = orc.radix_sort(ok, oi)
    np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_SORTED_KEYS), sk)
    np.testing.assert_array_equal(ctx.read_buffer(abi.GS_BUFFER_SORTED_IDS), si)
