"""The HIP library against the committed golden fixtures (tests/golden/*.npz):
visible set, keys, sorted order and GPS cells bit-exact, the composite within
1/1024 of the fixture's FP64 image."""
from pathlib import Path

import numpy as np
import pytest

from gamesplat_amd import abi

pytestmark = pytest.mark.gpu

GOLDEN = sorted((Path(__file__).resolve().parent / "golden").glob("*.npz"))


@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_library_reproduces_fixture(gs, path):
    f = np.load(path)
    cam = abi.GsCameraUniforms.from_array(f["camera"])
    W, H = int(cam.viewport[0]), int(cam.viewport[1])
    print(f"case: golden {path.stem}", flush=True)
    with gs.Context(0) as c:
        c.upload(f["gaussians"])
        c.cull(cam)
        np.testing.assert_array_equal(c.read_buffer(abi.GS_BUFFER_VISIBLE_IDS), f["visible_ids"])
        c.sort()
        n = len(f["gaussians"])
        np.testing.assert_array_equal(c.read_buffer(abi.GS_BUFFER_SORTED_IDS), f["sorted_ids"][:n])
        np.testing.assert_array_equal(c.read_buffer(abi.GS_BUFFER_SORTED_KEYS), f["sorted_keys"][:n])
        c.composite(16)
        img = c.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).astype(np.float64)
        assert np.abs(img - f["composite_fp64"]).max() <= 1.0 / 1024.0
        c.set_option(abi.GS_OPT_GPS_KEEP_CELLS, 1)
        c.set_gps_budget(1 << 20)
        c.gps_sample(frame_seed=1, supersample=1)
        cells = c.read_buffer(abi.GS_BUFFER_GPS_PACKED, shape=(H * W,))
        np.testing.assert_array_equal(cells, f["gps_cells"])
        assert c.stats()["gps_requested_work_items"] == int(f["gps_requested"])
        np.testing.assert_array_equal(c.read_image(abi.GS_IMAGE_GPS_SAMPLE_RGBA32F), f["gps_sample"])
        np.testing.assert_array_equal(c.read_image(abi.GS_IMAGE_GPS_DEPTH_R32F), f["gps_depth"])
