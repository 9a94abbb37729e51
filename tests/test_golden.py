"""The oracle against the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py): a change to the oracle that moves
any output fails here."""
from pathlib import Path

import numpy as np
import pytest

from gamesplat_amd import abi
from oracle import oracle as orc

GOLDEN = sorted((Path(__file__).resolve().parent / "golden").glob("*.npz"))


@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_oracle_reproduces_fixture(path):
    f = np.load(path)
    g = f["gaussians"]
    cam = abi.GsCameraUniforms.from_array(f["camera"])
    W, H = int(cam.viewport[0]), int(cam.viewport[1])
    keys, ids, proj, vis, _ = orc.cull(g, cam)
    np.testing.assert_array_equal(np.nonzero(ids != 0xFFFFFFFF)[0], f["visible_ids"])
    np.testing.assert_array_equal(keys, f["keys"])
    sk, si = orc.radix_sort(keys, ids)
    np.testing.assert_array_equal(sk, f["sorted_keys"])
    np.testing.assert_array_equal(si, f["sorted_ids"])
    np.testing.assert_array_equal(orc.composite(proj, si, W, H, mode=0), f["composite_fp64"])
    prepared, work, requested, _ = orc.gps_expand(proj, ids, 1 << 20, 1.0, 0.0)
    assert requested == int(f["gps_requested"])
    cells = orc.gps_splat(prepared, work, min(requested, 1 << 20), W, H, 1, 1.0)
    np.testing.assert_array_equal(cells, f["gps_cells"])


def test_fixtures_present():
    assert len(GOLDEN) >= 2
