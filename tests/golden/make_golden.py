"""Write the golden fixtures of this directory from the CPU oracle.

python tests/golden/make_golden.py

Each fixture (npz, no pickled objects) holds a small seeded scene, a camera
and the oracle's outputs for it: visible ids and depth keys, the stable
sorted order, the FP64 composite (top row first), and the GPS cells, sample
and depth of frame seed 1 (ss = 1, budget 2^20).  tests/test_golden.py pins
the oracle to them on the CPU; tests/test_gpu_golden.py pins the HIP
library to them on the GPU.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

CASES = {
    # name: (seed, gaussians, clusters, opacity logit mean, width, height, camera radius)
    "small_60x44": (0x5EED6001, 600, 4, 0.0, 60, 44, 15.0),
    "near_64x64": (0x5EED6002, 900, 3, 1.0, 64, 64, 9.0),
}


def make(name, seed, count, clusters, om, width, height, radius):
    import gamesplat_amd as gs
    from gamesplat_amd import scenes
    from oracle import oracle as orc

    g = gs.generate_scene(seed, count, clusters, om)
    cam = scenes.random_view_camera(seed, width, height, radius=radius)
    keys, ids, proj, vis, _ = orc.cull(g, cam)
    sk, si = orc.radix_sort(keys, ids)
    comp = orc.composite(proj, si, width, height, mode=0)
    prepared, work, requested, overflow = orc.gps_expand(proj, ids, 1 << 20, 1.0, 0.0)
    cells = orc.gps_splat(prepared, work, min(requested, 1 << 20), width, height, 1, 1.0)
    color, depth = orc.gps_resolve(cells, width, height, 1, 0.0)
    np.savez_compressed(
        HERE / f"{name}.npz",
        gaussians=g.astype(np.float32), camera=cam.as_array(),
        visible_ids=np.nonzero(ids != 0xFFFFFFFF)[0].astype(np.uint32), keys=keys.astype(np.uint32),
        sorted_keys=sk.astype(np.uint32), sorted_ids=si.astype(np.uint32),
        composite_fp64=comp.astype(np.float64),
        gps_cells=cells.astype(np.uint64), gps_requested=np.uint32(requested),
        gps_sample=color[::-1].astype(np.float32), gps_depth=depth[::-1].astype(np.float32))
    print(f"{name}: {count} Gaussians, {vis} visible, {requested} GPS items")


if __name__ == "__main__":
    for name, args in CASES.items():
        make(name, *args)
