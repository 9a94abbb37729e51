"""Pins the oracle's radix sort restatement (shaders/sort.glsl) against
numpy's stable sort and brute force."""
import itertools

import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 1000, 70_000])
def test_matches_stable_argsort(orc, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    keys[rng.random(n) < 0.2] = 0xFFFFFFFF  # culled sentinels
    keys[: n // 3] = keys[0]  # many ties
    ids = np.where(keys == 0xFFFFFFFF, 0xFFFFFFFF, np.arange(n)).astype(np.uint32)
    sk, si = orc.radix_sort(keys, ids)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(sk, keys[order])
    np.testing.assert_array_equal(si, ids[order])


def test_brute_force_tiny(orc):
    vals = [0, 1, 0x100, 0xFFFF, 0x01000000, 0xFFFFFFFF]
    for combo in itertools.product(vals, repeat=4):
        keys = np.array(combo, np.uint32)
        ids = np.arange(4, dtype=np.uint32)
        sk, si = orc.radix_sort(keys, ids)
        exp = sorted(range(4), key=lambda i: (combo[i], i))
        assert list(si) == exp


def test_depth_keys_sort_far_to_near(orc):
    depth = np.array([5.0, 0.5, 50.0, 5.0, 1.0], np.float32)
    keys = (np.uint32(0xFFFFFFFF) - depth.view(np.uint32)).astype(np.uint32)
    sk, si = orc.radix_sort(keys, np.arange(5, dtype=np.uint32))
    assert list(si) == [2, 0, 3, 4, 1]  # far first, equal depths by index
