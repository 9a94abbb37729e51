"""Pins the oracle's GPS restatement (shaders/gps.glsl) and running mean
(shaders/accum.glsl): expected work count 2 pi sqrt(det) Li2(alpha) (scipy),
per-pixel coverage probability alpha * G (the Poisson property the point
sampler is built on), budget bookkeeping, resolve and accumulation closed
forms."""
import numpy as np
import pytest
from scipy.special import spence

EMPTY = 0xFFFFFFFF00000000


def _rec(cx, cy, a, b, c, opacity, color=(0.5, 0.5, 0.5), ndc_z=0.5):
    r = np.zeros(16, np.float32)
    r[0:3] = color
    r[3] = opacity
    r[4:6] = (cx, cy)
    r[11] = ndc_z
    r[12:15] = (a, b, c)
    return r


def test_expected_work_count(orc):
    # second dilation: cov + 0.3; ss^2 cancels; count = floor(E) + Bernoulli(frac E)
    a, b, c, o = 40.0, 5.0, 25.0, 0.7
    n = 20000
    proj = np.tile(_rec(50, 50, a, b, c, o), (n, 1))
    ids = np.arange(n, dtype=np.uint32)
    for ss in (1.0, 2.0):
        prepared, work, requested, overflow = orc.gps_expand(proj, ids, 1 << 26, ss, 0.0)
        det = (a + 0.3) * (c + 0.3) - b * b
        E = 2 * np.pi * np.sqrt(det) * spence(1 - o)
        assert overflow == 0
        assert abs(requested / n - E) < 0.02 * E / np.sqrt(n) * 10 + 2e-4 * E
        assert abs(requested / n - E) / E < 3e-3


def _coverage_expectation(mean, L, alpha, size):
    """P(pixel (i, j) hit in a frame) = 1 - exp(-integral of lambda over its
    cell [i-1/2, i+1/2) x [j-1/2, j+1/2)), lambda = -ln(1 - alpha G)."""
    sub = 8
    off = (np.arange(sub) + 0.5) / sub - 0.5
    Li = np.linalg.inv(L)
    exp = np.zeros((size, size))
    for j in range(size):
        for i in range(size):
            xs = np.stack(np.meshgrid(i + off, j + off), -1).reshape(-1, 2) - mean
            u = xs @ Li.T
            lam = -np.log(1 - alpha * np.exp(-0.5 * (u * u).sum(1)))
            exp[j, i] = 1 - np.exp(-lam.mean())
    return exp


def test_coverage_probability_is_alpha_gaussian(orc):
    size = 40
    a, b, c, o = 30.0, 6.0, 18.0, 0.8
    proj = _rec(19.3, 20.6, a, b, c, o)[None]
    ids = np.array([0], np.uint32)
    prepared, work, requested, _ = orc.gps_expand(proj, ids, 1 << 20, 1.0, 0.0)
    frames = 600
    hits = np.zeros(size * size)
    for f in range(1, frames + 1):
        cells = orc.gps_splat(prepared, work, requested, size, size, f, 1.0)
        hits += cells != EMPTY
    emp = (hits / frames).reshape(size, size)
    ca, cc = a + 0.3, c + 0.3
    l00 = np.sqrt(ca)
    l10 = b / l00
    l11 = np.sqrt(cc - l10 * l10)
    exp = _coverage_expectation(np.array([19.3, 20.6]), np.array([[l00, 0], [l10, l11]]), o, size)
    sd = np.sqrt(np.maximum(exp * (1 - exp), 1e-4) / frames)
    z = np.abs(emp - exp) / sd
    assert z.max() < 6.0 and np.abs(emp - exp).mean() < 0.01
    assert exp.max() > 0.7  # centre coverage ~ alpha


def test_budget_bookkeeping(orc):
    n = 50
    rng = np.random.default_rng(10)
    proj = np.stack([_rec(10, 10, rng.uniform(1, 20), 0.0, rng.uniform(1, 20), rng.uniform(0.1, 0.9))
                     for _ in range(n)])
    ids = np.arange(n, dtype=np.uint32)
    ids[7] = 0xFFFFFFFF  # culled entries are skipped
    full, fwork, total, ov = orc.gps_expand(proj, ids, 1 << 20, 1.0, 0.0)
    assert ov == 0
    budget = total // 2
    prep, work, req, ov2 = orc.gps_expand(proj, ids, budget, 1.0, 0.0)
    assert req == total and ov2 == 1
    np.testing.assert_array_equal(work[:budget], fwork[:budget])  # reservation order = index order
    # each item record is (position, item index) with item indices 0..count-1 per position
    owners, items = fwork[:total, 0], fwork[:total, 1]
    for pos in np.unique(owners):
        it = items[owners == pos]
        assert list(it) == list(range(len(it)))
    assert 7 not in owners


def test_resolve_closed_form(orc):
    w, h, ss = 3, 2, 2
    cells = np.full(w * ss * h * ss, EMPTY, np.uint64)
    def put(x, y, key, rgb):
        r, g, b = rgb
        cells[y * w * ss + x] = (key << 32) | (r | (g << 8) | (b << 16) | 0xFF000000)
    put(0, 0, 100, (255, 0, 0))
    put(1, 1, 50, (0, 255, 0))
    put(5, 3, 7, (10, 20, 30))
    color, depth = orc.gps_resolve(cells, w, h, ss, 0.25)
    np.testing.assert_allclose(color[0, 0], [0.25, 0.25, 0.0, 0.5], atol=1e-7)
    assert depth[0, 0] == np.float32(50) / np.float32(4294967296.0)
    np.testing.assert_allclose(color[1, 2], [10 / 255 / 4, 20 / 255 / 4, 30 / 255 / 4, 0.25], atol=1e-7)
    np.testing.assert_array_equal(color[0, 1], [0.25, 0.25, 0.25, 0.0])
    assert depth[0, 1] == 1.0


def test_accumulate_is_running_mean(orc):
    rng = np.random.default_rng(11)
    samples = rng.uniform(0, 1, (20, 64)).astype(np.float32)
    h = samples[0]
    for n in range(1, 21):
        h = orc.accumulate(h, samples[n - 1], n)
    np.testing.assert_allclose(h, samples.mean(0), atol=2e-6)
