"""Pins the oracle's scalar maths: the deterministic log/exp/sin/cos against
libm (numpy float64), the Li2 fits against scipy's dilogarithm, the integer
hashes by inverting them, and float16 rounding against numpy."""
import numpy as np
import pytest
from scipy.special import spence

M32 = 0xFFFFFFFF


def _vec(orc, name, xs):
    f = getattr(orc.load(), name)
    return np.array([f(float(x)) for x in xs])


def test_log_exp_sin_cos_double(orc):
    rng = np.random.default_rng(1)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 4000), 1 + rng.uniform(-1e-3, 1e-3, 1000),
                        rng.uniform(0.5, 2.0, 2000)])
    got = _vec(orc, "or_log_d", x)
    ref = np.log(x)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    assert err.max() < 4e-15
    e = rng.uniform(-700, 700, 4000)
    got = _vec(orc, "or_exp_d", e)
    assert (np.abs(got - np.exp(e)) / np.exp(e)).max() < 4e-15
    s = np.concatenate([rng.uniform(0, 2 * np.pi, 3000), rng.uniform(-1e4, 1e4, 3000)])
    assert np.abs(_vec(orc, "or_sin_d", s) - np.sin(s)).max() < 2e-15
    assert np.abs(_vec(orc, "or_cos_d", s) - np.cos(s)).max() < 2e-15


def test_float_wrappers_round_libm(orc):
    rng = np.random.default_rng(2)
    x = rng.uniform(1e-7, 1.0, 20000).astype(np.float32)
    for name, fn in (("or_logf", np.log), ("or_sinf", np.sin), ("or_cosf", np.cos)):
        got = _vec(orc, name, x).astype(np.float32)
        ref = fn(x.astype(np.float64)).astype(np.float32)
        ulps = np.abs(got.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
        assert ulps.max() <= 1 and (ulps == 0).mean() > 0.9999, name
    e = rng.uniform(-80, 80, 20000).astype(np.float32)
    got = _vec(orc, "or_expf", e).astype(np.float32)
    ref = np.exp(e.astype(np.float64)).astype(np.float32)
    assert np.abs(got.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64)).max() <= 1
    assert orc.scalar("or_logf", 1.0) == 0.0
    assert orc.scalar("or_logf", 0.0) == -np.inf


def test_dilog_fit_matches_li2(orc):
    # Li2(x) = spence(1 - x); the reference's polynomial fit plus (1-x) log(1-x)
    x = np.linspace(0.0, 1.0, 2001)
    got = _vec(orc, "or_dilog", x)
    ref = spence(1.0 - x)
    assert np.abs(got - ref).max() < 2e-4
    assert abs(orc.scalar("or_dilog", 1.0) - np.pi ** 2 / 6) < 2e-4
    assert abs(orc.scalar("or_dilog", 0.0)) < 1e-4


def test_inv_dilog_inverts_li2(orc):
    t = np.linspace(0.0, 1.0, 2001)
    y = spence(1.0 - t)  # Li2(t) in [0, pi^2/6]
    got = _vec(orc, "or_inv_dilog", y)
    assert np.abs(got - t).max() < 3e-3


def _inv_xorshift(x, s):
    y = x
    for _ in range(32 // s + 1):
        y = x ^ (y >> s)
    return y & M32


def test_hash01_inverts(orc):
    rng = np.random.default_rng(3)
    inv1 = pow(0x846CA68B, -1, 1 << 32)
    inv0 = pow(0x7FEB352D, -1, 1 << 32)
    for x in list(rng.integers(0, 1 << 32, 2000, dtype=np.uint64)) + [0, 1, M32, 0x9E3779B9]:
        x = int(x)
        h = orc.scalar("or_hash01_bits", x)
        y = _inv_xorshift(h, 16)
        y = (y * inv1) & M32
        y = _inv_xorshift(y, 15)
        y = (y * inv0) & M32
        y = _inv_xorshift(y, 16)
        assert y == x
        assert orc.scalar("or_hash01", x) == np.float32(np.float32(h) * np.float32(2.0 ** -32))


def test_pcg2d_inverts(orc):
    rng = np.random.default_rng(4)
    a = 1664525
    inv_a = pow(a, -1, 1 << 32)
    for _ in range(2000):
        x0, y0 = (int(v) for v in rng.integers(0, 1 << 32, 2, dtype=np.uint64))
        x, y = orc.pcg2d(x0, y0)
        # undo: xorshift16 is an involution; additions are undone by subtraction
        x, y = x ^ (x >> 16), y ^ (y >> 16)
        y = (y - a * x) & M32
        x = (x - a * y) & M32
        x, y = x ^ (x >> 16), y ^ (y >> 16)
        y = (y - a * x) & M32
        x = (x - a * y) & M32
        x = ((x - 1013904223) * inv_a) & M32
        y = ((y - 1013904223) * inv_a) & M32
        assert (x, y) == (x0, y0)


def test_f16_conversion_matches_numpy(orc):
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 1 << 32, 200000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = vals[np.isfinite(vals)]
    # dense sample of the half range incl. subnormals and exact ties
    halves = np.arange(0, 1 << 16, dtype=np.uint16).view(np.float16)
    h32 = halves[np.isfinite(halves)].astype(np.float32)
    ties = ((h32[:-1] + h32[1:]) / 2).astype(np.float32)
    xs = np.concatenate([vals, h32, ties, rng.uniform(-70000, 70000, 50000).astype(np.float32)])
    f = orc.load().or_f32_to_f16
    got = np.array([f(float(v)) for v in xs], dtype=np.uint16)
    ref = xs.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(got, ref)
    g = orc.load().or_f16_to_f32
    allh = np.arange(0, 1 << 16, dtype=np.uint32)
    back = np.array([g(int(h)) for h in allh], dtype=np.float32)
    ref = allh.astype(np.uint16).view(np.float16).astype(np.float32)
    fin = np.isfinite(ref)
    np.testing.assert_array_equal(back[fin].view(np.uint32), ref[fin].view(np.uint32))


def _ulps(a, b):
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def test_logf_fma_against_libm(orc):
    """The binary32 FMA log (GPS sampler, StochasticSplats cutoff): every 61st float from the smallest
    normal to 1e30 and every 7th of its own domain [1e-7, 1] within 1 ulp of
    libm's log rounded to float, nearly all of them exact."""
    lo, hi = np.float32(1.17549435e-38).view(np.uint32), np.float32(1e30).view(np.uint32)
    bits = np.concatenate([np.arange(lo, hi, 61, dtype=np.uint32),
                           np.arange(np.float32(1e-7).view(np.uint32), np.float32(1.0).view(np.uint32) + 1, 7,
                                     dtype=np.uint32)])
    x = bits.view(np.float32)
    got = orc.logf_fma(x)
    ref = np.log(x.astype(np.float64)).astype(np.float32)
    u = _ulps(got, ref)
    assert u.max() <= 1 and (u == 0).mean() > 0.98
    k = np.arange(-126, 100)
    np.testing.assert_array_equal(orc.logf_fma(np.ldexp(np.float32(1), k).astype(np.float32)),
                                  (k * np.log(2.0)).astype(np.float32))
    assert orc.scalar("or_logf_fma", 1.0) == 0.0


def test_gps_sincosf_against_libm(orc):
    """The GPS sampler's binary32 sin/cos over every 37th float of [0, 2 pi)
    and the angles the sampler forms (2 pi times a 32-bit fraction): within
    1 ulp (sin) and 2 ulp (cos) of libm rounded to float and 1e-7 absolute,
    and the quadrant values exact."""
    two_pi = np.float32(6.28318530718)
    bits = np.arange(0, two_pi.view(np.uint32), 37, dtype=np.uint32)
    rng = np.random.default_rng(5)
    r1 = rng.integers(0, 2 ** 32, 2_000_000, dtype=np.uint64).astype(np.float32) * np.float32(2.32830643654e-10)
    x = np.concatenate([bits.view(np.float32), (two_pi * r1).astype(np.float32)])
    s, c = orc.gps_sincosf(x)
    xd = x.astype(np.float64)
    assert _ulps(s, np.sin(xd).astype(np.float32)).max() <= 1
    assert _ulps(c, np.cos(xd).astype(np.float32)).max() <= 2
    assert np.abs(s - np.sin(xd)).max() < 1e-7 and np.abs(c - np.cos(xd)).max() < 1e-7
    s0, c0 = orc.gps_sincosf(np.zeros(1, np.float32))
    assert s0[0] == 0.0 and c0[0] == 1.0
