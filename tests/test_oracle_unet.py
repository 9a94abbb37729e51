"""Pins of the denoiser oracle (oracle/unet.py), CPU only.

Each check is independent of the oracle's own formulas: brute-force loops,
adjoint identities, central finite differences of the exact float64 network,
and torch's bf16 rounding and Adam.
"""
import numpy as np
import pytest

from oracle import unet


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 20000),
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 0.0, -0.0],
                                 np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(unet.bf16(x), want)
    # ties go to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    assert unet.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0


def test_conv3x3_brute_force():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 5, 4, 3))
    w = rng.standard_normal((2, 9, 3))
    b = rng.standard_normal(2)
    got = unet.conv3x3(x, w, b)
    want = np.zeros((2, 5, 4, 2))
    for n in range(2):
        for i in range(5):
            for j in range(4):
                for o in range(2):
                    s = b[o]
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            ii, jj = i + dy, j + dx
                            if 0 <= ii < 5 and 0 <= jj < 4:
                                s += np.dot(w[o, (dy + 1) * 3 + dx + 1], x[n, ii, jj])
                    want[n, i, j, o] = s
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("op,adj,shape", [
    (unet._im2col, unet._col2im, (2, 6, 4, 3)),
    (unet.pool2, unet.pool2_adj, (2, 6, 4, 3)),
    (unet.up2, unet.up2_adj, (2, 3, 2, 3)),
])
def test_adjoint_pairs(op, adj, shape):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape)
    y = rng.standard_normal(op(x).shape)
    assert abs(np.vdot(op(x), y) - np.vdot(x, adj(y))) < 1e-10 * (1 + abs(np.vdot(op(x), y)))


def test_param_layout():
    lay, n = unet.param_layout()
    assert n == sum(int(np.prod(s)) for _, _, s in lay)
    assert lay[0] == ("e0a.w", 0, (32, 9, 8))
    # concat widths are upsampled + skip channels
    cin = {name: c for name, c, _ in unet.CONVS}
    assert cin["d2a"] == 96 + 64 and cin["d1a"] == 64 + 48 and cin["d0a"] == 48 + 32


def test_gradient_finite_differences():
    """Exact float64 network: analytic gradient == central differences."""
    rng = np.random.default_rng(4)
    flat = unet.init_params(5).astype(np.float64)
    flat += rng.standard_normal(flat.size) * 0.01  # non-zero biases, break symmetry
    inp = np.zeros((1, 8, 8, 8))
    inp[..., :5] = rng.uniform(0.0, 1.0, (1, 8, 8, 5))
    tgt = rng.uniform(0.0, 1.0, (1, 8, 8, 4))
    loss, out, grad = unet.forward_backward(flat, inp, tgt, bf16_points=False)
    lay, _ = unet.param_layout()

    def f(p):
        return unet.forward_backward(p, inp, tgt, bf16_points=False, want_grad=False)[0]

    # float32 parameter storage: perturb in float64 through a float64 copy
    picks = []
    for name, off, shape in lay:
        size = int(np.prod(shape))
        picks += list(off + rng.choice(size, size=min(2, size), replace=False))
    for k in picks:
        g = grad[k]
        h = 1e-4
        p1 = flat.copy(); p1[k] += h
        p2 = flat.copy(); p2[k] -= h
        fd = (f(p1) - f(p2)) / (2 * h)
        assert abs(fd - g) <= 1e-5 * max(1.0, abs(g)) + 1e-6, (k, g, fd)


def test_forward_backward_shapes_and_residual():
    rng = np.random.default_rng(6)
    flat = unet.init_params(7)
    lay, n = unet.param_layout()
    flat[[off for name, off, _ in lay if name == "head.w"][0]:] = 0.0  # zero head: output = input
    inp = np.zeros((2, 16, 8, 8))
    inp[..., :5] = rng.uniform(0.0, 1.0, (2, 16, 8, 5))
    loss, out, grad = unet.forward_backward(flat, inp, unet.bf16(inp)[..., :4])
    np.testing.assert_array_equal(out, unet.bf16(inp)[..., :4])
    assert loss == 0.0 and not grad.any()


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(8)
    p = rng.standard_normal(1000).astype(np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    tp = torch.nn.Parameter(torch.from_numpy(p.copy()))
    opt = torch.optim.Adam([tp], lr=3e-4, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 4):
        g = rng.standard_normal(1000).astype(np.float32)
        p, m, v = unet.adam(p, g, m, v, step, 3e-4)
        tp.grad = torch.from_numpy(g.copy())
        opt.step()
        np.testing.assert_allclose(p, tp.detach().numpy(), rtol=0, atol=2e-6)
