"""The learned denoiser on MFMA against the CPU oracle (oracle/unet.py), and
what training it on the renderer's own pairs gains.

Bars: the oracle mirrors every bf16 rounding point of the GPU, so the two
differ only by fp32-vs-fp64 accumulation order and the bf16 flips that
follows from it: output and loss to 1e-2 relative, each parameter tensor's
gradient to 3e-2 relative (L2) with cosine similarity >= 0.999; Adam to
1e-6 absolute.
"""
import numpy as np
import pytest

from gamesplat_amd import abi, scenes
from oracle import unet

pytestmark = pytest.mark.gpu


def _batch(seed, n, h, w):
    rng = np.random.default_rng(seed)
    inp = np.zeros((n, h, w, 8), np.float32)
    inp[..., :4] = rng.uniform(0.0, 1.0, (n, h, w, 4))
    inp[..., 4] = rng.uniform(0.3, 1.0, (n, h, w))
    tgt = np.clip(inp[..., :4] + rng.normal(0.0, 0.1, (n, h, w, 4)), 0.0, 1.0).astype(np.float32)
    return inp, tgt


def _params(seed):
    p = unet.init_params(seed)
    rng = np.random.default_rng(seed + 1)
    lay, _ = unet.param_layout()
    for name, off, shape in lay:
        if name.endswith(".b"):  # live ReLUs everywhere
            size = int(np.prod(shape))
            p[off:off + size] = rng.uniform(0.01, 0.05, size).astype(np.float32)
    return p


@pytest.fixture(scope="module")
def den(gs):
    ctx = gs.Context(0)
    d = gs.Denoiser(ctx, seed=3)
    yield ctx, d
    d.close()
    ctx.close()


def test_param_count(den):
    ctx, d = den
    assert d.P == unet.param_count()


def test_forward_and_gradients_match_oracle(den):
    ctx, d = den
    print("case: denoiser forward/backward n=2 32x32 vs oracle", flush=True)
    flat = _params(11)
    d.set_params(flat)
    assert np.array_equal(d.get_params(), flat)
    inp, tgt = _batch(12, 2, 32, 32)
    loss, out, grad = d.forward_backward(inp, tgt)
    rloss, rout, rgrad = unet.forward_backward(flat, inp, tgt)
    print(f"loss gpu {loss:.6e} oracle {rloss:.6e}", flush=True)
    assert abs(loss - rloss) <= 1e-2 * rloss
    oerr = np.linalg.norm(out - rout) / np.linalg.norm(rout)
    print(f"output rel L2 err {oerr:.3e}", flush=True)
    assert oerr < 1e-2
    lay, _ = unet.param_layout()
    worst = 0.0
    for name, off, shape in lay:
        size = int(np.prod(shape))
        g, r = grad[off:off + size].astype(np.float64), rgrad[off:off + size]
        nr = np.linalg.norm(r)
        if nr == 0.0:
            assert np.linalg.norm(g) == 0.0, name
            continue
        rel = np.linalg.norm(g - r) / nr
        cos = float(np.dot(g, r) / (np.linalg.norm(g) * nr))
        worst = max(worst, rel)
        assert rel < 3e-2 and cos > 0.999, (name, rel, cos)
    print(f"worst per-tensor gradient rel L2 err {worst:.3e}", flush=True)


def test_odd_sizes_and_batch(den):
    """Widths that are not multiples of the 32-pixel conv tile."""
    ctx, d = den
    print("case: denoiser forward n=3 24x40 vs oracle", flush=True)
    flat = _params(13)
    d.set_params(flat)
    inp, tgt = _batch(14, 3, 24, 40)
    loss, out, grad = d.forward_backward(inp, tgt)
    rloss, rout, rgrad = unet.forward_backward(flat, inp, tgt)
    assert np.linalg.norm(out - rout) / np.linalg.norm(rout) < 1e-2
    assert np.linalg.norm(grad - rgrad) / np.linalg.norm(rgrad) < 3e-2


def test_adam_step_matches_oracle(den):
    ctx, d = den
    print("case: denoiser Adam step", flush=True)
    flat = _params(15)
    d.set_params(flat)
    inp, tgt = _batch(16, 1, 16, 16)
    m = np.zeros_like(flat)
    v = np.zeros_like(flat)
    p = flat
    for step in range(1, 4):
        _, _, grad = d.forward_backward(inp, tgt, want_out=False)
        d.adam_step(1e-3)
        p, m, v = unet.adam(p, grad, m, v, step, 1e-3)
        np.testing.assert_allclose(d.get_params(), p, rtol=0, atol=1e-6)


def _frames(gs, ctx, d, scene, cams, spp, crops, seed):
    """Sorted-alpha target and an spp-sample GPS accumulation per camera."""
    out = []
    for k, cam in enumerate(cams):
        ctx.render_alpha(cam)
        alpha = ctx.read_image(abi.GS_IMAGE_ALPHA_RGBA16F).astype(np.float64)
        for f in range(1, spp + 1):
            ctx.gps_sample(frame_seed=seed * 1000 + k * 16 + f, supersample=1)
            ctx.accumulate(f)
        acc = ctx.read_image(abi.GS_IMAGE_ACCUM_RGBA32F).astype(np.float64)
        if crops:
            d.capture(crops, seed * 7919 + k)
        out.append((alpha, acc))
    return out


@pytest.mark.parametrize("width,height", [(64, 64), (61, 45)])
def test_denoise_matches_oracle_forward(gs, width, height):
    """Inference: the last convolution runs the 1x1 head and writes the
    width x height corner of the padded (multiple-of-8) frame."""
    print(f"case: gs_denoise {width}x{height} vs oracle forward", flush=True)
    g = gs.generate_scene(0x5EED0D01, 20000, 8, 0.0)
    cam = scenes.random_view_camera(0x5EED0D01, width, height)
    with gs.Context(0) as ctx, gs.Denoiser(ctx, seed=5) as d:
        ctx.upload(g)
        flat = _params(17)
        d.set_params(flat)
        ctx.cull(cam)
        ctx.gps_sample(frame_seed=3, supersample=1)
        ctx.accumulate(1)
        acc = ctx.read_image(abi.GS_IMAGE_ACCUM_RGBA32F)
        dep = ctx.read_image(abi.GS_IMAGE_GPS_DEPTH_R32F)
        d.denoise()
        img = d.read_image(width, height).astype(np.float64)
    # the network sees the device's row order (GL: bottom row first), zero
    # padded to multiples of 8; readback is top row first
    hp, wp = -(-height // 8) * 8, -(-width // 8) * 8
    inp = np.zeros((1, hp, wp, 8), np.float32)
    inp[0, :height, :width, :4] = acc[::-1]
    inp[0, :height, :width, 4] = dep[::-1]
    _, rout, _ = unet.forward_backward(flat, inp, np.zeros((1, hp, wp, 4)), want_grad=False)
    err = np.abs(img - rout[0][:height, :width][::-1]).max()
    print(f"denoise max abs err vs oracle {err:.3e}", flush=True)
    assert err < 2e-2


def test_training_gains_over_accumulation(gs):
    """Train on pairs the renderer makes (1 spp GPS mean + depth -> sorted
    alpha) and compare on a view it has not seen."""
    print("case: denoiser training gain, 1 spp", flush=True)
    g = gs.generate_scene(0x5EED0D02, 60000, 16, 0.0)
    rng = np.random.default_rng(9)
    cams = []
    for k in range(13):
        d_ = rng.standard_normal(3)
        d_ /= np.linalg.norm(d_)
        pose = scenes.look_at_origin_pose((d_ * rng.uniform(12.0, 18.0)).astype(np.float32))
        cams.append(scenes.camera_from_pose(pose, scenes.FOCAL_OVER_HEIGHT, 320, 288))
    with gs.Context(0) as ctx, gs.Denoiser(ctx, seed=7) as d:
        ctx.upload(g)
        _frames(gs, ctx, d, g, cams[:12], spp=1, crops=6, seed=1)
        assert d.pool_size() == 72
        losses = d.train(steps=300, batch=4, lr=1e-3, seed=5)
        print(f"loss first {losses[:10].mean():.4e} last {losses[-10:].mean():.4e}", flush=True)
        assert np.isfinite(losses).all()
        assert losses[-20:].mean() < 0.7 * losses[:20].mean()
        (alpha, acc), = _frames(gs, ctx, d, g, cams[12:], spp=1, crops=0, seed=2)
        d.denoise()
        den = d.read_image(320, 288).astype(np.float64)
    p_acc = unet.psnr(acc, alpha)
    p_den = unet.psnr(den, alpha)
    print(f"held-out view PSNR vs sorted alpha: 1 spp mean {p_acc:.2f} dB, denoised {p_den:.2f} dB", flush=True)
    assert p_den > p_acc + 1.0


def test_denoiser_errors(gs):
    print("case: denoiser argument errors", flush=True)
    with gs.Context(0) as ctx, gs.Denoiser(ctx, seed=1) as d:
        with pytest.raises(gs.GsError):
            d.denoise()  # no accumulation yet
        with pytest.raises(gs.GsError):
            d.capture(1, 0)  # no frames
        with pytest.raises(gs.GsError):
            d.train(1, 1, 1e-3, 0)  # empty pool
        with pytest.raises(gs.GsError):
            abi.check(d._lib.gs_denoiser_set_params(d._h, None, d.P))
        bad = np.zeros((1, 12, 16, 8), np.float32)
        with pytest.raises(gs.GsError):
            d.forward_backward(bad, np.zeros((1, 12, 16, 4), np.float32))  # h not a multiple of 8
