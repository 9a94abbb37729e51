"""CPU oracle of the learned denoiser (test infrastructure only).

A numpy restatement of the U-Net that libgamesplat_hip trains and runs on
MFMA, with the same parameter layout and the same bf16 rounding points, so
the GPU forward pass, loss and every weight gradient can be checked against
it.  Arithmetic is float64; each tensor the GPU stores in bf16 is rounded
here through float32 to bf16 (round to nearest even) at the same place.
With rounding off (`bf16=False`) the oracle is exact float64, which is what
the finite-difference and adjointness checks in tests/test_oracle_unet.py
pin.

Network (the learned part of the stochastic path; it has no counterpart in
the reference, whose stochastic output is shown as accumulated): input =
GPS running mean RGBA + GPS depth (5 channels, stored as 8), a 4-level U-Net
with 32/48/64/96 channels, two 3x3 convolutions + ReLU per level, 2x2
average pooling down, nearest 2x upsampling and skip concatenation
([upsampled, skip]) up, and a 1x1 head to RGBA added to the input mean.
Loss: mean squared error against the sorted-alpha RGBA.
"""
from __future__ import annotations

import numpy as np

IN_CH = 8  # 4 accum RGBA + depth + 3 zero channels
OUT_CH = 4

# (name, cin, cout): the 3x3 convolutions in order
CONVS = [
    ("e0a", 8, 32), ("e0b", 32, 32),
    ("e1a", 32, 48), ("e1b", 48, 48),
    ("e2a", 48, 64), ("e2b", 64, 64),
    ("ba", 64, 96), ("bb", 96, 96),
    ("d2a", 160, 64), ("d2b", 64, 64),
    ("d1a", 112, 48), ("d1b", 48, 48),
    ("d0a", 80, 32), ("d0b", 32, 32),
]
HEAD = ("head", 32, OUT_CH)


def param_layout():
    """[(name, offset, shape)] of the flat float32 parameter vector: per 3x3
    conv W[cout][tap][cin] (tap = (dy + 1) * 3 + dx + 1) then b[cout]; the
    head W[4][32], b[4]."""
    out, off = [], 0
    for name, cin, cout in CONVS:
        out.append((name + ".w", off, (cout, 9, cin)))
        off += cout * 9 * cin
        out.append((name + ".b", off, (cout,)))
        off += cout
    out.append(("head.w", off, (OUT_CH, 32)))
    off += OUT_CH * 32
    out.append(("head.b", off, (OUT_CH,)))
    off += OUT_CH
    return out, off


def param_count() -> int:
    return param_layout()[1]


def unpack(flat: np.ndarray) -> dict:
    lay, n = param_layout()
    assert flat.shape == (n,)
    return {name: flat[off:off + int(np.prod(shape))].reshape(shape) for name, off, shape in lay}


def init_params(seed: int) -> np.ndarray:
    """He-normal weights, zero biases; input channels 5-7 of e0a are zero."""
    rng = np.random.default_rng(seed)
    lay, n = param_layout()
    flat = np.zeros(n, np.float32)
    for name, off, shape in lay:
        if name.endswith(".w"):
            fan_in = int(np.prod(shape[1:]))
            w = rng.standard_normal(shape) * np.sqrt(2.0 / fan_in)
            if name == "e0a.w":
                w[:, :, 5:] = 0.0
            if name == "head.w":
                w *= 0.1
            flat[off:off + w.size] = w.astype(np.float32).ravel()
    return flat


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 through float32 (nearest, ties to even); float64 out."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


class _R:
    def __init__(self, on: bool):
        self.on = on

    def __call__(self, x):
        return bf16(x) if self.on else np.asarray(x, np.float64)


def _im2col(x):
    """(N, H, W, C) -> (N, H, W, 9, C), zero padding, tap = (dy+1)*3 + dx+1."""
    n, h, w, c = x.shape
    p = np.zeros((n, h + 2, w + 2, c), x.dtype)
    p[:, 1:-1, 1:-1] = x
    return np.stack([p[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] for dy in (-1, 0, 1) for dx in (-1, 0, 1)], axis=3)


def _col2im(cols):
    """Adjoint of _im2col: (N, H, W, 9, C) -> (N, H, W, C)."""
    n, h, w, _, c = cols.shape
    p = np.zeros((n, h + 2, w + 2, c), cols.dtype)
    k = 0
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            p[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] += cols[:, :, :, k]
            k += 1
    return p[:, 1:-1, 1:-1]


def conv3x3(x, w, b):
    """x (N,H,W,Cin), w (Cout,9,Cin), b (Cout) -> pre-activation (N,H,W,Cout)."""
    n, h, wd, cin = x.shape
    cols = _im2col(x).reshape(n * h * wd, 9 * cin)
    return (cols @ w.reshape(w.shape[0], -1).T + b).reshape(n, h, wd, -1)


def pool2(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool2_adj(dy):
    """Adjoint of pool2: each input receives 0.25 of its window's gradient."""
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) * 0.25


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_adj(dy):
    n, h, w, c = dy.shape
    return dy.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def forward_backward(flat: np.ndarray, inp: np.ndarray, target: np.ndarray, bf16_points: bool = True,
                     want_grad: bool = True):
    """inp (N,H,W,8), target (N,H,W,4); H, W multiples of 8.  Returns (loss,
    output (N,H,W,4), flat gradient or None)."""
    R = _R(bf16_points)
    P = unpack(np.asarray(flat, np.float32))
    Wc = {k: (R(v) if k.endswith(".w") and not k.startswith("head") else np.asarray(v, np.float64))
          for k, v in P.items()}
    x0 = R(inp)
    acts = {}

    def conv(name, x):
        z = conv3x3(x, Wc[name + ".w"], Wc[name + ".b"])
        y = R(np.maximum(z, 0.0))
        acts[name] = (x, y)
        return y

    s0 = conv("e0b", conv("e0a", x0))
    p1 = R(pool2(s0))
    s1 = conv("e1b", conv("e1a", p1))
    p2 = R(pool2(s1))
    s2 = conv("e2b", conv("e2a", p2))
    p3 = R(pool2(s2))
    u3 = conv("bb", conv("ba", p3))
    c2 = np.concatenate([up2(u3), s2], axis=3)
    u2 = conv("d2b", conv("d2a", c2))
    c1 = np.concatenate([up2(u2), s1], axis=3)
    u1 = conv("d1b", conv("d1a", c1))
    c0 = np.concatenate([up2(u1), s0], axis=3)
    u0 = conv("d0b", conv("d0a", c0))
    hw, hb = Wc["head.w"], Wc["head.b"]
    out = np.einsum("nhwc,oc->nhwo", u0, hw) + hb + x0[..., :OUT_CH]
    diff = out - np.asarray(target, np.float64)
    count = diff.size
    loss = float(np.mean(diff * diff))
    if not want_grad:
        return loss, out, None

    lay, n = param_layout()
    G = {name: np.zeros(shape) for name, _, shape in lay}
    dout = 2.0 * diff / count
    G["head.w"] = np.einsum("nhwo,nhwc->oc", dout, u0)
    G["head.b"] = dout.sum(axis=(0, 1, 2))
    du0 = R(np.einsum("nhwo,oc->nhwc", dout, hw))

    def conv_bwd(name, dy):
        """dy: gradient of the stored (post-ReLU) output; returns dx (bf16)."""
        x, y = acts[name]
        dz = dy * (y > 0.0)
        w = Wc[name + ".w"]
        nn, h, wd, cin = x.shape
        cols = _im2col(x).reshape(nn * h * wd, 9 * cin)
        dzm = dz.reshape(nn * h * wd, -1)
        G[name + ".w"] = (dzm.T @ cols).reshape(w.shape)
        G[name + ".b"] = dzm.sum(axis=0)
        dcols = (dzm @ w.reshape(w.shape[0], -1)).reshape(nn, h, wd, 9, cin)
        return R(_col2im(dcols))

    dc0 = conv_bwd("d0a", conv_bwd("d0b", du0))
    du1 = R(up2_adj(dc0[..., :48]))
    ds0 = dc0[..., 48:]
    dc1 = conv_bwd("d1a", conv_bwd("d1b", du1))
    du2 = R(up2_adj(dc1[..., :64]))
    ds1 = dc1[..., 64:]
    dc2 = conv_bwd("d2a", conv_bwd("d2b", du2))
    du3 = R(up2_adj(dc2[..., :96]))
    ds2 = dc2[..., 96:]
    dp3 = conv_bwd("ba", conv_bwd("bb", du3))
    ds2 = R(ds2 + pool2_adj(dp3))
    dp2 = conv_bwd("e2a", conv_bwd("e2b", ds2))
    ds1 = R(ds1 + pool2_adj(dp2))
    dp1 = conv_bwd("e1a", conv_bwd("e1b", ds1))
    ds0 = R(ds0 + pool2_adj(dp1))
    conv_bwd("e0a", conv_bwd("e0b", ds0))

    grad = np.zeros(n, np.float64)
    for name, off, shape in lay:
        grad[off:off + int(np.prod(shape))] = G[name].ravel()
    return loss, out, grad


def adam(flat, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (float32, the update the GPU applies)."""
    flat = np.asarray(flat, np.float32).copy()
    g = np.asarray(grad, np.float32)
    m = (np.float32(beta1) * m + np.float32(1 - beta1) * g).astype(np.float32)
    v = (np.float32(beta2) * v + np.float32(1 - beta2) * g * g).astype(np.float32)
    mh = m / np.float32(1 - beta1 ** step)
    vh = v / np.float32(1 - beta2 ** step)
    flat -= np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))
    return flat, m, v


def psnr(a, b, peak: float = 1.0) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0.0 else 10.0 * np.log10(peak * peak / mse)
