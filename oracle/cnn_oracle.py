"""CPU oracle of the learned warm start's network (test infrastructure only:
tests/ use it as the checker of the device kernels; the product never
imports it).

The network is this build's own (the reference has no learned component; its
contract is only that any initial guess is accepted and the DR iteration
decides the answer, stokes_2D_pure_shear_triangle.jl:159-163): a 4-level
encoder-decoder of 3x3 convolutions on NHWC images, FP64 here.

Conventions shared with the device code:
* images are [B][H][W][C]; a convolution weight is [Cout][3][3][Cin] with
  tap (ky, kx) reading pixel (y + ky - 1, x + kx - 1), zero outside the image;
* pool2 is the 2x2 average, up2 the 2x2 nearest-neighbour copy;
* the decoder adds the encoder's skip tensor to the upsampled one;
* every convolution but the head is followed by ReLU;
* the loss is the mean over the batch, the pixels and the three output
  channels of (y - t)^2.

`bf16` rounds to the nearest bfloat16 (ties to even) for the comparisons
that need the device's storage precision.

Pinned by tests/test_oracle_cnn.py: against torch's CPU conv2d/avg_pool2d/
interpolate (fixtures written by tests/golden/make_cnn_fixtures.py) and by
central finite differences of the loss.
"""
from __future__ import annotations

import numpy as np

# (name, Cin, Cout, resolution level) of the 14 convolutions, in forward order
LEVELS = (32, 64, 128, 128)
IN_CH = 16      # 4 features, zero-padded to 16 channels
OUT_CH = 32     # 3 outputs, zero-padded to 32 channels


def layer_table(in_ch=IN_CH, out_ch=OUT_CH, levels=LEVELS):
    c0, c1, c2, c3 = levels
    return [("e0a", in_ch, c0, 0), ("e0b", c0, c0, 0),
            ("e1a", c0, c1, 1), ("e1b", c1, c1, 1),
            ("e2a", c1, c2, 2), ("e2b", c2, c2, 2),
            ("e3a", c2, c3, 3), ("e3b", c3, c3, 3),
            ("d2a", c3, c2, 2), ("d2b", c2, c1, 2),
            ("d1a", c1, c1, 1), ("d1b", c1, c0, 1),
            ("d0a", c0, c0, 0), ("head", c0, out_ch, 0)]


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(a), np.nan, out)


def conv3x3(x, W, b):
    """y[b,y,x,k] = b[k] + sum_{ky,kx,c} W[k,ky,kx,c] x[b, y+ky-1, x+kx-1, c]."""
    B, H, Wd, C = x.shape
    xp = np.zeros((B, H + 2, Wd + 2, C))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((B, H, Wd, W.shape[0]))
    for ky in range(3):
        for kx in range(3):
            y += np.einsum("bhwc,kc->bhwk", xp[:, ky:ky + H, kx:kx + Wd], W[:, ky, kx, :])
    return y + b


def conv3x3_backward(x, W, gy):
    """Gradients of conv3x3 w.r.t. x, W and b for the output gradient gy."""
    B, H, Wd, C = x.shape
    xp = np.zeros((B, H + 2, Wd + 2, C))
    xp[:, 1:-1, 1:-1] = x
    gxp = np.zeros_like(xp)
    gW = np.zeros_like(W)
    for ky in range(3):
        for kx in range(3):
            gW[:, ky, kx, :] = np.einsum("bhwk,bhwc->kc", gy, xp[:, ky:ky + H, kx:kx + Wd])
            gxp[:, ky:ky + H, kx:kx + Wd] += np.einsum("bhwk,kc->bhwc", gy, W[:, ky, kx, :])
    return gxp[:, 1:-1, 1:-1], gW, gy.sum(axis=(0, 1, 2))


def pool2(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool2_backward(gy):
    return np.repeat(np.repeat(gy, 2, axis=1), 2, axis=2) / 4.0


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_backward(gy):
    B, H, W, C = gy.shape
    return gy.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def init_params(seed=0, layers=None):
    """He-normal weights (fan-in 9 Cin), zero biases; the padded input channels
    4.. of the first layer and the padded outputs 3.. of the head are zero."""
    rng = np.random.default_rng(seed)
    layers = layers or layer_table()
    params = []
    for name, cin, cout, _ in layers:
        W = rng.standard_normal((cout, 3, 3, cin)) * np.sqrt(2.0 / (9 * cin))
        if name == "e0a":
            W[..., 4:] = 0
        if name == "head":
            W *= 0.1
            W[3:] = 0
        params.append((W, np.zeros(cout)))
    return params


def forward(params, x, store=None, round_bf16=False, layers=None):
    """The network on x [B,H,W,IN_CH]; H, W multiples of 8. Returns the head
    output [B,H,W,OUT_CH]. store (a dict) receives every layer's input and
    output for the backward pass. round_bf16 rounds every stored activation
    (as the device stores them)."""
    layers = layers or layer_table()
    rd = bf16 if round_bf16 else (lambda a: a)
    st = store if store is not None else {}
    h = rd(x)
    skips = []
    li = 0

    def conv(h, relu):
        nonlocal li
        W, b = params[li]
        st[("in", li)] = h
        y = conv3x3(h, W, b)
        st[("pre", li)] = y
        if relu:
            y = np.maximum(y, 0.0)
        y = rd(y)
        st[("out", li)] = y
        li += 1
        return y

    for lev in range(4):
        h = conv(h, True)
        h = conv(h, True)
        if lev < 3:
            skips.append(h)
            h = rd(pool2(h))
    for lev in (2, 1, 0):
        h = rd(up2(h) + skips[lev])
        h = conv(h, True)
        h = conv(h, lev != 0 or li < len(layers) - 1)
    return h


def loss_and_grad(y, t, nout=3):
    """Mean-squared error over the batch, pixels and the first nout channels,
    and its gradient w.r.t. y (zero on the padded channels)."""
    d = y[..., :nout] - t
    n = d.size
    g = np.zeros_like(y)
    g[..., :nout] = 2.0 * d / n
    return float((d * d).sum() / n), g


def backward(params, store, gy, layers=None, round_bf16=False):
    """Gradients [(gW, gb)] of every layer for the output gradient gy, through
    the stored forward pass (FP64; ReLU masks from the stored outputs).
    round_bf16 rounds each layer's pre-activation gradient to bfloat16, as the
    device does before its matrix-core products."""
    layers = layers or layer_table()
    rd = bf16 if round_bf16 else (lambda a: a)
    L = len(layers)
    grads = [None] * L
    li = L - 1
    g = gy

    def conv_back(g, relu):
        nonlocal li
        W, _ = params[li]
        if relu:
            g = g * (store[("out", li)] > 0)
        g = rd(g)
        gx, gW, gb = conv3x3_backward(store[("in", li)], W, g)
        grads[li] = (gW, gb)
        li -= 1
        return gx

    gskip = {}
    for lev in (0, 1, 2):
        g = conv_back(g, lev != 0 or li < L - 1)
        g = conv_back(g, True)
        gskip[lev] = g          # d(loss)/d(skip_lev): the skip is added as is
        g = up2_backward(g)
    for lev in (3, 2, 1, 0):
        if lev < 3:
            g = pool2_backward(g) + gskip[lev]
        g = conv_back(g, True)
        g = conv_back(g, True)
    return grads


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) in place; t >= 1."""
    m *= b1
    m += (1 - b1) * g
    v *= b2
    v += (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    p -= lr * mh / (np.sqrt(vh) + eps)
