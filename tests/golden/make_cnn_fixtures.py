"""Writes tests/golden/cnn_small.npz: a small instance of the warm-start
network (levels 4/8/8/8 channels, 4 inputs, 4 outputs, 2 images of 16x16)
evaluated by torch on the CPU -- conv2d, avg_pool2d, nearest interpolate,
ReLU, mean-squared error and autograd -- an implementation independent of
the numpy oracle it pins. Run from the repository root:
    python tests/golden/make_cnn_fixtures.py
"""
import os
import sys

import numpy as np
import torch
import torch.nn.functional as F

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import cnn_oracle as CO  # noqa: E402  (only for the layer table and the initial weights)

LEVELS, IN_CH, OUT_CH = (4, 8, 8, 8), 4, 4


def main():
    layers = CO.layer_table(IN_CH, OUT_CH, LEVELS)
    params = CO.init_params(seed=3, layers=layers)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((2, 16, 16, IN_CH))
    t = rng.standard_normal((2, 16, 16, 3))
    tw = [(torch.tensor(W.transpose(0, 3, 1, 2).copy(), dtype=torch.float64, requires_grad=True),
           torch.tensor(b, dtype=torch.float64, requires_grad=True)) for W, b in params]
    h = torch.tensor(x.transpose(0, 3, 1, 2).copy(), dtype=torch.float64)
    li = 0
    skips = []

    def conv(h, relu):
        nonlocal li
        W, b = tw[li]
        li += 1
        y = F.conv2d(h, W, b, padding=1)
        return F.relu(y) if relu else y

    for lev in range(4):
        h = conv(h, True)
        h = conv(h, True)
        if lev < 3:
            skips.append(h)
            h = F.avg_pool2d(h, 2)
    for lev in (2, 1, 0):
        h = F.interpolate(h, scale_factor=2, mode="nearest") + skips[lev]
        h = conv(h, True)
        h = conv(h, lev != 0)
    y = h
    tt = torch.tensor(t.transpose(0, 3, 1, 2).copy(), dtype=torch.float64)
    loss = ((y[:, :3] - tt) ** 2).mean()
    loss.backward()
    out = {"x": x, "t": t, "y": y.detach().numpy().transpose(0, 2, 3, 1), "loss": np.array(loss.item())}
    for i, ((W, b), (tW, tb)) in enumerate(zip(params, tw)):
        out[f"W{i}"] = W
        out[f"b{i}"] = b
        out[f"gW{i}"] = tW.grad.numpy().transpose(0, 2, 3, 1)
        out[f"gb{i}"] = tb.grad.numpy()
    np.savez_compressed(os.path.join(HERE, "cnn_small.npz"), **out)
    print("wrote", os.path.join(HERE, "cnn_small.npz"), "loss", loss.item())


if __name__ == "__main__":
    main()
