"""The warm-start network's CPU oracle (oracle/cnn_oracle.py) pinned against
torch's CPU implementation (fixture tests/golden/cnn_small.npz written by
tests/golden/make_cnn_fixtures.py), central finite differences of the loss,
the adjoint identities of pooling/upsampling, closed forms of the bfloat16
rounding and of the first Adam step."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import cnn_oracle as CO  # noqa: E402

FIX = os.path.join(ROOT, "tests", "golden", "cnn_small.npz")
SMALL = CO.layer_table(4, 4, (4, 8, 8, 8))


def _fixture():
    z = np.load(FIX)
    params = [(z[f"W{i}"], z[f"b{i}"]) for i in range(len(SMALL))]
    return z, params


def test_forward_matches_torch():
    z, params = _fixture()
    y = CO.forward(params, z["x"], layers=SMALL)
    assert np.abs(y - z["y"]).max() <= 1e-12 * np.abs(z["y"]).max()
    loss, _ = CO.loss_and_grad(y, z["t"])
    assert loss == pytest.approx(float(z["loss"]), rel=1e-13)


def test_backward_matches_torch_autograd():
    z, params = _fixture()
    st = {}
    y = CO.forward(params, z["x"], store=st, layers=SMALL)
    _, gy = CO.loss_and_grad(y, z["t"])
    grads = CO.backward(params, st, gy, layers=SMALL)
    for i, (gW, gb) in enumerate(grads):
        assert np.abs(gW - z[f"gW{i}"]).max() <= 1e-12 * max(np.abs(z[f"gW{i}"]).max(), 1e-300), i
        assert np.abs(gb - z[f"gb{i}"]).max() <= 1e-12 * max(np.abs(z[f"gb{i}"]).max(), 1e-300), i


def test_full_network_gradient_by_central_differences():
    layers = CO.layer_table()
    params = CO.init_params(seed=5)
    rng = np.random.default_rng(2)
    x = np.zeros((1, 8, 8, CO.IN_CH))
    x[..., :4] = rng.standard_normal((1, 8, 8, 4))
    t = rng.standard_normal((1, 8, 8, 3))
    st = {}
    y = CO.forward(params, x, store=st)
    _, gy = CO.loss_and_grad(y, t)
    grads = CO.backward(params, st, gy)
    h = 1e-6
    for li in (0, 3, 7, 9, 13):
        W, b = params[li]
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in W.shape)
            if li == 0:
                idx = idx[:3] + (int(rng.integers(0, 4)),)
            if li == 13:
                idx = (int(rng.integers(0, 3)),) + idx[1:]
            w0 = W[idx]
            W[idx] = w0 + h
            lp = CO.loss_and_grad(CO.forward(params, x), t)[0]
            W[idx] = w0 - h
            lm = CO.loss_and_grad(CO.forward(params, x), t)[0]
            W[idx] = w0
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[li][0][idx]) <= 1e-5 * abs(fd) + 1e-9, (layers[li][0], idx)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 8, 6, 3))
    y = rng.standard_normal((2, 4, 3, 3))
    assert np.vdot(CO.pool2(x), y) == pytest.approx(np.vdot(x, CO.pool2_backward(y)), rel=1e-13)
    assert np.vdot(CO.up2(y), x) == pytest.approx(np.vdot(y, CO.up2_backward(x)), rel=1e-13)
    assert np.allclose(CO.pool2(CO.up2(y)), y)


def test_conv_is_a_shifted_sum():
    # a single unit tap (ky, kx) = (0, 2) reads pixel (y - 1, x + 1)
    x = np.random.default_rng(4).standard_normal((1, 5, 5, 1))
    W = np.zeros((1, 3, 3, 1))
    W[0, 0, 2, 0] = 1.0
    y = CO.conv3x3(x, W, np.zeros(1))
    assert y[0, 2, 2, 0] == x[0, 1, 3, 0] and y[0, 0, 2, 0] == 0.0 and y[0, 3, 4, 0] == 0.0


def test_bf16_rounding():
    assert CO.bf16(1.0) == 1.0
    assert CO.bf16(1 + 2 ** -8) == 1.0                     # tie -> even mantissa
    assert CO.bf16(1 + 3 * 2 ** -8) == 1 + 2 ** -6          # tie -> even mantissa
    assert CO.bf16(1 + 2 ** -8 + 2 ** -12) == 1 + 2 ** -7   # above the tie
    assert CO.bf16(-3.0) == -3.0
    r = np.random.default_rng(0).standard_normal(1000)
    assert np.all(np.abs(CO.bf16(r) - r) <= 2 ** -8 * np.abs(r))   # half an ulp of 8 significant bits


def test_first_adam_step_moves_by_lr_times_sign():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -1e-3, 4.0])
    m, v = np.zeros(3), np.zeros(3)
    CO.adam_step(p, g, m, v, 1, lr=0.01, eps=0.0)
    assert np.allclose(p, [1.0 - 0.01, -2.0 + 0.01, 0.5 - 0.01], rtol=0, atol=1e-15)
