"""The warm-start network on the device (femcnn.hip) against its CPU oracle
(oracle/cnn_oracle.py, itself pinned to torch and finite differences):

* the matrix-core product (bfloat16 operands, FP32 accumulation) against the
  FP64 product of the same bfloat16 values, ragged shapes included;
* the forward pass against the oracle with the device's bfloat16 storage
  emulated (weights, activations): tolerance 2e-2 of max |y|, a few bfloat16
  ulps, since a rounding boundary can fall either way after FP32 sums;
* every layer's weight and bias gradient against the oracle's backward with
  the same rounding: relative L2 error 1e-2 per layer on a network whose ReLU
  masks cannot switch, and the bfloat16 sensitivity scale (0.15) on a
  He-initialised one;
* Adam steps on one fixed batch lower the loss.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import cnn_oracle as CO  # noqa: E402

pytestmark = pytest.mark.gpu


def _ws():
    import femtools_amd as fa
    from femtools_amd import warmstart as WS
    return fa, WS


@pytest.mark.parametrize("M,N,K", [(32, 32, 16), (100, 70, 48), (1000, 32, 144), (257, 128, 1152), (5, 3, 32)])
def test_gemm_bf16_matches_fp64_product(M, N, K):
    fa, WS = _ws()
    rng = np.random.default_rng(M * 7 + N)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((N, K)).astype(np.float32)
    Ab, Bb = WS.bf16_bits(A), WS.bf16_bits(B)
    got = WS.gemm_bf16(fa.default_context(), Ab, Bb)
    ref = WS.bits_to_f32(Ab).astype(np.float64) @ WS.bits_to_f32(Bb).astype(np.float64).T
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def _setup(B=2, H=16, seed=3):
    fa, WS = _ws()
    net = WS.Network(fa.default_context(), max_batch=B, H=H, W=H)
    params = WS.init_params(seed)
    rng = np.random.default_rng(seed)
    for W, b in params:                       # nonzero biases exercise the bias path
        b[:] = rng.standard_normal(b.shape) * 0.05
    params[-1][1][3:] = 0
    net.set_params(params)
    x = np.zeros((B, H, H, 16), np.float32)
    x[..., :4] = rng.standard_normal((B, H, H, 4))
    t = rng.standard_normal((B, H, H, 3)).astype(np.float32)
    # the oracle sees the device's bfloat16 weights
    oparams = [(CO.bf16(W).astype(np.float64), b.astype(np.float64)) for W, b in params]
    return fa, WS, net, params, oparams, x, t


def test_forward_matches_oracle():
    fa, WS, net, params, oparams, x, t = _setup()
    y = net.forward(x)
    yo = CO.forward(oparams, x.astype(np.float64), round_bf16=True)[..., :3]
    err = np.abs(y - yo).max() / np.abs(yo).max()
    print("forward max rel err", err)
    assert err < 2e-2


def _grad_errors(net, WS, oparams, x, t, fa):
    ctx = fa.default_context()
    xd = fa.device_array(ctx, x, dtype=np.float32)
    td = fa.device_array(ctx, t, dtype=np.float32)
    loss = net.train_step(xd, td, x.shape[0], 0.0, 0)
    st = {}
    yo = CO.forward(oparams, x.astype(np.float64), store=st, round_bf16=True)
    lo, gy = CO.loss_and_grad(yo, t.astype(np.float64))
    go = CO.backward(oparams, st, gy, round_bf16=True)
    errs = []
    for li, ((gW, gb), (oW, ob)) in enumerate(zip(net.get_grads(), go)):
        eW = np.linalg.norm(gW - oW) / max(np.linalg.norm(oW), 1e-30)
        eb = np.linalg.norm(gb - ob) / max(np.linalg.norm(ob), 1e-30)
        print(WS.LAYERS[li][0], f"gW rel {eW:.2e} gb rel {eb:.2e}")
        errs.append((eW, eb))
    return loss, lo, errs


def test_gradients_match_oracle_without_relu_switching():
    # all-positive weights, biases and inputs keep every pre-activation
    # positive: the ReLU masks cannot differ between device and oracle, so the
    # comparison pins the arithmetic of every forward and backward kernel
    # (convolutions, dgrad, wgrad, pooling, upsampling, skips, biases)
    fa, WS = _ws()
    B, H = 2, 16
    net = WS.Network(fa.default_context(), max_batch=B, H=H, W=H)
    rng = np.random.default_rng(8)
    params = []
    for name, ci, co, _ in WS.LAYERS:
        W = (np.abs(rng.standard_normal((co, 3, 3, ci))) / (9 * ci)).astype(np.float32)
        b = rng.uniform(0.05, 0.2, co).astype(np.float32)
        if name == "e0a":
            W[..., 4:] = 0
        if name == "head":
            W[3:] = 0
            b[3:] = 0
        params.append((W, b))
    net.set_params(params)
    x = np.zeros((B, H, H, 16), np.float32)
    x[..., :4] = rng.uniform(0.1, 1.0, (B, H, H, 4))
    t = rng.standard_normal((B, H, H, 3)).astype(np.float32)
    oparams = [(CO.bf16(W).astype(np.float64), b.astype(np.float64)) for W, b in params]
    loss, lo, errs = _grad_errors(net, WS, oparams, x, t, fa)
    assert loss == pytest.approx(lo, rel=1e-3)
    for (eW, eb), layer in zip(errs, WS.LAYERS):
        assert eW < 1e-2 and eb < 1e-2, layer[0]


def test_gradients_match_oracle_random_network():
    # a He-initialised network: about a third of the pre-activations sit
    # below zero and bfloat16 rounding switches some ReLU masks, which the
    # backward pass amplifies -- the FP64 oracle with and without bfloat16
    # rounding differs by 2-10 % per layer on this case, so the device is held
    # to that scale
    fa, WS, net, params, oparams, x, t = _setup()
    loss, lo, errs = _grad_errors(net, WS, oparams, x, t, fa)
    assert loss == pytest.approx(lo, rel=2e-2)
    for (eW, eb), layer in zip(errs, WS.LAYERS):
        assert eW < 0.15 and eb < 0.15, layer[0]


def test_adam_steps_lower_the_loss():
    fa, WS, net, params, oparams, x, t = _setup()
    ctx = fa.default_context()
    xd = fa.device_array(ctx, x, dtype=np.float32)
    td = fa.device_array(ctx, t * 0.1, dtype=np.float32)
    losses = [net.train_step(xd, td, x.shape[0], 2e-3, k + 1) for k in range(40)]
    print("losses", losses[0], losses[-1])
    assert losses[-1] < 0.5 * losses[0]
    # the padded outputs of the head stay zero
    assert np.all(net.get_params()[-1][0][3:] == 0)


def test_trained_warm_start_gains_over_the_zero_guess():
    # the committed weights (trained on viscous inclusion problems) on the
    # first step of the visco-elasto-plastic series generator (20 seeded
    # inclusions, C = 1, the series' solver keywords) at 128 x 128 and
    # 256 x 256 cells, seeds 0-2, solved in the fast form to 1e-8 from zero
    # velocity, from the background pure shear and from the network's
    # prediction: the prediction must converge on every problem and take fewer
    # iterations in total than either start. Step 1, where the reference's
    # time loop starts from the previous step's solution, is printed.
    from femtools_amd import warmstart as WS
    net = WS.load_default()
    rows = []
    for n in (128, 256):
        rows += WS.evaluate_gain(net, n, [0, 1, 2], tolerance=1e-8)
    for r in rows:
        print({k: (v if not isinstance(v, dict) else (v["iterations"], v["converged"])) for k, v in r.items()},
              flush=True)
    summary = WS.summarize_gain(rows)
    print(summary, flush=True)
    assert all(r["learned"]["converged"] for r in rows)
    assert summary["learned"]["iterations"] < summary["pure_shear"]["iterations"]
    assert summary["learned"]["iterations"] < summary["zero"]["iterations"]
