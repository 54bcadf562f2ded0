"""GPU tests of the warm-start network's kernels (pecs_mlp_*): FP32 loss and
gradients against the NumPy reference (tests/mlp_reference.py, itself pinned
by finite differences), Adam training lowering the loss, BF16 inference
against FP32 inference, and the rejection of bad sizes."""
import numpy as np
import pytest

import pecs_amd
from mlp_reference import forward, layout, loss_and_grad

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


# (nin, nhidden, nlayers, nout, batch): tile edges and non-multiples of 64 / 16
@pytest.mark.parametrize("dims", [(4, 16, 3, 24, 5), (3, 70, 2, 130, 67), (4, 64, 3, 1000, 33)])
def test_gradients_match_reference(dims):
    nin, nh, nl, nout, b = dims
    _say(f"[gpu] mlp gradients {dims}")
    net = pecs_amd.WarmStartNet(nin, nh, nl, nout, seed=3)
    rng = np.random.default_rng(11)
    x = rng.normal(size=(b, nin)).astype(np.float32)
    y = rng.normal(size=(b, nout)).astype(np.float32)
    p = net.params().astype(np.float64)
    lay, npar = layout(nin, nh, nl, nout)
    assert npar == net.nparams
    loss, g = net.gradients(x, y)
    rloss, rg = loss_and_grad(p, lay, x, y)
    assert abs(loss - rloss) <= 1e-5 * abs(rloss)
    scale = np.abs(rg).max()
    assert np.abs(g - rg).max() <= 1e-4 * scale, np.abs(g - rg).max() / scale
    # FP32 inference against the reference forward pass
    yp = net.predict(x, bf16=False)
    yr = forward(p, lay, x)[-1]
    assert np.abs(yp - yr).max() <= 1e-5 * np.abs(yr).max()
    net.close()


def test_training_and_bf16_inference():
    _say("[gpu] mlp training + bf16 inference")
    rng = np.random.default_rng(2)
    nin, nout = 4, 200
    A = rng.normal(size=(nin, nout))
    x = rng.uniform(-1, 1, size=(256, nin)).astype(np.float32)
    y = (np.tanh(x @ A) + 0.1 * x[:, :1]).astype(np.float32)
    net = pecs_amd.WarmStartNet(nin, 64, 3, nout, seed=1)
    first = net.train_step(x[:64], y[:64], 3e-3)
    loss = first
    for it in range(300):
        i = (it * 64) % 256
        loss = net.train_step(x[i:i + 64], y[i:i + 64], 3e-3)
    assert loss < 0.1 * first, (first, loss)
    y32 = net.predict(x, bf16=False)
    y16 = net.predict(x, bf16=True)
    rel = np.abs(y16 - y32).max() / np.abs(y32).max()
    assert 0 < rel < 3e-2, rel   # BF16 weights and activations: a few 1e-3
    # the same after parameters are replaced (the BF16 copy is refreshed)
    p = net.params()
    net.set_params(p * 0.5)
    assert not np.allclose(net.predict(x, bf16=True), y16)
    net.close()


def test_rejects_bad_sizes():
    with pytest.raises(RuntimeError, match="sizes"):
        pecs_amd.WarmStartNet(0, 8, 2, 4)


# encoder + convolutional decoder: (nin, nh, nl, c0, g0, stages, batch); the
# shapes cross tile edges (64-row / 64-column tiles, 16-deep k steps)
DECODERS = [
    (4, 16, 2, 4, 2, [(8, 1), (3, 0)], 3),
    (4, 40, 3, 8, 4, [(16, 1), (8, 1), (4, 0)], 5),
    (3, 24, 2, 5, 3, [(7, 1), (6, 1), (2, 1)], 2),
    (4, 32, 2, 8, 4, [(8, 1), (8, 1), (2, 1)], 16),  # weight gradients with split k (K = 16 x 32^2)
]


@pytest.mark.parametrize("spec", DECODERS)
def test_decoder_gradients_match_reference(spec):
    from mlp_reference import forward_decoder, layout_decoder, loss_and_grad_decoder
    nin, nh, nl, c0, g0, stages, b = spec
    _say(f"[gpu] decoder gradients {spec}")
    net = pecs_amd.WarmStartNet(nin, nh, nl, 0, seed=4, decoder=(c0, g0, stages))
    lay, npar = layout_decoder(nin, nh, nl, c0, g0, stages)
    assert npar == net.nparams
    G = g0 * 2 ** sum(u for _, u in stages)
    assert net.nout == stages[-1][0] * G * G
    rng = np.random.default_rng(12)
    x = rng.normal(size=(b, nin)).astype(np.float32)
    y = rng.normal(size=(b, net.nout)).astype(np.float32)
    p = net.params().astype(np.float64)
    loss, g = net.gradients(x, y)
    rloss, rg = loss_and_grad_decoder(p, lay, x, y)
    assert abs(loss - rloss) <= 1e-5 * abs(rloss)
    # per parameter block: every layer's weights and biases
    for yl in lay["layers"]:
        for lo, hi in ((yl["w"], yl["b"]), (yl["b"], yl["b"] + (yl["cout"] if yl["conv"] else yl["nout"]))):
            scale = np.abs(rg[lo:hi]).max()
            assert np.abs(g[lo:hi] - rg[lo:hi]).max() <= 1e-4 * scale + 1e-12, (yl, lo, hi)
    yp = net.predict(x, bf16=False)
    yr = forward_decoder(p, lay, x)[-1].reshape(b, -1)
    assert np.abs(yp - yr).max() <= 1e-5 * np.abs(yr).max()
    y16 = net.predict(x, bf16=True)
    rel = np.abs(y16 - yr).max() / np.abs(yr).max()
    assert 0 < rel < 3e-2, rel
    net.close()


def test_decoder_training():
    _say("[gpu] decoder training")
    rng = np.random.default_rng(7)
    G = 16
    yy, xx = np.meshgrid(np.arange(G) / G, np.arange(G) / G, indexing="ij")
    x = rng.uniform(-1, 1, size=(128, 3)).astype(np.float32)
    # smooth fields whose shape depends on the inputs
    y = np.stack([np.sin(np.pi * (xx[None] * (1 + 0.5 * x[:, :1, None]) + 0.3 * x[:, 1:2, None])) *
                  np.cos(np.pi * yy[None] * (1 + 0.3 * x[:, 2:3, None])),
                  np.exp(-((xx[None] - 0.5 - 0.2 * x[:, :1, None]) ** 2 + (yy[None] - 0.5) ** 2) * 8)], 1)
    y = y.reshape(128, -1).astype(np.float32)
    net = pecs_amd.WarmStartNet(3, 64, 2, 0, seed=2, decoder=(16, 4, [(16, 1), (2, 1)]))
    first = net.train_step(x[:32], y[:32], 2e-3)
    loss = first
    for it in range(400):
        i = (it * 32) % 128
        loss = net.train_step(x[i:i + 32], y[i:i + 32], 2e-3)
    assert loss < 0.1 * first, (first, loss)
    net.close()
