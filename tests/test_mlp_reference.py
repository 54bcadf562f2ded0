"""Pins the NumPy reference of the warm-start network (tests/mlp_reference.py)
that the GPU gradient test compares against: its gradients against central
finite differences of its own loss, and its layout against a hand-built
two-layer case."""
import numpy as np

from mlp_reference import forward, layout, loss_and_grad


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(5)
    lay, npar = layout(3, 7, 2, 5)
    p = rng.normal(size=npar) * 0.5
    x = rng.normal(size=(4, 3))
    y = rng.normal(size=(4, 5))
    loss, g = loss_and_grad(p, lay, x, y)
    h = 1e-6
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd = (loss_and_grad(p + e, lay, x, y)[0] - loss_and_grad(p - e, lay, x, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_layout_and_forward_by_hand():
    lay, npar = layout(2, 3, 1, 1)
    assert npar == 2 * 3 + 3 + 3 * 1 + 1
    W1 = np.array([[1.0, -1.0, 0.5], [2.0, 0.0, -1.0]])
    b1 = np.array([0.0, 1.0, -0.25])
    W2 = np.array([[1.0], [2.0], [-3.0]])
    b2 = np.array([0.5])
    p = np.concatenate([W1.ravel(), b1, W2.ravel(), b2])
    x = np.array([[1.0, 1.0]])
    h = np.maximum(x @ W1 + b1, 0)          # [3, 0, -0.75] -> [3, 0, 0]
    assert np.allclose(h, [[3.0, 0.0, 0.0]])
    assert np.allclose(forward(p, lay, x)[-1], h @ W2 + b2)   # 3.5
    assert np.allclose(forward(p, lay, x)[-1], [[3.5]])


def test_decoder_gradients_match_finite_differences():
    from mlp_reference import layout_decoder, loss_and_grad_decoder
    rng = np.random.default_rng(6)
    lay, npar = layout_decoder(3, 5, 1, 2, 2, [(3, 1), (2, 0)])
    p = rng.normal(size=npar) * 0.5
    x = rng.normal(size=(2, 3))
    y = rng.normal(size=(2, 2 * 4 * 4))
    loss, g = loss_and_grad_decoder(p, lay, x, y)
    h = 1e-6
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd = (loss_and_grad_decoder(p + e, lay, x, y)[0] - loss_and_grad_decoder(p - e, lay, x, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_decoder_convolution_by_hand():
    # one 3x3 convolution of a 2x2 map, no upsampling: each output is the
    # weighted sum of the input pixels its taps land on (zero outside)
    from mlp_reference import forward_decoder, layout_decoder
    lay, npar = layout_decoder(1, 1, 1, 1, 2, [(1, 0)])
    p = np.zeros(npar)
    y0 = lay["layers"][0]
    y1 = lay["layers"][1]
    y2 = lay["layers"][2]
    p[y0["w"]] = 1.0                      # hidden = relu(x)
    p[y1["w"]:y1["w"] + 4] = [1.0, 2.0, 3.0, 4.0]  # feature map a = hidden * [[1, 2], [3, 4]]
    Wc = np.arange(1.0, 10.0).reshape(3, 3)
    p[y2["w"]:y2["b"]] = Wc.ravel()
    p[y2["b"]] = 0.5
    out = forward_decoder(p, lay, np.array([[1.0]]))[-1][0, 0]
    a = np.array([[1.0, 2.0], [3.0, 4.0]])
    for yy in range(2):
        for xx in range(2):
            s = 0.5
            for ty in range(3):
                for tx in range(3):
                    Y, X = yy + ty - 1, xx + tx - 1
                    if 0 <= Y < 2 and 0 <= X < 2:
                        s += Wc[ty, tx] * a[Y, X]
            assert abs(out[yy, xx] - s) < 1e-12
