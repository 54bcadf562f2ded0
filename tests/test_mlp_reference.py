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
