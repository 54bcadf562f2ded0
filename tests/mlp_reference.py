"""NumPy reference of the warm-start network (test infrastructure): the
forward pass, mean-squared-error loss and backpropagated gradients of the
fully connected ReLU network in pecs_amd/csrc/mlp.hip, in float64, with the
same flat parameter layout (W_1 (nin x nh, row-major), b_1, W_2, b_2, ...)."""
import numpy as np


def layout(nin, nh, nl, nout):
    dims = [nin] + [nh] * nl + [nout]
    out, off = [], 0
    for l in range(1, len(dims)):
        w = (off, dims[l - 1], dims[l])
        off += dims[l - 1] * dims[l]
        b = (off, dims[l])
        off += dims[l]
        out.append((w, b))
    return out, off


def unpack(params, lay):
    Ws, bs = [], []
    for (wo, fi, fo), (bo, n) in lay:
        Ws.append(params[wo:wo + fi * fo].reshape(fi, fo))
        bs.append(params[bo:bo + n])
    return Ws, bs


def forward(params, lay, x):
    Ws, bs = unpack(np.asarray(params, np.float64), lay)
    acts = [np.asarray(x, np.float64)]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = acts[-1] @ W + b
        acts.append(np.maximum(z, 0.0) if l < len(Ws) - 1 else z)
    return acts


def loss_and_grad(params, lay, x, y):
    params = np.asarray(params, np.float64)
    Ws, _ = unpack(params, lay)
    acts = forward(params, lay, x)
    diff = acts[-1] - np.asarray(y, np.float64)
    n = diff.size
    loss = float(np.sum(diff * diff) / n)
    grad = np.zeros_like(params)
    d = 2.0 * diff / n
    for l in range(len(Ws) - 1, -1, -1):
        (wo, fi, fo), (bo, nb) = lay[l]
        grad[wo:wo + fi * fo] = (acts[l].T @ d).reshape(-1)
        grad[bo:bo + nb] = d.sum(axis=0)
        if l > 0:
            d = (d @ Ws[l].T) * (acts[l] > 0)
    return loss, grad
