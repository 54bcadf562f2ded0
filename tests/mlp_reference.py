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


# ---- the encoder + convolutional decoder (pecs_mlp_create_decoder) ----------
def layout_decoder(nin, nh, nl, c0, g0, stages):
    """Layers as dicts with parameter offsets, and the parameter count."""
    layers, off = [], 0
    widths = [nin] + [nh] * nl + [c0 * g0 * g0]
    for i in range(1, len(widths)):
        layers.append(dict(conv=False, nin=widths[i - 1], nout=widths[i], w=off, b=off + widths[i - 1] * widths[i]))
        off += widths[i - 1] * widths[i] + widths[i]
    cin, side = c0, g0
    for cout, up in stages:
        R = side * (2 if up else 1)
        layers.append(dict(conv=True, cin=cin, cout=cout, R=R, up=bool(up), w=off, b=off + cout * cin * 9))
        off += cout * cin * 9 + cout
        cin, side = cout, R
    return dict(layers=layers, c0=c0, g0=g0), off


def _conv(a, W, b):
    """3x3 convolution with zero padding: a [B, ci, R, R] -> [B, co, R, R]."""
    R = a.shape[2]
    pad = np.pad(a, ((0, 0), (0, 0), (1, 1), (1, 1)))
    out = np.zeros((a.shape[0], W.shape[0], R, R))
    for ty in range(3):
        for tx in range(3):
            out += np.einsum("oc,bcyx->boyx", W[:, :, ty, tx], pad[:, :, ty:ty + R, tx:tx + R])
    return out + b[None, :, None, None], pad


def forward_decoder(params, lay, x):
    """Activations of every layer (the inputs of each layer, then the output);
    maps as [B, C, R, R]."""
    p = np.asarray(params, np.float64)
    a = np.asarray(x, np.float64)
    acts = [a]
    L = lay["layers"]
    for i, y in enumerate(L):
        last = i == len(L) - 1
        if not y["conv"]:
            W = p[y["w"]:y["w"] + y["nin"] * y["nout"]].reshape(y["nin"], y["nout"])
            z = a.reshape(a.shape[0], -1) @ W + p[y["b"]:y["b"] + y["nout"]]
        else:
            if a.ndim == 2:
                a = a.reshape(a.shape[0], lay["c0"], lay["g0"], lay["g0"])
            up = np.repeat(np.repeat(a, 2, axis=2), 2, axis=3) if y["up"] else a
            W = p[y["w"]:y["b"]].reshape(y["cout"], y["cin"], 3, 3)
            z, _ = _conv(up, W, p[y["b"]:y["b"] + y["cout"]])
        a = z if last else np.maximum(z, 0.0)
        acts.append(a)
    return acts


def loss_and_grad_decoder(params, lay, x, y_true):
    p = np.asarray(params, np.float64)
    acts = forward_decoder(p, lay, x)
    out = acts[-1].reshape(acts[-1].shape[0], -1)
    diff = out - np.asarray(y_true, np.float64).reshape(out.shape)
    n = diff.size
    loss = float(np.sum(diff * diff) / n)
    grad = np.zeros_like(p)
    d = (2.0 * diff / n).reshape(acts[-1].shape)
    L = lay["layers"]
    for i in range(len(L) - 1, -1, -1):
        y = L[i]
        a_in = acts[i]
        if not y["conv"]:
            a2 = a_in.reshape(a_in.shape[0], -1)
            d2 = d.reshape(d.shape[0], -1)
            W = p[y["w"]:y["b"]].reshape(y["nin"], y["nout"])
            grad[y["w"]:y["b"]] = (a2.T @ d2).ravel()
            grad[y["b"]:y["b"] + y["nout"]] = d2.sum(axis=0)
            d = d2 @ W.T
        else:
            if a_in.ndim == 2:
                a_in = a_in.reshape(a_in.shape[0], lay["c0"], lay["g0"], lay["g0"])
            up = np.repeat(np.repeat(a_in, 2, axis=2), 2, axis=3) if y["up"] else a_in
            R = y["R"]
            pad = np.pad(up, ((0, 0), (0, 0), (1, 1), (1, 1)))
            W = p[y["w"]:y["b"]].reshape(y["cout"], y["cin"], 3, 3)
            gW = np.zeros_like(W)
            dpad = np.zeros_like(pad)
            for ty in range(3):
                for tx in range(3):
                    gW[:, :, ty, tx] = np.einsum("boyx,bcyx->oc", d, pad[:, :, ty:ty + R, tx:tx + R])
                    dpad[:, :, ty:ty + R, tx:tx + R] += np.einsum("oc,boyx->bcyx", W[:, :, ty, tx], d)
            grad[y["w"]:y["b"]] = gW.ravel()
            grad[y["b"]:y["b"] + y["cout"]] = d.sum(axis=(0, 2, 3))
            dup = dpad[:, :, 1:-1, 1:-1]
            if y["up"]:
                B, C, _, _ = dup.shape
                dup = dup.reshape(B, C, R // 2, 2, R // 2, 2).sum(axis=(3, 5))
            d = dup
            if i > 0 and not L[i - 1]["conv"]:
                d = d.reshape(d.shape[0], -1)
        if i > 0:
            d = d * (acts[i].reshape(d.shape) > 0)
    return loss, grad
