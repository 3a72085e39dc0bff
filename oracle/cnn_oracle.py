"""CPU oracle of the learned closure network (TEST INFRASTRUCTURE ONLY: used by tests/ as the
checker of the HIP kernels in nse2d-channel-gpu_amd/csrc/closure.hip; never by the product path).

The network: six 3x3 convolutions, 4 -> 64 -> 64 -> 64 -> 64 -> 64 -> 1 channels, GELU (tanh form,
Hendrycks and Gimpel) after the first five; images [B][H][W][C] with x periodic and zero padding
beyond the walls in y. Parameters are one flat float32 array: per layer W[9][cout][cin] (tap
t = 3(dy+1) + (dx+1)) then b[cout]; the first layer's cin is 32 (the 4 inputs padded with zeros).

emulate=True rounds to bf16 (round to nearest even) where the GPU stores bf16: the packed inputs,
the conv weights of layers 0..4, every pre-activation z and activation a, and every back-propagated
pre-activation gradient. Accumulation is in float64 (the GPU accumulates in fp32). With
emulate=False everything is float64: the finite-difference reference for the gradients.
"""
import numpy as np

C = 64
CIN0 = 32
NIN = 4
S = 0.7978845608028654  # sqrt(2/pi)


def bf16(x):
    """Round to bf16 (nearest, ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(S * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    th = np.tanh(S * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * S * (1.0 + 3.0 * 0.044715 * z * z)


def layout():
    """[(offset_W, cout, cin, offset_b)] of the flat parameter array, and its length."""
    out, o = [], 0
    for l in range(6):
        cin = CIN0 if l == 0 else C
        cout = 1 if l == 5 else C
        ow = o
        o += 9 * cout * cin
        out.append((ow, cout, cin, o))
        o += cout
    return out, o


def unpack(p):
    lay, _ = layout()
    return [(p[ow:ow + 9 * co * ci].reshape(9, co, ci).astype(np.float64), p[ob:ob + co].astype(np.float64))
            for ow, co, ci, ob in lay]


def shift(x, dy, dx):
    """s[b, y, x] = x[b, y+dy, x+dx] with x periodic and zeros beyond y in [0, H)."""
    H = x.shape[1]
    s = np.roll(x, -dx, axis=2)
    out = np.zeros_like(s)
    if dy == 0:
        out[:] = s
    elif dy > 0:
        out[:, :H - dy] = s[:, dy:]
    else:
        out[:, -dy:] = s[:, :H + dy]
    return out


def conv(x, w):
    """z[b,y,x,co] = sum_t sum_ci x[b, y+dy, x+dx, ci] w[t, co, ci]."""
    z = 0.0
    for t in range(9):
        z = z + shift(x, t // 3 - 1, t % 3 - 1) @ w[t].T
    return z


def conv_t(g, w):
    """Adjoint of conv: d[b,y,x,ci] = sum_t sum_co g[b, y-dy, x-dx, co] w[t, co, ci]."""
    d = 0.0
    for t in range(9):
        d = d + shift(g, -(t // 3 - 1), -(t % 3 - 1)) @ w[t]
    return d


def conv_wgrad(x, g):
    """dW[t, co, ci] = sum_{b,y,x} g[b,y,x,co] x[b, y+dy, x+dx, ci]."""
    co, ci = g.shape[-1], x.shape[-1]
    return np.stack([g.reshape(-1, co).T @ shift(x, t // 3 - 1, t % 3 - 1).reshape(-1, ci) for t in range(9)])


def forward(p, inp, emulate=True):
    """inp [B][H][W][4] -> (out [B][H][W], cache)."""
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(p)
    B, H, W, _ = inp.shape
    x = np.zeros((B, H, W, CIN0))
    x[..., :NIN] = inp
    x = r(x)
    cache = {"x0": x, "z": [], "a": []}
    a = x
    for l in range(5):
        w, b = layers[l]
        z = r(conv(a, r(w)) + b)
        a = r(gelu(z))
        cache["z"].append(z)
        cache["a"].append(a)
    w5, b5 = layers[5]
    out = conv(a, w5)[..., 0] + b5[0]
    return out, cache


def backward(p, cache, gout, emulate=True):
    """Flat gradient of sum(gout * out)."""
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(p)
    lay, n = layout()
    grad = np.zeros(n)
    w5, _ = layers[5]
    a4 = cache["a"][4]
    g = gout[..., None]
    ow, co, ci, ob = lay[5]
    grad[ow:ow + 9 * co * ci] = conv_wgrad(a4, g).ravel()
    grad[ob] = gout.sum()
    dz = r(conv_t(g, w5) * gelu_grad(cache["z"][4]))
    for l in range(4, -1, -1):
        xin = cache["a"][l - 1] if l > 0 else cache["x0"]
        ow, co, ci, ob = lay[l]
        grad[ow:ow + 9 * co * ci] = conv_wgrad(xin, dz).ravel()
        grad[ob:ob + co] = dz.reshape(-1, co).sum(axis=0)
        if l > 0:
            w, _ = layers[l]
            dz = r(conv_t(dz, r(w)) * gelu_grad(cache["z"][l - 1]))
    return grad


def mse(out, target):
    """Mean squared error over rows 1..H-2 and its gradient with respect to out."""
    B, H, W = out.shape
    d = np.zeros_like(out)
    d[:, 1:-1] = out[:, 1:-1] - target[:, 1:-1]
    n = B * (H - 2) * W
    return float(np.sum(d * d) / n), 2.0 * d / n


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
