"""NumPy float64 oracle of the learned U-Net preconditioner (test
infrastructure only; the network has no reference counterpart).

Tensors are channels-last ``(H, W, C)``; conv weights ``(3, 3, Cout, Cin)``
(tap-major).  Architecture (bias-free, so positively homogeneous):

    e_l = relu(conv(relu(conv(x_l))))        x_0 = r,  x_{l+1} = avgpool2(e_l)
    d_L = e_L
    d_l = relu(conv(relu(conv_u(up2(d_{l+1})) + e_l)))
    y   = conv_out(d_0)

Backward is hand-written and pinned by finite differences in
``tests/test_oracle_unet.py``.

``bf16=True`` rounds to bfloat16 (round-to-nearest-even) at the points where
the device network stores activations and activation gradients (every
hidden conv output after its ReLU, pooled inputs, and each masked gradient
tensor), so that the GPU kernels can be compared against the same storage
precision; with ``bf16=False`` the network is exact FP64.
"""

from __future__ import annotations

import numpy as np


def conv3x3(x, w):
    """Zero-padded 3x3 convolution. x (H, W, Ci), w (3, 3, Co, Ci) -> (H, W, Co)."""
    H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    out = np.zeros((H, W, w.shape[2]))
    for ky in range(3):
        for kx in range(3):
            out += xp[ky:ky + H, kx:kx + W, :] @ w[ky, kx].T
    return out


def conv3x3_bwd(x, w, gy):
    """Gradients of conv3x3 w.r.t. x and w."""
    H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w)
    for ky in range(3):
        for kx in range(3):
            gxp[ky:ky + H, kx:kx + W, :] += gy @ w[ky, kx]
            gw[ky, kx] = gy.reshape(-1, gy.shape[2]).T @ xp[ky:ky + H, kx:kx + W, :].reshape(-1, x.shape[2])
    return gxp[1:-1, 1:-1, :], gw


def pool2(x):
    H, W, C = x.shape
    return x.reshape(H // 2, 2, W // 2, 2, C).mean(axis=(1, 3))


def pool2_bwd(gy):
    return np.repeat(np.repeat(gy, 2, axis=0), 2, axis=1) / 4.0


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up2_bwd(gy):
    H, W, C = gy.shape
    return gy.reshape(H // 2, 2, W // 2, 2, C).sum(axis=(1, 3))


def relu(x):
    return np.maximum(x, 0.0)


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even) via float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _identity(x):
    return x


def layer_names(nlev):
    names = []
    for l in range(nlev):
        names += [f"e{l}a", f"e{l}b"]
    for l in range(nlev - 2, -1, -1):
        names += [f"u{l}", f"d{l}"]
    return names + ["out"]


def weight_shapes(channels):
    """Conv weight shapes for channels (C0, ..., C_{L-1}); input has 1 channel."""
    L = len(channels)
    shapes = {}
    cin = 1
    for l, c in enumerate(channels):
        shapes[f"e{l}a"] = (3, 3, c, cin)
        shapes[f"e{l}b"] = (3, 3, c, c)
        cin = c
    for l in range(L - 2, -1, -1):
        shapes[f"u{l}"] = (3, 3, channels[l], channels[l + 1])
        shapes[f"d{l}"] = (3, 3, channels[l], channels[l])
    shapes["out"] = (3, 3, 1, channels[0])
    return shapes


def init_weights(channels, seed=0):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal(s) * np.sqrt(2.0 / (9 * s[3]))
            for k, s in weight_shapes(channels).items()}


def forward(Wt, r, channels, bf16=False):
    """Returns the output (H, W) and the cache needed by ``backward``."""
    q = bf16_round if bf16 else _identity
    L = len(channels)
    cache = {}
    x = r[:, :, None]
    e = []
    for l in range(L):
        if l > 0:
            x = q(pool2(e[-1]))
        a1 = conv3x3(x, Wt[f"e{l}a"])
        h1 = q(relu(a1))
        a2 = conv3x3(h1, Wt[f"e{l}b"])
        h2 = q(relu(a2))
        cache[f"e{l}"] = (x, h1, h2)
        e.append(h2)
    d = e[-1]
    for l in range(L - 2, -1, -1):
        u = up2(d)
        a1 = conv3x3(u, Wt[f"u{l}"]) + e[l]
        h1 = q(relu(a1))
        a2 = conv3x3(h1, Wt[f"d{l}"])
        h2 = q(relu(a2))
        cache[f"d{l}"] = (u, h1, h2)
        d = h2
    y = conv3x3(d, Wt["out"])
    cache["out"] = d
    return y[:, :, 0], cache


def backward(Wt, cache, gy, channels, bf16=False):
    """Weight gradients of sum(gy * y)."""
    q = bf16_round if bf16 else _identity
    L = len(channels)
    g = {}
    gd, g["out"] = conv3x3_bwd(cache["out"], Wt["out"], gy[:, :, None])
    ge = [None] * L
    # decoder, from the finest level back to the coarsest
    grad_d = gd
    for l in range(0, L - 1):
        u, h1, h2 = cache[f"d{l}"]
        ga2 = q(grad_d * (h2 > 0))
        gh1, g[f"d{l}"] = conv3x3_bwd(h1, Wt[f"d{l}"], ga2)
        ga1 = q(gh1 * (h1 > 0))
        ge[l] = ga1 if ge[l] is None else ge[l] + ga1
        gu, g[f"u{l}"] = conv3x3_bwd(u, Wt[f"u{l}"], ga1)
        grad_d = up2_bwd(q(gu))
    # encoder, from the coarsest level back to the finest
    gcarry = grad_d  # gradient into e_{L-1}
    for l in range(L - 1, -1, -1):
        x, h1, h2 = cache[f"e{l}"]
        gh2 = gcarry if l == L - 1 else ge[l] + gcarry
        ga2 = q(gh2 * (h2 > 0))
        gh1, g[f"e{l}b"] = conv3x3_bwd(h1, Wt[f"e{l}b"], ga2)
        ga1 = q(gh1 * (h1 > 0))
        gx, g[f"e{l}a"] = conv3x3_bwd(x, Wt[f"e{l}a"], ga1)
        if l > 0:
            gcarry = pool2_bwd(q(gx))
    return g


def adam_step(Wt, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015), in place; returns nothing."""
    for k in Wt:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] ** 2
        mh = m[k] / (1 - b1**t)
        vh = v[k] / (1 - b2**t)
        Wt[k] = Wt[k] - lr * mh / (np.sqrt(vh) + eps)
