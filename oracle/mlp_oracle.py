"""FP64 restatement of the surrogate network (test infrastructure only).

Same parameter layout as the HIP library: for each layer, W (d_in x d_out,
row-major) followed by b (d_out); SiLU after every hidden layer; loss = mean over
batch and outputs of (y - (t - mu_out)/sd_out)^2 with inputs normalised by
(x - mu_in)/sd_in.  Forward, loss and back-propagated gradients in float64; the
gradients are pinned by central finite differences in tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np


def layer_dims(n_in, width, depth, n_out):
    dims = [n_in] + [width] * depth + [n_out]
    return list(zip(dims[:-1], dims[1:]))


def unpack(p, n_in, width, depth, n_out):
    out, o = [], 0
    for a, b in layer_dims(n_in, width, depth, n_out):
        W = p[o:o + a * b].reshape(a, b)
        o += a * b
        bb = p[o:o + b]
        o += b
        out.append((W, bb))
    return out


def bf16_round(x):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32).copy()
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def forward(p, x, norm, dims, bf16=False):
    n_in, width, depth, n_out = dims
    mu, sd = norm[:n_in], norm[n_in:2 * n_in]
    h = (np.asarray(x, np.float64) - mu) / sd
    if bf16:
        h = bf16_round(h)
    layers = unpack(np.asarray(p, np.float64), *dims)
    cache = [h]
    zs = []
    for li, (W, b) in enumerate(layers):
        Wl = bf16_round(W) if bf16 else W
        z = h @ Wl + b
        zs.append(z)
        if li < len(layers) - 1:
            h = silu(z)
            if bf16:
                h = bf16_round(h)
        else:
            h = z
        cache.append(h)
    return h, cache, zs


def loss_and_grad(p, x, t, norm, dims, bf16=False):
    n_in, width, depth, n_out = dims
    omu, osd = norm[2 * n_in:2 * n_in + n_out], norm[2 * n_in + n_out:]
    y, cache, zs = forward(p, x, norm, dims, bf16)
    tn = (np.asarray(t, np.float64) - omu) / osd
    r = y - tn
    B = y.shape[0]
    loss = float((r ** 2).mean())
    layers = unpack(np.asarray(p, np.float64), *dims)
    grads = []
    dz = 2.0 * r / r.size
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        Wl = bf16_round(W) if bf16 else W
        h_prev = cache[li]
        gW = h_prev.T @ dz
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if li > 0:
            dh = dz @ Wl.T
            z = zs[li - 1]
            s = 1.0 / (1.0 + np.exp(-z))
            dz = dh * s * (1.0 + z * (1.0 - s))
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat, B


def init_params(dims, seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    ld = layer_dims(*dims)
    for i, (a, b) in enumerate(ld):
        lim = np.sqrt(6.0 / a) * (0.5 if i == len(ld) - 1 else 1.0)
        parts.append(rng.uniform(-lim, lim, a * b))
        parts.append(np.zeros(b))
    return np.concatenate(parts)
