"""Float64 numpy restatement of the surrogate MLP (test infrastructure):
forward and analytic backward of 64 -> 4 x 256 GELU -> 2 with the mean
weighted squared-error loss used by the library's trainer."""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(layers, X):
    """layers: [(W[out][in], b[out])]; returns (Y, cache)."""
    a = X.astype(np.float64)
    cache = [a]
    zs = []
    for W, b in layers[:-1]:
        z = a @ W.astype(np.float64).T + b
        zs.append(z)
        a = gelu(z)
        cache.append(a)
    W, b = layers[-1]
    y = a @ W.astype(np.float64).T + b
    return y, (cache, zs)


def _bf16(x):
    """Round to bfloat16 (nearest even) and back, as the GPU stores activations."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(layers, X):
    """forward() with every hidden activation rounded to bfloat16, as the library keeps them
    between layers; products and sums stay float64."""
    a = X.astype(np.float64)
    for W, b in layers[:-1]:
        a = _bf16(gelu(a @ W.astype(np.float64).T + b))
    W, b = layers[-1]
    return a @ W.astype(np.float64).T + b


def loss_and_grad(layers, X, T, wt=(1.0, 1.0)):
    B = X.shape[0]
    y, (acts, zs) = forward(layers, X)
    w = np.asarray(wt, np.float64)
    e = y - T
    loss = float((w * e * e).sum() / B)
    dy = 2.0 * e * w / B
    grads = [None] * len(layers)
    W, b = layers[-1]
    grads[-1] = (dy.T @ acts[-1], dy.sum(0))
    dh = dy @ W.astype(np.float64)
    for l in range(len(layers) - 2, -1, -1):
        dz = dh * gelu_grad(zs[l])
        Wl, _ = layers[l]
        grads[l] = (dz.T @ acts[l], dz.sum(0))
        dh = dz @ Wl.astype(np.float64)
    return loss, grads


def flat(grads):
    return np.concatenate([np.concatenate([gW.ravel(), gb.ravel()]) for gW, gb in grads])


def loss_and_grad_bf16(layers, X, T, wt=(1.0, 1.0)):
    """loss_and_grad() with bfloat16 rounding where the library's trainer stores values: the
    hidden pre-activations Z and activations H after each layer, and every dZ of the backward
    pass (dZ = dH * gelu'(Z) from the stored Z); weights enter as given (bf16 copies), products
    and sums stay float64.  The weight gradient of layer l is dZ_l^T H_(l-1) with both rounded."""
    B = X.shape[0]
    a = X.astype(np.float64)
    acts, zbs = [a], []
    for W, b in layers[:-1]:
        z = a @ W.astype(np.float64).T + b
        zbs.append(_bf16(z))
        a = _bf16(gelu(z))
        acts.append(a)
    W5, b5 = layers[-1]
    y = a @ W5.astype(np.float64).T + b5
    w = np.asarray(wt, np.float64)
    e = y - T
    loss = float((w * e * e).sum() / B)
    g = 2.0 * e * w / B
    grads = [None] * len(layers)
    grads[-1] = (g.T @ acts[-1], g.sum(0))
    dz = _bf16((g @ W5.astype(np.float64)) * gelu_grad(zbs[-1]))
    for l in range(len(layers) - 2, -1, -1):
        grads[l] = (dz.T @ acts[l], dz.sum(0))
        if l > 0:
            dz = _bf16((dz @ layers[l][0].astype(np.float64)) * gelu_grad(zbs[l - 1]))
    return loss, grads
