"""Numpy oracle of the perceptron decoder (forward, loss, gradients, Adam).

TEST INFRASTRUCTURE ONLY, like the rest of this package: tests/ compare the
HIP decoder with it; the product never imports it.

The network and its numerics restate the decoder kernels: ndet -> h1 -> h2 ->
h3 -> 1, GELU (erf form) between layers, binary cross-entropy on the logit.
With `emulate=True` the values the device keeps in BF16 are rounded to BF16
at the same points (weights, activations after GELU, the GELU derivative at
each pre-activation that the forward pass keeps for the backward one,
back-propagated pre-activation gradients); sums are float64 here and float32
on the device.
With `emulate=False` everything stays float64, which is what the finite-
difference check of the gradient formulas uses.

Parameter layout (flat): W1[h1][kin] b1 W2[h2][h1] b2 W3[h3][h2] b3 w4[h3] b4,
kin = ndet rounded up to 128.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float32."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def layout(ndet: int, h1: int, h2: int, h3: int):
    kin = (ndet + 127) // 128 * 128
    shapes = [(h1, kin), (h1,), (h2, h1), (h2,), (h3, h2), (h3,), (h3,), (1,)]
    offs, o = [], 0
    for s in shapes:
        offs.append(o)
        o += int(np.prod(s))
    return kin, shapes, offs, o


def unpack(params: np.ndarray, ndet, h1, h2, h3):
    _, shapes, offs, _ = layout(ndet, h1, h2, h3)
    return [params[o:o + int(np.prod(s))].reshape(s) for o, s in zip(offs, shapes)]


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def dgelu(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def forward(params, X, dims, emulate=True):
    """Logits and the intermediate values; X is (rows, kin) 0/1."""
    ndet, h1, h2, h3 = dims
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(np.asarray(params, np.float64), *dims)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    W1, W2, W3, w4 = (r(w).astype(np.float64) for w in (W1, W2, W3, w4))
    X = np.asarray(X, np.float64)
    pre1 = X @ W1.T + b1
    a1 = r(gelu(pre1)).astype(np.float64)
    pre2 = a1 @ W2.T + b2
    a2 = r(gelu(pre2)).astype(np.float64)
    pre3 = a2 @ W3.T + b3
    a3 = r(gelu(pre3)).astype(np.float64)
    z = a3 @ w4 + b4[0]
    return z, dict(W=(W1, W2, W3, w4), pre=(pre1, pre2, pre3), act=(a1, a2, a3))


def loss_and_grads(params, X, y, dims, emulate=True):
    ndet, h1, h2, h3 = dims
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    z, c = forward(params, X, dims, emulate)
    W1, W2, W3, w4 = c["W"]
    pre1, pre2, pre3 = c["pre"]
    a1, a2, a3 = c["act"]
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    B = X.shape[0]
    sp = np.where(z > 0, z + np.log1p(np.exp(-np.abs(z))), np.log1p(np.exp(z)))
    loss = float(np.mean(sp - y * z))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / B
    dg1, dg2, dg3 = (r(dgelu(p)).astype(np.float64) for p in (pre1, pre2, pre3))
    d3 = r(dz[:, None] * w4[None, :] * dg3).astype(np.float64)
    gw4 = a3.T @ dz
    gb4 = np.array([dz.sum()])
    gW3 = d3.T @ a2
    gb3 = d3.sum(0)
    d2 = r((d3 @ W3) * dg2).astype(np.float64)
    gW2 = d2.T @ a1
    gb2 = d2.sum(0)
    d1 = r((d2 @ W2) * dg1).astype(np.float64)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    g = np.concatenate([a.reshape(-1) for a in (gW1, gb1, gW2, gb2, gW3, gb3, gw4, gb4)])
    return loss, g


def adam(params, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected); returns new arrays."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
