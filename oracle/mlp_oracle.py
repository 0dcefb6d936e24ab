"""TEST INFRASTRUCTURE ONLY: numpy restatement of the schedule selector's
network (16 -> 64 -> 64 -> 8, ReLU, softmax cross-entropy) used to check the
GPU forward/backward kernels. Weight matrices are rounded to BF16
(round-to-nearest-even) exactly as the kernels read them; all arithmetic here
is float64. Parameter order: W1[64][16], b1[64], W2[64][64], b2[64], W3[8][64], b3[8].
"""
from __future__ import annotations

import numpy as np

SHAPES = [("W1", (64, 16)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (8, 64)), ("b3", (8,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16_round(a):
    """float32 -> nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, (u >> 16) << 16, r)
    return r.astype(np.uint32).view(np.float32)


def unpack(p):
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], np.float64).reshape(shape)
        k += n
    return out


def pack(d):
    return np.concatenate([d[name].reshape(-1) for name, _ in SHAPES])


def effective(p, bf16=True):
    d = unpack(p)
    if bf16:
        for w in ("W1", "W2", "W3"):
            d[w] = bf16_round(d[w].astype(np.float32)).astype(np.float64)
    return d


def forward(p, x, bf16=True):
    d = effective(p, bf16)
    x = np.asarray(x, np.float64)
    z1 = x @ d["W1"].T + d["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ d["W2"].T + d["b2"]
    h2 = np.maximum(z2, 0)
    return h2 @ d["W3"].T + d["b3"], (x, z1, h1, z2, h2, d)


def loss_and_grad(p, x, y, bf16=True, plant=0):
    """Mean cross-entropy and its gradient w.r.t. the parameters (the BF16
    rounding is treated as identity in the backward pass, as the kernels do).
    plant 1 drops the ReLU derivative of the second hidden layer."""
    logits, (x, z1, h1, z2, h2, d) = forward(p, x, bf16)
    y = np.asarray(y)
    n = len(y)
    m = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - m)
    z = e.sum(axis=1, keepdims=True)
    loss = float(np.mean(np.log(z[:, 0]) + m[:, 0] - logits[np.arange(n), y]))
    dl = e / z
    dl[np.arange(n), y] -= 1
    dl /= n
    g = {"W3": dl.T @ h2, "b3": dl.sum(0)}
    d2 = dl @ d["W3"]
    if plant != 1:
        d2 = d2 * (z2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ d["W2"]) * (z1 > 0)
    g["W1"] = d1.T @ x
    g["b1"] = d1.sum(0)
    return loss, pack(g)
