"""CPU oracle of the surrogate perceptron -- TEST INFRASTRUCTURE ONLY.

float64 forward and backward pass of the four-layer SiLU perceptron the
device trains (x -> width -> width -> width -> [mu, log-variance]) on the
Gaussian negative log-likelihood mean_b sum_a 0.5 [(y - mu)^2 exp(-lv) + lv],
lv clamped to [-12, 12]. With ``bf16=True`` it rounds to bfloat16 (nearest
even, from float32) exactly where the device does: the inputs, the weights,
each hidden activation and each back-propagated gradient as they enter a
matrix product; everything else stays float64. With ``bf16=False`` it is the
plain float64 network, which the tests pin against central finite
differences. No counterpart in the reference (the surrogate is new).

Flat parameter layout (the library's): for each layer l, W_l [K_l][N_l]
row-major then b_l [N_l]; K_0 = p_in and N_3 = 2 p_out padded to
multiples of 64.
"""
from __future__ import annotations

import numpy as np


def pad64(x: int) -> int:
    return (x + 63) // 64 * 64


def shapes(p_in: int, p_out: int, width: int) -> list:
    return [(pad64(p_in), width), (width, width), (width, width), (width, pad64(2 * p_out))]


def unflatten(flat: np.ndarray, p_in: int, p_out: int, width: int) -> list:
    out, o = [], 0
    for K, N in shapes(p_in, p_out, width):
        W = flat[o:o + K * N].reshape(K, N).astype(np.float64)
        o += K * N
        b = flat[o:o + N].astype(np.float64)
        o += N
        out.append((W, b))
    return out


def flatten(grads: list) -> np.ndarray:
    return np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) via float32, returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward_backward(flat: np.ndarray, x: np.ndarray, y: np.ndarray, p_in: int, p_out: int,
                     width: int, bf16: bool = True) -> dict:
    """Outputs (B x N_3), mean loss, flat gradients."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unflatten(np.asarray(flat), p_in, p_out, width)
    B = x.shape[0]
    K0 = layers[0][0].shape[0]
    h = np.zeros((B, K0))
    h[:, :p_in] = x
    h = r(h)
    ins, zs = [], []
    for li, (W, b) in enumerate(layers):
        ins.append(h)
        z = h @ r(W) + b
        if li < len(layers) - 1:
            zs.append(z)
            h = r(silu(z))
        else:
            out = z
    p = p_out
    mu, lvr = out[:, :p], out[:, p:2 * p]
    lv = np.clip(lvr, -12.0, 12.0)
    e = np.exp(-lv)
    res = y - mu
    loss = float(np.mean(np.sum(0.5 * (res * res * e + lv), axis=1)))
    g = np.zeros_like(out)
    g[:, :p] = -res * e / B
    g[:, p:2 * p] = np.where((lvr > -12.0) & (lvr < 12.0), 0.5 * (1.0 - res * res * e) / B, 0.0)
    g = r(g)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        grads[li] = (ins[li].T @ g, g.sum(axis=0))
        if li == 0:
            break
        g = r((g @ r(W).T) * dsilu(zs[li - 1]))
    return {"out": out, "loss": loss, "grads": flatten(grads)}


def init_flat(p_in: int, p_out: int, width: int, rng: np.random.Generator,
              scale: float = 1.0) -> np.ndarray:
    """Random parameters in the flat layout with zero padding."""
    parts = []
    real = [(p_in, width), (width, width), (width, width), (width, 2 * p_out)]
    for (K, N), (k, nn) in zip(shapes(p_in, p_out, width), real):
        W = np.zeros((K, N))
        W[:k, :nn] = rng.uniform(-1, 1, (k, nn)) * scale * np.sqrt(6.0 / (k + nn))
        b = np.zeros(N)
        b[:nn] = rng.normal(0, 0.1, nn)
        parts += [W.ravel(), b]
    return np.concatenate(parts)
