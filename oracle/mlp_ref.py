"""CPU oracle of the closure network -- TEST INFRASTRUCTURE ONLY.

A numpy statement of the engine's MLP (csrc/mlp.hip): layers W [out][in], b [out],
SiLU on hidden layers, identity output, MSE loss mean((y - t)^2).  Two modes:

* exact=True  -- float64 throughout (checked against finite differences);
* exact=False -- the engine's mixed precision: inputs, weights and hidden
  activations rounded to bf16 (round to nearest even), dZ rounded to bf16 before
  the weight/input-gradient products, products accumulated exactly (float64),
  biases and the output in full precision.
"""
from __future__ import annotations

import numpy as np


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bf16 (nearest even) and return them as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, widths: list) -> list:
    layers, o = [], 0
    for i in range(len(widths) - 1):
        w = params[o:o + widths[i] * widths[i + 1]].reshape(widths[i + 1], widths[i])
        o += widths[i] * widths[i + 1]
        b = params[o:o + widths[i + 1]]
        o += widths[i + 1]
        layers.append((w.astype(np.float64), b.astype(np.float64)))
    return layers


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, widths, x, exact=True):
    """Returns (output, cache)."""
    rnd = (lambda a: a) if exact else (lambda a: bf16(a.astype(np.float32)))
    layers = unpack(params, widths)
    a = rnd(np.asarray(x, dtype=np.float64))
    acts, zs = [a], []
    for i, (w, b) in enumerate(layers):
        z = a @ rnd(w).T + b
        zs.append(z)
        if i + 1 < len(layers):
            a = rnd(silu(z))
            acts.append(a)
        else:
            a = z
    return a, (acts, zs, layers)


LOG2E = 1.4426950408889634


def forward_log2_scaled(params, widths, x, exact=True):
    """The fused inference kernel's arithmetic (csrc/mlp.hip, mlp_fused_forward):
    hidden layers carry a = log2(e) SiLU(z), rounded to bf16 (exact=False), their
    biases enter as log2(e) b (float32), and SiLU is z' / (1 + 2^-z') on the
    log2-scaled sums z'; layer 0's bias is the bf16 high + low split of b against
    two constant-1 inputs; the output is (W_out a + log2(e) b_out) / log2(e).
    With exact=True every rounding is off and it equals forward(exact=True)."""
    rnd = (lambda a: a) if exact else (lambda a: bf16(a.astype(np.float32)))
    f32 = (lambda a: a) if exact else (lambda a: np.asarray(a, np.float32).astype(np.float64))
    layers = unpack(params, widths)
    a = rnd(np.asarray(x, dtype=np.float64))
    for i, (w, b) in enumerate(layers):
        if i == 0:
            b_hi = rnd(b)
            b_lo = rnd(b - b_hi)
            zs = LOG2E * (a @ rnd(w).T + (b_hi + b_lo))
        else:
            zs = a @ rnd(w).T + f32(LOG2E * b)
        if i + 1 < len(layers):
            a = rnd(zs / (1.0 + np.exp2(-zs)))
        else:
            a = zs / LOG2E
    return a


def loss_and_grad(params, widths, x, t, exact=True):
    rnd = (lambda a: a) if exact else (lambda a: bf16(a.astype(np.float32)))
    y, (acts, zs, layers) = forward(params, widths, x, exact)
    t = np.asarray(t, dtype=np.float64)
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    dy = 2.0 * (y - t) / n
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        w, _ = layers[i]
        dz = dy if i + 1 == len(layers) else dy * silu_grad(zs[i])
        db = dz.sum(axis=0)
        dzr = rnd(dz)
        dw = dzr.T @ acts[i]
        grads.append((dw, db))
        dy = dzr @ rnd(w)
    grads.reverse()
    flat = np.concatenate([np.concatenate([dw.reshape(-1), db]) for dw, db in grads])
    return loss, flat


def glorot_params(widths, rng):
    parts = []
    for i in range(len(widths) - 1):
        lim = np.sqrt(6.0 / (widths[i] + widths[i + 1]))
        parts.append(rng.uniform(-lim, lim, widths[i] * widths[i + 1]))
        parts.append(rng.uniform(-0.1, 0.1, widths[i + 1]))
    return np.concatenate(parts)
