"""CPU ORACLE for the learned-predictor network — test infrastructure only.

A numpy restatement of the MLP that the HIP library trains and serves
(GELU hidden layers, linear output, loss = sum(mask * (y - t)^2) / count),
with the forward pass, the backward pass and one Adam step.  It has no
counterpart in the reference (the network proposes the Newton initial guess
that src/TimeIntegrationPKG/N_TIA_OneStep.C:135 computes by polynomial
extrapolation).

``bf16=True`` rounds what the GPU holds in BF16 (inputs, weights, hidden
activations, output and hidden-layer gradients) to BF16 with round-to-
nearest-even, and accumulates in float64, so GPU/oracle differences come
only from the GPU's FP32 accumulation order.  ``bf16=False`` is plain
float64 arithmetic: the tests pin that mode against central finite
differences and the closed forms of GELU.

Parameter layout: per layer W [out][in] then b [out], concatenated.
"""
from __future__ import annotations

import math

import numpy as np

_ERF = np.vectorize(math.erf, otypes=[np.float64])


def bf16_round(a):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    f = np.ascontiguousarray(np.asarray(a, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def gelu(z):
    return 0.5 * z * (1.0 + _ERF(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _ERF(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def layer_dims(in_dim, width, hidden, out_dim):
    dims = []
    for l in range(hidden + 1):
        di = in_dim if l == 0 else width
        do = width if l < hidden else out_dim
        dims.append((di, do))
    return dims


def unpack(params, dims):
    out, off = [], 0
    for di, do in dims:
        W = params[off: off + di * do].reshape(do, di)
        off += di * do
        b = params[off: off + do]
        off += do
        out.append((W, b))
    return out


def forward(params, X, dims, bf16=True):
    """Returns (Y, cache): Y [rows][out] and the per-layer (input, Z)."""
    rd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(np.asarray(params, np.float64), dims)
    h = rd(X)
    cache = []
    for l, (W, b) in enumerate(layers):
        z = h @ rd(W).T + b
        cache.append((h, z))
        if l + 1 < len(layers):
            h = rd(gelu(z))
        else:
            return z, cache
    raise AssertionError("unreachable")


def loss_grads(params, X, T, mask, count, dims, bf16=True):
    """Loss and gradient (same layout as params) of sum(mask (y-t)^2)/count."""
    rd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    Y, cache = forward(params, X, dims, bf16)
    d = (Y - T) * mask
    loss = float(np.sum(d * d) / count)
    layers = unpack(np.asarray(params, np.float64), dims)
    grads = []
    dz = rd(2.0 * d / count)
    for l in range(len(layers) - 1, -1, -1):
        h, _ = cache[l]
        W, _ = layers[l]
        gW = dz.T @ h
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if l == 0:
            break
        _, zprev = cache[l - 1]
        dz = rd((dz @ rd(W)) * gelu_d(zprev))
    flat = []
    for gW, gb in reversed(grads):
        flat.append(gW.ravel())
        flat.append(gb.ravel())
    return loss, np.concatenate(flat)


def adam_step(params, grads, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    p = params - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
