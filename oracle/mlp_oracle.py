"""CPU oracle for the circulation MLP (test infrastructure only: the product never imports it).

The network has no counterpart in the reference (which solves every step with LAPACK,
python/pyvl/solver.py:223-227); this module restates, in float64 numpy, what the device
kernels compute:

    A_0 = X,  Z_l = A_l W_l^T + b_l,  A_{l+1} = SiLU(Z_l)  (l < L-1),  y = Z_{L-1}[:, 0]
    loss = mean((y - t)^2)

and its exact gradient by back-propagation.  ``bf16=True`` rounds, as the device does,
the parameters, the network input, every stored activation A_l and every back-propagated
error to bfloat16 (round to nearest even) while all sums stay in float64; the remaining
difference to the device is its FP32 accumulation order and its exp.

Pins: the float64 gradient is checked against central finite differences of the loss
(tests/test_mlp_oracle.py), the bfloat16 rounding against exactly representable values
and ties, and Adam against its closed form after one step (m/(sqrt(v)+eps) = sign(g)).
"""

from __future__ import annotations

import numpy as np


def to_bf16(a) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) through float32; returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray, dims) -> list:
    """[(W (out x in), b (out))] from the flat layout: per layer W row-major, then b."""
    out, off = [], 0
    for l in range(len(dims) - 1):
        fi, fo = dims[l], dims[l + 1]
        w = np.asarray(p[off:off + fi * fo], np.float64).reshape(fo, fi)
        off += fi * fo
        out.append((w, np.asarray(p[off:off + fo], np.float64)))
        off += fo
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, dims, X, bf16: bool = False):
    """Returns (y, [A_l], [Z_l])."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(p, dims)
    a = r(X)
    acts, zs = [], []
    for l, (w, b) in enumerate(layers):
        acts.append(a)
        z = a @ r(w).T + b
        zs.append(z)
        if l < len(layers) - 1:
            a = r(_silu(z))
    return zs[-1][:, 0], acts, zs


def loss_and_gradient(p, dims, X, t, bf16: bool = False):
    """Mean squared error and its gradient with respect to the flat parameter vector."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(p, dims)
    y, acts, zs = forward(p, dims, X, bf16)
    n = y.size
    e = y - np.asarray(t, np.float64)
    loss = float(np.mean(e * e))
    delta = r((2.0 * e / n)[:, None])
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        w, _ = layers[l]
        grads[l] = (delta.T @ acts[l], delta.sum(axis=0))
        if l > 0:
            delta = r((delta @ r(w)) * _silu_grad(zs[l - 1]))
    return loss, np.concatenate([np.concatenate([g.ravel(), gb]) for g, gb in grads])


def adam(p, g, m, v, step: int, lr: float, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step with the weight decay added to the gradient; returns (p, m, v)."""
    gi = g + wd * p
    m = b1 * m + (1.0 - b1) * gi
    v = b2 * v + (1.0 - b2) * gi * gi
    c1, c2 = 1.0 - b1**step, 1.0 - b2**step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
