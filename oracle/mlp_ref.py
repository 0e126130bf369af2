"""CPU restatement of the learned predictor's network and its gradient (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the GPU network
kernels (``pelelm_amd/csrc/mlp.hip``); the product never imports it.  The reference has
no learned component, so there is no reference file to follow; this module restates
the network the C ABI documents (``include/pelelm_mlp.h``):

    h0 = x;  z_l = h_{l-1} W_l^T + b_l;  h_l = gelu(z_l)  (l = 1..3);  y = h_3 W_4^T + b_4
    L = sum (y - t)^2 / (B n_out)

with parameters in the flat order of the ABI (per layer W[fan_out][fan_in], then b).
``bf16=True`` rounds, as the GPU kernels do, the network input, the weights, every
hidden activation and every back-propagated pre-activation gradient to bfloat16
(round to nearest even); accumulation stays FP64 here (FP32 on the GPU).
With ``bf16=False`` the gradient is the exact derivative of the FP64 forward, which
tests pin by central finite differences.
"""
from __future__ import annotations

import math

import numpy as np

HIDDEN = 128


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _erf(x):
    return np.vectorize(math.erf, otypes=[np.float64])(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def layout(n_in: int, n_out: int):
    """[(fan_in, fan_out, w_offset, b_offset)] per layer and the parameter count."""
    dims = [(n_in, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, n_out)]
    out, off = [], 0
    for fi, fo in dims:
        out.append((fi, fo, off, off + fi * fo))
        off += fi * fo + fo
    return out, off


def unpack(p: np.ndarray, n_in: int, n_out: int):
    lay, n = layout(n_in, n_out)
    assert p.size == n
    return [(p[ow:ow + fi * fo].reshape(fo, fi).astype(np.float64), p[ob:ob + fo].astype(np.float64))
            for fi, fo, ow, ob in lay]


def forward(p, x, n_out, bf16=False):
    """Outputs and the per-layer inputs / pre-activations needed by ``grad``."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(p, x.shape[1], n_out)
    h = rnd(x)
    hs, zs = [h], []
    for li, (W, b) in enumerate(layers):
        z = h @ rnd(W).T + b
        if li < 3:
            zs.append(z)
            h = rnd(gelu(z))
            hs.append(h)
        else:
            y = z
    return y, hs, zs


def loss_and_grad(p, x, t, bf16=False):
    """Mean-squared error and its gradient in the flat parameter order."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    B, n_out = t.shape
    y, hs, zs = forward(p, x, n_out, bf16)
    d = y - t
    loss = float(np.sum(d * d) / (B * n_out))
    g = rnd(d * (2.0 / (B * n_out)))
    layers = unpack(p, x.shape[1], n_out)
    lay, n = layout(x.shape[1], n_out)
    grad = np.zeros(n)
    for li in range(3, -1, -1):
        fi, fo, ow, ob = lay[li]
        grad[ow:ow + fi * fo] = (g.T @ hs[li]).reshape(-1)
        grad[ob:ob + fo] = g.sum(axis=0)
        if li > 0:
            g = rnd((g @ rnd(layers[li][0])) * gelu_d(zs[li - 1]))
    return loss, grad
