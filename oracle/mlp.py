"""CPU restatement of the learned-guess network kernels -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` import this module, as the checker of the BF16 MFMA kernels of
``chronusq_public_amd/csrc/mlp.hip`` (no reference counterpart: the reference has no learned
guess; include/singleslater/singleslater_guess.h offers SAD, CORE, RANDOM and READ).

The arithmetic follows the kernels: bf16 = the upper 16 bits of an IEEE float rounded to
nearest even (inf/nan kept); matrix products of bf16 operands accumulate in float32; biases,
ReLU masks, bias gradients, the loss and Adam stay float32.  With ``bf16=False`` every
rounding is skipped and the arithmetic is float64: that form is pinned by finite differences
(tests/test_oracle_mlp.py), the bf16 rounding by known answers, so the rounded form inherits
both.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bf16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(special, u & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32)


def _rd(x, bf16: bool):
    return bf16_round(x) if bf16 else x


def forward(X, W: Sequence[np.ndarray], b: Sequence[np.ndarray], bf16: bool = True):
    """ReLU network; returns the output and the cache (layer inputs as used, pre-activations)."""
    dt = np.float32 if bf16 else np.float64
    h = _rd(np.asarray(X, dtype=dt), bf16)
    cache: List[Tuple[np.ndarray, np.ndarray]] = []
    Z = None
    for li, (w, bias) in enumerate(zip(W, b)):
        Z = (h @ _rd(np.asarray(w, dtype=dt), bf16)).astype(dt) + np.asarray(bias, dtype=dt)
        cache.append((h, Z))
        if li < len(W) - 1:
            h = _rd(np.maximum(Z, 0), bf16)
    return Z, cache


def backward(dY, W: Sequence[np.ndarray], cache, bf16: bool = True):
    """Parameter gradients for dL/dY."""
    dt = np.float32 if bf16 else np.float64
    nl = len(W)
    gW: List = [None] * nl
    gb: List = [None] * nl
    dZ = np.asarray(dY, dtype=dt)
    for li in range(nl - 1, -1, -1):
        h = cache[li][0]
        dZr = _rd(dZ, bf16)
        gW[li] = (h.T @ dZr).astype(dt)
        gb[li] = dZ.sum(axis=0, dtype=dt)
        if li > 0:
            dH = (dZr @ _rd(np.asarray(W[li], dtype=dt), bf16).T).astype(dt)
            dZ = np.where(cache[li - 1][1] > 0, dH, 0).astype(dt)
    return gW, gb


def mse(Y, T, M, w=None, norm=None):
    """loss = sum m w (y - t)^2 / norm, dY = 2 m w (y - t) / norm; norm defaults to sum m."""
    Y, T, M = (np.asarray(a, dtype=np.float64) for a in (Y, T, M))
    m = M if w is None else M * np.asarray(w, dtype=np.float64)[:, None]
    norm = float(max(1.0, M.sum())) if norm is None else norm
    d = Y - T
    return float((m * d * d).sum() / norm), 2.0 * m * d / norm


def adam(p, g, m, v, step: int, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam update with bias correction (decoupled form not used: wd adds wd p to g)."""
    g = g + wd * p
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    p = p - lr * (m / c1) / (np.sqrt(v / c2) + eps)
    return p, m, v
