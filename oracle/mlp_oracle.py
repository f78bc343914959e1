"""CPU FP64 restatement of the warm-start network (test infrastructure only:
imported by tests/ and the benchmark's checks, never by the product path).

The network has no reference counterpart (the reference solver has no learned
component); the arithmetic it restates is the contract of include/pdlpx_mlp.h:

  H0 = bf16(x);  Z_l = H_{l-1} bf16(W_l)^T + b_l;  H_l = bf16(relu(Z_l)) (l < L);  y = Z_L
  loss = mean (y - t)^2;  G_L = 2 (y - t) / (batch d_L)
  dW_l = bf16(G_l)^T H_{l-1};  db_l = sum_batch G_l;  G_{l-1} = (bf16(G_l) bf16(W_l)) * [H_{l-1} > 0]

with every product and sum in FP64 here (FP32 accumulation on the GPU).
``rounding=False`` drops every bf16() -- the plain FP64 network, whose
backward pass is pinned by central finite differences and by the closed form
of a single linear layer (least squares: dW = 2/(n d) (X W^T + b - T)^T X).
"""
from __future__ import annotations

import numpy as np


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def _r(a, rounding):
    return bf16(a) if rounding else np.asarray(a, dtype=np.float64)


def forward(params, x, rounding: bool = True):
    """params: list of (W (d_l x d_{l-1}), b (d_l)).  Returns (y, H list)."""
    H = [_r(x, rounding)]
    L = len(params)
    y = None
    for l, (W, b) in enumerate(params, start=1):
        Z = H[-1] @ _r(W, rounding).T + np.asarray(b, dtype=np.float64)
        if l < L:
            H.append(_r(np.maximum(Z, 0.0), rounding))
        else:
            y = Z
    return y, H


def loss_and_grads(params, x, t, rounding: bool = True):
    """(loss, [dW_l], [db_l]) of one batch."""
    y, H = forward(params, x, rounding)
    t = np.asarray(t, dtype=np.float64)
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    G = 2.0 * (y - t) / n
    if rounding:   # the GPU forms G in FP32 before rounding it to bf16
        G = G.astype(np.float32).astype(np.float64)
    L = len(params)
    dW = [None] * L
    db = [None] * L
    for l in range(L, 0, -1):
        W, _ = params[l - 1]
        Gr = _r(G, rounding)
        dW[l - 1] = Gr.T @ H[l - 1]
        db[l - 1] = G.sum(axis=0)
        if l > 1:
            G = (Gr @ _r(W, rounding)) * (H[l - 1] > 0.0)
            if rounding:
                G = G.astype(np.float32).astype(np.float64)
    return loss, dW, db


def finite_difference(params, x, t, layer: int, index: tuple[int, int] | int, h: float = 1e-6) -> float:
    """Central difference of the unrounded loss in one parameter (a W entry
    for a (row, col) index, a bias entry for an int)."""
    def f(delta):
        ps = [(np.array(W, dtype=np.float64), np.array(b, dtype=np.float64)) for W, b in params]
        if isinstance(index, tuple):
            ps[layer][0][index] += delta
        else:
            ps[layer][1][index] += delta
        y, _ = forward(ps, x, rounding=False)
        return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))
    return (f(h) - f(-h)) / (2.0 * h)
