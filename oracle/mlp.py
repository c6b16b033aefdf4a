"""CPU oracle of the surrogate MLP (test infrastructure only): FP64 torch with
BF16 rounding at exactly the points where the GPU kernels round (GEMM
operands), so GPU and oracle differ only by FP32 accumulation order.

Pinned by: torch autograd on the unrounded FP64 network (the manual backward
formulas), and central finite differences of the loss (tests/test_oracle_mlp.py).
"""

from __future__ import annotations

import torch


def bf16(t: torch.Tensor) -> torch.Tensor:
    """Round to bfloat16 (round-to-nearest-even), back to float64."""
    return t.to(torch.bfloat16).to(torch.float64)


def silu(z):
    return z * torch.sigmoid(z)


def silu_grad(z):
    s = torch.sigmoid(z)
    return s * (1 + z * (1 - s))


def forward(Ws, bs, x, *, round_bf16: bool = True):
    """Returns (y, cache) with cache = [(h_l, z_l)]; SiLU after every layer but the last."""
    r = bf16 if round_bf16 else (lambda t: t)
    h = x.to(torch.float64)
    cache = []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = r(h) @ r(W.to(torch.float64)).T + b.to(torch.float64)
        cache.append((h, z))
        h = silu(z) if l < len(Ws) - 1 else z
    return h, cache


def backward(Ws, cache, dy, *, round_bf16: bool = True):
    """Manual backward: dz = dy * silu'(z); dW = dz^T h; db = sum dz; dh = dz W.
    With round_bf16 the GEMM operands are rounded as on the GPU."""
    r = bf16 if round_bf16 else (lambda t: t)
    grads = [None] * len(Ws)
    g = dy.to(torch.float64)
    for l in range(len(Ws) - 1, -1, -1):
        h, z = cache[l]
        dz = g * silu_grad(z) if l < len(Ws) - 1 else g
        grads[l] = (r(dz).T @ r(h), dz.sum(0))
        g = r(dz) @ r(Ws[l].to(torch.float64))
    return grads


def mse_grad(y, t):
    d = y - t.to(torch.float64)
    return 2.0 * d / d.numel(), float((d * d).sum())


def adam(w, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** step)) / (torch.sqrt(v / (1 - b2 ** step)) + eps)
    return w, m, v
