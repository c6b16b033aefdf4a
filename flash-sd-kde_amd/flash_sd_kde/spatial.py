"""Spatial order of a point set (no reference counterpart).

Kernels that leave out tiles whose every term is negligible (the LOO
bandwidth grid, the score forms with FP16 phi) skip the more, the closer
together the points of a tile lie. `spatial_order` returns the permutation
that sorts the rows by the Morton (Z-order) code of their coordinates along
the three leading principal axes, so neighbours in the order are mostly
neighbours in space. Sums over all pairs change only by rounding under it.
"""

from __future__ import annotations

import numpy as np
import torch


def _spread3(v: torch.Tensor) -> torch.Tensor:
    """Bit b of a 10-bit value moved to bit 3b (the magic-mask bit spread)."""
    v = v & 0x3FF
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    return (v | (v << 2)) & 0x09249249


def spatial_order(x: torch.Tensor, bits: int = 10) -> torch.Tensor:
    """Permutation of the rows of x (n, d) that sorts them by the Morton code
    of their coordinates along the three leading principal axes (`bits` bits
    per axis, at most 10): bit b of axis a goes to code bit 3b + a."""
    n = x.shape[0]
    if n < 3:
        return torch.arange(n, device=x.device)
    if not 1 <= bits <= 10:
        raise ValueError(f"bits must be in [1, 10], got {bits}")
    xf = x.float()
    y = xf - xf.mean(0)
    cov = (y.T @ y).double().cpu().numpy() / (n - 1)
    _, vecs = np.linalg.eigh(cov)
    k = min(3, x.shape[1])
    axes = torch.as_tensor(np.ascontiguousarray(vecs[:, -k:]), dtype=torch.float32,
                           device=x.device)
    p = y @ axes
    lo, hi = torch.aminmax(p, dim=0)
    scale = (1 << bits) - 1
    q = ((p - lo) * (scale / (hi - lo).clamp_min(1e-30))).to(torch.int32).clamp_(0, scale)
    code = torch.zeros(n, dtype=torch.int32, device=x.device)
    for a in range(k):
        code |= _spread3(q[:, a]) << a
    return torch.argsort(code, stable=True)
