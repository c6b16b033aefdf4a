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


_SPREAD_TABLES: dict = {}


def _spread_table(device: torch.device) -> torch.Tensor:
    """_spread3 of 0..1023 on `device` (one gather per axis instead of the
    fifteen elementwise steps of the bit spread)."""
    key = str(device)
    t = _SPREAD_TABLES.get(key)
    if t is None:
        t = _spread3(torch.arange(1024, dtype=torch.int32)).to(device)
        _SPREAD_TABLES[key] = t
    return t


def spatial_order(x: torch.Tensor, bits: int = 10) -> torch.Tensor:
    """Permutation of the rows of x (n, d) that sorts them by the Morton code
    of their coordinates along the three leading principal axes (`bits` bits
    per axis, at most 10, over the mean +- 4 standard deviations along each;
    bit b of axis a goes to code bit 3b + a; an axis whose standard deviation
    is below 1e-6 of the leading one, rounding noise in rank-deficient data,
    contributes zeros). The axes come from the
    covariance of a strided sample of at most ~65k rows. A float32 (n, 16)
    CUDA tensor takes the library's codes instead (fsk_spatial_codes16: the
    axes from a 4096-point sample by power iteration on the GPU, no host
    synchronization), then one argsort."""
    n = x.shape[0]
    if n < 3:
        return torch.arange(n, device=x.device)
    if not 1 <= bits <= 10:
        raise ValueError(f"bits must be in [1, 10], got {bits}")
    if x.is_cuda and x.dim() == 2 and x.shape[1] == 16 and x.dtype == torch.float32:
        from flash_sd_kde import _ops
        return torch.argsort(_ops.spatial_codes16(x, bits), stable=True)
    xf = x.float()
    # the frame from a strided sample of at most ~65k rows (it only orders)
    ys = xf[::max(1, n // 65536)]
    mu = ys.mean(0)
    yc = ys - mu
    cov = (yc.T @ yc).double().cpu().numpy() / max(ys.shape[0] - 1, 1)
    ev, vecs = np.linalg.eigh(cov)
    k = min(3, x.shape[1])
    axes = torch.as_tensor(np.ascontiguousarray(vecs[:, -k:]), dtype=torch.float32,
                           device=x.device)
    # cells over mu +- 4 standard deviations per axis; rows beyond fall into
    # the end cells
    sd = np.sqrt(np.maximum(ev[-k:], 0.0))
    half = torch.as_tensor(np.maximum(4.0 * sd, 1e-30), dtype=torch.float32, device=x.device)
    p = (xf - mu) @ axes
    scale = (1 << bits) - 1
    q = ((p + half) * (scale / (2.0 * half))).clamp_(0, scale).to(torch.int32)
    flat = torch.as_tensor(sd <= 1e-6 * sd.max(), device=x.device)
    q[:, flat] = 0
    lut = _spread_table(x.device)
    code = lut[q[:, 0]]
    for a in range(1, k):
        code |= lut[q[:, a]] << a
    return torch.argsort(code, stable=True)
