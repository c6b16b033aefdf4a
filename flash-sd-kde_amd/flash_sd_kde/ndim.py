"""KDE and one-step SD-KDE for any dimension d <= 16 by zero padding.

The reference fixes d = 16 for its multivariate kernels (globals.py:25,
flash_sd_kde/kde.py:120); its paper's rebuttal describes other dimensions as
padding to the kernel width. Appending zero coordinates to every point
leaves all pairwise distances, hence every kernel value, unchanged, so the
16-D kernels give the d-dimensional sums exactly; only the Gaussian normaliser
(2 pi)^(-d/2) h^(-d) depends on d. The empirical score of a padded point has
zero padded components, so the debiased points are the padded debiased points
with the padding removed. Dimensions above 16 are rejected.
"""

from __future__ import annotations

import math

import torch

from flash_sd_kde import _native as nat
from flash_sd_kde import _ops
from globals import DEFAULT_EPS, ND_FEATURES


def pad16(points, device) -> tuple[torch.Tensor, int]:
    """(rows, 16) float32 device tensor with zero columns after the d given, and d."""
    t = points if isinstance(points, torch.Tensor) else torch.as_tensor(points)
    if t.ndim == 1:
        t = t[:, None]
    if t.ndim != 2:
        raise ValueError("expected a 2D (rows, d) array.")
    d = int(t.shape[1])
    if not 1 <= d <= ND_FEATURES:
        raise ValueError(f"dimension must be between 1 and {ND_FEATURES}, got {d}.")
    if t.shape[0] == 0:
        raise ValueError("points must be non-empty.")
    dev = torch.device(device)
    nat.require_gpu(dev, "pad16")
    out = torch.zeros((t.shape[0], ND_FEATURES), dtype=torch.float32, device=dev)
    out[:, :d] = t.to(device=dev, dtype=torch.float32)
    return out, d


def kde_eval_nd(data, queries, bandwidth: float, *, device="cuda") -> torch.Tensor:
    """Gaussian KDE densities (m,) float32 for d-dimensional points, d <= 16."""
    h = float(bandwidth)
    if not h > 0.0:
        raise ValueError("bandwidth must be positive.")
    x, d = pad16(data, device)
    q, dq = pad16(queries, device)
    if dq != d:
        raise ValueError(f"data has {d} features but queries have {dq}.")
    sums = _ops.kde16_sums(x, q, h)
    norm = 1.0 / (x.shape[0] * (2.0 * math.pi) ** (d / 2.0) * h ** d)
    return sums * norm


def kde_logdensity_nd(data, queries, bandwidth: float, *, device="cuda",
                      precision: str = "fp32") -> torch.Tensor:
    """Log-densities (m,) float64 by log-sum-exp for d <= 16."""
    h = float(bandwidth)
    if not h > 0.0:
        raise ValueError("bandwidth must be positive.")
    x, d = pad16(data, device)
    q, dq = pad16(queries, device)
    if dq != d:
        raise ValueError(f"data has {d} features but queries have {dq}.")
    lp16 = _ops.kde16_logdensity(x, q, h, precision=precision)
    # the kernel normalises with d = 16; put back the (16 - d) missing factors
    k = ND_FEATURES - d
    return lp16 + k * (0.5 * math.log(2.0 * math.pi) + math.log(h))


def sd_kde_fit_transform_nd(data, bandwidth: float, *, device="cuda",
                            eps: float = DEFAULT_EPS) -> torch.Tensor:
    """One-step empirical SD-KDE debiased points (n, d) for d <= 16: exact
    ordered score pass (FP32 MFMA, FP64 combine) and FP64 debias."""
    h = float(bandwidth)
    if not h > 0.0:
        raise ValueError("bandwidth must be positive.")
    x, d = pad16(data, device)
    pdf, w = _ops.score16(x, h, variant="ordered")
    return _ops.debias(x, pdf, w, h, eps)[:, :d].contiguous()
