"""KDE and one-step SD-KDE for any dimension d <= 64 by zero padding.

The reference fixes d = 16 for its multivariate kernels (globals.py:25,
flash_sd_kde/kde.py:120); its paper's rebuttal describes other dimensions as
padding to the kernel width (paper/rebuttal.md:253,505). Appending zero
coordinates to every point leaves all pairwise distances, hence every kernel
value, unchanged, so padded sums are the d-dimensional sums exactly; only the
Gaussian normaliser (2 pi)^(-d/2) h^(-d) depends on d, and the empirical score
of a padded point has zero padded components.

* d <= 16: the points are padded to 16 columns here and run through the 16-D
  kernels (the drop-in path);
* 16 < d <= 64: the library's d-dimensional entry points pad inside, to
  16 ceil(d/16) columns, and run every 16-wide slice in one MFMA chain.
"""

from __future__ import annotations

import math

import torch

from flash_sd_kde import _native as nat
from flash_sd_kde import _ops
from globals import DEFAULT_EPS, ND_FEATURES

MAX_DIM = nat.ND_MAX_DIM


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


def _dim(points) -> int:
    shape = tuple(points.shape) if hasattr(points, "shape") else (len(points), len(points[0]))
    if len(shape) == 1:
        return 1
    if len(shape) != 2:
        raise ValueError("expected a 2D (rows, d) array.")
    d = int(shape[1])
    if not 1 <= d <= MAX_DIM:
        raise ValueError(f"dimension must be between 1 and {MAX_DIM}, got {d}.")
    return d


def _wide(points, device) -> torch.Tensor:
    dev = torch.device(device)
    nat.require_gpu(dev, "d-dimensional KDE")
    t = points if isinstance(points, torch.Tensor) else torch.as_tensor(points)
    return _ops.as_points_nd(t[:, None] if t.ndim == 1 else t, dev)


def _check_h(bandwidth: float) -> float:
    h = float(bandwidth)
    if not h > 0.0:
        raise ValueError("bandwidth must be positive.")
    return h


def kde_sums_nd(data, queries, bandwidth: float, *, device="cuda",
                laplacian: bool = False) -> tuple[torch.Tensor, int, int]:
    """(unnormalised sums (m,), n, d) for d-dimensional points."""
    h = _check_h(bandwidth)
    d, dq = _dim(data), _dim(queries)
    if dq != d:
        raise ValueError(f"data has {d} features but queries have {dq}.")
    if d <= ND_FEATURES and not laplacian:
        x, _ = pad16(data, device)
        q, _ = pad16(queries, device)
        return _ops.kde16_sums(x, q, h), x.shape[0], d
    x, q = _wide(data, device), _wide(queries, device)
    return _ops.kde_nd_sums(x, q, h, laplacian=laplacian), x.shape[0], d


def kde_eval_nd(data, queries, bandwidth: float, *, device="cuda") -> torch.Tensor:
    """Gaussian KDE densities (m,) float32 for d-dimensional points, d <= 64.
    In high dimensions the normalised density can fall below FP32's range
    (((2 pi)^(d/2) h^d n)^-1 ~ 1e-58 for d = 64, h = 2.8): use
    kde_logdensity_nd there."""
    sums, n, d = kde_sums_nd(data, queries, bandwidth, device=device)
    h = float(bandwidth)
    return sums * (1.0 / (n * (2.0 * math.pi) ** (d / 2.0) * h ** d))


def kde_eval_linearized_nd(data, queries, bandwidth: float, *, device="cuda") -> torch.Tensor:
    """Laplacian-corrected densities, terms phi (1 + d/2 - |q-x|^2/(2h^2)),
    d <= 64 (the d-dimensional kernels carry the 1 + d/2 of the dimension; at
    d = 16 this is flash_sd_kde.kde_eval_linearized's sum)."""
    sums, n, d = kde_sums_nd(data, queries, bandwidth, device=device, laplacian=True)
    h = float(bandwidth)
    return sums * (1.0 / (n * (2.0 * math.pi) ** (d / 2.0) * h ** d))


def kde_logdensity_nd(data, queries, bandwidth: float, *, device="cuda",
                      precision: str = "fp32") -> torch.Tensor:
    """Log-densities (m,) float64. d <= 16: log-sum-exp kernels with the chosen
    cross-term precision; d > 16: log of the FP32 sums (FP32 precision only;
    a query whose every kernel value underflows FP32 gives -inf)."""
    h = _check_h(bandwidth)
    d = _dim(data)
    if d <= ND_FEATURES:
        x, d = pad16(data, device)
        q, dq = pad16(queries, device)
        if dq != d:
            raise ValueError(f"data has {d} features but queries have {dq}.")
        lp16 = _ops.kde16_logdensity(x, q, h, precision=precision)
        # the kernel normalises with d = 16; put back the (16 - d) missing factors
        k = ND_FEATURES - d
        return lp16 + k * (0.5 * math.log(2.0 * math.pi) + math.log(h))
    if precision != "fp32":
        raise ValueError(f"d = {d}: only precision='fp32' is available above 16 features")
    sums, n, d = kde_sums_nd(data, queries, h, device=device)
    return torch.log(sums.double()) - math.log(n) - 0.5 * d * math.log(2.0 * math.pi) \
        - d * math.log(h)


def sd_kde_fit_transform_nd(data, bandwidth: float, *, device="cuda",
                            eps: float = DEFAULT_EPS) -> torch.Tensor:
    """One-step empirical SD-KDE debiased points (n, d) for d <= 64: exact
    ordered score pass (FP32 MFMA, FP64 combine) and FP64 debias."""
    h = _check_h(bandwidth)
    d = _dim(data)
    if d <= ND_FEATURES:
        x, d = pad16(data, device)
        pdf, w = _ops.score16(x, h, variant="ordered")
        return _ops.debias(x, pdf, w, h, eps)[:, :d].contiguous()
    x = _wide(data, device)
    pdf, w = _ops.score_nd(x, h)
    return _ops.debias_nd(x, pdf, w, h, eps)
