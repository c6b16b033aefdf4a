"""The "flash" KDE API (reference kernels/flash_sd_kde/kde.py).

gaussian_kde_triton / gaussian_kde_triton_nd / gaussian_kde_score_triton keep
the reference signatures; the Triton launch knobs (block_m, block_n,
num_warps, num_stages) are accepted and ignored, since the HIP kernels fix
their own MFMA tiling. Results are computed by the gfx950 HIP library.
"""

from __future__ import annotations

import math
from typing import Sequence

import numpy as np
import torch

from flash_sd_kde import _native, _ops

from .common import _ND_FEATURES, _to_torch_tensor


def _device(device) -> torch.device:
    device = torch.device(device)
    return device


def gaussian_kde_triton(
    data: Sequence[float] | torch.Tensor,
    queries: Sequence[float] | torch.Tensor,
    bandwidth: float,
    *,
    block_m: int = 128,
    block_n: int = 128,
    num_warps: int = 4,
    num_stages: int = 2,
    device: str | torch.device = "cuda",
    synchronize: bool = True,
) -> torch.Tensor:
    """1-D Gaussian KDE densities, shape (n_query,) (reference :147-225)."""
    del block_m, block_n, num_warps, num_stages
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = _device(device)
    _native.require_gpu(device, "gaussian_kde_triton")
    train = _to_torch_tensor(data, device).reshape(-1)
    query = _to_torch_tensor(queries, device).reshape(-1)
    n_data = train.numel()
    # reference atomic kernel K11: atomics per tile, then / (h n) with 1/sqrt(2 pi) inside
    out = _ops.kde1_sums(train, query, bandwidth, atomic=True)
    if synchronize:
        torch.cuda.synchronize(device)
    out.mul_(1.0 / (math.sqrt(2.0 * math.pi) * bandwidth * n_data))
    return out


def gaussian_kde_triton_numpy(data, queries, bandwidth: float, *, device="cuda") -> np.ndarray:
    return gaussian_kde_triton(data=data, queries=queries, bandwidth=bandwidth,
                               device=device).detach().cpu().numpy()


def gaussian_kde_triton_nd(
    data: Sequence[Sequence[float]] | torch.Tensor,
    queries: Sequence[Sequence[float]] | torch.Tensor,
    bandwidth: float,
    *,
    block_m: int = 64,
    block_n: int = 64,
    num_warps: int = 4,
    num_stages: int = 2,
    device: str | torch.device = "cuda",
    synchronize: bool = True,
) -> torch.Tensor:
    """16-D Gaussian KDE densities, shape (n_query,) (reference :242-331)."""
    del block_m, block_n, num_warps, num_stages
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = _device(device)
    _native.require_gpu(device, "gaussian_kde_triton_nd")
    train = _ops.as_points16(data, device)
    query = _ops.as_points16(queries, device)
    n_data, n_query = train.shape[0], query.shape[0]
    if n_data == 0 or n_query == 0:
        raise ValueError("data and queries must contain at least one sample.")
    out = _ops.kde16_sums(train, query, bandwidth, atomic=True)
    if synchronize:
        torch.cuda.synchronize(device)
    inv_h = 1.0 / bandwidth
    norm = (inv_h ** _ND_FEATURES) / (((2.0 * math.pi) ** (_ND_FEATURES / 2.0)) * n_data)
    out.mul_(norm)
    return out


def gaussian_kde_triton_nd_numpy(data, queries, bandwidth: float, *, device="cuda") -> np.ndarray:
    return gaussian_kde_triton_nd(data=data, queries=queries, bandwidth=bandwidth,
                                  device=device).detach().cpu().numpy()


def gaussian_kde_score_triton(
    data: Sequence[float] | torch.Tensor,
    queries: Sequence[float] | torch.Tensor,
    bandwidth: float,
    *,
    block_m: int = 128,
    block_n: int = 128,
    num_warps: int = 4,
    num_stages: int = 2,
    device: str | torch.device = "cuda",
    synchronize: bool = True,
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """1-D KDE score d/dx log p(x) with its pdf and derivative (reference :348-438).

    Returns (score, pdf, deriv) with pdf = sum phi / (n h) and
    deriv = sum(-z phi) / (n h^2), phi = exp(-z^2/2) / sqrt(2 pi), z = (q-x)/h,
    score = deriv / (pdf + 1e-12).
    """
    del block_m, block_n, num_warps, num_stages
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = _device(device)
    _native.require_gpu(device, "gaussian_kde_score_triton")
    train = _to_torch_tensor(data, device).reshape(-1)
    query = _to_torch_tensor(queries, device).reshape(-1)
    n_data = train.numel()
    pdf_acc, deriv_acc = _ops.kde1_score_sums(train, query, bandwidth)
    if synchronize:
        torch.cuda.synchronize(device)
    inv_sqrt_2pi = 0.3989422804014327
    inv_norm = inv_sqrt_2pi / (n_data * bandwidth)
    pdf = pdf_acc * inv_norm
    deriv = deriv_acc * (inv_norm / bandwidth)
    score = deriv / (pdf + 1e-12)
    return score, pdf, deriv
