"""The "flash" empirical SD-KDE API (reference kernels/flash_sd_kde/empirical.py).

emp_score_16d_flash_sd_kde returns (pdf_sum (n,), weighted_sum (n, 16)) with
weighted_sum a transposed view of a (16, n) buffer, as the reference does
(:180,223). The accumulators come from the symmetric HIP score kernel (each
unordered pair once: FP32 MFMA cross term and phi.X on both sides, FP64
atomic accumulation), the fastest FP32 score on MI355X; like the reference's
atomic kernel it is reproducible to rounding only. The debias step runs in
FP64 on the GPU.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from flash_sd_kde import _native, _ops

from .common import _to_torch_tensor
from .kde import gaussian_kde_score_triton


def emp_score_16d_flash_sd_kde(
    data: Sequence[Sequence[float]] | torch.Tensor,
    bandwidth: float,
    *,
    block_m: int = 64,
    block_n: int = 2048,
    num_warps: int = 2,
    num_stages: int = 2,
    device: str | torch.device = "cuda",
    synchronize: bool = True,
) -> tuple[torch.Tensor, torch.Tensor]:
    """pdf_sum[i] = sum_j exp(-|x_i - x_j|^2/(2h^2)); weighted_sum[i] = sum_j (...) x_j."""
    del block_m, block_n, num_warps, num_stages
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = torch.device(device)
    _native.require_gpu(device, "emp_score_16d_flash_sd_kde")
    train = _ops.as_points16(data, device)
    if train.shape[0] == 0:
        raise ValueError("data must contain at least one element.")
    pdf, weighted = _ops.score16(train, bandwidth, variant="symmetric", transposed_w=True)
    if synchronize:
        torch.cuda.synchronize(device)
    return pdf, weighted


def empirical_sd_kde_triton_nd(
    data: Sequence[Sequence[float]] | torch.Tensor,
    bandwidth: float,
    *,
    block_m: int = 64,
    block_n: int = 2048,
    num_warps: int = 2,
    num_stages: int = 2,
    device: str | torch.device = "cuda",
    return_tensor: bool = False,
    synchronize: bool = True,
) -> tuple[torch.Tensor | np.ndarray, float]:
    """One-step 16-D empirical SD-KDE debiasing; returns (debiased, h)."""
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = torch.device(device)
    _native.require_gpu(device, "empirical_sd_kde_triton_nd")
    train = _ops.as_points16(data, device)
    if train.shape[0] == 0:
        raise ValueError("data must contain at least one element.")
    pdf_sum, weighted_sum = emp_score_16d_flash_sd_kde(
        train, bandwidth, block_m=block_m, block_n=block_n, num_warps=num_warps,
        num_stages=num_stages, device=device, synchronize=False)
    debiased = _ops.debias(train, pdf_sum, weighted_sum, bandwidth, 1e-12)
    if synchronize:
        torch.cuda.synchronize(device)
    if return_tensor:
        return debiased, bandwidth
    return debiased.detach().cpu().numpy(), bandwidth


def empirical_sd_kde_triton(
    data: Sequence[float] | torch.Tensor,
    *,
    block_m: int = 64,
    block_n: int = 128,
    num_warps: int = 1,
    num_stages: int = 2,
    device: str | torch.device = "cuda",
    return_tensor: bool = False,
    synchronize: bool = True,
) -> tuple[torch.Tensor | np.ndarray, float]:
    """One-step 1-D empirical SD-KDE with its own bandwidth
    h = 0.4 * min(std, IQR/1.34) * n^(-1/9) (reference :290-362)."""
    if isinstance(data, torch.Tensor):
        host = data.detach().cpu().numpy().astype(np.float32, copy=False)
    else:
        host = np.asarray(data, dtype=np.float32)
    if host.size == 0:
        raise ValueError("data must contain at least one element.")
    n = host.size
    std_dev = float(host.std())
    iqr = float(np.percentile(host, 75) - np.percentile(host, 25))
    sigma = min(std_dev, iqr / 1.34)
    h = 0.4 * sigma * (n ** (-1 / 9))
    delta = 0.5 * (h ** 2)
    device = torch.device(device)
    _native.require_gpu(device, "empirical_sd_kde_triton")
    train = _to_torch_tensor(host, device)
    scores, _, _ = gaussian_kde_score_triton(train, train, h, block_m=block_m, block_n=block_n,
                                             num_warps=num_warps, num_stages=num_stages,
                                             device=device, synchronize=synchronize)
    debiased = train + delta * scores
    if synchronize:
        torch.cuda.synchronize(device)
    if return_tensor:
        return debiased, h
    return debiased.detach().cpu().numpy(), h
