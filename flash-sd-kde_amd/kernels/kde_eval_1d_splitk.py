"""1-D KDE sums: split-K and atomic forms.

Drop-in for the reference module kernels/kde_eval_1d_splitk.py (wrappers
:141-219 and :222-256), computed by the gfx950 HIP library.
"""

from __future__ import annotations

from typing import Sequence

import torch

from flash_sd_kde import _ops


def kde_eval_1d_splitk(
    data: Sequence[float] | torch.Tensor,
    queries: Sequence[float] | torch.Tensor,
    bandwidth: float,
    *,
    device: torch.device,
    autotune: bool,
    apply_laplacian_correction: bool = False,
) -> torch.Tensor:
    del autotune
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = torch.device(device)
    train = _ops.as_points1(data, device)
    query = _ops.as_points1(queries, device)
    if train.numel() == 0 or query.numel() == 0:
        raise ValueError("data and queries must be non-empty.")
    return _ops.kde1_sums(train, query, bandwidth, laplacian=apply_laplacian_correction,
                          atomic=False)


def kde_eval_1d_atomic(
    data: Sequence[float] | torch.Tensor,
    queries: Sequence[float] | torch.Tensor,
    bandwidth: float,
    *,
    device: torch.device,
    block_m: int = 128,
    block_n: int = 128,
    apply_laplacian_correction: bool = False,
) -> torch.Tensor:
    del block_m, block_n
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = torch.device(device)
    train = _ops.as_points1(data, device)
    query = _ops.as_points1(queries, device)
    return _ops.kde1_sums(train, query, bandwidth, laplacian=apply_laplacian_correction,
                          atomic=True)
