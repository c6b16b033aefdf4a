"""Empirical-score accumulators, deterministic ordered split-K.

Drop-in for the reference module kernels/emp_score_16d_ordered_splitk.py
(wrapper :162-263). The number of data splits depends only on n and the FP32
partial workspace is capped (the reference's grows as n^2/64), so the result
is bitwise reproducible and fits memory at n = 2^20 and beyond.
"""

from __future__ import annotations

from typing import Sequence, Tuple

import torch

from flash_sd_kde import _ops
from globals import PRECISION_FAST_TF32, PRECISION_FP32_IEEE


def emp_score_16d_ordered_splitk(
    data: Sequence[Sequence[float]] | torch.Tensor,
    bandwidth: float,
    *,
    device: torch.device,
    precision_mode: str,
    use_precomputed_norms: bool,
    autotune: bool,
) -> Tuple[torch.Tensor, torch.Tensor]:
    del use_precomputed_norms, autotune
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = torch.device(device)
    train = _ops.as_points16(data, device)
    if train.shape[0] == 0:
        raise ValueError("data must be non-empty.")
    if precision_mode not in {PRECISION_FAST_TF32, PRECISION_FP32_IEEE}:
        raise ValueError(f"invalid precision_mode {precision_mode}.")
    return _ops.score16(train, bandwidth, variant="ordered")
