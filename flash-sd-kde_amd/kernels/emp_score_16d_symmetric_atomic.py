"""Empirical-score accumulators, symmetric tile pairs with atomic scatter.

Drop-in for the reference module kernels/emp_score_16d_symmetric_atomic.py
(wrapper :163-234). Each unordered 512x512 block pair is evaluated once on the
GPU (a workgroup of four waves holds 512 points of one block); both sides are
scattered with FP64 atomics (order not fixed) and rounded to FP32 at the end. Outputs are produced by the call (no caller zeroing).
"""

from __future__ import annotations

from typing import Sequence, Tuple

import torch

from flash_sd_kde import _ops
from globals import PRECISION_FAST_TF32, PRECISION_FP32_IEEE


def emp_score_16d_symmetric_atomic(
    data: Sequence[Sequence[float]] | torch.Tensor,
    bandwidth: float,
    *,
    device: torch.device,
    precision_mode: str,
    use_precomputed_norms: bool,
    block_size: int = 64,
    autotune: bool = False,
) -> Tuple[torch.Tensor, torch.Tensor]:
    """(pdf_sum (n,), weighted_sum (n, 16)); j = i is included (phi_ii = 1)."""
    del use_precomputed_norms, block_size, autotune
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = torch.device(device)
    train = _ops.as_points16(data, device)
    if train.shape[0] == 0:
        raise ValueError("data must be non-empty.")
    if precision_mode not in {PRECISION_FAST_TF32, PRECISION_FP32_IEEE}:
        raise ValueError(f"invalid precision_mode {precision_mode}.")
    return _ops.score16(train, bandwidth, variant="symmetric")
