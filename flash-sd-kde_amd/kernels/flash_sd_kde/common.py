"""Host helpers of the "flash" API (reference kernels/flash_sd_kde/common.py).

The reference promotes Triton block sizes to respect CUDA's 65 535 grid limit
(:50-107). HIP launches here use linear grids of (query block, split) tasks
with no such limit, so `_resolve_launch_shape` only validates and reports the
tile grid it was asked for.
"""

from __future__ import annotations

import math
from typing import Sequence

import torch

from flash_sd_kde import _ops

_CUDA_MAX_GRID_DIM_X = 65_535
_CUDA_MAX_GRID_DIM_Y = 65_535
_MAX_BLOCK_TILE = 1024
_ND_FEATURES = 16


def _to_torch_tensor(array_like: Sequence[float] | torch.Tensor, device: torch.device) -> torch.Tensor:
    if isinstance(array_like, torch.Tensor):
        tensor = array_like.to(device=device, dtype=torch.float32, copy=False)
    else:
        tensor = torch.as_tensor(array_like, dtype=torch.float32, device=device)
    return tensor.contiguous()


def _to_matrix_tensor(array_like, device: torch.device, *, dim: int) -> torch.Tensor:
    if dim != _ND_FEATURES:
        raise ValueError(f"only dim={_ND_FEATURES} is supported, got {dim}")
    return _ops.as_points16(array_like, device)


def _next_power_of_two(value: int) -> int:
    if value <= 1:
        return 1
    return 1 << ((value - 1).bit_length())


def _resolve_launch_shape(*, n_query: int, n_data: int, block_m: int, block_n: int,
                          kernel_name: str) -> tuple[int, int, int, int]:
    bm, bn = int(block_m), int(block_n)
    if bm <= 0 or bn <= 0:
        raise ValueError(f"{kernel_name}: block sizes must be positive, got ({bm}, {bn}).")
    return bm, bn, math.ceil(n_query / bm), math.ceil(n_data / bn)
