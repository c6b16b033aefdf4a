"""Host helpers shared by the "flash" API modules (the reference's
kernels/flash_sd_kde/common.py converts inputs and promotes Triton block
sizes to respect CUDA's 65 535 grid limit, :28-107). The HIP launches here
use linear task grids with no such limit, so only the input conversion is
needed: array-likes become contiguous float32 tensors on the device, with
non-finite coordinates refused (flash_sd_kde._ops.check_finite).
"""

from __future__ import annotations

from typing import Sequence

import torch

from flash_sd_kde import _ops

_ND_FEATURES = 16


def _to_torch_tensor(array_like: Sequence[float] | torch.Tensor, device: torch.device) -> torch.Tensor:
    if isinstance(array_like, torch.Tensor):
        tensor = array_like.to(device=device, dtype=torch.float32, copy=False)
    else:
        tensor = torch.as_tensor(array_like, dtype=torch.float32, device=device)
    return _ops.check_finite(tensor.contiguous())
