"""Split-K combine of per-split partial sums, in FP64 on the GPU.

Drop-in for the reference module kernels/reduce_partials.py (:61-108). The
reference sums in FP32 and falls back to torch on CPU tensors; here the sum is
FP64 (rounded once to FP32) and CPU tensors are rejected: this package has no
CPU compute path.
"""

from __future__ import annotations

import torch

from flash_sd_kde import _ops


def reduce_pdf_partials(partials: torch.Tensor, *, block_m: int = 256) -> torch.Tensor:
    """(S, m) -> (m,): out[i] = sum_s partials[s, i]."""
    del block_m
    return _ops.reduce_partials(partials)


def reduce_weighted_partials(partials: torch.Tensor, *, block_m: int = 128) -> torch.Tensor:
    """(S, m, 16) -> (m, 16): out[i, f] = sum_s partials[s, i, f]."""
    del block_m
    return _ops.reduce_weighted(partials)
