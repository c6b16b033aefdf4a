"""16-D KDE sums: split-K and atomic forms.

Drop-in for the reference module kernels/kde_eval_16d_stream_splitk.py
(wrappers :263-370 and :373-452). The sums are computed by the gfx950 HIP
library: FP32 MFMA cross term, fused clamp/exp2/row-sum, FP32 partials per data
split and an FP64 combine (split-K), or per-tile atomics (atomic form).
"""

from __future__ import annotations

from typing import Sequence

import torch

from flash_sd_kde import _ops
from globals import PRECISION_FAST_TF32, PRECISION_FP32_IEEE


def _check_precision(precision_mode: str) -> None:
    if precision_mode not in {PRECISION_FAST_TF32, PRECISION_FP32_IEEE}:
        raise ValueError(f"invalid precision_mode {precision_mode}.")


def kde_eval_16d_splitk(
    data: Sequence[Sequence[float]] | torch.Tensor,
    queries: Sequence[Sequence[float]] | torch.Tensor,
    bandwidth: float,
    *,
    device: torch.device,
    precision_mode: str,
    use_precomputed_norms: bool,
    autotune: bool,
    apply_laplacian_correction: bool = False,
) -> torch.Tensor:
    """Unnormalised sums sum_j phi(q_i, x_j), shape (m,), float32.

    `use_precomputed_norms` and `autotune` are accepted for signature
    compatibility: norms are always precomputed on the device (in FP64), and
    the tile plan depends only on (n, m), which keeps the result bitwise
    reproducible.
    """
    del use_precomputed_norms, autotune
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = torch.device(device)
    train = _ops.as_points16(data, device)
    query = _ops.as_points16(queries, device)
    if train.shape[0] == 0 or query.shape[0] == 0:
        raise ValueError("data and queries must be non-empty.")
    _check_precision(precision_mode)
    return _ops.kde16_sums(train, query, bandwidth, precision_mode=precision_mode,
                           laplacian=apply_laplacian_correction, atomic=False)


def kde_eval_16d_atomic(
    data: Sequence[Sequence[float]] | torch.Tensor,
    queries: Sequence[Sequence[float]] | torch.Tensor,
    bandwidth: float,
    *,
    device: torch.device,
    precision_mode: str,
    use_precomputed_norms: bool,
    block_m: int = 64,
    block_n: int = 64,
    autotune: bool = False,
    apply_laplacian_correction: bool = False,
) -> torch.Tensor:
    """Same sums accumulated with atomics (summation order not fixed).

    `block_m`/`block_n` are Triton tile hints of the reference; the HIP kernel
    uses its own 16x16 MFMA tiles and ignores them.
    """
    del use_precomputed_norms, block_m, block_n, autotune
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = torch.device(device)
    train = _ops.as_points16(data, device)
    query = _ops.as_points16(queries, device)
    _check_precision(precision_mode)
    return _ops.kde16_sums(train, query, bandwidth, precision_mode=precision_mode,
                           laplacian=apply_laplacian_correction, atomic=True)
