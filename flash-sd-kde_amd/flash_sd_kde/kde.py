"""KDE evaluation and empirical SD-KDE (reference flash_sd_kde/kde.py).

Same signatures, validation order and normalisations as the reference
(kde_eval :49-124, kde_eval_linearized :127-149, kde_eval_linearized_nonfused
:152-255, emp_sd_kde_fit_transform :258-326). All pair sums run in the gfx950
HIP library; the SD-KDE debias x + (h^2/2) score runs in FP64 on the GPU.
"""

from __future__ import annotations

import math
from typing import Sequence, Tuple

import torch

from flash_sd_kde import _native, _ops
from flash_sd_kde.config import (
    DEFAULT_HEURISTICS,
    select_emp_score_backend,
    validate_emp_score_backend,
    validate_kde_backend,
    validate_precision_mode,
)
from globals import (
    DEFAULT_EPS,
    DEFAULT_KDE_BACKEND,
    DEFAULT_PRECISION_MODE,
    EMP_SCORE_BACKEND_FLASH_SD_KDE,
    EMP_SCORE_BACKEND_ORDERED_SPLITK,
    KDE_BACKEND_SPLITK_STREAM,
    ND_FEATURES,
)
from kernels.emp_score_16d_ordered_splitk import emp_score_16d_ordered_splitk
from kernels.emp_score_16d_symmetric_atomic import emp_score_16d_symmetric_atomic
from kernels.flash_sd_kde import emp_score_16d_flash_sd_kde
from kernels.kde_eval_16d_stream_splitk import kde_eval_16d_atomic, kde_eval_16d_splitk
from kernels.kde_eval_1d_splitk import kde_eval_1d_atomic, kde_eval_1d_splitk


def _n_rows(data) -> int:
    return data.shape[0] if isinstance(data, torch.Tensor) else len(data)


def _require_cuda(device: torch.device, what: str) -> None:
    if device.type != "cuda":
        raise ValueError(f"{what} currently expects a CUDA device.")
    if not torch.cuda.is_available():
        raise RuntimeError("CUDA is not available but was requested.")
    _native.load_library()


def kde_eval(
    data: Sequence[float] | Sequence[Sequence[float]] | torch.Tensor,
    queries: Sequence[float] | Sequence[Sequence[float]] | torch.Tensor,
    bandwidth: float,
    *,
    device: str | torch.device = "cuda",
    precision_mode: str = DEFAULT_PRECISION_MODE,
    kde_backend: str = DEFAULT_KDE_BACKEND,
    use_precomputed_norms: bool = True,
    autotune: bool = True,
    apply_laplacian_correction: bool = False,
) -> torch.Tensor:
    """Gaussian KDE densities (or the Laplacian-corrected linearized form).

    1-D data: sum * 1/(sqrt(2 pi) h n); 2-D data (16 features): sum * 1/((2 pi)^8 h^16 n).
    """
    validate_precision_mode(precision_mode)
    validate_kde_backend(kde_backend)
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = torch.device(device)
    _require_cuda(device, "kde_eval")

    ndim = _ops.infer_ndim(data)
    if ndim == 1:
        n_data = _n_rows(data)
        if kde_backend == KDE_BACKEND_SPLITK_STREAM:
            pdf_sum = kde_eval_1d_splitk(data, queries, bandwidth, device=device,
                                         autotune=autotune,
                                         apply_laplacian_correction=apply_laplacian_correction)
        else:
            pdf_sum = kde_eval_1d_atomic(data, queries, bandwidth, device=device,
                                         apply_laplacian_correction=apply_laplacian_correction)
        return pdf_sum * ((1.0 / (math.sqrt(2.0 * math.pi) * bandwidth)) / n_data)

    if ndim == 2:
        if kde_backend == KDE_BACKEND_SPLITK_STREAM:
            pdf_sum = kde_eval_16d_splitk(data, queries, bandwidth, device=device,
                                          precision_mode=precision_mode,
                                          use_precomputed_norms=use_precomputed_norms,
                                          autotune=autotune,
                                          apply_laplacian_correction=apply_laplacian_correction)
        else:
            pdf_sum = kde_eval_16d_atomic(data, queries, bandwidth, device=device,
                                          precision_mode=precision_mode,
                                          use_precomputed_norms=use_precomputed_norms,
                                          autotune=autotune,
                                          apply_laplacian_correction=apply_laplacian_correction)
        n_data = _n_rows(data)
        dim = ND_FEATURES
        norm = 1.0 / ((2.0 * math.pi) ** (dim / 2.0) * (bandwidth ** dim) * n_data)
        return pdf_sum * norm

    raise ValueError("data must be 1D or 2D.")


def kde_eval_linearized(data, queries, bandwidth: float, *, device="cuda",
                        precision_mode: str = DEFAULT_PRECISION_MODE,
                        kde_backend: str = DEFAULT_KDE_BACKEND,
                        use_precomputed_norms: bool = True, autotune: bool = True) -> torch.Tensor:
    """Linearized Emp-SD-KDE density K - (h^2/2) Laplacian K, fused into the pair kernel."""
    return kde_eval(data, queries, bandwidth, device=device, precision_mode=precision_mode,
                    kde_backend=kde_backend, use_precomputed_norms=use_precomputed_norms,
                    autotune=autotune, apply_laplacian_correction=True)


def kde_eval_linearized_nonfused(
    data, queries, bandwidth: float, *, device="cuda",
    precision_mode: str = DEFAULT_PRECISION_MODE, kde_backend: str = DEFAULT_KDE_BACKEND,
    use_precomputed_norms: bool = True, autotune: bool = True, chunk_size: int | None = None,
) -> torch.Tensor:
    """Linearized density from two separate passes, as the reference's
    non-fused form (flash_sd_kde/kde.py:152-255) for timing against the fused
    kernel: the base density, then a second pair pass for the correction's
    moment S = sum_j phi |q - x_j|^2 / h^2 (FSK_FLAG_PHI_S), combined in FP64
    as base (1 + d/2) - norm S / 2. `chunk_size` is accepted for signature
    compatibility (the reference chunks its torch pass over the data; the
    moment pass here streams the data through the kernel's own splits)."""
    del chunk_size
    validate_precision_mode(precision_mode)
    validate_kde_backend(kde_backend)
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    device = torch.device(device)
    _require_cuda(device, "kde_eval_linearized_nonfused")
    base = kde_eval(data, queries, bandwidth, device=device, precision_mode=precision_mode,
                    kde_backend=kde_backend, use_precomputed_norms=use_precomputed_norms,
                    autotune=autotune)
    atomic = kde_backend != KDE_BACKEND_SPLITK_STREAM
    ndim = _ops.infer_ndim(data)
    if ndim == 1:
        train, query = _ops.as_points1(data, device), _ops.as_points1(queries, device)
        s_sum = _ops.kde1_sums(train, query, bandwidth, atomic=atomic, moment=True)
        dim = 1.0
    elif ndim == 2:
        train, query = _ops.as_points16(data, device), _ops.as_points16(queries, device)
        s_sum = _ops.kde16_sums(train, query, bandwidth, atomic=atomic, moment=True)
        dim = float(ND_FEATURES)
    else:
        raise ValueError("data must be 1D or 2D.")
    norm = 1.0 / ((2.0 * math.pi) ** (dim / 2.0) * bandwidth ** dim * train.shape[0])
    return (base.double() * (1.0 + 0.5 * dim) - 0.5 * norm * s_sum.double()).float()


def emp_sd_kde_fit_transform(
    data: Sequence[Sequence[float]] | torch.Tensor,
    bandwidth: float,
    *,
    device: str | torch.device = "cuda",
    precision_mode: str = DEFAULT_PRECISION_MODE,
    emp_score_backend: str | None = None,
    use_precomputed_norms: bool = True,
    autotune: bool = True,
    eps: float = DEFAULT_EPS,
    return_debug: bool = False,
) -> torch.Tensor | Tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """One-step empirical SD-KDE: x + (h^2/2) * (w/(pdf + eps) - x)/h^2."""
    if bandwidth <= 0:
        raise ValueError("bandwidth must be positive.")
    validate_precision_mode(precision_mode)
    if emp_score_backend is None:
        emp_score_backend = select_emp_score_backend(_n_rows(data), heuristics=DEFAULT_HEURISTICS)
    validate_emp_score_backend(emp_score_backend)
    device = torch.device(device)
    _require_cuda(device, "emp_sd_kde_fit_transform")

    if emp_score_backend == EMP_SCORE_BACKEND_FLASH_SD_KDE:
        pdf_sum, weighted_sum = emp_score_16d_flash_sd_kde(data, bandwidth, device=device,
                                                           synchronize=False)
    elif emp_score_backend == EMP_SCORE_BACKEND_ORDERED_SPLITK:
        pdf_sum, weighted_sum = emp_score_16d_ordered_splitk(
            data, bandwidth, device=device, precision_mode=precision_mode,
            use_precomputed_norms=use_precomputed_norms, autotune=autotune)
    else:
        pdf_sum, weighted_sum = emp_score_16d_symmetric_atomic(
            data, bandwidth, device=device, precision_mode=precision_mode,
            use_precomputed_norms=use_precomputed_norms, autotune=autotune)

    train = _ops.as_points16(data, device)
    if return_debug:
        debiased, score = _ops.debias(train, pdf_sum, weighted_sum, bandwidth, eps,
                                      return_score=True)
        return debiased, pdf_sum, weighted_sum, score
    return _ops.debias(train, pdf_sum, weighted_sum, bandwidth, eps)
